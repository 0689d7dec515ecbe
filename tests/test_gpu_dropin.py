"""The C++ drop-in adapter (include/rrsph_hip.hpp) on the GPU: the reference's
loop shape `while (time <= simulation_time)` with RRSPHOutput-style loaders,
driven by tests/cpp/dropin_driver, against the FP64 oracle."""
import os

import numpy as np
import pytest

import dropin
import oracle
from sph2d_amd import generators as G

pytestmark = pytest.mark.gpu


def _oracle_run(prm, lay, nsteps):
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=prm.start_simulation_time, dt=prm.dt)
    assert ora.step(nsteps) == 0
    return ora.state()


def test_dropin_dam_break_fixed_dt(tmp_path):
    """Five fixed-dt steps of the C0 dam break: the loop runs exactly five
    steps, the final arrays and the loaders' arrays (probed at step 5) equal
    the oracle's state to 1e-12."""
    assert os.path.exists(dropin.DRIVER), "dropin_driver not built"
    lay = G.dam_break()
    prm = lay.params
    print("dropin dam break n", prm.ntotal, flush=True)
    res = dropin.run(tmp_path, prm, lay, nsteps=5, probe_every=5)
    assert res.returncode == 0, res.stdout + res.stderr
    out = dropin.read_output(tmp_path, prm.ntotal)
    assert out["steps"] == 5 and out["probe_step"] == 5
    ref = _oracle_run(prm, lay, 5)
    assert abs(out["time"] - ref["time"]) < 1e-15
    assert np.max(np.abs(out["r"] - ref["r"])) < 1e-13
    assert np.max(np.abs(out["rho"] - ref["rho"]) / ref["rho"]) < 1e-12
    assert np.array_equal(out["itype"], ref["itype"])
    assert np.array_equal(out["probe_rho"], out["rho"])
    assert np.array_equal(out["probe_p"], out["p"])


def test_dropin_piston_dynamic_dt(tmp_path):
    """Coarse piston flume with the dynamic CFL dt: the adapter keeps
    params.dt current; step count, time and state match the oracle run for
    the same number of steps."""
    lay = G.piston_tank(delta=0.02, length=4.0, depth=0.4)
    prm = lay.params
    print("dropin piston n", prm.ntotal, flush=True)
    res = dropin.run(tmp_path, prm, lay, nsteps=8)
    assert res.returncode == 0, res.stdout + res.stderr
    out = dropin.read_output(tmp_path, prm.ntotal)
    assert out["steps"] >= 1
    ref = _oracle_run(prm, lay, out["steps"])
    assert abs(out["time"] - ref["time"]) <= 1e-12 * ref["time"]
    assert abs(out["dt"] - ref["dt"]) <= 1e-12 * ref["dt"]
    assert np.max(np.abs(out["r"] - ref["r"])) < 1e-11
    assert np.max(np.abs(out["rho"] - ref["rho"]) / ref["rho"]) < 1e-11


def test_dropin_stop_crash_dump(tmp_path):
    """A fluid particle thrown out of the domain: the device check halts the
    step (STOP), the adapter writes the crash dump through the output and
    rethrows the reference's message (Output.cpp:395-412, ConsistencyCheck.cpp:184-195)."""
    lay = G.dam_break()
    prm = lay.params
    v = lay.v.copy()
    v[0] = [1.0e4, 0.0]
    lay2 = G.Layout(prm, lay.r, v, lay.rho, lay.itype)
    # the step the oracle leaves the domain on (the particle stays inside
    # after the first half step and crosses the floor on the second)
    ora = oracle.Oracle(prm)
    ora.set_state(lay2.r, lay2.v, lay2.rho, lay2.itype, time=prm.start_simulation_time, dt=prm.dt)
    crash = next(k for k in range(1, 6) if ora.step(1) != 0)
    t_crash = ora.state()["time"]
    res = dropin.run(tmp_path, prm, lay2, nsteps=5)
    assert res.returncode == 1, res.stdout
    assert f"crash dump at step {crash} time {t_crash:g}" in res.stdout, res.stdout
    assert "Catch exception: Encounter particle outside boundaries" in res.stdout, res.stdout
