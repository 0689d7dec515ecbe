"""Neighbour-list steps (the cubic configuration, 2D and 3D) against the FP64
oracle and against the sorted step of every step (RRSPH_HIP_LIST=0).

Between builds the step reads lists of the particles within k h + skin and
applies the reference's exact cut-off test, so the neighbour sets are the
cell walk's; only the summation order inside a particle's sums differs
(the list order is the walk order at the build).  A tiny skin makes the
lists go stale within a step or two, so the device halts list steps and the
host re-runs them as builds: the results stay the oracle's.
"""
import copy

import numpy as np
import pytest

import oracle
from sph2d_amd import generators as G
from sph2d_amd import hip as H
from sph2d_amd import params as P

pytestmark = pytest.mark.gpu


def _flume():
    return G.piston_tank(delta=0.02, length=4.0, depth=0.4)


def _run(lay, prm, nsteps, precision, env=None, monkeypatch=None):
    if env:
        for k, v in env.items():
            monkeypatch.setenv(k, v)
    gpu = H.HipStep(prm, precision=precision)
    if env:
        for k in env:
            monkeypatch.delenv(k)
    gpu.upload(lay.r, lay.v, lay.rho, lay.itype)
    gpu.step(nsteps)
    out = gpu.download()
    out["time"] = gpu.time()
    out["lists"] = gpu.list_stats()
    gpu.close()
    return out


def test_list_steps_match_the_oracle():
    """300 FP64 steps of the piston flume (continuity, dynamic CFL dt,
    boundary motion): the lists were used and rebuilt, and the state equals
    the oracle's to FP64 summation-order precision."""
    lay = _flume()
    prm = lay.params
    g = _run(lay, prm, 300, H.FP64)
    ls = g["lists"]
    print("list stats", ls, flush=True)
    assert ls["list_steps"] > 200 and ls["builds"] >= 2 and ls["interval"] > 0
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(300) == 0
    r = ora.state()
    t, dt, step = g["time"]
    assert step == 300 and abs(t - r["time"]) <= 1e-11 * r["time"] and abs(dt - r["dt"]) <= 1e-10 * r["dt"]
    assert np.array_equal(g["itype"], r["itype"])
    err_r = float(np.max(np.abs(g["r"] - r["r"])))
    err_rho = float(np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]))
    print("300 steps: max |dr|", err_r, "max rel drho", err_rho, flush=True)
    assert err_r < 1e-10 and err_rho < 1e-10


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_list_steps_equal_the_sorted_steps(monkeypatch, precision):
    """The same 200 steps with lists and with the sort every step."""
    lay = _flume()
    prm = lay.params
    a = _run(lay, prm, 200, precision)
    b = _run(lay, prm, 200, precision, {"RRSPH_HIP_LIST": "0"}, monkeypatch)
    assert a["lists"]["list_steps"] > 0 and b["lists"]["list_steps"] == 0 and b["lists"]["interval"] == 0
    assert np.array_equal(a["itype"], b["itype"])
    tol_r, tol_rho = (1e-11, 1e-12) if precision == H.FP64 else (1e-7, 1e-6)
    assert np.max(np.abs(a["r"] - b["r"])) < tol_r
    assert np.max(np.abs(a["rho"] - b["rho"]) / b["rho"]) < tol_rho


def test_stale_lists_halt_and_rebuild(monkeypatch):
    """Skin 0.2% of the cut-off: the lists go stale every step or two; the
    device halts those list steps before they read the lists and the host
    re-runs them as builds.  60 FP64 steps still equal the oracle's."""
    lay = _flume()
    prm = lay.params
    g = _run(lay, prm, 60, H.FP64, {"RRSPH_HIP_LIST_SKIN": "0.002"}, monkeypatch)
    ls = g["lists"]
    print("tiny skin list stats", ls, flush=True)
    assert ls["stale_halts"] >= 1 and ls["builds"] >= 10
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(60) == 0
    r = ora.state()
    assert g["time"][2] == 60
    assert np.max(np.abs(g["r"] - r["r"])) < 1e-11
    assert np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]) < 1e-11


@pytest.mark.parametrize("density_treatment", [P.DENSITY_SUMMATION, P.DENSITY_CONTINUITY])
def test_list_steps_dam_break_matches_the_oracle(density_treatment):
    """C0 dam break (fixed dt, collapsing column: the fastest flow of the
    tested cases), 150 FP64 steps with lists, summation and continuity."""
    lay = G.dam_break(density_treatment=density_treatment)
    prm = lay.params
    g = _run(lay, prm, 150, H.FP64)
    print("dam break lists", g["lists"], flush=True)
    assert g["lists"]["list_steps"] > 0
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(150) == 0
    r = ora.state()
    assert np.array_equal(g["itype"], r["itype"])
    assert np.max(np.abs(g["r"] - r["r"])) < 1e-11
    assert np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]) < 1e-11


def test_run_until_then_step_rebuilds_the_lists():
    """A run stopped at t_end leaves launched steps unexecuted, builds among
    them; the next call builds again before it reads any list."""
    lay = _flume()
    prm = lay.params
    gpu = H.HipStep(prm, precision=H.FP64)
    gpu.upload(lay.r, lay.v, lay.rho, lay.itype)
    gpu.step(40)
    t = gpu.time()[0]
    taken = gpu.run_until(t + 25 * gpu.time()[1])
    gpu.step(30)
    g = gpu.download()
    step = gpu.time()[2]
    gpu.close()
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(step) == 0
    r = ora.state()
    assert 20 <= taken <= 30 and step == 70 + taken
    assert np.max(np.abs(g["r"] - r["r"])) < 1e-11


def _dam3d(density_treatment=P.DENSITY_SUMMATION):
    """The small 3D dam break (delta 0.02 m); summation density renormalised
    every 5th step."""
    lay = G.dam_break_3d(delta=0.02, density_treatment=density_treatment)
    prm = copy.copy(lay.params)
    if density_treatment == P.DENSITY_SUMMATION:
        prm.density_normalization = 1
        prm.density_normalization_every = 5
    return lay, prm


@pytest.mark.parametrize("density_treatment", [P.DENSITY_SUMMATION, P.DENSITY_CONTINUITY])
def test_3d_list_steps_match_the_oracle(density_treatment):
    """150 FP64 steps of the collapsing 3D column with lists (summation with
    renormalisation, and continuity) equal the oracle's state."""
    lay, prm = _dam3d(density_treatment)
    print("3D dam break n", prm.ntotal, "density", density_treatment, flush=True)
    g = _run(lay, prm, 150, H.FP64)
    ls = g["lists"]
    print("3D list stats", ls, flush=True)
    assert ls["list_steps"] > 100 and ls["builds"] >= 2 and ls["interval"] > 0
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(150) == 0
    r = ora.state()
    assert g["time"][2] == 150
    assert np.array_equal(g["itype"], r["itype"])
    err_r = float(np.max(np.abs(g["r"] - r["r"])))
    err_rho = float(np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]))
    print("150 steps: max |dr|", err_r, "max rel drho", err_rho, flush=True)
    assert err_r < 1e-11 and err_rho < 1e-11


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_3d_list_steps_equal_the_sorted_steps(monkeypatch, precision):
    """The same 100 steps of the 3D column with lists and with the sort every step."""
    lay, prm = _dam3d()
    a = _run(lay, prm, 100, precision)
    b = _run(lay, prm, 100, precision, {"RRSPH_HIP_LIST": "0"}, monkeypatch)
    assert a["lists"]["list_steps"] > 0 and b["lists"]["list_steps"] == 0
    assert np.array_equal(a["itype"], b["itype"])
    tol_r, tol_rho = (1e-12, 1e-12) if precision == H.FP64 else (1e-7, 1e-5)
    assert np.max(np.abs(a["r"] - b["r"])) < tol_r
    assert np.max(np.abs(a["rho"] - b["rho"]) / b["rho"]) < tol_rho


def test_3d_stale_lists_halt_and_rebuild(monkeypatch):
    """3D with a skin of 0.2% of the cut-off: stale-list halts and re-runs as
    builds; 40 FP64 steps still equal the oracle's."""
    lay, prm = _dam3d()
    g = _run(lay, prm, 40, H.FP64, {"RRSPH_HIP_LIST_SKIN": "0.002"}, monkeypatch)
    ls = g["lists"]
    print("3D tiny skin list stats", ls, flush=True)
    assert ls["builds"] >= 5
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(40) == 0
    r = ora.state()
    assert g["time"][2] == 40
    assert np.max(np.abs(g["r"] - r["r"])) < 1e-12
    assert np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]) < 1e-12
