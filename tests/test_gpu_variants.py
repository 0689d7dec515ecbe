"""GPU parity of the reference behaviours outside the default configuration,
against the FP64 oracle through the C-ABI:

* the reference's own wave makers with boundary motion (RRSPH_CL):
  first- and second-order pistons (NWM_METHOD_DYNAMIC_1/2), the wall that
  disappears (NWM_METHOD_WALL_DISAPPEAR) and the solitary-wave paddle
  (NWM_METHOD_SOLITARY_RAYLEIGH), cl/TimeIntegration.cl:196-292 and
  CLTimeIntegration.cpp:369-413;
* the Desbrun kernel for every term (SmoothingKernel.h:220-240) through the
  general pair kernel;
* the "never ending simulation" stop of the dynamic time step
  (UpdateAcceleration.cpp:26-28), forced by a huge Lennard-Jones force.
"""
import copy

import numpy as np
import pytest

import oracle
from sph2d_amd import generators as G
from sph2d_amd import hip as H
from sph2d_amd import params as P

pytestmark = pytest.mark.gpu


def _scale_err(a, b):
    s = float(np.max(np.abs(b))) or 1.0
    return float(np.max(np.abs(a - b))) / s


def _wave_maker_case(nwm):
    """The coarse piston flume (4 m x 0.4 m, delta 0.02 m, dynamic CFL dt)
    with one of the reference's wave makers on its left wall."""
    lay = G.piston_tank(delta=0.02, length=4.0, depth=0.4)
    prm = copy.copy(lay.params)
    prm.nwm = nwm
    prm.piston_a, prm.piston_b = (), ()
    prm.nwm_wave_length = 2.0
    prm.nwm_wave_magnitude = 0.06
    prm.nwm_phase = 0.2
    prm.nwm_time_start = 0.0
    prm.start_simulation_time = 0.3
    if nwm == P.NWM_METHOD_SOLITARY_RAYLEIGH:
        # start at the paddle's fastest instant, t = Tf
        H_, h = prm.nwm_wave_magnitude, prm.depth
        k = np.sqrt(3 * H_ / (4 * h * h * (H_ + h)))
        c = np.sqrt(prm.g * (H_ + h))
        prm.start_simulation_time = 2 / (k * c) * (3.8 + H_ / h)
    P.post_fill_in_model_params(prm)
    return lay, prm


@pytest.mark.parametrize("nwm", [P.NWM_METHOD_DYNAMIC_1, P.NWM_METHOD_DYNAMIC_2, P.NWM_METHOD_WALL_DISAPPEAR,
                                 P.NWM_METHOD_SOLITARY_RAYLEIGH])
def test_reference_wave_makers_match_oracle(nwm):
    """20 FP64 steps with boundary motion: positions, velocities, densities,
    types, time and dt equal the oracle's; the wave-maker wall moved (or,
    for the disappearing wall, is gone)."""
    lay, prm = _wave_maker_case(nwm)
    print("wave maker nwm", nwm, "n", prm.ntotal, flush=True)
    gpu = H.HipStep(prm, precision=H.FP64)
    gpu.upload(lay.r, lay.v, lay.rho, lay.itype)
    gpu.step(20)
    g = gpu.download()
    t, dt, step = gpu.time()
    gpu.close()
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=prm.start_simulation_time, dt=prm.dt)
    assert ora.step(20) == 0
    r = ora.state()
    assert step == 20
    assert abs(t - r["time"]) <= 1e-12 * r["time"] and abs(dt - r["dt"]) <= 1e-12 * r["dt"]
    assert np.array_equal(g["itype"], r["itype"])
    assert np.max(np.abs(g["r"] - r["r"])) < 1e-11
    assert _scale_err(g["v"], r["v"]) < 1e-9
    assert np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]) < 1e-11
    wm = slice(prm.nwm_particles_start, prm.nwm_particles_end)
    if nwm == P.NWM_METHOD_WALL_DISAPPEAR:
        assert np.all(g["itype"][wm] == 0)
        assert np.all(g["itype"][prm.nwm_particles_end:] == P.TYPE_BOUNDARY)
    else:
        moved = g["r"][wm, 0] - lay.r[wm, 0]
        assert np.min(np.abs(moved)) > 1e-6, moved  # the wall moved along x ...
        assert np.max(np.abs(g["r"][wm, 1] - lay.r[wm, 1])) == 0.0  # ... and only along x
        assert np.all(g["v"][wm, 0] == g["v"][prm.nwm_particles_start, 0])


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_desbrun_kernel_matches_oracle(precision):
    """Desbrun's spiky kernel for density, pressure, viscosity and XSPH
    (summation and continuity): one step's fields and five steps' state."""
    lay = G.dam_break()
    for dens in (P.DENSITY_CONTINUITY, P.DENSITY_SUMMATION):
        prm = copy.copy(lay.params)
        prm.density_treatment = dens
        prm.density_skf = prm.intf_skf = prm.artificial_viscosity_skf = prm.average_velocity_skf = P.SKF_DESBRUN
        prm.artificial_pressure_skf = P.SKF_DESBRUN
        P.fill_in_computing_params(prm)
        assert prm.cell_scale_k == 2
        print("desbrun density", dens, "precision", precision, flush=True)
        gpu = H.HipStep(prm, precision=precision)
        gpu.set_export_fields(True)
        gpu.upload(lay.r, lay.v, lay.rho, lay.itype)
        gpu.step(1)
        f = gpu.fields()
        ora = oracle.Oracle(prm)
        ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
        assert ora.step(1) == 0
        fr = ora.fields()
        tol = 1e-12 if precision == H.FP64 else 2e-3
        assert _scale_err(f["a"], fr["a"]) < tol
        assert _scale_err(f["av"], fr["av"]) < tol
        if dens == P.DENSITY_CONTINUITY:
            assert _scale_err(f["drho"], fr["drho"]) < tol
        gpu.step(4)
        assert ora.step(4) == 0
        g, r = gpu.download(), ora.state()
        gpu.close()
        assert np.max(np.abs(g["r"] - r["r"])) < (1e-12 if precision == H.FP64 else 1e-7)
        assert np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]) < (1e-12 if precision == H.FP64 else 1e-5)


def test_never_ending_simulation_stops_like_the_reference():
    """Dynamic dt with the reductions on: a fluid particle 1e-6 m from a wall
    particle gets a Lennard-Jones acceleration of ~1e60 m/s^2, so dt falls
    below 1e-6 h / c and the run stops with "never ending simulation"
    before whole_step, as the oracle does; the state is the uploaded one."""
    lay = G.dam_break()
    prm = copy.copy(lay.params)
    prm.dt_correction_method = P.DT_CORRECTION_DYNAMIC
    prm.dt_dynamic_mode = 1
    prm.CFL_coef = 0.2
    r = lay.r.copy()
    wall = prm.nfluid + 200
    r[0] = r[wall] + np.array([1e-6, 7e-7])
    gpu = H.HipStep(prm, precision=H.FP64)
    gpu.upload(r, lay.v, lay.rho, lay.itype)
    with pytest.raises(H.RrsphError) as e:
        gpu.step(3)
    assert e.value.status == 4 and "never ending simulation" in str(e.value)
    t, _dt, step = gpu.time()
    g = gpu.download()
    gpu.close()
    ora = oracle.Oracle(prm)
    ora.set_state(r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(3) == 2  # the oracle's "never ending simulation"
    ref = ora.state()
    assert step == ref["step"] == 0 and t == 0.0
    assert np.array_equal(g["r"], r) and np.array_equal(ref["r"], r)
    assert np.array_equal(g["v"], lay.v) and np.array_equal(g["rho"], lay.rho)
