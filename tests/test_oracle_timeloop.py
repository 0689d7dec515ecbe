"""Pins the oracle's time loop (TimeIntegration.cpp:60-102) where no golden
vector of the reference reaches: predict_half_step, whole_step with the
first-step dt/2, the dt rules, the piston wave maker, FIX and STOP.

Each check is independent of the oracle's own loop code: closed-form
kinematics of particles that feel gravity only, a finite difference of the
piston displacement, and a restatement of the loop body composed from the
oracle's golden-pinned modules (grid_find, density, the force terms).
"""
import copy
import ctypes

import numpy as np
import pytest

import oracle
from sph2d_amd import generators as G
from sph2d_amd import params as PP


def _lonely_case(boundary_motion=False):
    """One fluid particle and one wall particle far apart (no neighbours)."""
    prm = copy.copy(G.dam_break().params)
    prm.ntotal, prm.nfluid, prm.nvirt = 2, 1, 1
    prm.boundary_motion = boundary_motion
    r = np.array([[1.0, 0.8], [2.5, 0.3]])
    v = np.array([[0.3, 0.1], [0.05, 0.02]])
    rho = np.array([1000.0, 1000.0])
    itype = np.array([2, -2], dtype=np.int32)
    return prm, r, v, rho, itype


@pytest.mark.parametrize("boundary_motion", [False, True])
def test_free_fall_kinematics(boundary_motion):
    """No neighbours: a = (0, -g), drho = 0, av = 0.  whole_step: the first
    step advances v by a dt/2, later steps by a dt; r += v_new dt; walls stay
    put (RRSPH_OMP) or drift with their velocity (RRSPH_CL boundary motion)."""
    prm, r, v, rho, itype = _lonely_case(boundary_motion)
    ora = oracle.Oracle(prm)
    ora.set_state(r, v, rho, itype, time=0.0, dt=prm.dt)
    assert ora.step(3) == 0
    st = ora.state()
    g = np.array([0.0, -prm.g])
    dt = prm.dt
    vf, rf = v[0].copy(), r[0].copy()
    for k in range(3):
        vf = vf + g * (dt * 0.5 if k == 0 else dt)
        rf = rf + vf * dt
    assert np.allclose(st["v"][0], vf, rtol=0, atol=1e-15)
    assert np.allclose(st["r"][0], rf, rtol=0, atol=1e-15)
    assert np.array_equal(st["v"][1], v[1])
    want_wall = r[1] + 3 * dt * v[1] if boundary_motion else r[1]
    assert np.allclose(st["r"][1], want_wall, rtol=0, atol=1e-15)
    assert np.array_equal(st["rho"], rho)
    assert st["step"] == 3 and abs(st["time"] - 3 * dt) < 1e-18
    f = ora.fields()
    assert np.allclose(f["a"], np.tile(g, (2, 1)))
    assert not f["av"].any() and not f["drho"].any()


@pytest.mark.parametrize("boundary_motion", [False, True])
def test_3d_free_fall_kinematics(boundary_motion):
    """The same in 3D (gravity along -y, the vertical axis of the 3D cases):
    every component of v and r follows the leapfrog closed form."""
    prm = copy.copy(G.dam_break_3d().params)
    prm.ntotal, prm.nfluid, prm.nvirt = 2, 1, 1
    prm.boundary_motion = boundary_motion
    prm.density_treatment = 1  # continuity: with no neighbours drho = 0
    r = np.array([[0.4, 0.3, 0.2], [0.7, 0.05, 0.35]])
    v = np.array([[0.3, 0.1, -0.2], [0.05, 0.02, 0.04]])
    rho = np.array([1000.0, 1000.0])
    itype = np.array([2, -2], dtype=np.int32)
    ora = oracle.Oracle(prm)
    ora.set_state(r, v, rho, itype, time=0.0, dt=prm.dt)
    assert ora.step(3) == 0
    st = ora.state()
    g = np.array([0.0, -prm.g, 0.0])
    dt = prm.dt
    vf, rf = v[0].copy(), r[0].copy()
    for k in range(3):
        vf = vf + g * (dt * 0.5 if k == 0 else dt)
        rf = rf + vf * dt
    assert np.allclose(st["v"][0], vf, rtol=0, atol=1e-15)
    assert np.allclose(st["r"][0], rf, rtol=0, atol=1e-15)
    want_wall = r[1] + 3 * dt * v[1] if boundary_motion else r[1]
    assert np.allclose(st["r"][1], want_wall, rtol=0, atol=1e-15)
    assert np.array_equal(st["v"][1], v[1]) and np.array_equal(st["rho"], rho)
    assert np.allclose(ora.fields()["a"], np.tile(g, (2, 1)))


def test_fix_kills_and_stop_halts():
    """A fluid particle leaving the box: FIX turns it into itype 0 and keeps
    its position (WholeStep.cpp:46-53); STOP returns the outside status
    (ConsistencyCheck.cpp:184-195); PRINT keeps going."""
    prm, r, v, rho, itype = _lonely_case()
    r[0] = [prm.x_maxgeom - 1e-4, 0.8]
    v[0] = [10.0, 0.0]
    for treatment, status in [(2, 0), (1, 3), (0, 0)]:
        p = copy.copy(prm)
        p.consistency_treatment = treatment
        p.consistency_check = True
        ora = oracle.Oracle(p)
        ora.set_state(r, v, rho, itype, time=0.0, dt=p.dt)
        assert ora.step(1) == status
        st = ora.state()
        if treatment == 2:
            assert st["itype"][0] == 0 and np.array_equal(st["r"][0], r[0])
            assert ora.outside() == 0
        else:
            assert st["itype"][0] == 2 and st["r"][0][0] > p.x_maxgeom
            assert ora.outside() == 1


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_3d_fix_kills_and_stop_halts_on_every_face(axis):
    """3D: a fluid particle leaving the box through the upper face of each
    axis (z included): FIX turns it into itype 0 and keeps its position,
    STOP returns the outside status, PRINT keeps going with it outside."""
    prm = copy.copy(G.dam_break_3d().params)
    prm.ntotal, prm.nfluid, prm.nvirt = 2, 1, 1
    lo = np.array([prm.x_mingeom, prm.y_mingeom, prm.z_mingeom])
    hi = np.array([prm.x_maxgeom, prm.y_maxgeom, prm.z_maxgeom])
    mid = 0.5 * (lo + hi)
    r = np.array([mid, lo + 0.1 * (hi - lo)])
    r[0, axis] = hi[axis] - 1e-4
    v = np.zeros((2, 3))
    v[0, axis] = 10.0
    rho = np.array([1000.0, 1000.0])
    itype = np.array([2, -2], dtype=np.int32)
    prm.g = 0.0  # only the push across the face moves the particle
    for treatment, status in [(2, 0), (1, 3), (0, 0)]:
        p = copy.copy(prm)
        p.consistency_treatment = treatment
        p.consistency_check = True
        ora = oracle.Oracle(p)
        ora.set_state(r, v, rho, itype, time=0.0, dt=p.dt)
        assert ora.step(1) == status
        st = ora.state()
        if treatment == 2:
            assert st["itype"][0] == 0 and np.array_equal(st["r"][0], r[0])
            assert ora.outside() == 0
        else:
            assert st["itype"][0] == 2 and st["r"][0][axis] > hi[axis]
            assert ora.outside() == 1


def test_dt_rules():
    """dt_correction_method 0 keeps params.dt; method 2 in the OpenMP build's
    mode (disabled reductions) gives CFL h / c; in the reduction mode
    dt = CFL min(sqrt(h / max|a|), h / (c + max mu)) (UpdateAcceleration.cpp:15-29)
    of the accelerations of the same step."""
    lay = G.dam_break()
    prm = copy.copy(lay.params)
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(2) == 0 and ora.state()["dt"] == prm.dt
    prm.dt_correction_method = 2
    prm.CFL_coef = 0.2
    prm.dt_dynamic_mode = 0
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(2) == 0
    assert ora.state()["dt"] == prm.CFL_coef * (prm.hsml / prm.eos_c)
    prm.dt_dynamic_mode = 1
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(1) == 0
    f, st = ora.fields(), ora.state()
    amax = float(np.max(np.sqrt(f["a"][:, 0] ** 2 + f["a"][:, 1] ** 2)))
    mumax = float(np.max(f["arvmu"]))
    want = prm.CFL_coef * min(np.sqrt(prm.hsml / amax), prm.hsml / (prm.eos_c + mumax))
    assert abs(st["dt"] - want) <= 1e-15 * want
    # time advanced by the dt of the step itself
    assert st["time"] == st["dt"]


def test_fourier_piston_velocity_is_the_derivative_of_its_displacement():
    """nwm = 6: the wall velocity equals d/dt of x(t) = sum a_k sin(k w t) +
    b_k cos(k w t) (central difference, step 1e-6 s)."""
    prm = copy.copy(G.dam_break().params)
    rng = np.random.default_rng(3)
    prm.nwm = 6
    prm.piston_omega = 2 * np.pi / 1.6
    prm.piston_a = tuple(rng.normal(0, 0.005, 16))
    prm.piston_b = tuple(rng.normal(0, 0.005, 16))
    op = oracle.from_params(prm)
    L = oracle.lib()
    L.oracle_piston_velocity.restype = ctypes.c_double
    L.oracle_piston_velocity.argtypes = [ctypes.POINTER(oracle.OracleParams), ctypes.c_double]
    k = np.arange(1, 17)

    def x(t):
        return float(np.sum(np.array(prm.piston_a) * np.sin(k * prm.piston_omega * t)
                            + np.array(prm.piston_b) * np.cos(k * prm.piston_omega * t)))

    for t in [0.0, 0.137, 0.9, 2.5]:
        h = 1e-6
        fd = (x(t + h) - x(t - h)) / (2 * h)
        assert abs(L.oracle_piston_velocity(ctypes.byref(op), t) - fd) < 1e-8


def test_first_order_piston_velocity():
    """nwm = 2: v = S0 w / 2 cos(w t + phase - w t_start), the derivative of
    the paddle stroke S0/2 sin(...) (ModelParams.tex piston wave maker;
    cl/TimeIntegration.cl:206-215), with S0 = H / m1 from Input.cpp:125-137."""
    prm = copy.copy(G.dam_break().params)
    prm.nwm = 2
    prm.nwm_wave_length = 2.0
    prm.nwm_wave_magnitude = 0.05
    prm.depth = 0.5
    prm.nwm_phase = 0.3
    prm.nwm_time_start = 0.1
    PP.post_fill_in_model_params(prm)
    kd = prm.nwm_wave_number * prm.depth
    m1 = 2 * np.sinh(kd) ** 2 / (np.sinh(kd) * np.cosh(kd) + kd)
    S0 = prm.nwm_wave_magnitude / m1
    w = prm.nwm_freq
    assert abs(w ** 2 - prm.g * prm.nwm_wave_number * np.tanh(kd)) < 1e-12 * w ** 2
    op = oracle.from_params(prm)
    L = oracle.lib()
    L.oracle_piston_velocity.restype = ctypes.c_double
    L.oracle_piston_velocity.argtypes = [ctypes.POINTER(oracle.OracleParams), ctypes.c_double]
    for t in [0.1, 0.4, 1.3]:
        stroke = lambda s: 0.5 * S0 * np.sin(w * s + prm.nwm_phase - w * prm.nwm_time_start)
        h = 1e-6
        fd = (stroke(t + h) - stroke(t - h)) / (2 * h)
        assert abs(L.oracle_piston_velocity(ctypes.byref(op), t) - fd) < 1e-8


def _composed_steps(prm, lay, nsteps):
    """The loop body of TimeIntegration.cpp:60-102 composed from the oracle's
    module functions (grid_find, density_con/sum, update_acceleration), with
    predict_half_step (PredictHalfStep.cpp:4-23), update_dt
    (UpdateAcceleration.cpp:15-29) and whole_step (WholeStep.cpp:19-61)
    written out here."""
    L = oracle.lib()
    op = oracle.from_params(prm)
    n, D = prm.ntotal, prm.dim
    r, v = lay.r.copy(), lay.v.copy()
    rho, itype = lay.rho.copy(), lay.itype.copy()
    a = np.zeros((n, D)); drho = np.zeros(n)
    dt, time = prm.dt, 0.0
    cont = prm.density_treatment != 0
    for step in range(nsteps):
        rho_pred = rho + drho * dt * 0.5 if cont else None
        v_pred = v + a * dt * 0.5
        st, nb = oracle.grid_find(prm, r, itype)
        assert st == 0
        nbf = np.ascontiguousarray(nb.reshape(-1))
        p = np.zeros(n)
        if cont:
            drho = np.zeros(n)
            L.oracle_density_con(ctypes.byref(op), oracle.ptr(r), oracle.ptr(v), oracle.ptr(nbf, ctypes.c_uint32),
                                 oracle.ptr(rho), oracle.ptr(drho), oracle.ptr(p))
        else:
            L.oracle_density_sum(ctypes.byref(op), oracle.ptr(r), oracle.ptr(nbf, ctypes.c_uint32),
                                 oracle.ptr(rho), oracle.ptr(p))
        a = np.zeros((n, D)); av = np.zeros((n, D)); mu = np.zeros(n)
        amax, mumax = ctypes.c_double(), ctypes.c_double()
        L.oracle_update_acceleration(ctypes.byref(op), oracle.ptr(itype, ctypes.c_int32), oracle.ptr(r),
                                     oracle.ptr(v_pred), oracle.ptr(nbf, ctypes.c_uint32),
                                     oracle.ptr(rho_pred if cont else rho), oracle.ptr(p), oracle.ptr(a),
                                     oracle.ptr(av), oracle.ptr(mu), ctypes.byref(amax), ctypes.byref(mumax))
        if prm.dt_correction_method == 2:
            am, mm = (amax.value, mumax.value) if prm.dt_dynamic_mode == 1 else (0.0, 0.0)
            with np.errstate(divide="ignore"):
                dt = prm.CFL_coef * min(np.sqrt(prm.hsml / am) if am > 0 else np.inf, prm.hsml / (prm.eos_c + mm))
        vdt = dt if step else dt * 0.5
        if cont:
            rho = rho + drho * vdt
        fl = itype > 0
        v[fl] = v[fl] + a[fl] * vdt
        if prm.average_velocity:
            v[fl] = v[fl] + av[fl]
        r[fl] = r[fl] + v[fl] * dt
        time += dt
    return {"r": r, "v": v, "rho": rho, "time": time, "dt": dt, "a": a, "drho": drho}


@pytest.mark.parametrize("density_treatment,dt_method", [(1, 0), (0, 0), (1, 2)])
def test_loop_matches_composed_modules(density_treatment, dt_method):
    """Three steps of the oracle's loop equal the loop body composed from its
    golden-pinned modules (same arithmetic, so to the last bit or nearly)."""
    lay = G.dam_break(density_treatment=density_treatment)
    prm = copy.copy(lay.params)
    if dt_method == 2:
        prm.dt_correction_method = 2
        prm.CFL_coef = 0.2
        prm.dt_dynamic_mode = 1
    want = _composed_steps(prm, lay, 3)
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(3) == 0
    st, f = ora.state(), ora.fields()
    assert np.max(np.abs(st["r"] - want["r"])) < 1e-15
    assert np.max(np.abs(st["v"] - want["v"])) < 1e-12
    assert np.max(np.abs(st["rho"] - want["rho"]) / want["rho"]) < 1e-15
    assert abs(st["time"] - want["time"]) < 1e-18 and st["dt"] == want["dt"]
    assert np.max(np.abs(f["a"] - want["a"])) < 1e-9


@pytest.mark.parametrize("density_treatment", [1, 0])
def test_3d_loop_matches_composed_modules(density_treatment):
    """The same for the 3D dam break (16k particles): three steps of the
    oracle's 3D loop equal its 3D modules composed, whose pair terms the
    planar lifts of the golden case pin."""
    lay = G.dam_break_3d()
    prm = copy.copy(lay.params)
    prm.density_treatment = density_treatment
    want = _composed_steps(prm, lay, 3)
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(3) == 0
    st, f = ora.state(), ora.fields()
    assert np.max(np.abs(st["r"] - want["r"])) < 1e-15
    assert np.max(np.abs(st["v"] - want["v"])) < 1e-12
    assert np.max(np.abs(st["rho"] - want["rho"]) / want["rho"]) < 1e-15
    assert abs(st["time"] - want["time"]) < 1e-18 and st["dt"] == want["dt"]
    assert np.max(np.abs(f["a"] - want["a"])) < 1e-9
    assert np.max(np.abs(st["v"] - lay.v)) > 1e-4  # the particles moved


@pytest.mark.parametrize("dim", [2, 3])
def test_update_acceleration_is_the_sum_of_the_pinned_terms(dim):
    """update_acceleration's a = external (gravity + LJ) + internal
    (pressure) + artificial viscosity, each of which matches the reference's
    golden vectors (in 3D: through the planar lifts of the golden case); its
    XSPH output equals the pinned average-velocity term and its mu the
    pinned AV mu (UpdateAcceleration.cpp:59-100)."""
    lay = G.dam_break() if dim == 2 else G.dam_break_3d()
    prm = lay.params
    L = oracle.lib()
    op = oracle.from_params(prm)
    n, D = prm.ntotal, prm.dim
    r, rho, it = lay.r, lay.rho, lay.itype
    v = np.random.default_rng(1).normal(0, 0.05, (n, D))
    rho = rho * (1 + np.random.default_rng(2).normal(0, 1e-3, n))
    st, nb = oracle.grid_find(prm, r, it)
    nbf = np.ascontiguousarray(nb.reshape(-1))
    p = np.zeros(n); drho = np.zeros(n)
    L.oracle_density_con(ctypes.byref(op), oracle.ptr(r), oracle.ptr(v), oracle.ptr(nbf, ctypes.c_uint32),
                         oracle.ptr(rho), oracle.ptr(drho), oracle.ptr(p))
    a = np.zeros((n, D)); av = np.zeros((n, D)); mu = np.zeros(n)
    amax, mumax = ctypes.c_double(), ctypes.c_double()
    L.oracle_update_acceleration(ctypes.byref(op), oracle.ptr(it, ctypes.c_int32), oracle.ptr(r), oracle.ptr(v),
                                 oracle.ptr(nbf, ctypes.c_uint32), oracle.ptr(rho), oracle.ptr(p), oracle.ptr(a),
                                 oracle.ptr(av), oracle.ptr(mu), ctypes.byref(amax), ctypes.byref(mumax))
    ext = np.zeros((n, D)); inf = np.zeros((n, D)); arv = np.zeros((n, D)); mu2 = np.zeros(n)
    xs = np.zeros((n, D))
    L.oracle_term_external(ctypes.byref(op), oracle.ptr(r), oracle.ptr(it, ctypes.c_int32),
                           oracle.ptr(nbf, ctypes.c_uint32), oracle.ptr(ext))
    L.oracle_term_internal(ctypes.byref(op), oracle.ptr(r), oracle.ptr(nbf, ctypes.c_uint32), oracle.ptr(p),
                           oracle.ptr(rho), oracle.ptr(inf))
    L.oracle_term_art_visc(ctypes.byref(op), oracle.ptr(r), oracle.ptr(v), oracle.ptr(nbf, ctypes.c_uint32),
                           oracle.ptr(rho), oracle.ptr(arv), oracle.ptr(mu2))
    L.oracle_term_average_velocity(ctypes.byref(op), oracle.ptr(r), oracle.ptr(v), oracle.ptr(it, ctypes.c_int32),
                                   oracle.ptr(nbf, ctypes.c_uint32), oracle.ptr(rho), n, oracle.ptr(xs))
    want = ext + inf + arv  # the external term carries gravity, as the reference's exdvdt targets do
    scale = np.max(np.abs(want))
    assert np.max(np.abs(a - want)) <= 1e-12 * scale
    assert np.max(np.abs(mu - mu2)) <= 1e-12 * max(np.max(mu2), 1e-30)
    assert np.max(np.abs(av - xs)) <= 1e-12 * max(np.max(np.abs(xs)), 1e-30)
    assert amax.value == pytest.approx(np.max(np.linalg.norm(a, axis=1)), rel=1e-14)
    assert mumax.value == pytest.approx(np.max(mu), rel=1e-14)
