"""The segmented Fourier piston and the adjoint of the FP64 3D step
(rrsph_hip_adjoint_* on a 3D context): paddle velocities against the
closed form, the 3D step with paddles against the oracle, the (x, z) smooth
gauges against numpy, and gradients against central differences of the GPU
forward and of the FP64 oracle."""
import copy

import numpy as np
import pytest

import oracle
from helpers import smooth_gauges_np
from sph2d_amd import generators as G
from sph2d_amd import hip as H
from sph2d_amd.adjoint import DifferentiableRun, central_difference

pytestmark = pytest.mark.gpu

DELTA = 0.04
# one gauge in front of each of the 8 paddles (0.25 m wide across z)
GX, GZ = [0.12] * 8, [0.125 + 0.25 * g for g in range(8)]


def _basin(mod=None):
    lay = G.piston_basin_3d(delta=DELTA, segments=8, nharm=4, sigma=0.01, seed=9)
    prm = lay.params
    if mod:
        mod(prm)
    return lay, prm


def _spec(prm):
    return H.gauge_spec(GX, radius=3 * DELTA, tau=DELTA / 4, y_ref=prm.depth, zs=GZ)


def _case(nsteps=30, every=8, mod=None):
    lay, prm = _basin(mod)
    sim = H.HipStep(prm, precision=H.FP64)
    run = DifferentiableRun(sim, dict(lay.state), nsteps, every, _spec(prm), prm.piston_omega)
    return lay, run


def _paddle_velocity(prm, t, g):
    H_ = len(prm.piston_a) // prm.piston_segments
    a = np.array(prm.piston_a[g * H_:(g + 1) * H_])
    b = np.array(prm.piston_b[g * H_:(g + 1) * H_])
    k = np.arange(1, H_ + 1) * prm.piston_omega
    return float(np.sum(k * (a * np.cos(k * t) - b * np.sin(k * t))))


def test_paddles_move_with_their_own_series_and_match_the_oracle():
    """After three steps every piston particle moves along x with the velocity
    of its paddle (floor(z * 8 / W)) at the last step's start time, and the
    whole FP64 state equals the oracle's (which assigns paddles itself)."""
    lay, prm = _basin()
    print("basin n", prm.ntotal, flush=True)
    sim = H.HipStep(prm, precision=H.FP64)
    sim.upload(lay.r, lay.v, lay.rho, lay.itype)
    sim.step(3)
    st = sim.download()
    sim.close()
    lo, hi = prm.nwm_particles_start, prm.nwm_particles_end
    z = lay.r[lo:hi, 2]
    seg = np.clip(np.floor((z - prm.piston_seg_zmin) * (8 / (prm.piston_seg_zmax - prm.piston_seg_zmin))), 0, 7)
    t_last = 2 * prm.dt
    want = np.array([_paddle_velocity(prm, t_last, int(g)) for g in seg])
    assert len(np.unique(seg)) == 8
    assert np.allclose(st["v"][lo:hi, 0], want, rtol=1e-12, atol=1e-15)
    assert np.all(st["v"][lo:hi, 1:] == 0.0)
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(3) == 0
    ref = ora.state()
    assert np.max(np.abs(st["r"] - ref["r"])) < 1e-12
    assert np.max(np.abs(st["v"] - ref["v"])) < 1e-10
    assert np.max(np.abs(st["rho"] - ref["rho"]) / ref["rho"]) < 1e-12


def test_3d_smooth_gauges_match_numpy():
    lay, prm = _basin()
    sim = H.HipStep(prm, precision=H.FP64)
    sim.upload(lay.r, lay.v, lay.rho, lay.itype)
    sim.step(5)
    st = sim.download()
    got = sim.smooth_gauges(_spec(prm))
    sim.close()
    want = smooth_gauges_np(st["r"], st["itype"], GX, 3 * DELTA, DELTA / 4, prm.depth, zs=GZ)
    assert np.all(want > 0)
    assert np.allclose(got, want, rtol=0, atol=1e-13), (got, want)


def test_3d_paddle_gradient_matches_central_differences():
    """dL/da, dL/db of L = sum c * eta over 30 steps (4 checkpoint segments)
    for coefficients of every paddle: adjoint against central differences
    of the GPU forward."""
    lay, run = _case()
    prm = lay.params
    a, b = np.array(prm.piston_a), np.array(prm.piston_b)
    eta = run.forward(a, b)
    c = np.random.default_rng(11).normal(0.0, 1.0, eta.shape)
    g = run.backward(c, len(a))
    assert g["a"].shape == (32,) and g["r"].shape == (prm.ntotal, 3)
    loss = lambda e: float(np.sum(c * e))
    for which, k in (("a", 0), ("b", 1), ("a", 9), ("b", 14), ("a", 19), ("b", 23), ("a", 28), ("b", 31)):
        fd = central_difference(run, a, b, loss, which, k, 1e-6)
        ad = g[which][k]
        print(which, k, ad, fd, flush=True)
        assert abs(fd) > 1e-4, (which, k, fd)  # the gauges see the paddle: the check is not vacuous
        # 1e-5: the FD of a 3D run carries ~1e-6 relative rounding noise at eps = 1e-6
        assert abs(ad - fd) <= 1e-5 * abs(fd), (which, k, ad, fd)
    run.sim.close()


def _smooth_av(p):
    p.artificial_shear_visc = 0.0
    p.artificial_bulk_visc = 0.5


@pytest.mark.parametrize("mod,wall_noise", [(_smooth_av, False), (None, True)], ids=["alpha0", "alpha_moving_walls"])
def test_3d_initial_state_gradient_dot_test(mod, wall_noise):
    """<dL/dS0, d> equals the directional difference of L along d (positions,
    fluid velocities, densities) from a moving start: to 1e-6 with alpha = 0
    (no switch); with alpha > 0 and the walls moving slightly (no pair
    exactly on the viscosity switch v_ij . r_ij = 0) to 5e-3, the size of
    the switch kinks the difference at eps = 1e-4 crosses."""
    lay, run = _case(nsteps=12, every=5, mod=mod)
    prm = lay.params
    a, b = np.array(prm.piston_a), np.array(prm.piston_b)
    run.sim.upload(lay.r, lay.v, lay.rho, lay.itype)
    run.sim.set_piston(prm.piston_omega, a, b)
    run.sim.step(10)
    st = run.sim.download()
    if wall_noise:
        walls = st["itype"] < 0
        st["v"][walls] = np.random.default_rng(13).normal(0.0, 1e-3, (int(walls.sum()), 3))
    run.state = {"r": st["r"], "v": st["v"], "rho": st["rho"], "itype": st["itype"]}
    eta = run.forward(a, b)
    c = np.random.default_rng(3).normal(0.0, 1.0, eta.shape)
    g = run.backward(c, len(a))
    rng = np.random.default_rng(7)
    n = prm.ntotal
    d = {"r": rng.normal(0, 1e-3 * DELTA, (n, 3)), "v": rng.normal(0, 1e-2, (n, 3)), "rho": rng.normal(0, 1e-1, n)}
    d["v"][lay.itype < 0] = 0.0
    pred = float(sum(np.sum(g[k] * d[k]) for k in d))
    base = dict(run.state)

    def loss_at(eps):
        run.state = {"r": base["r"] + eps * d["r"], "v": base["v"] + eps * d["v"], "rho": base["rho"] + eps * d["rho"],
                     "itype": base["itype"]}
        return float(np.sum(c * run.forward(a, b)))

    fd = (loss_at(1e-4) - loss_at(-1e-4)) / 2e-4
    print("dot", pred, fd, flush=True)
    assert abs(fd) > 1e-6, fd
    assert abs(pred - fd) <= (5e-3 if wall_noise else 1e-6) * abs(fd), (pred, fd)
    run.sim.close()


def test_3d_paddle_gradient_matches_oracle_finite_differences():
    """The adjoint gradient against central differences of the FP64 oracle's
    3D run with paddles (its own forward, numpy gauges): within 1e-3."""
    lay, run = _case(nsteps=16, every=6)
    prm = lay.params
    a, b = np.array(prm.piston_a), np.array(prm.piston_b)
    eta = run.forward(a, b)
    c = np.random.default_rng(5).normal(0.0, 1.0, eta.shape)
    g = run.backward(c, len(a))

    def oracle_loss(aa, bb):
        p = copy.copy(prm)
        p.piston_a, p.piston_b = tuple(aa), tuple(bb)
        ora = oracle.Oracle(p)
        ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=p.dt)
        total = 0.0
        for k in range(run.nsteps):
            assert ora.step(1) == 0
            st = ora.state()
            total += float(np.sum(c[k] * smooth_gauges_np(st["r"], st["itype"], GX, 3 * DELTA, DELTA / 4,
                                                          prm.depth, zs=GZ)))
        return total

    eps = 1e-6
    for which, k in (("a", 4), ("a", 27)):
        aa, bb = a.copy(), b.copy()
        (aa if which == "a" else bb)[k] += eps
        lp = oracle_loss(aa, bb)
        (aa if which == "a" else bb)[k] -= 2 * eps
        lm = oracle_loss(aa, bb)
        fd = (lp - lm) / (2 * eps)
        print(which, k, g[which][k], fd, flush=True)
        assert abs(fd) > 1e-4, (which, k, fd)
        assert abs(g[which][k] - fd) <= 1e-3 * abs(fd), (which, k, g[which][k], fd)
    run.sim.close()


def test_3d_fp32_trajectory_adjoint_matches_the_fp64_gradient():
    """Mixed precision in 3D (the basin with paddles): the adjoint on the FP32
    trajectory against the FP64 adjoint, within 1e-3 relative on every
    coefficient above a thousandth of the largest."""
    lay, run64 = _case(nsteps=20, every=8)
    prm = lay.params
    a, b = np.array(prm.piston_a), np.array(prm.piston_b)
    eta64 = run64.forward(a, b)
    c = np.random.default_rng(17).normal(0.0, 1.0, eta64.shape)
    g64 = run64.backward(c, len(a))
    run64.sim.close()
    sim32 = H.HipStep(prm, precision=H.FP32)
    run32 = DifferentiableRun(sim32, dict(lay.state), run64.nsteps, run64.every, run64.spec, prm.piston_omega)
    eta32 = run32.forward(a, b)
    assert np.max(np.abs(eta32 - eta64)) < 1e-5
    g32 = run32.backward(c, len(a))
    sim32.close()
    for which in ("a", "b"):
        big = np.abs(g64[which]) > 1e-3 * np.max(np.abs(g64[which]))
        err = np.abs(g32[which] - g64[which])[big] / np.abs(g64[which])[big]
        print(which, "fp32 vs fp64 max rel", float(err.max()), flush=True)
        assert err.max() < 1e-3, (which, err.max())


@pytest.mark.parametrize("mod", [_smooth_av, None], ids=["alpha0", "alpha"])
def test_3d_adjoint_with_lists_equals_the_sorted_adjoint(monkeypatch, mod):
    """The tape of list steps (builds at every checkpoint, the forward's
    schedule repeated by the replays, the reverse steps on each build's
    lists) and the tape that sorts every step give the same gauges and
    gradients: the neighbour sets are the same, only summation orders
    differ.  With alpha > 0 the step has the viscosity switch as a kink,
    and pairs of the two slightly different trajectories can sit on either
    side of it: the initial-state gradients are then compared only with
    the smooth step (alpha = 0)."""
    out = {}
    for lists in ("1", "0"):
        monkeypatch.setenv("RRSPH_HIP_LIST", lists)
        lay, run = _case(nsteps=30, every=8, mod=mod)
        prm = lay.params
        a, b = np.array(prm.piston_a), np.array(prm.piston_b)
        eta = run.forward(a, b)
        c = np.random.default_rng(5).normal(0.0, 1.0, eta.shape)
        g = run.backward(c, len(a))
        out[lists] = (eta, g, run.sim.list_stats())
        run.sim.close()
    print("list stats", out["1"][2], out["0"][2], flush=True)
    assert out["1"][2]["builds"] >= 4 and out["1"][2]["list_steps"] > 0 and out["0"][2]["interval"] == 0
    e1, e0 = out["1"][0], out["0"][0]
    assert np.max(np.abs(e1 - e0)) <= 1e-12 * np.max(np.abs(e0))
    for key in ("a", "b", "r", "v", "rho") if mod else ("a", "b"):
        g1, g0 = out["1"][1][key], out["0"][1][key]
        err = float(np.max(np.abs(g1 - g0))) / max(float(np.max(np.abs(g0))), 1e-300)
        print(key, "max rel difference", err, flush=True)
        assert err < 1e-9, key
