"""The hand-written adjoint of the FP64 step (rrsph_hip_adjoint_*) against
central finite differences of the GPU forward and of the FP64 oracle, and
the smooth gauge surrogate against the max-in-window gauge."""
import copy

import numpy as np
import pytest

import oracle
from helpers import smooth_gauges_np
from sph2d_amd import generators as G
from sph2d_amd import hip as H
from sph2d_amd.adjoint import DifferentiableRun, PistonGauges, central_difference

pytestmark = pytest.mark.gpu

DELTA = 0.02


def _case(nsteps=60, every=16, nharm=4, mod=None):
    lay = G.adjoint_tank(delta=DELTA, length=1.2, depth=0.3, wall_height=0.6, nharm=nharm, sigma=0.01, seed=5)
    prm = lay.params
    if mod:
        mod(prm)
    sim = H.HipStep(prm, precision=H.FP64)
    spec = H.gauge_spec([0.08, 0.16, 0.3], radius=3 * DELTA, tau=DELTA / 4, y_ref=0.3)
    run = DifferentiableRun(sim, dict(lay.state), nsteps, every, spec, prm.piston_omega)
    return lay, run


def _weights(shape, seed=11):
    return np.random.default_rng(seed).normal(0.0, 1.0, shape)


def test_smooth_gauge_tends_to_the_max_gauge():
    """tau -> 0: the surrogate lies between max_i (y_i + tau log(w_i / sum w))
    and the max-in-window gauge (HeightTesting.h:25-54), a sandwich whose
    width tau log(sum w / w_top) vanishes with tau; at tau = delta / 4 it
    equals the numpy surrogate."""
    lay = G.piston_tank(delta=DELTA, length=1.2, depth=0.3, wall_height=0.6)
    sim = H.HipStep(lay.params, precision=H.FP64)
    sim.upload(lay.r, lay.v, lay.rho, lay.itype)
    sim.step(20)
    xs = [0.2, 0.5, 0.9]
    R = 3 * DELTA
    search_n = R / lay.params.hsml
    mx = sim.gauges(xs, search_n=search_n)
    tau = 1e-5
    spec = H.gauge_spec(xs, radius=R, tau=tau, y_ref=float(np.max(mx)) + 1e-4)
    sm = sim.smooth_gauges(spec)
    st = sim.download()
    fl = st["itype"] == 2
    x, y = st["r"][fl, 0], st["r"][fl, 1]
    for g, xg in enumerate(xs):
        u = 1.0 - (x - xg) ** 2 / R ** 2
        w = np.where(u > 0, u ** 3, 0.0)
        inside = w > 0
        lower = np.max(y[inside] + tau * np.log(w[inside] / w.sum()))
        assert lower - 1e-12 <= sm[g] <= mx[g] + 1e-12, (g, lower, sm[g], mx[g])
        assert mx[g] - lower < 1e-3  # the sandwich is narrow: the surrogate tracks the max
    spec2 = H.gauge_spec(xs, radius=R, tau=DELTA / 4, y_ref=0.3)
    assert np.allclose(sim.smooth_gauges(spec2), smooth_gauges_np(st["r"], st["itype"], xs, R, DELTA / 4, 0.3),
                       rtol=0, atol=1e-13)
    sim.close()


def test_piston_gradient_matches_central_differences():
    """dL/da_k, dL/db_k of L = sum c * eta over 60 steps (4 checkpoint
    segments): adjoint against central differences of the GPU forward."""
    lay, run = _case()
    prm = lay.params
    a, b = np.array(prm.piston_a), np.array(prm.piston_b)
    eta = run.forward(a, b)
    c = _weights(eta.shape)
    g = run.backward(c, len(a))
    loss = lambda e: float(np.sum(c * e))
    for which in ("a", "b"):
        for k in range(len(a)):
            fd = central_difference(run, a, b, loss, which, k, 1e-6)
            ad = g[which][k]
            print(which, k, ad, fd, flush=True)
            assert abs(fd) > 1e-4, (which, k, fd)  # the gauges see the piston: the check is not vacuous
            assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-3) + 1e-9, (which, k, ad, fd)
    run.sim.close()


def _moving_dot_case(mod=None, nsteps=24, every=5, wall_noise=False):
    """A moving start (10 steps in) and a random direction d in positions,
    fluid velocities and densities of every particle.  wall_noise: the walls
    of the start move with small random velocities, so that no pair sits
    exactly on the viscosity switch v_ij . r_ij = 0 (walls at rest do)."""
    lay, run = _case(nsteps=nsteps, every=every, mod=mod)
    prm = lay.params
    a, b = np.array(prm.piston_a), np.array(prm.piston_b)
    run.sim.upload(lay.r, lay.v, lay.rho, lay.itype)
    run.sim.set_piston(prm.piston_omega, a, b)
    run.sim.step(10)
    st = run.sim.download()
    if wall_noise:
        walls = st["itype"] < 0
        st["v"][walls] = np.random.default_rng(13).normal(0.0, 1e-3, (int(walls.sum()), 2))
    run.state = {"r": st["r"], "v": st["v"], "rho": st["rho"], "itype": st["itype"]}
    eta = run.forward(a, b)
    c = _weights(eta.shape, seed=3)
    g = run.backward(c, len(a))
    rng = np.random.default_rng(7)
    n = prm.ntotal
    d = {"r": rng.normal(0, 1e-3 * DELTA, (n, 2)), "v": rng.normal(0, 1e-2, (n, 2)), "rho": rng.normal(0, 1e-1, n)}
    d["v"][lay.itype < 0] = 0.0
    pred = float(sum(np.sum(g[k] * d[k]) for k in d))
    base = dict(run.state)

    def loss_at(eps):
        run.state = {"r": base["r"] + eps * d["r"], "v": base["v"] + eps * d["v"], "rho": base["rho"] + eps * d["rho"],
                     "itype": base["itype"]}
        return float(np.sum(c * run.forward(a, b)))

    return run, pred, loss_at


def test_initial_state_gradient_dot_test():
    """<dL/dS0, d> equals the directional difference of L along d, with a
    continuously differentiable artificial viscosity (alpha = 0, beta > 0:
    the mu^2 term has no kink where v_ij . r_ij changes sign)."""
    def smooth_av(p):
        p.artificial_shear_visc = 0.0
        p.artificial_bulk_visc = 0.5

    run, pred, loss_at = _moving_dot_case(smooth_av)
    fd = (loss_at(1e-4) - loss_at(-1e-4)) / 2e-4
    print("dot", pred, fd, flush=True)
    assert abs(fd) > 1e-6, fd
    assert abs(pred - fd) <= 1e-6 * abs(fd), (pred, fd)
    run.sim.close()


def test_initial_state_gradient_with_the_viscosity_switch():
    """alpha > 0: the term -alpha c mu switches off where v_ij . r_ij = 0, a
    kink of the step.  Walls at rest sit exactly on the switch: the adjoint
    takes the switched-off branch there and the one-sided differences along
    d differ (the kink is real).  With the walls moving slightly no pair sits
    on it, but a random direction at eps = 1e-4 still carries some of the
    ~1e6 pair-steps across the switch, so the central difference agrees with
    the adjoint only to the size of those kinks (measured 2e-3 relative; the
    piston gradients, whose perturbation crosses almost no switch, agree to
    1e-6 with alpha > 0)."""
    run, pred, loss_at = _moving_dot_case(wall_noise=True)
    fd = (loss_at(1e-4) - loss_at(-1e-4)) / 2e-4
    print("dot alpha", pred, fd, flush=True)
    assert abs(fd) > 1e-6, fd
    assert abs(pred - fd) <= 5e-3 * abs(fd), (pred, fd)
    run.sim.close()
    run, pred, loss_at = _moving_dot_case()
    L0 = loss_at(0.0)
    plus = (loss_at(1e-5) - L0) / 1e-5
    minus = (L0 - loss_at(-1e-5)) / 1e-5
    print("one-sided", plus, minus, pred, flush=True)
    assert abs(plus - minus) > 1e-6 * abs(pred)
    run.sim.close()


def test_torch_autograd_function():
    """PistonGauges is a torch.autograd.Function whose backward is the adjoint."""
    torch = pytest.importorskip("torch")
    lay, run = _case(nsteps=20, every=6)
    prm = lay.params
    a = torch.tensor(prm.piston_a, dtype=torch.float64, requires_grad=True)
    b = torch.tensor(prm.piston_b, dtype=torch.float64, requires_grad=True)
    eta = PistonGauges.apply(a, b, run)
    target = eta.detach() + 0.001
    loss = ((eta - target) ** 2).sum()
    loss.backward()
    g = run.backward(2 * (eta.detach().numpy() - target.numpy()), len(prm.piston_a))
    assert np.allclose(a.grad.numpy(), g["a"], rtol=1e-12, atol=0)
    assert np.allclose(b.grad.numpy(), g["b"], rtol=1e-12, atol=0)
    run.sim.close()


def test_piston_gradient_matches_oracle_finite_differences():
    """The adjoint gradient against central differences of the FP64 oracle
    (its own forward and numpy gauges): agreement within 1e-3 relative."""
    lay, run = _case(nsteps=40, every=10)
    prm = lay.params
    a, b = np.array(prm.piston_a), np.array(prm.piston_b)
    eta = run.forward(a, b)
    c = _weights(eta.shape, seed=5)
    g = run.backward(c, len(a))
    spec = run.spec
    xs = [spec.x[i] for i in range(spec.n)]

    def oracle_loss(aa, bb):
        p = copy.copy(prm)
        p.piston_a, p.piston_b = tuple(aa), tuple(bb)
        ora = oracle.Oracle(p)
        ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=p.dt)
        total = 0.0
        for k in range(run.nsteps):
            assert ora.step(1) == 0
            st = ora.state()
            total += float(np.sum(c[k] * smooth_gauges_np(st["r"], st["itype"], xs, spec.radius, spec.tau,
                                                          spec.y_ref)))
        return total

    eps = 1e-6
    for which, k in (("a", 0), ("b", 1)):
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


def test_fp32_trajectory_adjoint_matches_the_fp64_gradient():
    """Mixed precision: the adjoint on the FP32 step's trajectory (FP32 pair
    arithmetic, FP64 positions; adjoint arithmetic FP64) against the FP64
    adjoint, which central differences pin to 1e-6: within 1e-3 relative on
    every coefficient above a thousandth of the largest."""
    lay, run64 = _case(nsteps=40, every=10)
    prm = lay.params
    a, b = np.array(prm.piston_a), np.array(prm.piston_b)
    eta64 = run64.forward(a, b)
    c = _weights(eta64.shape, seed=5)
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


def test_adjoint_rejects_summation():
    lay, _ = _case()
    prm = copy.copy(lay.params)
    prm.density_treatment = 0
    sim = H.HipStep(prm, precision=H.FP64)
    sim.upload(lay.r, lay.v, lay.rho, lay.itype)
    spec = H.gauge_spec([0.1], radius=0.06, tau=0.005, y_ref=0.3)
    with pytest.raises(H.RrsphError) as e:
        sim.adjoint_forward(5, 5, spec)
    assert e.value.status == 7
    sim.close()


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_state_after_backward_is_the_forward_final_state(precision):
    """After a reverse pass over several checkpoint segments the context
    holds the forward's final state again: download() (pressure included)
    and probe('p') return what they returned after the forward; fields()
    refuses until a step has run, then works."""
    lay, run = _case(nsteps=40, every=12)
    run.sim.close()
    sim = H.HipStep(lay.params, precision=precision)
    sim.set_export_fields(True)
    run.sim = sim
    prm = lay.params
    a, b = np.array(prm.piston_a), np.array(prm.piston_b)
    eta = run.forward(a, b)
    pts = np.array([[0.3, 0.1], [0.6, 0.2], [0.9, 0.15]])
    before, pb = sim.download(), sim.probe(pts, "p")
    run.backward(_weights(eta.shape), len(a))
    after, pa = sim.download(), sim.probe(pts, "p")
    for k in ("r", "v", "rho", "p", "itype"):
        assert np.array_equal(before[k], after[k]), k
    assert np.array_equal(pb, pa)
    with pytest.raises(H.RrsphError):
        sim.fields()
    sim.step(1)
    assert np.all(np.isfinite(sim.fields()["a"]))
    sim.close()


def test_autograd_backward_after_another_forward_uses_its_own_tape():
    """Two PistonGauges.apply calls on one context summed into one loss:
    each backward gets the gradient of its own trajectory (the second
    forward replaced the tape of the first, which is re-recorded)."""
    import torch

    lay, run = _case(nsteps=24, every=8)
    prm = lay.params
    a0, b0 = np.array(prm.piston_a), np.array(prm.piston_b)
    a1, b1 = 1.5 * a0, -b0
    c = _weights((24, 3))
    ct = torch.from_numpy(c)
    ta0 = torch.tensor(a0, requires_grad=True)
    tb0 = torch.tensor(b0, requires_grad=True)
    ta1 = torch.tensor(a1, requires_grad=True)
    tb1 = torch.tensor(b1, requires_grad=True)
    loss = (ct * PistonGauges.apply(ta0, tb0, run)).sum() + (ct * PistonGauges.apply(ta1, tb1, run)).sum()
    loss.backward()
    for a, b, ta, tb in ((a0, b0, ta0, tb0), (a1, b1, ta1, tb1)):
        run.forward(a, b)
        g = run.backward(c, len(a))
        assert np.allclose(ta.grad.numpy(), g["a"], rtol=1e-12, atol=0)
        assert np.allclose(tb.grad.numpy(), g["b"], rtol=1e-12, atol=0)
    assert not np.allclose(ta0.grad.numpy(), ta1.grad.numpy())
    run.sim.close()


@pytest.mark.parametrize("cap", [None, "4"])
def test_taped_steps_in_one_graph_equal_the_launches_one_by_one(monkeypatch, cap):
    """With graphs on, each taped forward step with the incremental re-sort
    runs as one graph (the step, its record and its gauges, which take
    their rows from the control block's step count); with graphs off the
    same kernels are launched one by one.  The gauge histories are equal bit
    for bit and so are the gradients of the initial state (the piston
    gradients sum the wave maker's adjoints with atomics: 1e-12).  With room
    for 4 changed slots in the re-sort, steps halt and resume with the full
    sort, their records and gauges written again."""
    if cap:
        monkeypatch.setenv("RRSPH_HIP_RESORT_CAP", cap)
    res = []
    for graph in (True, False):
        lay = G.adjoint_tank(delta=DELTA, length=1.2, depth=0.3, wall_height=0.6, nharm=4, sigma=0.01, seed=5)
        prm = lay.params
        sim = H.HipStep(prm, precision=H.FP64, graph=graph)
        spec = H.gauge_spec([0.08, 0.16, 0.3], radius=3 * DELTA, tau=DELTA / 4, y_ref=0.3)
        run = DifferentiableRun(sim, dict(lay.state), 60, 16, spec, prm.piston_omega)
        a, b = np.array(prm.piston_a), np.array(prm.piston_b)
        eta = run.forward(a, b)
        grads = run.backward(_weights(eta.shape), len(a))
        res.append((eta, grads, sim.sort_stats()))
        sim.close()
    (e1, g1, s1), (e0, g0, s0) = res
    print("sort stats", s1, s0)
    assert s1["incremental"] > 0
    if cap:
        assert s1["fallbacks"] > 0
    assert np.array_equal(e1, e0)
    for k in ("r", "v", "rho"):
        assert np.array_equal(g1[k], g0[k]), k
    for k in ("a", "b"):
        assert np.allclose(g1[k], g0[k], rtol=1e-12, atol=1e-12 * np.abs(g0[k]).max()), k


@pytest.mark.parametrize("cap", [None, "4"])
def test_kept_states_equal_the_replayed_segments(monkeypatch, cap):
    """When every step's state fits in half the free device memory the
    forward keeps them all and the reverse pass replays nothing;
    RRSPH_HIP_ADJ_STORE=segments keeps one segment's and replays the others
    from their checkpoints.  A replay repeats its steps exactly, so the
    reverse steps see the same states and sorts: equal gauges and state
    gradients bit for bit (piston gradients: atomics, 1e-12), with and
    without re-sort halts in the forward and the replays."""
    if cap:
        monkeypatch.setenv("RRSPH_HIP_RESORT_CAP", cap)
    res = {}
    for store in ("all", "segments"):
        if store == "segments":
            monkeypatch.setenv("RRSPH_HIP_ADJ_STORE", "segments")
        else:
            monkeypatch.delenv("RRSPH_HIP_ADJ_STORE", raising=False)
        lay, run = _case(nsteps=60, every=16)
        prm = lay.params
        a, b = np.array(prm.piston_a), np.array(prm.piston_b)
        eta = run.forward(a, b)
        grads = run.backward(_weights(eta.shape), len(a))
        after = run.sim.download()
        res[store] = (eta, grads, after, run.sim.sort_stats())
        run.sim.close()
    (e1, g1, s1, q1), (e0, g0, s0, q0) = res["all"], res["segments"]
    print("sort stats", q1, q0)
    if cap:
        assert q1["fallbacks"] > 0
    assert q0["incremental"] + q0["full"] > q1["incremental"] + q1["full"]  # the replays ran
    assert np.array_equal(e1, e0)
    for k in ("r", "v", "rho"):
        assert np.array_equal(g1[k], g0[k]), k
        assert np.array_equal(s1[k], s0[k]), k  # the context ends on the forward's final state
    for k in ("a", "b"):
        assert np.allclose(g1[k], g0[k], rtol=1e-12, atol=1e-12 * np.abs(g0[k]).max()), k
