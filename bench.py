"""Benchmark of the RRSPH_HIP WCSPH step on one MI355X.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5]

Workload (default): the 2D piston wave flume at ~1M particles ("C1-1M"):
20 m x 0.5 m of water at spacing 0.0031623 m (999,350 fluid particles,
1,020,263 in total with 3-layer Lennard-Jones walls), continuity density,
Tait EOS, artificial viscosity, XSPH, piston x(t) = 0.05 sin(2 pi t / 1.2),
dynamic CFL dt from the max-|a| / max-mu reductions.  A "step" is one whole
time step (sort, predict, density, forces, dt, integrate) over all particles.

Prints ONE JSON line.  `value` is particle-steps/s of the FP32 path (FP32
pair arithmetic with FP64 positions, dt and reductions) with the state
resident in HBM; the FP64 path is timed beside it.  `roofline` is the
dominant kernel's algorithmic bytes per launch over its average duration
(HIP events on the context's stream); `cpu_baseline` is the FP64 CPU oracle
(a port of the reference's RRSPH_OMP step) on a bounded sample of the same
workload, on this host's cores.  Beside the headline: the 3D dam break
(6.6M particles, FP32 and FP64), the 2D differentiable wave tank and the 3D
piston basin with 8 paddles (adjoint / forward cost, gradient against
central differences).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
FP32_VECTOR_PEAK_TFLOPS = 157.3  # MI355X FP32 vector peak (packed)
VALU_SIMDS = 1024      # 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9       # MI355X peak engine clock
# Algorithmic bytes per particle for the pair pass (continuity, XSPH, AV,
# dynamic dt) and for the whole step: the canonical pipeline's compulsory
# traffic, one read of every input and one write of every output per kernel.
B_ALG_PAIR = {"fp32": 64, "fp64": 108}
B_ALG_STEP = {"fp32": 304, "fp64": 440}


def progress(*msg):
    """A progress line on stderr (long legs print one per stage)."""
    print("[bench]", *msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000, help="timed steps (the flume's stated run: 20,000)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--delta", type=float, default=None, help="particle spacing (default: the ~1M flume)")
    ap.add_argument("--cpu-steps", type=int, default=180, help="oracle steps in the CPU baseline sample (~10 s)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--error-steps", type=int, default=100, help="steps before the FP32-vs-FP64 comparison")
    ap.add_argument("--adjoint-steps", type=int, default=2000, help="steps of the differentiable wave tank")
    ap.add_argument("--adjoint-every", type=int, default=100, help="checkpoint interval of the adjoint")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint3d-steps", type=int, default=100, help="steps of the 3D basin adjoint (0: skip)")
    ap.add_argument("--adjoint3d-every", type=int, default=50)
    ap.add_argument("--adjoint3d-delta", type=float, default=0.01)
    ap.add_argument("--adjoint3d-fd", type=int, default=8, help="coefficients checked by central differences")
    ap.add_argument("--steps-3d", type=int, default=500, help="steps of the 3D dam break (0: skip)")
    ap.add_argument("--delta-3d", type=float, default=0.0025, help="spacing of the 3D dam break")
    ap.add_argument("--configs", default="", help="run the configurations at their stated lengths instead: "
                                                  "all, or a comma list of c1, c2, c3, c4 (one JSON line each)")
    return ap.parse_args()


def gpu_leg(layout, precision, steps, warmup, profile_steps):
    import torch

    from sph2d_amd import hip as H

    prm = layout.params
    g = H.HipStep(prm, precision=precision)
    g.upload(layout.r, layout.v, layout.rho, layout.itype)
    if warmup:
        g.step(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.step(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall_s = t1 - t0
    # steps between two HIP events on the context's stream (at most 2,000)
    ev_steps = min(steps, 2000)
    ev_ms = g.time_steps(ev_steps)
    # per-kernel averages: HIP events around every launch on the context's stream
    per_kernel, prof_total = g.profile_steps(profile_steps)
    sort_stats = g.sort_stats()
    list_stats = g.list_stats()
    t, dt, step = g.time()
    diag = g.diagnostics()
    # one more step with the pair-pass fields exported, for the FP32-vs-FP64 comparison
    g.set_export_fields(True)
    g.step(1)
    state = g.download()
    state["a"] = g.fields()["a"]
    state["step"] = g.time()[2]
    g.close()
    return {
        "wall_s": wall_s,
        "event_ms": ev_ms,
        "event_steps": ev_steps,
        "per_kernel_ms": per_kernel,
        "profiled_step_ms": prof_total / max(profile_steps, 1),
        "time": t,
        "dt": dt,
        "step": step,
        "diag": diag,
        "sort_stats": sort_stats,
        "list_stats": list_stats,
        "state": state,
    }


def error_leg(layout, steps):
    """FP32 against FP64 on the GPU after `steps` steps from the same state:
    the fields of the next step's pair pass and the state."""
    from sph2d_amd import hip as H

    states = {}
    for name, prec in (("fp32", H.FP32), ("fp64", H.FP64)):
        g = H.HipStep(layout.params, precision=prec)
        g.upload(layout.r, layout.v, layout.rho, layout.itype)
        g.step(steps)
        g.set_export_fields(True)
        g.step(1)
        states[name] = g.download()
        states[name]["a"] = g.fields()["a"]
        states[name]["step"] = g.time()[2]
        g.close()
    return precision_error(states["fp32"], states["fp64"], layout.params.delta)


TARGET_SEED = 20240607  # the second coefficient set, whose gauge history is the target


def adjoint_leg(steps, every, gauges=(2.0, 3.0, 4.0, 5.0), fd=False):
    """The differentiable 2D wave tank (8 m x 1 m flume, depth 0.4 m,
    delta 0.005 m, 16-harmonic Fourier piston, FP64): plain forward steps,
    the forward with checkpoints and gauges, and the reverse pass, each timed
    on the wall clock.  The loss is the squared error of the gauge
    elevations against the history of a forward run with a second seeded
    coefficient set.  With fd, every one of the 32 coefficients is checked
    against central differences with step 1e-6 (and 2e-6, whose difference
    from the 1e-6 value estimates the differencing noise)."""
    import numpy as np
    import torch

    from sph2d_amd import generators as G
    from sph2d_amd import hip as H
    from sph2d_amd.adjoint import DifferentiableRun, central_difference

    lay = G.adjoint_tank()
    prm = lay.params
    tgt = G.adjoint_tank(seed=TARGET_SEED).params
    g = H.HipStep(prm, precision=H.FP64)
    delta = prm.delta
    spec = H.gauge_spec(list(gauges), radius=3 * delta, tau=delta / 4, y_ref=prm.depth)
    # plain forward (graph replays)
    g.upload(lay.r, lay.v, lay.rho, lay.itype)
    g.step(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.step(steps)
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t0
    run = DifferentiableRun(g, dict(lay.state), steps, every, spec, prm.piston_omega)
    a, b = np.array(prm.piston_a), np.array(prm.piston_b)
    target = run.forward(np.array(tgt.piston_a), np.array(tgt.piston_b))
    loss = lambda e: float(np.sum((e - target) ** 2))
    # forward with checkpoints and gauges, then the reverse pass
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eta = run.forward(a, b)
    t1 = time.perf_counter()
    grads = run.backward(2.0 * (eta - target), len(a))
    t2 = time.perf_counter()
    out = {
        "workload": "2D differentiable wave tank 8 m x 1 m, depth 0.4 m, delta 0.005 m, %d particles (%d fluid), "
                    "16-harmonic Fourier piston, gauges at x = %s m, %d steps, checkpoint every %d, FP64; loss: "
                    "squared gauge error against the history of coefficient seed %d"
                    % (prm.ntotal, prm.nfluid, ", ".join("%g" % x for x in gauges), steps, every, TARGET_SEED),
        "forward_s": round(fwd, 4),
        "taped_forward_s": round(t1 - t0, 4),
        "backward_s": round(t2 - t1, 4),
        "reverse_over_forward": round((t2 - t1) / fwd, 3),
        "taped_and_reverse_over_forward": round((t2 - t0) / fwd, 3),
        "loss": loss(eta),
        "grad_norm": float((grads["a"] ** 2).sum() + (grads["b"] ** 2).sum()) ** 0.5,
    }
    if fd:
        rows = []
        for which in ("a", "b"):
            progress("wave tank: central differences of", which)
            for k in range(len(a)):
                fd1 = central_difference(run, a, b, loss, which, k, 1e-6)
                fd2 = central_difference(run, a, b, loss, which, k, 2e-6)
                ad = float(grads[which][k])
                noise = abs(fd1 - fd2)
                rows.append({"coef": f"{which}[{k}]", "adjoint": ad, "fd": fd1, "fd_noise": noise,
                             "rel_err": abs(ad - fd1) / max(abs(fd1), 1e-300),
                             "above_noise": abs(fd1) > 10 * noise})
        ok = [r for r in rows if r["above_noise"]]
        out["fd_check"] = rows
        out["fd_coefficients_above_noise"] = len(ok)
        out["fd_max_rel_err_above_noise"] = max((r["rel_err"] for r in ok), default=None)
        out["fd_max_abs_err"] = max(abs(r["adjoint"] - r["fd"]) for r in rows)
        out["fd_max_abs"] = max(abs(r["fd"]) for r in rows)
    g.close()
    return out


def adjoint3d_leg(steps, every, delta, nfd):
    """The differentiable 3D piston basin (6 m x 2 m, depth 0.5 m, 8 paddles
    x 4 harmonics = 64 coefficients): for the FP32 step (FP32 pairs, FP64
    positions; adjoint arithmetic FP64 on its trajectory) and the FP64 step,
    plain forward steps, the taped forward and the reverse pass on the wall
    clock; then both gradients of the gauge loss against FP64 central
    differences on `nfd` sampled coefficients.  The 4 x 4 gauge grid sits
    within the distance sound travels in the run, so every gauge sees the
    paddles."""
    import numpy as np
    import torch

    from sph2d_amd import generators as G
    from sph2d_amd import hip as H
    from sph2d_amd.adjoint import DifferentiableRun, central_difference

    lay = G.piston_basin_3d(delta=delta)
    prm = lay.params
    reach = max(prm.eos_c * steps * prm.dt, 8 * delta)
    xs = [2 * delta + reach * (i + 1) / 5 for i in range(4)]
    zs = [0.25, 0.75, 1.25, 1.75]
    gx, gz = [x for x in xs for _ in zs], [z for _ in xs for z in zs]
    spec = H.gauge_spec(gx, radius=3 * delta, tau=delta / 4, y_ref=prm.depth, zs=gz)
    a, b = np.array(prm.piston_a), np.array(prm.piston_b)
    loss = lambda e: float(np.sum((e - prm.depth) ** 2))
    out = {"workload": "3D piston basin 6 m x 2 m, depth 0.5 m, delta %g m, %d particles (%d fluid), 8 paddles x 4 "
                       "harmonics, 4 x 4 gauges, %d steps, checkpoint every %d" % (delta, prm.ntotal, prm.nfluid,
                                                                                steps, every)}
    grads = {}
    for name, prec in (("fp32", H.FP32), ("fp64", H.FP64)):
        progress("3D basin", name, "forward, taped forward, reverse")
        g = H.HipStep(prm, precision=prec)
        g.upload(lay.r, lay.v, lay.rho, lay.itype)
        g.step(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.step(steps)
        torch.cuda.synchronize()
        fwd = time.perf_counter() - t0
        run = DifferentiableRun(g, dict(lay.state), steps, every, spec, prm.piston_omega)
        t0 = time.perf_counter()
        eta = run.forward(a, b)
        t1 = time.perf_counter()
        grads[name] = run.backward(2.0 * (eta - prm.depth), len(a))
        t2 = time.perf_counter()
        out[name] = {"forward_s": round(fwd, 4), "taped_forward_s": round(t1 - t0, 4),
                     "backward_s": round(t2 - t1, 4), "adjoint_over_forward": round((t2 - t1) / fwd, 3),
                     "taped_and_reverse_over_forward": round((t2 - t0) / fwd, 3)}
        if name == "fp64":
            errs = []
            rng = np.random.default_rng(1)
            for j in rng.choice(2 * len(a), size=nfd, replace=False):
                which, k = ("a", int(j)) if j < len(a) else ("b", int(j) - len(a))
                progress("3D basin: central difference of", which, k)
                fd = central_difference(run, a, b, loss, which, k, 1e-6)
                errs.append({"coef": f"{which}[{k}]", "fd_fp64": fd, "adjoint_fp64": grads["fp64"][which][k],
                             "adjoint_fp32": grads["fp32"][which][k],
                             "rel_err_fp64": abs(grads["fp64"][which][k] - fd) / max(abs(fd), 1e-300),
                             "rel_err_fp32": abs(grads["fp32"][which][k] - fd) / max(abs(fd), 1e-300)})
            out["fd_check"] = errs
            out["fd_max_rel_err_fp64"] = max(e["rel_err_fp64"] for e in errs) if errs else None
            out["fd_max_rel_err_fp32"] = max(e["rel_err_fp32"] for e in errs) if errs else None
        g.close()
    return out


def leg_3d(steps, delta, renorm_every=30):
    """3D dam break (y up): box 1.6 x 0.8 x 0.6 m, fluid block 0.4 x 0.3 x 0.6 m,
    h = 1.3 delta, summation density renormalised every 30th step, 3-layer LJ
    walls; FP32 and FP64, `steps` steps timed from rest after 2 (the start-up
    with its short list-build intervals included)."""
    import torch

    from sph2d_amd import generators as G
    from sph2d_amd import hip as H

    lay = G.dam_break_3d(delta=delta, box=(1.6, 0.8, 0.6), column=(0.4, 0.3, 0.6), jitter=0.0)
    prm = lay.params
    prm.density_normalization = 1
    prm.density_normalization_every = renorm_every
    out = {"workload": "3D dam break 1.6 x 0.8 x 0.6 m box, 0.4 x 0.3 x 0.6 m fluid block, delta=%g m, h=1.3 delta, "
                       "%d particles (%d fluid), summation density renormalised every %dth step, cubic kernels, "
                       "fixed dt" % (delta, prm.ntotal, prm.nfluid, renorm_every),
           "particles": prm.ntotal}
    states = {}
    for name, prec in (("fp32", H.FP32), ("fp64", H.FP64)):
        g = H.HipStep(prm, precision=prec)
        g.upload(lay.r, lay.v, lay.rho, lay.itype)
        g.step(2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.step(steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        per_kernel, _ = g.profile_steps(2)
        sort_stats = g.sort_stats()
        list_stats = g.list_stats()
        g.set_export_fields(True)
        g.step(1)
        states[name] = g.download()
        states[name]["a"] = g.fields()["a"]
        states[name]["step"] = g.time()[2]
        g.close()
        out[name] = {"value": round(prm.ntotal * steps / wall, 1), "ms_per_step": round(wall / steps * 1e3, 4),
                     "pair_ms": round(per_kernel["pair_kernel"], 4),
                     "density_sum_ms": round(per_kernel["density_sum_kernel"], 4), "list_stats": list_stats,
                     "sort_stats": sort_stats}
    out["fp32_over_fp64"] = round(out["fp64"]["ms_per_step"] / out["fp32"]["ms_per_step"], 3)
    out["fp32_error_vs_fp64"] = precision_error(states["fp32"], states["fp64"], delta)
    return out


def copy_bandwidth():
    """Device-to-device copy of 2 GiB (read + write counted), best of 5, on
    the box itself: the attainable HBM bandwidth beside the published peak."""
    import torch

    n = 1 << 31
    src = torch.empty(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = 0.0
    for _ in range(5):
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        dst.copy_(src)
        t1.record()
        t1.synchronize()
        best = max(best, 2 * n / (t0.elapsed_time(t1) * 1e-3) / 1e9)
    del src, dst
    torch.cuda.empty_cache()
    return best


def precision_error(s32, s64, delta):
    """Error of the FP32 run against the FP64 run after the same steps from
    the same initial state: max relative density error, max position
    difference (m and in units of the particle spacing), and the max
    acceleration difference relative to the largest FP64 acceleration."""
    import numpy as np

    rho = float(np.max(np.abs(s32["rho"] - s64["rho"]) / s64["rho"]))
    dr = float(np.max(np.abs(s32["r"] - s64["r"])))
    a64 = np.linalg.norm(s64["a"], axis=1)
    da = float(np.max(np.linalg.norm(s32["a"] - s64["a"], axis=1)) / max(float(a64.max()), 1e-300))
    return {"steps": int(s64["step"]), "rho_rel_max": rho, "r_max_m": dr, "r_max_over_delta": dr / delta,
            "a_max_rel_to_scale": da}


def cpu_leg(layout, steps, samples=3):
    """The FP64 OpenMP oracle on the same flume: after 2 warm-up steps,
    `samples` consecutive runs of steps // samples steps each.  Returns the
    median rate (particle-steps/s), the threads, the seconds of all runs and
    the rate of each run."""
    import statistics

    import oracle

    prm = layout.params
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    o = oracle.Oracle(prm, num_threads=threads)
    o.set_state(layout.r, layout.v, layout.rho, layout.itype, time=0.0, dt=prm.dt)
    o.step(2)
    per = max(steps // samples, 1)
    rates, secs = [], 0.0
    for _ in range(samples):
        t0 = time.perf_counter()
        st = o.step(per)
        t1 = time.perf_counter()
        if st != 0:
            raise RuntimeError(f"oracle status {st}")
        rates.append(prm.ntotal * per / (t1 - t0))
        secs += t1 - t0
    return statistics.median(rates), threads, secs, rates


def c1_full(steps):
    """C1-1M over its stated 20,000 steps, FP32 and FP64: ms per step over the
    whole run, the list and sort statistics, and the FP32 error against the
    FP64 run at the end."""
    import torch

    from sph2d_amd import generators as G
    from sph2d_amd import hip as H

    layout = G.piston_tank(delta=G.c1_1m_delta())
    prm = layout.params
    out = {"workload": "C1-1M 20 m x 0.5 m piston flume, %d particles, %d steps from rest" % (prm.ntotal, steps)}
    states = {}
    for name, prec in (("fp32", H.FP32), ("fp64", H.FP64)):
        g = H.HipStep(prm, precision=prec)
        g.upload(layout.r, layout.v, layout.rho, layout.itype)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.step(steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        g.set_export_fields(True)
        g.step(1)
        states[name] = g.download()
        states[name]["a"] = g.fields()["a"]
        states[name]["step"] = g.time()[2]
        out[name] = {"ms_per_step": round(wall / steps * 1e3, 5), "value": round(prm.ntotal * steps / wall, 1),
                     "simulated_time_s": g.time()[0], "lists": g.list_stats(), "sort_stats": g.sort_stats(),
                     "diag": g.diagnostics()}
        g.close()
    out["fp32_over_fp64"] = round(out["fp64"]["ms_per_step"] / out["fp32"]["ms_per_step"], 3)
    out["fp32_error_vs_fp64"] = precision_error(states["fp32"], states["fp64"], prm.delta)
    return out


def c2_full(steps, delta):
    """C2, the 3D dam break with summation density renormalised every 30th
    step, over its stated 5,000 steps: FP32 and FP64 ms per step and the FP32
    error against FP64 at the end."""
    import torch

    from sph2d_amd import generators as G
    from sph2d_amd import hip as H

    lay = G.dam_break_3d(delta=delta, box=(1.6, 0.8, 0.6), column=(0.4, 0.3, 0.6), jitter=0.0)
    prm = lay.params
    prm.density_normalization = 1
    prm.density_normalization_every = 30
    out = {"workload": "C2 3D dam break, delta %g m, %d particles, summation renormalised every 30th step, %d steps"
                       % (delta, prm.ntotal, steps)}
    states = {}
    for name, prec in (("fp32", H.FP32), ("fp64", H.FP64)):
        progress("C2", name, steps, "steps")
        g = H.HipStep(prm, precision=prec)
        g.upload(lay.r, lay.v, lay.rho, lay.itype)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(0, steps, 500):
            g.step(min(500, steps - _))
            progress("C2", name, "step", g.time()[2])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        g.set_export_fields(True)
        g.step(1)
        states[name] = g.download()
        states[name]["a"] = g.fields()["a"]
        states[name]["step"] = g.time()[2]
        out[name] = {"ms_per_step": round(wall / steps * 1e3, 4), "value": round(prm.ntotal * steps / wall, 1),
                     "sort_stats": g.sort_stats(), "diag": g.diagnostics()}
        g.close()
    out["fp32_over_fp64"] = round(out["fp64"]["ms_per_step"] / out["fp32"]["ms_per_step"], 3)
    out["fp32_error_vs_fp64"] = precision_error(states["fp32"], states["fp64"], delta)
    return out


def configs_main(args):
    """The configurations at their stated lengths (about 10 minutes):
    C1-1M 20,000 steps, C2 5,000 steps, C3 with the 32-coefficient
    finite-difference check (the stated gauges and gauges within the signal's
    reach), C4 1,000 steps with a checkpoint every 50."""
    import torch

    torch.cuda.set_device(0)
    only = set(args.configs.split(",")) if args.configs != "all" else {"c1", "c2", "c3", "c4"}
    out = {}
    if "c1" in only:
        out["c1_1m"] = c1_full(20000)
        print(json.dumps({"c1_1m": out["c1_1m"]}), flush=True)
    if "c2" in only:
        out["c2"] = c2_full(5000, args.delta_3d)
        print(json.dumps({"c2": out["c2"]}), flush=True)
    if "c3" in only:
        out["c3"] = adjoint_leg(2000, 100, fd=True)
        print(json.dumps({"c3": out["c3"]}), flush=True)
        out["c3_gauges_in_reach"] = adjoint_leg(2000, 100, gauges=(0.05, 0.1, 0.15, 0.2), fd=True)
        print(json.dumps({"c3_gauges_in_reach": out["c3_gauges_in_reach"]}), flush=True)
    if "c4" in only:
        out["c4"] = adjoint3d_leg(1000, 50, 0.01, 8)
        print(json.dumps({"c4": out["c4"]}), flush=True)
    return out


def main():
    args = parse()
    if args.configs:
        if args.gpus != 1:
            print("bench.py: one GPU only", file=sys.stderr)
            sys.exit(2)
        configs_main(args)
        return
    if args.gpus != 1:
        print(f"bench.py: RRSPH_HIP runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    import torch

    from sph2d_amd import generators as G
    from sph2d_amd import hip as H

    torch.cuda.set_device(0)
    delta = args.delta or G.c1_1m_delta()
    layout = G.piston_tank(delta=delta)
    prm = layout.params
    n = prm.ntotal

    r32 = gpu_leg(layout, H.FP32, args.steps, args.warmup, args.profile_steps)
    value = n * args.steps / r32["wall_s"]
    r64 = None
    if not args.no_fp64:
        r64 = gpu_leg(layout, H.FP64, args.steps, args.warmup, args.profile_steps)

    # dominant kernel of the FP32 step
    pk = r32["per_kernel_ms"]
    # time per step of each kernel class: the sort (three launches of each
    # radix kernel), reorder and list build run on build steps only when
    # the step uses neighbour lists
    ls = r32["list_stats"]
    build_frac = ls["builds"] / max(ls["builds"] + ls["list_steps"], 1) if ls["interval"] else 1.0
    per_step = {k: pk[k] * (3 if "radix" in k else 1) * (build_frac if ("radix" in k or k in (
        "reorder_kernel", "list_build_kernel")) else 1.0) for k in pk if pk[k] > 0}
    dom = max(per_step, key=per_step.get)
    pair_ms = pk["pair_kernel"]
    achieved = B_ALG_PAIR["fp32"] * n / (pair_ms * 1e-3) / 1e9
    # the pair pass: pair_list_kernel on the list steps, pair_fast_kernel when
    # the step sorts and walks the cells every step
    roofline = {
        "kernel": "pair_list_kernel<float>" if r32["list_stats"]["interval"] else "pair_fast_kernel<float>",
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "bytes_per_particle": B_ALG_PAIR["fp32"],
        "avg_launch_ms": round(pair_ms, 5),
        "step_frac": round(value * B_ALG_STEP["fp32"] / 1e9 / HBM_PEAK_GBS, 4),
        "dominant_by_time": dom,
        "per_step_ms": {k: round(v, 5) for k, v in per_step.items()},
    }
    traffic_file = os.path.join(ROOT, "profiles", "pair_pass_c1_1m_fp32_traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            prof = json.load(f)
        if prof.get("kernel", "").split("<")[0] != "rrsph::" + roofline["kernel"].split("<")[0]:
            prof = {}  # profiled on the other pair pass
        roofline["traffic"] = prof.get("hbm_bytes_per_launch")
        # the pair pass is bound by VALU issue: the profiled VALU instructions
        # of one launch at one instruction per 4 cycles on each of the 1,024
        # SIMDs (wave64 on 16 lanes), over this run's average launch time
        valu = prof.get("valu_insts_per_launch")
        if valu:
            valu_ms = valu * 4 / (VALU_SIMDS * CLOCK_HZ) * 1e3
            roofline["valu_issue_bound_ms"] = round(valu_ms, 5)
            roofline["valu_issue_frac"] = round(valu_ms / pair_ms, 4)
        # FP32 VALU throughput of the pair pass: the profiled floating-point
        # operations of one launch (masked-off lanes included, an upper
        # bound) over this run's average launch time, against the vector
        # peak (157.3 TFLOP/s with packed FP32, which this build does not
        # emit; 78.6 without)
        flop = prof.get("fp32_flop_per_launch")
        if flop:
            tf = flop / (pair_ms * 1e-3) / 1e12
            roofline["valu_fp32_tflops"] = round(tf, 2)
            roofline["valu_fp32_frac_of_peak"] = round(tf / FP32_VECTOR_PEAK_TFLOPS, 4)

    roofline["copy_measured_gbs"] = round(copy_bandwidth(), 1)
    roofline["frac_of_measured_copy"] = round(achieved / roofline["copy_measured_gbs"], 4)

    err_early = None if args.no_fp64 else error_leg(layout, args.error_steps)
    r3d = leg_3d(args.steps_3d, args.delta_3d) if args.steps_3d > 0 else None

    adj = None
    if not args.no_adjoint:
        adj = adjoint_leg(args.adjoint_steps, args.adjoint_every)
    adj3 = None
    if args.adjoint3d_steps > 0:
        adj3 = adjoint3d_leg(args.adjoint3d_steps, args.adjoint3d_every, args.adjoint3d_delta, args.adjoint3d_fd)

    cpu = None
    if not args.no_cpu:
        cpu_rate, cores, secs, rates = cpu_leg(layout, args.cpu_steps)
        cpu = {
            "value": round(cpu_rate, 1),
            "unit": "particle-steps/s",
            "cores": cores,
            "kind": "port",
            "sample": f"median of {len(rates)} consecutive runs of {max(args.cpu_steps // len(rates), 1)} steps "
                      f"of the same {n}-particle flume after 2 warm-up steps ({secs:.1f} s in all; runs at "
                      + ", ".join(f"{r:.3g}" for r in rates) + " particle-steps/s), FP64 OpenMP oracle",
        }

    out = {
        "metric": "particle-steps/s at 1M 2D (FP32+comp vs FP64); adjoint/forward cost ratio",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["wall_s"] / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "C1-1M: 2D piston wave flume 20 m x 0.5 m, delta=%.7g m, %d particles (%d fluid), "
                        "continuity density, dynamic CFL dt, FP32 pair arithmetic with FP64 positions"
                        % (delta, n, prm.nfluid),
            "particles": n,
            "precision": "fp32 pairs, fp64 positions/dt/reductions",
        },
        "event_ms_per_step": round(r32["event_ms"] / r32["event_steps"], 5),
        "event_steps": r32["event_steps"],
        "per_kernel_ms": {k: round(v, 5) for k, v in pk.items()},
        "sort_stats": r32["sort_stats"],
        "list_stats": r32["list_stats"],
        "fp64": None if r64 is None else {
            "value": round(n * args.steps / r64["wall_s"], 1),
            "ms_per_step": round(r64["wall_s"] / args.steps * 1e3, 5),
            "per_kernel_ms": {k: round(v, 5) for k, v in r64["per_kernel_ms"].items()},
            "list_stats": r64["list_stats"],
        },
        "fp32_over_fp64": None if r64 is None else round(r64["wall_s"] / r32["wall_s"], 3),
        # after --error-steps steps from the same state (the per-step error) ...
        "fp32_error_vs_fp64": err_early,
        # ... and after the whole timed run, where the two free-surface flows
        # have drifted apart as chaotic flows do
        "fp32_error_vs_fp64_end": None if r64 is None else precision_error(r32["state"], r64["state"], delta),
        "adjoint_over_forward": None if adj is None else adj["reverse_over_forward"],
        "adjoint_taped_and_reverse_over_forward": None if adj is None else adj["taped_and_reverse_over_forward"],
        "adjoint": adj,
        "adjoint_3d": adj3,
        "dam_break_3d": r3d,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
