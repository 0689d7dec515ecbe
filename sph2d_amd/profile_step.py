"""Profiling driver: runs a few RRSPH_HIP steps of a workload, nothing else.

    python -m sph2d_amd.profile_step [--steps 3] [--warmup 2] [--fp64] [--delta D] [--dam-break-3d]
    python -m sph2d_amd.profile_step --copy-calibration

Default workload: the 2D piston flume at ~1M particles.  --dam-break-3d:
the 3D dam break (1.6 x 0.8 x 0.6 m box, 0.4 x 0.3 x 0.6 m fluid block,
delta 0.0025 m unless --delta is given, ~6.6M particles).

Meant to sit under rocprofv3 (kernel traces, PMC counter passes).  The
calibration mode copies a 256 MiB device buffer five times so that the
FETCH_SIZE / WRITE_SIZE counters can be scaled against a copy of known size.
"""
import argparse
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fp64", action="store_true")
    ap.add_argument("--delta", type=float, default=None)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--dam-break-3d", action="store_true")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        x = torch.ones(256 * 1024 * 1024 // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print("copied", 5 * x.numel() * 4, "bytes", flush=True)
        return
    from sph2d_amd import generators as G
    from sph2d_amd import hip as H

    if args.dam_break_3d:
        lay = G.dam_break_3d(delta=args.delta or 0.0025, box=(1.6, 0.8, 0.6), column=(0.4, 0.3, 0.6), jitter=0.0)
    else:
        lay = G.piston_tank(delta=args.delta or G.c1_1m_delta())
    g = H.HipStep(lay.params, precision=H.FP64 if args.fp64 else H.FP32)
    g.upload(lay.r, lay.v, lay.rho, lay.itype)
    if args.warmup:
        g.step(args.warmup)
    g.step(args.steps)
    t, dt, step = g.time()
    print(f"steps {step} time {t:.6g} dt {dt:.6g} particles {lay.params.ntotal}", flush=True)
    g.close()


if __name__ == "__main__":
    main()
