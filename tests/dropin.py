"""Runs the C++ drop-in driver (tests/cpp/dropin_driver) on a layout:
writes params.txt and state.bin, reads out.bin back."""
import dataclasses
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "dropin_driver")


def write_inputs(d, prm, lay, nsteps):
    """simulation_time is placed half a step past the last wanted step so the
    loop `while (time <= simulation_time)` runs exactly nsteps fixed-dt steps."""
    p = os.path.join(d, "params.txt")
    with open(p, "w") as f:
        for fld in dataclasses.fields(prm):
            v = getattr(prm, fld.name)
            if fld.name in ("piston_a", "piston_b"):
                if len(v):
                    f.write(f"{fld.name} {','.join(repr(float(x)) for x in v)}\n")
                continue
            if fld.name == "simulation_time":
                v = prm.start_simulation_time + (nsteps - 0.5) * prm.dt
            if isinstance(v, bool):
                v = int(v)
            f.write(f"{fld.name} {float(v)!r}\n")
    s = os.path.join(d, "state.bin")
    n = prm.ntotal
    with open(s, "wb") as f:
        np.array([n], dtype=np.uint32).tofile(f)
        np.ascontiguousarray(lay.r, dtype=np.float64).tofile(f)
        np.ascontiguousarray(lay.v, dtype=np.float64).tofile(f)
        np.ascontiguousarray(lay.rho, dtype=np.float64).tofile(f)
        np.ascontiguousarray(lay.itype, dtype=np.int32).tofile(f)
    return p, s


def run(d, prm, lay, nsteps, fp32=False, probe_every=0, timeout=120):
    d = str(d)
    p, s = write_inputs(d, prm, lay, nsteps)
    out = os.path.join(d, "out.bin")
    cmd = [DRIVER, p, s, out, "fp32" if fp32 else "fp64", str(probe_every)]
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)


def read_output(d, n, dim=2):
    buf = open(os.path.join(str(d), "out.bin"), "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(buf, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a

    steps = int(take(np.uint64, 1)[0])
    t, dt = take(np.float64, 2)
    out = {"steps": steps, "time": float(t), "dt": float(dt)}
    out["r"] = take(np.float64, dim * n).reshape(n, dim)
    out["v"] = take(np.float64, dim * n).reshape(n, dim)
    out["rho"] = take(np.float64, n)
    out["p"] = take(np.float64, n)
    out["itype"] = take(np.int32, n)
    out["probe_step"] = int(take(np.uint64, 1)[0])
    out["probe_rho"] = take(np.float64, n)
    out["probe_p"] = take(np.float64, n)
    return out
