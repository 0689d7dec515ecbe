"""Shared helpers for the tests: loading the golden dam-break case, the
reference's checkDouble tolerance, and small comparisons."""
from __future__ import annotations

import json
import os

import numpy as np

from sph2d_amd import params as P

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_dambreak_case():
    """Golden case of the reference's module tests: params + dump arrays (as parsed)."""
    d = np.load(os.path.join(GOLDEN, "case_dambreak.npz"))
    mp = P.parse_model_params(json.loads(bytes(d["model_params_json"]).decode()))
    pp = P.parse_particle_params(json.loads(bytes(d["particle_params_json"]).decode()))
    prm = P.make_params(mp, pp, start_time=0.0)
    r = np.stack([d["dump_x"], d["dump_y"]], axis=1)
    v = np.stack([d["dump_vx"], d["dump_vy"]], axis=1)
    rho = d["dump_rho"].copy()
    itype = d["dump_itype"].astype(np.int32)
    P.finalize(prm, r)
    targets = {k: d[k] for k in d.files if k.startswith("OMP_")}
    return prm, {"r": r, "v": v, "rho": rho, "itype": itype}, targets


def check_double(v1, v2, digits: int) -> np.ndarray:
    """testSPH.h:51-97 (checkDouble<N>): relative 10^-N, absolute below 1e-3."""
    v1 = np.asarray(v1, dtype=np.float64)
    v2 = np.asarray(v2, dtype=np.float64)
    mx = np.maximum(np.abs(v1), np.abs(v2))
    diff = np.abs(v1 - v2)
    rel = np.where(mx < 1e-3, diff, diff / np.where(mx == 0, 1, mx))
    return rel < 10.0 ** (-digits)


def rel_err(a, b, floor: float = 0.0) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(b), floor)
    den = np.where(den == 0, 1, den)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0
