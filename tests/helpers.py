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


def near_cutoff_case(seed=7, centres=(12, 6), sats=64):
    """Particles whose pair distances sit within a few ulps of the cut-off
    k h: satellites on circles of radius k h around widely spaced centres.
    Separate rounding of dx*dx + dy*dy and a fused multiply-add decide a
    share of these pairs differently, so the neighbour sets pin the exact
    FP64 arithmetic of the reference's test (GridFind.cpp:144-176)."""
    import copy

    from sph2d_amd import generators as G

    lay = G.dam_break()
    prm = copy.copy(lay.params)
    cut = prm.cell_scale_k * prm.hsml
    rng = np.random.default_rng(seed)
    pts = []
    for a in range(centres[0]):
        for b in range(centres[1]):
            c = np.array([0.2 + 3.1 * cut * a, 0.2 + 3.1 * cut * b]) + rng.uniform(0, 1e-3, 2)
            pts.append(c)
            th = rng.uniform(0, 2 * np.pi, sats)
            for t in th:
                pts.append(c + cut * np.array([np.cos(t), np.sin(t)]))
    r = np.array(pts, dtype=np.float64)
    n = len(r)
    prm.ntotal = prm.nfluid = n
    prm.nvirt = 0
    prm.maxn = 1 << int(np.ceil(np.log2(n)) + 1)
    prm.max_neighbours = 128
    itype = np.full(n, 2, dtype=np.int32)
    return prm, r, itype


def brute_force_sets(r, cut2, itype):
    """O(N^2) neighbour sets with separately rounded FP64 products (numpy
    does not fuse), the definition of GridFind.cpp:144-176."""
    n = len(r)
    out = []
    for j in range(n):
        if itype[j] == 0:
            out.append(np.zeros(0, dtype=np.uint32))
            continue
        dx = r[:, 0] - r[j, 0]
        dy = r[:, 1] - r[j, 1]
        d2 = dx * dx + dy * dy
        m = (d2 < cut2) & (itype != 0)
        m[j] = False
        out.append(np.nonzero(m)[0].astype(np.uint32))
    return out


def smooth_gauges_np(r, itype, xs, radius, tau, y_ref, ptype=2, zs=None):
    """The smooth gauge surrogate in numpy (the definition in include/rrsph_hip.h):
    eta = y_ref + tau log(sum w e / sum w), w = (1 - d^2 / R^2)^3, e = exp((y - y_ref) / tau),
    d = x - x_g (2D) or the distance from (x_g, z_g) in the x-z plane (3D, zs given)."""
    sel = itype == ptype
    x, y = r[sel, 0], r[sel, 1]
    out = np.zeros(len(xs))
    for g, xg in enumerate(xs):
        d2 = (x - xg) ** 2 if zs is None else (x - xg) ** 2 + (r[sel, 2] - zs[g]) ** 2
        u = 1.0 - d2 / radius ** 2
        w = np.where(u > 0, u ** 3, 0.0)
        ws = w.sum()
        if ws > 0:
            out[g] = y_ref + tau * np.log((w * np.exp((y - y_ref) / tau)).sum() / ws)
    return out


def corner_cutoff_case(dim=2, seed=11, per_centre=48, stride=3):
    """Particles across the corners of their cells at the cut-off.  In every
    stride-th cell a centre sits at distance rho from a corner of its cell
    (in 3D a corner or, with one axis free, an edge), rho = cut (1 + e) with
    e from -1e-3 to +1e-10, and per_centre satellites lie at distance
    cut (1 + k ulp), |k| <= 8, in directions near the corner's.  With e < 0
    the satellites sit in the cell diagonally across the corner and are
    neighbours or not by the last bits of the FP64 test; with e > 0 that
    cell holds no neighbour.  A walk that skips corner cells or columns
    (GridFind.cpp:144-176 visits all 3^dim cells) must keep every neighbour."""
    import copy
    import itertools

    from sph2d_amd import generators as G

    lay = G.dam_break() if dim == 2 else G.dam_break_3d()
    prm = copy.copy(lay.params)
    cut = prm.cell_scale_k * prm.hsml
    lo = np.array([prm.x_mingeom, prm.y_mingeom, prm.z_mingeom][:dim])
    hi = np.array([prm.x_maxgeom, prm.y_maxgeom, prm.z_maxgeom][:dim])
    ncell = np.floor((hi - lo) / cut).astype(int) + 1
    rng = np.random.default_rng(seed)
    offsets = [-1e-3, -1e-6, -1e-10, -1e-12, -1e-14, -4.4e-16, 0.0, 4.4e-16, 1e-14, 1e-12, 1e-11, 1e-10]
    pts = []
    for cidx in itertools.product(*[range(1, n - 2, stride) for n in ncell]):
        origin = lo + np.array(cidx) * cut
        naxes = dim if dim == 2 else int(rng.integers(2, 4))
        axes = rng.permutation(dim)[:naxes]
        s = rng.choice([-1.0, 1.0], dim)
        dirv = np.zeros(dim)
        dirv[axes] = rng.uniform(0.25, 1.0, naxes)
        dirv /= np.linalg.norm(dirv)
        rho = cut * (1.0 + offsets[int(rng.integers(len(offsets)))])
        corner = origin + np.where(s > 0, cut, 0.0)
        centre = corner - s * rho * dirv
        free = np.setdiff1d(np.arange(dim), axes)
        centre[free] = origin[free] + rng.uniform(0.3, 0.7, len(free)) * cut
        pts.append(centre)
        for _ in range(per_centre):
            d = dirv + rng.normal(0.0, rng.choice([1e-6, 1e-4, 1e-2]), dim)
            d /= np.linalg.norm(d)
            r = cut * (1.0 + int(rng.integers(-8, 9)) * 2.220446049250313e-16)
            pts.append(centre + s * r * d)
    r = np.array(pts, dtype=np.float64)
    n = len(r)
    prm.ntotal = prm.nfluid = n
    prm.nvirt = 0
    prm.maxn = 1 << int(np.ceil(np.log2(n)) + 1)
    prm.max_neighbours = 128
    itype = np.full(n, 2, dtype=np.int32)
    return prm, r, itype


def corner_pairs_at_cutoff(prm, r, sets, rel=1e-12):
    """Neighbour pairs whose cells differ along two or more axes (diagonal
    across a corner or an edge) at distance within rel of the cut-off."""
    dim = r.shape[1]
    cut = prm.cell_scale_k * prm.hsml
    lo = np.array([prm.x_mingeom, prm.y_mingeom, prm.z_mingeom][:dim])
    key = np.floor((r - lo) / cut).astype(np.int64)
    count = 0
    for j, s in enumerate(sets):
        if len(s) == 0:
            continue
        s = np.asarray(s, dtype=np.int64)
        d = np.sqrt(np.sum((r[s] - r[j]) ** 2, axis=1))
        diag = np.sum(key[s] != key[j], axis=1) >= 2
        count += int(np.sum(diag & (np.abs(d - cut) <= rel * cut)))
    return count
