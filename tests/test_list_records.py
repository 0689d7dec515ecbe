"""The FP32 cut-off decision of the list steps, restated in numpy on the CPU.

Between two builds a list step decides each entry in FP32 from the near
records: xi = (float)(x - origin of the home cell), so that
r_i - r_j = (xi_i - xi_j) + (float)((dcx, dcy) * cell) with (dcx, dcy) the
offset of the two home cells.  It takes the FP64 positions and the
reference's test (|r|^2 < (k h)^2 with separate roundings,
GridFind.cpp:144-176) only when the FP32 squared distance lies within 1e-5
of the cut-off.  The build keeps every listed position within a cell of its
home cell, and a list step runs only while no particle has moved half the
skin since the build, so |xi| stays below 2.1 cells.  These tests check on the
C1-1M geometry (20 m flume, delta 3.16 mm) that the FP32 squared distance
stays far inside that band, with and without fused multiply-adds, for every
entry of the lists of a build followed by displacements up to the largest a
list step allows (half the skin), so that every decision outside the band is
the FP64 test's.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from sph2d_amd import generators as G

SKIN = 0.125  # the default skin, a fraction of the cut-off
BAND = 1e-5


def _f32(x):
    return np.asarray(x, dtype=np.float32)


def _fma32(a, b, c):
    """float32 fma: the product of two floats is exact in float64."""
    return _f32(np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64))


def _entries(seed, x0=9.9, x1=10.1, near_cut=0, edge=None):
    lay = G.piston_tank(delta=G.c1_1m_delta())
    prm = lay.params
    cell = prm.cell_scale_k * prm.hsml
    xmin, ymin = prm.x_mingeom, prm.y_mingeom
    sel = (lay.r[:, 0] > x0) & (lay.r[:, 0] < x1)
    rb = lay.r[sel].copy()  # positions at the build
    rng = np.random.default_rng(seed)
    # cell keys of the build (GridUtils.h:135-144 coordinate rule)
    hx = np.floor((rb[:, 0] - xmin) / cell).astype(np.int64)
    hy = np.floor((rb[:, 1] - ymin) / cell).astype(np.int64)
    reach = cell * (1 + SKIN)
    pairs = np.array(sorted(cKDTree(rb).query_pairs(reach * (1 + 1e-9))), dtype=np.int64)
    # displacements of at most half the skin since the build
    ang = rng.uniform(0, 2 * np.pi, len(rb))
    mag = rng.uniform(0, 1, len(rb)) * 0.5 * SKIN * cell * (1 - 1e-9)
    r = rb + np.stack([mag * np.cos(ang), mag * np.sin(ang)], axis=1)
    if near_cut:
        # move the second particle of some pairs to within a few ulps of the
        # cut-off, or of an edge of the band (its FP32 value) with edge="lo"/"hi"
        dist = cell
        if edge:
            dist = np.sqrt(float(np.float32(cell * cell * (1 - BAND if edge == "lo" else 1 + BAND))))
        idx = rng.choice(len(pairs), size=near_cut, replace=False)
        for i, j in pairs[idx]:
            d = r[j] - r[i]
            d /= np.linalg.norm(d)
            r[j] = r[i] + d * dist * (1 + rng.uniform(-1e-12, 1e-12))
    return r, hx, hy, pairs, cell, xmin, ymin


def _decide(r, hx, hy, pairs, cell, xmin, ymin, fused):
    ox = xmin + hx * cell
    oy = ymin + hy * cell
    xi = _f32(r[:, 0] - ox)
    yi = _f32(r[:, 1] - oy)
    i, j = pairs[:, 0], pairs[:, 1]
    # the table of home-cell offsets: (float)(dc * cell), dc from -3 to 4
    offx = _f32((hx[i] - hx[j]) * cell)
    offy = _f32((hy[i] - hy[j]) * cell)
    dx = _f32(_f32(xi[i] - xi[j]) + offx)
    dy = _f32(_f32(yi[i] - yi[j]) + offy)
    if fused:
        d2 = _fma32(dx, dx, _f32(dy * dy))
    else:
        d2 = _f32(_f32(dx * dx) + _f32(dy * dy))
    far = np.maximum(np.maximum(np.abs(xi[i]), np.abs(yi[i])), np.maximum(np.abs(xi[j]), np.abs(yi[j]))) > _f32(
        2.1 * cell)
    ex = r[i, 0] - r[j, 0]
    ey = r[i, 1] - r[j, 1]
    d2e = ex * ex + ey * ey  # numpy does not contract: the reference's separate roundings
    cut2 = cell * cell
    return d2.astype(np.float64), d2e, far, cut2


@pytest.mark.parametrize("fused", [False, True])
def test_fp32_entry_decisions_outside_the_band_are_the_fp64_tests(fused):
    r, hx, hy, pairs, cell, xmin, ymin = _entries(seed=7)
    assert len(pairs) > 50_000
    d2, d2e, far, cut2 = _decide(r, hx, hy, pairs, cell, xmin, ymin, fused)
    assert not far.any()
    rel = np.abs(d2 - d2e) / cut2
    print("entries", len(pairs), "max |d2_fp32 - d2_fp64| / cut2", rel.max())
    assert rel.max() < BAND / 10  # a tenth of the band at most
    lo, hi = np.float32(cut2 * (1 - BAND)), np.float32(cut2 * (1 + BAND))
    outside = ~((d2 > lo) & (d2 < hi))
    assert np.array_equal((d2 <= lo)[outside], (d2e < cut2)[outside])


def test_pairs_at_the_cut_off_fall_in_the_band():
    """Pairs within 1e-12 of the cut-off: every one is sent to the FP64 test."""
    r, hx, hy, pairs, cell, xmin, ymin = _entries(seed=11, near_cut=2000)
    for fused in (False, True):
        d2, d2e, far, cut2 = _decide(r, hx, hy, pairs, cell, xmin, ymin, fused)
        at_cut = np.abs(d2e - cut2) < 1e-9 * cut2
        assert at_cut.sum() >= 1000
        lo, hi = np.float32(cut2 * (1 - BAND)), np.float32(cut2 * (1 + BAND))
        assert np.all((d2[at_cut] > lo) & (d2[at_cut] < hi))


@pytest.mark.parametrize("edge", ["lo", "hi"])
def test_pairs_on_the_band_edges_are_decided_as_the_fp64_test(edge):
    """Pairs placed at the FP32 value of a band edge: some FP32 squared
    distances round exactly onto the edge, which lies outside the open band
    (lo, hi), so the FP32 value decides them.  d2 == lo must count as within
    the cut-off and d2 == hi as outside (the exact value lies within a
    rounding of the edge, a 1e-5 of the cut-off away from it): the decision
    is d2 <= lo, as the list steps take it."""
    r, hx, hy, pairs, cell, xmin, ymin = _entries(seed=13, near_cut=4000, edge=edge)
    for fused in (False, True):
        d2, d2e, far, cut2 = _decide(r, hx, hy, pairs, cell, xmin, ymin, fused)
        lo, hi = np.float32(cut2 * (1 - BAND)), np.float32(cut2 * (1 + BAND))
        on = d2 == np.float64(lo if edge == "lo" else hi)
        print(edge, "fused" if fused else "separate", "pairs on the edge", int(on.sum()))
        assert on.sum() >= 20
        assert np.all((d2e < cut2)[on] == (edge == "lo"))
        outside = ~((d2 > lo) & (d2 < hi))
        assert np.array_equal((d2 <= lo)[outside], (d2e < cut2)[outside])
