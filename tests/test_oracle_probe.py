"""The oracle's FuncAtPoint (tools/FuncAtPoint/lib/src/fatp.cpp:14-58,
:121-144) pinned by closed forms the definition reduces to, with the cubic
spline (SmoothingKernel.h:20-45) for both the density and the probe, so that
its normalisation f cancels:

* one particle alone within the cut-off: rho = m W(0), value = f w(q) / w(0);
* two particles a distance s apart, nothing else near: rho = m (W(0) + W(s))
  for both, value = (f_a w(q_a) + f_b w(q_b)) / (w(0) + w(s / h));
* a square lattice, probe at an interior site: every particle within reach
  has the same summation density m S, S the lattice sum of W, and the probe
  sums the same W over the same offsets, so a constant field gives exactly
  the constant (partition of unity);
* points past the cut-off of every particle give 0.
w(q) = 2/3 - q^2 + q^3 / 2 for q <= 1, (2 - q)^3 / 6 for q <= 2, 0 beyond.
The normalisation cancels in every one of these, so the same closed forms
hold in 3D (the 3D cubic spline has the same shape w), checked on a lone
particle, a pair and a cubic lattice as well."""
import copy

import numpy as np

import oracle
from sph2d_amd import generators as G


def w(q):
    q = np.asarray(q, dtype=np.float64)
    return np.where(q <= 1, 2 / 3 - q ** 2 + 0.5 * q ** 3, np.where(q <= 2, (2 - q) ** 3 / 6, 0.0))


def _params(n):
    prm = copy.copy(G.dam_break().params)
    prm.ntotal = prm.nfluid = n
    prm.nvirt = 0
    prm.maxn = max(prm.maxn, n)
    return prm


def _state(r, f):
    n = len(r)
    return {"r": np.asarray(r, dtype=np.float64), "v": np.zeros((n, 2)), "rho": np.zeros(n),
            "p": np.asarray(f, dtype=np.float64), "itype": np.full(n, 2, dtype=np.int32)}


def test_isolated_particle():
    prm = _params(3)
    h = prm.hsml
    r = [[1.0, 1.0], [2.0, 1.0], [1.0, 1.8]]  # the other two are far from the probes
    f = [2.5, 7.0, -3.0]
    qs = np.array([0.0, 0.3, 0.9, 1.0, 1.4, 1.99, 2.0, 2.5])
    pts = np.stack([1.0 + qs * h * 0.6, 1.0 + qs * h * 0.8], axis=1)  # |x - r_0| = q h
    st, got = oracle.probe(prm, _state(r, f), pts, "p")
    assert st == 0
    want = 2.5 * w(qs) / w(0.0)
    want[qs >= 2.0] = 0.0  # |r - x|^2 < (2h)^2 is strict
    assert np.allclose(got, want, rtol=1e-13, atol=1e-15), (got, want)


def test_particle_pair():
    prm = _params(2)
    h = prm.hsml
    s = 1.3 * h
    r = np.array([[1.0, 1.0], [1.0 + s, 1.0]])
    f = [4.0, -1.5]
    pts = np.array([[1.0 + 0.5 * s, 1.0], [1.0 + 0.2 * s, 1.0 + 0.3 * h], [1.0 - 0.7 * h, 0.9]])
    st, got = oracle.probe(prm, _state(r, f), pts, "p")
    assert st == 0
    qa = np.linalg.norm(pts - r[0], axis=1) / h
    qb = np.linalg.norm(pts - r[1], axis=1) / h
    want = (4.0 * w(qa) - 1.5 * w(qb)) / (w(0.0) + w(s / h))
    assert np.allclose(got, want, rtol=1e-13, atol=1e-15), (got, want)


def test_partition_of_unity_on_a_lattice():
    prm = _params(0)
    d = prm.delta
    k = np.arange(40)
    xx, yy = np.meshgrid(0.5 + d * k, 0.5 + d * k, indexing="ij")
    r = np.stack([xx.ravel(), yy.ravel()], axis=1)
    prm.ntotal = prm.nfluid = len(r)
    prm.maxn = max(prm.maxn, len(r))
    interior = r[[820, 861, 900, 1015]]  # sites (i, j) = i * 40 + j, at least 15 sites from every edge
    st_, got = oracle.probe(prm, _state(r, np.full(len(r), 3.0)), interior, "p")
    assert st_ == 0
    assert np.allclose(got, 3.0, rtol=1e-13), got
    # between sites the interpolant of a constant deviates from it, a little
    mid = r[[820, 861]] + 0.5 * d
    st_, off = oracle.probe(prm, _state(r, np.full(len(r), 3.0)), mid, "p")
    assert np.all(np.abs(off - 3.0) < 0.1) and np.all(off != 3.0), off


def test_velocity_component_and_stride():
    prm = _params(1)
    h = prm.hsml
    state = _state([[1.0, 1.0]], [0.0])
    state["v"] = np.array([[0.25, -0.75]])
    pts = np.array([[1.0 + 0.5 * h, 1.0]])
    for field, f in (("vx", 0.25), ("vy", -0.75)):
        st, got = oracle.probe(prm, state, pts, field)
        assert st == 0 and np.isclose(got[0], f * w(0.5) / w(0.0), rtol=1e-13), (field, got)


def _params3(n):
    prm = copy.copy(G.dam_break_3d().params)
    prm.ntotal = prm.nfluid = n
    return prm


def _state3(r, f):
    n = len(r)
    return {"r": np.asarray(r, dtype=np.float64), "v": np.zeros((n, 3)), "rho": np.zeros(n),
            "p": np.asarray(f, dtype=np.float64), "itype": np.full(n, 2, dtype=np.int32)}


def test_3d_isolated_particle_and_pair():
    prm = _params3(3)
    h = prm.hsml
    u = np.array([0.48, 0.6, 0.64])  # unit vector
    r = [[0.3, 0.2, 0.15], [0.6, 0.2, 0.15], [0.3, 0.45, 0.3]]  # the other two are far from the probes
    qs = np.array([0.0, 0.3, 0.9, 1.0, 1.4, 1.99, 2.0, 2.5])
    pts = np.array(r[0]) + np.outer(qs * h, u)
    st, got = oracle.probe(prm, _state3(r, [2.5, 7.0, -3.0]), pts, "p")
    assert st == 0
    want = 2.5 * w(qs) / w(0.0)
    want[qs >= 2.0] = 0.0
    assert np.allclose(got, want, rtol=1e-13, atol=1e-15), (got, want)
    prm = _params3(2)
    s = 1.3 * h
    rr = np.array([[0.3, 0.2, 0.15], [0.3, 0.2, 0.15]]) + np.array([[0.0, 0.0, 0.0], s * u])
    pts = np.array([rr[0] + 0.5 * s * u, rr[0] + np.array([0.2, -0.3, 0.5]) * h, rr[1] + np.array([0.7, 0.1, 0.0]) * h])
    st, got = oracle.probe(prm, _state3(rr, [4.0, -1.5]), pts, "p")
    assert st == 0
    qa = np.linalg.norm(pts - rr[0], axis=1) / h
    qb = np.linalg.norm(pts - rr[1], axis=1) / h
    want = (4.0 * w(qa) - 1.5 * w(qb)) / (w(0.0) + w(s / h))
    assert np.allclose(got, want, rtol=1e-13, atol=1e-15), (got, want)


def test_3d_partition_of_unity_on_a_cubic_lattice():
    prm = _params3(0)
    d = prm.delta
    k = np.arange(14)
    xx, yy, zz = np.meshgrid(0.05 + d * k, 0.05 + d * k, 0.02 + d * k, indexing="ij")
    r = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1)
    prm.ntotal = prm.nfluid = len(r)
    idx = lambda i, j, l: (i * 14 + j) * 14 + l
    interior = r[[idx(7, 7, 7), idx(6, 7, 8), idx(7, 6, 6)]]  # at least 5 sites (> 2h) from every face
    assert 2 * prm.hsml < 5 * d
    st_, got = oracle.probe(prm, _state3(r, np.full(len(r), 3.0)), interior, "p")
    assert st_ == 0
    assert np.allclose(got, 3.0, rtol=1e-13), got
    mid = interior[:2] + 0.5 * d
    st_, off = oracle.probe(prm, _state3(r, np.full(len(r), 3.0)), mid, "p")
    assert np.all(np.abs(off - 3.0) < 0.1) and np.all(off != 3.0), off
