"""Pins the oracle's 3D pair terms through its 2D ones, which the
reference's golden vectors pin (tests/test_oracle_golden.py).

The reference's 2D module-test dam break (23,030 particles), lifted into a
thin 3D domain at z = 0 with zero z velocity, has the same neighbour sets in
3D as in 2D.  Every kernel-weighted pair sum then differs only by the ratio
of the 3D to the 2D kernel factor (the cubic spline's shape function is the
same, SmoothingKernel.h:14-67; the 3D factor is pinned by its integral in
tests/test_oracle_3d.py): the summation and continuity densities, the
pressure force (both approximations), the artificial viscosity and the
average velocity.  The external force (gravity and the Lennard-Jones wall
force) has no kernel and is the same, as are the viscosity's mu and the
pressures; every z component is 0.  The same case lifted into the x-z
plane instead, (x, y) -> (x, 0, y), exercises the z parts: its squared
distances are the same sums bit for bit, so the neighbour sets are too, and
the (x, z) components of the kernel-weighted terms equal the scaled 2D
(x, y) ones with y exactly 0 (gravity is not rotated: the external force
is compared whole in the x-y lift, its wall part in both).
"""
import copy
import ctypes

import numpy as np
import pytest

import oracle
from helpers import load_dambreak_case


@pytest.fixture(scope="module", params=["xy", "xz"])
def planar(request):
    prm2, st, _tg = load_dambreak_case()
    prm3 = copy.copy(prm2)
    prm3.dim = 3
    n = prm2.ntotal
    r2 = np.ascontiguousarray(st["r"], dtype=np.float64)
    v2 = np.ascontiguousarray(st["v"], dtype=np.float64)
    z = np.zeros(n)
    if request.param == "xy":
        prm3.z_mingeom, prm3.z_maxgeom = -0.1, 0.1
        r3 = np.ascontiguousarray(np.column_stack([r2[:, 0], r2[:, 1], z]))
        v3 = np.ascontiguousarray(np.column_stack([v2[:, 0], v2[:, 1], z]))
    else:
        prm3.z_mingeom, prm3.z_maxgeom = prm2.y_mingeom, prm2.y_maxgeom
        prm3.y_mingeom, prm3.y_maxgeom = -0.1, 0.1
        r3 = np.ascontiguousarray(np.column_stack([r2[:, 0], z, r2[:, 1]]))
        v3 = np.ascontiguousarray(np.column_stack([v2[:, 0], z, v2[:, 1]]))
    s2, nb2 = oracle.grid_find(prm2, r2, st["itype"])
    s3, nb3 = oracle.grid_find(prm3, r3, st["itype"])
    assert s2 == 0 and s3 == 0
    L = oracle.lib()
    L.oracle_kernel_w.restype = ctypes.c_double
    L.oracle_kernel_w.argtypes = [ctypes.POINTER(oracle.OracleParams), ctypes.c_double, ctypes.c_int32]
    w3 = L.oracle_kernel_w(ctypes.byref(oracle.from_params(prm3)), 0.0, 1)
    w2 = L.oracle_kernel_w(ctypes.byref(oracle.from_params(prm2)), 0.0, 1)
    ratio = w3 / w2
    # (1 / (pi h^3)) / (10 / (7 pi h^2)), up to the single-precision pi of the 2D factor
    assert ratio == pytest.approx(7.0 / (10.0 * prm2.hsml), rel=1e-6)
    c2 = (prm2, r2, v2, nb2)
    c3 = (prm3, r3, v3, nb3)
    return c2, c3, st, ratio, request.param


def _sets(nb, n):
    out = []
    for row in nb:
        end = int(np.argmax(row == n))
        out.append(sorted(row[:end].tolist()))
    return out


def test_planar_neighbour_sets_are_the_2d_sets(planar):
    (prm2, _r2, _v2, nb2), (_p3, _r3, _v3, nb3), _st, _ratio, _lift = planar
    n = prm2.ntotal
    s2, s3 = _sets(nb2, n), _sets(nb3, n)
    assert s2 == s3
    assert sum(len(s) for s in s2) > 10 * n  # a real neighbourhood


def _run(case, st, which, p=None, **kw):
    prm, r, v, nb = case
    q = copy.copy(prm)
    for k, val in kw.items():
        setattr(q, k, val)
    op = oracle.from_params(q)
    n, D = prm.ntotal, prm.dim
    L = oracle.lib()
    NB = oracle.ptr(np.ascontiguousarray(nb.reshape(-1)), ctypes.c_uint32)
    rho = np.ascontiguousarray(st["rho"], dtype=np.float64)
    it = np.ascontiguousarray(st["itype"], dtype=np.int32)
    if which == "density_con":
        drho, pres = np.zeros(n), np.zeros(n)
        L.oracle_density_con(ctypes.byref(op), oracle.ptr(r), oracle.ptr(v), NB, oracle.ptr(rho), oracle.ptr(drho),
                             oracle.ptr(pres))
        return drho, pres
    if which == "density_sum":
        out, pres = np.zeros(n), np.zeros(n)
        L.oracle_density_sum(ctypes.byref(op), oracle.ptr(r), NB, oracle.ptr(out), oracle.ptr(pres))
        return out, pres
    out = np.zeros((n, D))
    if which == "internal":
        L.oracle_term_internal(ctypes.byref(op), oracle.ptr(r), NB, oracle.ptr(p), oracle.ptr(rho), oracle.ptr(out))
        return out
    if which == "external":
        L.oracle_term_external(ctypes.byref(op), oracle.ptr(r), oracle.ptr(it, ctypes.c_int32), NB, oracle.ptr(out))
        return out
    if which == "art_visc":
        mu = np.zeros(n)
        L.oracle_term_art_visc(ctypes.byref(op), oracle.ptr(r), oracle.ptr(v), NB, oracle.ptr(rho), oracle.ptr(out),
                               oracle.ptr(mu))
        return out, mu
    if which == "average_velocity":
        L.oracle_term_average_velocity(ctypes.byref(op), oracle.ptr(r), oracle.ptr(v), oracle.ptr(it, ctypes.c_int32),
                                       NB, oracle.ptr(rho), prm.nfluid, oracle.ptr(out))
        return out
    raise ValueError(which)


def _scaled(out3, out2, scale, lift="xy"):
    """out3 = scale * out2 to 1e-12 of the largest value (the 3D walk sums
    the same pairs in another order), the 2D axes mapped by the lift; the
    third component is exactly 0."""
    if out3.ndim == 2:
        axes, zero = ((0, 1), 2) if lift == "xy" else ((0, 2), 1)
        assert np.all(out3[:, zero] == 0.0)
        out3 = out3[:, list(axes)]
    want = scale * out2
    top = float(np.max(np.abs(want)))
    assert top > 0.0
    err = float(np.max(np.abs(out3 - want)))
    assert err <= 1e-12 * top, err / top


def test_planar_densities(planar):
    c2, c3, st, ratio, lift = planar
    d2, p2 = _run(c2, st, "density_con")
    d3, p3 = _run(c3, st, "density_con")
    _scaled(d3, d2, ratio, lift)
    assert np.array_equal(p3, p2)  # the EOS of the same densities
    s2, _ = _run(c2, st, "density_sum")
    s3, _ = _run(c3, st, "density_sum")
    _scaled(s3, s2, ratio, lift)


@pytest.mark.parametrize("approx", [1, 2])
def test_planar_pressure_force(planar, approx):
    c2, c3, st, ratio, lift = planar
    _d, p = _run(c2, st, "density_con")
    a2 = _run(c2, st, "internal", p=p, intf_sph_approximation=approx, artificial_pressure=0)
    a3 = _run(c3, st, "internal", p=p, intf_sph_approximation=approx, artificial_pressure=0)
    _scaled(a3, a2, ratio, lift)


def test_planar_wall_force(planar):
    """The Lennard-Jones wall force (repulsive boundaries minus the dynamic
    ones, which leaves gravity out) rotates with the lift; in the x-y lift
    the whole external force, gravity along -y included, is the 2D one."""
    c2, c3, st, _ratio, lift = planar
    ext = {(d, sbt): _run(c, st, "external", boundary_treatment=sbt) for d, c in ((2, c2), (3, c3)) for sbt in (0, 1)}
    lj2, lj3 = ext[2, 1] - ext[2, 0], ext[3, 1] - ext[3, 0]
    assert np.any(lj2[:, 0] != 0.0)  # the walls act
    _scaled(lj3, lj2, 1.0, lift)
    if lift == "xy":
        for sbt in (0, 1):
            _scaled(ext[3, sbt], ext[2, sbt], 1.0)


def test_planar_artificial_viscosity(planar):
    c2, c3, st, ratio, lift = planar
    a2, mu2 = _run(c2, st, "art_visc", dt_correction_method=2, artificial_viscosity=True)
    a3, mu3 = _run(c3, st, "art_visc", dt_correction_method=2, artificial_viscosity=True)
    _scaled(a3, a2, ratio, lift)
    assert np.max(np.abs(mu3 - mu2)) <= 1e-12 * np.max(np.abs(mu2))


def test_planar_average_velocity(planar):
    c2, c3, st, ratio, lift = planar
    kw = dict(average_velocity=True, average_velocity_skf=1, average_velocity_coef=0.05)
    a2 = _run(c2, st, "average_velocity", **kw)
    a3 = _run(c3, st, "average_velocity", **kw)
    _scaled(a3, a2, ratio, lift)
