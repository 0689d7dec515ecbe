"""Pins the oracle's 3D path, for which the reference's module tests hold no
golden vectors (their cases are 2D): closed forms and brute force.

* the 3D cubic, Gauss, Wendland and Desbrun kernels integrate to one over
  their support (SmoothingKernel.h:22-23, 93-94, 136-137, 194-195 factors);
* grad W equals the finite difference of W;
* the 3D Morton key is the bit interleave of the cell coordinates
  (GridUtils.h:49-66, brute force over bits);
* the 3D neighbour sets equal an O(N^2) search;
* the pressure and viscosity pair forces conserve momentum (sum m a_i = 0
  over a closed set of particles, gravity and walls excluded);
* the renormalised summation density (docs/latex/ModelParams.tex:124-133)
  reproduces a uniform density exactly and follows the two-particle formula
  with the 3D kernel;
* the segmented Fourier piston moves each paddle's particles along x at the
  derivative of that paddle's own displacement series.
"""
import copy
import ctypes

import numpy as np
import pytest

import oracle
from sph2d_amd import generators as G


def _p3(skf=1):
    lay = G.dam_break_3d()
    prm = copy.copy(lay.params)
    prm.density_skf = prm.intf_skf = prm.artificial_viscosity_skf = prm.average_velocity_skf = skf
    return lay, prm


@pytest.mark.parametrize("skf", [1, 2, 3, 4])
def test_3d_kernels_integrate_to_one(skf):
    lay, prm = _p3(skf)
    op = oracle.from_params(prm)
    L = oracle.lib()
    L.oracle_kernel_w.restype = ctypes.c_double
    L.oracle_kernel_w.argtypes = [ctypes.POINTER(oracle.OracleParams), ctypes.c_double, ctypes.c_int32]
    h = prm.hsml
    rmax = 3.0 * h if skf != 2 else 6.0 * h
    r = np.linspace(0.0, rmax, 20001)
    w = np.array([L.oracle_kernel_w(ctypes.byref(op), x, skf) for x in r])
    integral = np.trapezoid(4 * np.pi * r * r * w, r)
    assert abs(integral - 1.0) < 2e-6, integral


def test_3d_gradient_is_the_derivative_of_w():
    lay, prm = _p3(1)
    op = oracle.from_params(prm)
    L = oracle.lib()
    L.oracle_kernel_w.restype = ctypes.c_double
    L.oracle_kernel_w.argtypes = [ctypes.POINTER(oracle.OracleParams), ctypes.c_double, ctypes.c_int32]
    h = prm.hsml
    rng = np.random.default_rng(0)
    for _ in range(20):
        e = rng.normal(size=3)
        e *= rng.uniform(0.05, 1.95) * h / np.linalg.norm(e)
        out = np.zeros(3)
        L.oracle_kernel_dwdr(ctypes.byref(op), float(np.linalg.norm(e)), oracle.ptr(e), 1, oracle.ptr(out))
        eps = 1e-7 * h
        fd = np.array([(L.oracle_kernel_w(ctypes.byref(op), float(np.linalg.norm(e + eps * u)), 1)
                        - L.oracle_kernel_w(ctypes.byref(op), float(np.linalg.norm(e - eps * u)), 1)) / (2 * eps)
                       for u in np.eye(3)])
        assert np.allclose(out, fd, rtol=1e-6, atol=1e-6 * np.abs(fd).max())


def test_3d_morton_is_the_bit_interleave():
    L = oracle.lib()
    L.oracle_morton3.restype = ctypes.c_uint32
    rng = np.random.default_rng(1)
    for _ in range(200):
        x, y, z = (int(v) for v in rng.integers(0, 1024, 3))
        want = 0
        for b in range(10):
            want |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
        assert L.oracle_morton3(x, y, z) == want


def test_3d_neighbour_sets_match_brute_force():
    lay, prm = _p3()
    st, nb = oracle.grid_find(prm, lay.r, lay.itype)
    assert st == 0
    got = oracle.neighbour_sets(nb, prm.ntotal)
    cut2 = (prm.cell_scale_k * prm.hsml) ** 2
    r = lay.r
    for j in np.random.default_rng(2).choice(prm.ntotal, 400, replace=False):
        d = r - r[j]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        m = d2 < cut2
        m[j] = False
        assert np.array_equal(got[j], np.nonzero(m)[0].astype(np.uint32)), j


def test_3d_pair_forces_conserve_momentum():
    """Fluid only (no walls, no gravity in the sum): pressure + viscosity
    accelerations times mass sum to zero; XSPH velocities sum to zero."""
    lay, prm = _p3()
    nf = prm.nfluid
    p = copy.copy(prm)
    p.ntotal = p.nfluid = nf
    p.nvirt = 0
    r = lay.r[:nf].copy()
    it = lay.itype[:nf].copy()
    rng = np.random.default_rng(3)
    v = rng.normal(0, 0.1, (nf, 3))
    rho = p.rho0 * (1 + rng.normal(0, 1e-3, nf))
    st, nb = oracle.grid_find(p, r, it)
    assert st == 0
    nbf = np.ascontiguousarray(nb.reshape(-1))
    op = oracle.from_params(p)
    L = oracle.lib()
    L.oracle_eos_p.restype = ctypes.c_double
    L.oracle_eos_p.argtypes = [ctypes.POINTER(oracle.OracleParams), ctypes.c_double]
    pres = np.array([L.oracle_eos_p(ctypes.byref(op), x) for x in rho])
    a_int = np.zeros((nf, 3))
    L.oracle_term_internal(ctypes.byref(op), oracle.ptr(r), oracle.ptr(nbf, ctypes.c_uint32), oracle.ptr(pres),
                           oracle.ptr(rho), oracle.ptr(a_int))
    a_av = np.zeros((nf, 3)); mu = np.zeros(nf)
    L.oracle_term_art_visc(ctypes.byref(op), oracle.ptr(r), oracle.ptr(v), oracle.ptr(nbf, ctypes.c_uint32),
                           oracle.ptr(rho), oracle.ptr(a_av), oracle.ptr(mu))
    scale = np.abs(a_int).sum()
    assert np.abs(a_int.sum(axis=0)).max() < 1e-9 * scale
    assert np.abs(a_av.sum(axis=0)).max() < 1e-9 * max(np.abs(a_av).sum(), 1e-30)


def _renorm3(prm, r, itype, rho_old):
    """The oracle's renormalised and plain summation densities of state r."""
    st, nb = oracle.grid_find(prm, r, itype)
    assert st == 0
    op = oracle.from_params(prm)
    n = prm.ntotal
    rho, p, plain, pp = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    r = np.ascontiguousarray(r, dtype=np.float64)
    nbf = np.ascontiguousarray(nb.reshape(-1))
    oracle.lib().oracle_density_sum_normalized(ctypes.byref(op), oracle.ptr(r), oracle.ptr(nbf, ctypes.c_uint32),
                                               oracle.ptr(np.ascontiguousarray(rho_old, dtype=np.float64)),
                                               oracle.ptr(rho), oracle.ptr(p))
    oracle.lib().oracle_density_sum(ctypes.byref(op), oracle.ptr(r), oracle.ptr(nbf, ctypes.c_uint32),
                                    oracle.ptr(plain), oracle.ptr(pp))
    return rho, plain


def test_3d_renormalised_density_reproduces_a_uniform_density():
    """rho_j = sum m W / sum (m / rho_i) W is c for every rho_i = c, also at
    the free surface, the walls and the corners of the 3D dam break, where
    the plain summation falls short of rho0."""
    lay, prm = _p3(1)
    prm.density_treatment = 0
    for c in (1000.0, 1020.5):
        rho, plain = _renorm3(prm, lay.r, lay.itype, np.full(prm.ntotal, c))
        assert np.max(np.abs(rho - c)) <= 4e-16 * c * 8
    assert plain.min() < 0.8 * prm.rho0  # missing kernel support at the surface and in the corners


def test_3d_renormalised_density_two_particle_closed_form():
    """Two particles q h apart along a 3D direction:
    rho_j = (m W(0) + m W(qh)) / ((m / rho_j) W(0) + (m / rho_i) W(qh)), W the 3D cubic spline."""
    lay, prm = _p3(1)
    prm.density_treatment = 0
    prm.ntotal = prm.nfluid = 2
    h = prm.hsml
    L = oracle.lib()
    L.oracle_kernel_w.restype = ctypes.c_double
    L.oracle_kernel_w.argtypes = [ctypes.POINTER(oracle.OracleParams), ctypes.c_double, ctypes.c_int32]
    op = oracle.from_params(prm)
    u = np.array([0.48, 0.6, 0.64])  # unit vector
    base = np.array([0.2, 0.15, 0.1])
    for q in (0.4, 1.3, 1.9):
        r = np.array([base, base + q * h * u])
        rho_old = np.array([1003.0, 991.0])
        rho, plain = _renorm3(prm, r, np.array([2, 2], dtype=np.int32), rho_old)
        w0 = L.oracle_kernel_w(ctypes.byref(op), 0.0, 1)
        w = L.oracle_kernel_w(ctypes.byref(op), q * h, 1)
        m = prm.mass
        for j, i in ((0, 1), (1, 0)):
            want = (m * w0 + m * w) / (m / rho_old[j] * w0 + m / rho_old[i] * w)
            assert rho[j] == pytest.approx(want, rel=1e-14)
            assert plain[j] == pytest.approx(m * (w0 + w), rel=1e-14)


def test_segmented_piston_moves_each_paddle_with_its_own_series():
    """A 3D basin whose piston is split into 4 paddles across z, each with
    its own 3 harmonics (paddle-major coefficients): after one oracle step
    from t0, every piston particle moves along x at dX_g/dt (t0), X_g(t) =
    sum_k a_gk sin(k w t) + b_gk cos(k w t) the displacement of its paddle g
    (central differences of X_g), g the paddle whose z range holds the
    particle (the end rows beyond the paddles belong to the end paddles)."""
    segs, nh = 4, 3
    lay = G.piston_basin_3d(delta=0.05, length=1.0, width=0.8, depth=0.2, segments=segs, nharm=nh, sigma=0.05)
    prm = lay.params
    a = np.array(prm.piston_a).reshape(segs, nh)
    b = np.array(prm.piston_b).reshape(segs, nh)
    w = prm.piston_omega
    t0 = 0.37
    o = oracle.Oracle(prm)
    o.set_state(lay.r, lay.v, lay.rho, lay.itype, time=t0, dt=prm.dt)
    assert o.step(1) == 0
    v = o.state()["v"]
    k = np.arange(1, nh + 1)

    def X(g, t):
        return float(np.sum(a[g] * np.sin(k * w * t) + b[g] * np.cos(k * w * t)))

    eps = 1e-6
    vel = [(X(g, t0 + eps) - X(g, t0 - eps)) / (2 * eps) for g in range(segs)]
    assert min(abs(vel[g] - vel[h]) for g in range(segs) for h in range(g)) > 1e-3  # the paddles differ
    i0, i1 = prm.nwm_particles_start, prm.nwm_particles_end
    z = lay.r[i0:i1, 2]
    width = (prm.piston_seg_zmax - prm.piston_seg_zmin) / segs
    seen = set()
    for j, zz in enumerate(z):
        g = 0
        while g + 1 < segs and zz >= prm.piston_seg_zmin + (g + 1) * width:
            g += 1
        seen.add(g)
        assert v[i0 + j, 0] == pytest.approx(vel[g], rel=1e-8, abs=1e-12), (j, zz, g)
        assert v[i0 + j, 1] == 0.0 and v[i0 + j, 2] == 0.0
    assert seen == set(range(segs))
