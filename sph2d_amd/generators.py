"""Seeded particle layouts for the workloads the step is measured on.

Conventions of the reference's generators (tools/Scripts/Piston2DGenerator.py:96-176,
sources/RRSPH_COMMON/tests/caseDamBreak/Generator.py:97-256):
  * fluid particles first, then boundary particles;
  * boundary walls of `layers` rows at spacing delta, odd layers shifted by
    delta/2 along the wall ("chess order"), going away from the fluid;
  * the piston wall is the first block of boundary particles and
    [nwm_particles_start, nwm_particles_end) marks it;
  * fluid starts at rest; density rho0 (dam break) or hydrostatic from the
    Tait EOS (piston tank, Piston2DGenerator.py:78-94);
  * itype 2 fluid, -2 boundary.
Fluid positions get a uniform jitter in [-jitter, +jitter] from a fixed seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, Tuple

import numpy as np

from . import params as P


@dataclass
class Layout:
    params: P.ExperimentParams
    r: np.ndarray
    v: np.ndarray
    rho: np.ndarray
    itype: np.ndarray

    @property
    def state(self) -> Dict[str, np.ndarray]:
        return {"r": self.r, "v": self.v, "rho": self.rho, "itype": self.itype}


def _wall_line(p0, direction, outward, length, delta, layers) -> np.ndarray:
    """Points of a straight wall of `layers` rows (chess order)."""
    p0 = np.asarray(p0, dtype=np.float64)
    d = np.asarray(direction, dtype=np.float64)
    o = np.asarray(outward, dtype=np.float64)
    n = int(round(length / delta)) + 1
    pts = []
    for layer in range(layers):
        shift = 0.5 * delta if layer % 2 else 0.0
        for k in range(n):
            pts.append(p0 + d * (k * delta + shift) + o * (layer * delta))
    return np.array(pts)


def _base_params(delta: float, hsml_coef: float, ntotal: int, nfluid: int, geom, depth: float) -> P.ExperimentParams:
    prm = P.ExperimentParams()
    prm.dim = 2
    prm.delta = delta
    prm.intf_hsml_coef = hsml_coef
    prm.ntotal = ntotal
    prm.nfluid = nfluid
    prm.nvirt = ntotal - nfluid
    prm.x_mingeom, prm.x_maxgeom, prm.y_mingeom, prm.y_maxgeom = geom
    prm.depth = depth
    prm.rho0 = 1000.0
    prm.density_skf = prm.intf_skf = prm.artificial_viscosity_skf = prm.average_velocity_skf = P.SKF_CUBIC
    prm.artificial_pressure_skf = P.SKF_CUBIC
    prm.intf_sph_approximation = P.INTF_SPH_APPROXIMATION_2
    prm.eos_sound_vel_method = P.EOS_SOUND_VEL_DAM_BREAK
    prm.eos_sound_vel_coef = 1.0
    prm.artificial_viscosity = True
    prm.artificial_shear_visc = 0.1
    prm.artificial_bulk_visc = 0.0
    prm.average_velocity = True
    prm.average_velocity_coef = 0.3
    prm.boundary_treatment = P.SBT_REPULSIVE
    prm.visc = False
    prm.max_neighbours = 64
    prm.consistency_check = True
    prm.consistency_treatment = P.CONSISTENCY_STOP
    prm.nwm_particles_start = ntotal
    prm.nwm_particles_end = ntotal
    return prm


def dam_break(delta: float = 0.01, tank=(3.2, 1.6), column=(1.0, 0.5), layers: int = 3, jitter: float = 1e-4,
              seed: int = 20240601, density_treatment: int = P.DENSITY_CONTINUITY, dt: float = 5e-5,
              hsml_coef: float = 1.2) -> Layout:
    """2D dam break: tank 3.2 x 1.6 m, fluid column 1.0 x 0.5 m at spacing delta,
    3-layer LJ walls (left, right, bottom), fixed dt."""
    W, H = tank
    cw, ch = column
    nx, ny = int(round(cw / delta)), int(round(ch / delta))
    # fluid starts at the Lennard-Jones zero-force distance r0 = 2h from the
    # wall rows (ExtForce.h:26: the wall force acts below r0)
    gap = 2.0 * hsml_coef * delta
    xs = gap + delta * np.arange(nx)
    ys = gap + delta * np.arange(ny)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    fluid = np.stack([X.ravel(), Y.ravel()], axis=1)
    rng = np.random.default_rng(seed)
    fluid = fluid + rng.uniform(-jitter, jitter, size=fluid.shape)
    bottom = _wall_line((0.0, 0.0), (1, 0), (0, -1), W, delta, layers)
    left = _wall_line((0.0, delta), (0, 1), (-1, 0), H - delta, delta, layers)
    right = _wall_line((W, delta), (0, 1), (1, 0), H - delta, delta, layers)
    walls = np.concatenate([left, bottom, right])
    r = np.concatenate([fluid, walls])
    nfluid = len(fluid)
    n = len(r)
    margin = (layers + 2) * delta
    prm = _base_params(delta, hsml_coef, n, nfluid, (-margin, W + margin, -margin, H + 0.4), ch)
    prm.density_treatment = density_treatment
    prm.dt_correction_method = P.DT_CORRECTION_CONST_VALUE
    prm.dt = dt
    prm.simulation_time = 2000 * dt
    v = np.zeros_like(r)
    rho = np.full(n, prm.rho0)
    itype = np.concatenate([np.full(nfluid, P.TYPE_WATER), np.full(n - nfluid, P.TYPE_BOUNDARY)]).astype(np.int32)
    P.fill_in_computing_params(prm)
    P.post_fill_in_model_params(prm)
    return Layout(prm, r, v, rho, itype)


def _wall_plane(axis: int, level: float, outward: int, extent, delta: float, layers: int) -> np.ndarray:
    """`layers` planes of wall particles normal to `axis` (0 x, 1 y, 2 z) on
    the lattice of spacing delta covering `extent` ((lo, hi) per other axis),
    the first at `level`, the next ones further out along `outward` (+1/-1);
    odd layers are shifted by delta / 2 in both in-plane directions."""
    others = [a for a in range(3) if a != axis]
    pts = []
    for layer in range(layers):
        shift = 0.5 * delta if layer % 2 else 0.0
        u = np.arange(extent[0][0] + shift, extent[0][1] + 1e-9, delta)
        w = np.arange(extent[1][0] + shift, extent[1][1] + 1e-9, delta)
        U, Wg = np.meshgrid(u, w, indexing="ij")
        q = np.zeros((U.size, 3))
        q[:, axis] = level + outward * layer * delta
        q[:, others[0]] = U.ravel()
        q[:, others[1]] = Wg.ravel()
        pts.append(q)
    return np.concatenate(pts)


def dam_break_3d(delta: float = 0.02, box=(0.8, 0.6, 0.4), column=(0.3, 0.3, 0.4), layers: int = 3,
                 jitter: float = 1e-4, seed: int = 20240604, density_treatment: int = P.DENSITY_SUMMATION,
                 hsml_coef: float = 1.3, cfl: float = 0.2, max_neighbours: int = 128) -> Layout:
    """3D dam break (y up): box x 0..L, y 0..H, z 0..W with 3-layer Lennard-Jones
    walls on the floor and the four sides, a fluid block (x, y, z) = column
    in the corner x = 0, fluid starting at the LJ zero-force distance 2h from
    the walls, at rest with rho = rho0, cubic kernels, fixed dt = cfl h / c."""
    L, H, W = box
    cl, chh, cw = column
    gap = 2.0 * hsml_coef * delta
    nx = int(round((cl - gap) / delta)) if cl > gap else 1
    ny = int(round(chh / delta))
    nz = int(round((cw - 2 * gap) / delta)) + 1
    xs = gap + delta * np.arange(nx)
    ys = gap + delta * np.arange(ny)
    zs = gap + delta * np.arange(nz)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    fluid = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    rng = np.random.default_rng(seed)
    fluid = fluid + rng.uniform(-jitter, jitter, size=fluid.shape)
    top = H
    floor = _wall_plane(1, 0.0, -1, ((0.0, L), (0.0, W)), delta, layers)
    left = _wall_plane(0, 0.0, -1, ((delta, top), (0.0, W)), delta, layers)
    right = _wall_plane(0, L, +1, ((delta, top), (0.0, W)), delta, layers)
    back = _wall_plane(2, 0.0, -1, ((delta, L - delta), (delta, top)), delta, layers)
    front = _wall_plane(2, W, +1, ((delta, L - delta), (delta, top)), delta, layers)
    walls = np.concatenate([floor, left, right, back, front])
    r = np.concatenate([fluid, walls])
    nfluid, n = len(fluid), len(fluid) + len(walls)
    margin = (layers + 2) * delta
    prm = _base_params(delta, hsml_coef, n, nfluid, (-margin, L + margin, -margin, H + 0.4), chh)
    prm.dim = 3
    prm.z_mingeom, prm.z_maxgeom = -margin, W + margin
    prm.max_neighbours = max_neighbours
    prm.density_treatment = density_treatment
    prm.dt_correction_method = P.DT_CORRECTION_CONST_VALUE
    P.fill_in_computing_params(prm)
    P.post_fill_in_model_params(prm)
    prm.dt = cfl * prm.hsml / prm.eos_c
    prm.simulation_time = 5000 * prm.dt
    v = np.zeros_like(r)
    rho = np.full(n, prm.rho0)
    itype = np.concatenate([np.full(nfluid, P.TYPE_WATER), np.full(n - nfluid, P.TYPE_BOUNDARY)]).astype(np.int32)
    return Layout(prm, r, v, rho, itype)


def hydrostatic_rho(y: np.ndarray, surface: float, rho0: float, c: float, g: float = P.G) -> np.ndarray:
    """rho from the Tait EOS at hydrostatic pressure (Piston2DGenerator.py:78-94)."""
    p = rho0 * g * np.maximum(surface - y, 0.0)
    B = c * c * rho0 / 7.0
    return rho0 * np.power(p / B + 1.0, 1.0 / 7.0)


def piston_tank(delta: float = 0.0075, length: float = 20.0, depth: float = 0.5, layers: int = 3,
                amplitude: float = 0.05, period: float = 1.2, wall_height: float = 1.0, jitter: float = 0.0,
                seed: int = 20240602, hsml_coef: float = 1.2, cfl: float = 0.2,
                fourier: Tuple[Tuple[float, ...], Tuple[float, ...], float] = None) -> Layout:
    """2D piston wave flume: fluid length x depth at rest with hydrostatic density,
    a vertical piston wall on the left moving as x(t) = A sin(2 pi t / T) (or a
    Fourier series), 3-layer LJ walls, continuity density, dynamic CFL dt.  The
    fluid starts 2 delta from the walls, where the reference's piston
    generator puts it (tools/Scripts/Piston2DGenerator.py:108), inside the 2h
    interaction range of the piston and the floor."""
    nx, ny = int(round(length / delta)), int(round(depth / delta))
    gap = 2.0 * delta
    xs = gap + delta * np.arange(nx)
    ys = gap + delta * np.arange(ny)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    fluid = np.stack([X.ravel(), Y.ravel()], axis=1)
    if jitter:
        rng = np.random.default_rng(seed)
        fluid = fluid + rng.uniform(-jitter, jitter, size=fluid.shape)
    L_end = xs[-1] + gap
    surface = ys[-1] + 0.5 * delta
    piston = _wall_line((0.0, delta), (0, 1), (-1, 0), wall_height, delta, layers)
    bottom = _wall_line((-(layers + 1) * delta, 0.0), (1, 0), (0, -1),
                        L_end + 2 * (layers + 1) * delta, delta, layers)
    right = _wall_line((L_end, delta), (0, 1), (1, 0), wall_height, delta, layers)
    nfluid = len(fluid)
    r = np.concatenate([fluid, piston, bottom, right])
    n = len(r)
    margin = (layers + 3) * delta
    geom = (-margin - 2 * amplitude, L_end + margin, -margin, wall_height + 0.6)
    prm = _base_params(delta, hsml_coef, n, nfluid, geom, depth)
    prm.density_treatment = P.DENSITY_CONTINUITY
    prm.dt_correction_method = P.DT_CORRECTION_DYNAMIC
    prm.CFL_coef = cfl
    prm.dt = 0.0
    prm.dt_dynamic_mode = 1
    prm.boundary_motion = True
    prm.nwm = P.NWM_METHOD_FOURIER
    prm.nwm_direction = P.NWM_DIRECTION_AXIS_X
    prm.nwm_particles_start = nfluid
    prm.nwm_particles_end = nfluid + len(piston)
    prm.nwm_time_start = 0.0
    if fourier is None:
        prm.piston_omega = 2 * math.pi / period
        prm.piston_a = (amplitude,)
        prm.piston_b = (0.0,)
    else:
        prm.piston_a, prm.piston_b, prm.piston_omega = tuple(fourier[0]), tuple(fourier[1]), float(fourier[2])
    prm.simulation_time = 1e9
    P.fill_in_computing_params(prm)
    P.post_fill_in_model_params(prm)
    v = np.zeros_like(r)
    rho = hydrostatic_rho(r[:, 1], surface, prm.rho0, prm.eos_c)
    itype = np.concatenate([np.full(nfluid, P.TYPE_WATER), np.full(n - nfluid, P.TYPE_BOUNDARY)]).astype(np.int32)
    return Layout(prm, r, v, rho, itype)


def adjoint_tank(delta: float = 0.005, length: float = 8.0, depth: float = 0.4, wall_height: float = 1.0,
                 nharm: int = 16, period: float = 1.6, sigma: float = 0.005, seed: int = 20240603,
                 cfl: float = 0.2) -> Layout:
    """Differentiable wave tank: the piston flume with a Fourier piston of
    `nharm` harmonics at w = 2 pi / period, coefficients drawn from
    N(0, sigma^2) with `seed`, and a fixed dt = cfl h / c (dt independent of
    the state, as the adjoint requires)."""
    rng = np.random.default_rng(seed)
    a = tuple(rng.normal(0.0, sigma, nharm))
    b = tuple(rng.normal(0.0, sigma, nharm))
    lay = piston_tank(delta=delta, length=length, depth=depth, wall_height=wall_height,
                      fourier=(a, b, 2 * math.pi / period))
    prm = lay.params
    prm.dt_correction_method = P.DT_CORRECTION_CONST_VALUE
    prm.dt_dynamic_mode = 0
    prm.dt = cfl * prm.hsml / prm.eos_c
    return lay


def piston_basin_3d(delta: float = 0.01, length: float = 6.0, width: float = 2.0, depth: float = 0.5,
                    wall_height: float = 1.0, layers: int = 3, segments: int = 8, nharm: int = 4,
                    period: float = 1.6, sigma: float = 0.01, seed: int = 20240605, hsml_coef: float = 1.2,
                    cfl: float = 0.2, max_neighbours: int = 128) -> Layout:
    """3D differentiable piston wave basin (y up): fluid length x depth x width
    at rest with hydrostatic density; the wall at x = 0 is a piston split into
    `segments` equal paddles across z, each moving as its own Fourier series
    of `nharm` harmonics at w = 2 pi / period (coefficients N(0, sigma^2) from
    `seed`, paddle-major); 3-layer LJ walls on the floor, the two sides and
    the far end; continuity density; fixed dt = cfl h / c.  The fluid starts
    2 delta from every wall, as the reference's piston generator places it
    (tools/Scripts/Piston2DGenerator.py:108), inside the 2h interaction
    range, so every paddle acts on the fluid from the first step."""
    gap = 2.0 * delta
    nx, ny = int(round(length / delta)), int(round(depth / delta))
    nz = int(round((width - 2 * gap) / delta)) + 1
    xs = gap + delta * np.arange(nx)
    ys = gap + delta * np.arange(ny)
    zs = gap + delta * np.arange(nz)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    fluid = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    L_end = xs[-1] + gap
    W = zs[-1] + gap
    surface = ys[-1] + 0.5 * delta
    x0, x1 = -(layers + 1) * delta, L_end + (layers + 1) * delta
    piston = _wall_plane(0, 0.0, -1, ((delta, wall_height), (delta, W - delta)), delta, layers)
    floor = _wall_plane(1, 0.0, -1, ((x0, x1), (0.0, W)), delta, layers)
    side0 = _wall_plane(2, 0.0, -1, ((x0, x1), (delta, wall_height)), delta, layers)
    side1 = _wall_plane(2, W, +1, ((x0, x1), (delta, wall_height)), delta, layers)
    end = _wall_plane(0, L_end, +1, ((delta, wall_height), (delta, W - delta)), delta, layers)
    nfluid = len(fluid)
    r = np.concatenate([fluid, piston, floor, side0, side1, end])
    n = len(r)
    margin = (layers + 3) * delta
    prm = _base_params(delta, hsml_coef, n, nfluid, (x0 - margin - 0.2, x1 + margin, -margin, wall_height + 0.6),
                       depth)
    prm.dim = 3
    prm.z_mingeom, prm.z_maxgeom = -margin - layers * delta, W + margin + layers * delta
    prm.max_neighbours = max_neighbours
    prm.density_treatment = P.DENSITY_CONTINUITY
    prm.boundary_motion = True
    prm.nwm = P.NWM_METHOD_FOURIER
    prm.nwm_direction = P.NWM_DIRECTION_AXIS_X
    prm.nwm_particles_start = nfluid
    prm.nwm_particles_end = nfluid + len(piston)
    prm.nwm_time_start = 0.0
    rng = np.random.default_rng(seed)
    prm.piston_a = tuple(rng.normal(0.0, sigma, segments * nharm))
    prm.piston_b = tuple(rng.normal(0.0, sigma, segments * nharm))
    prm.piston_omega = 2 * math.pi / period
    # paddle boundaries midway between the rows of wall particles (z on a
    # delta / 2 lattice), so that no particle sits on a boundary
    half = 0.5 * delta
    w = max(round(W / segments / half), 1) * half
    z0 = 0.5 * half + round(0.5 * (W - segments * w) / half) * half
    prm.piston_segments = segments
    prm.piston_seg_zmin, prm.piston_seg_zmax = z0, z0 + segments * w
    prm.dt_correction_method = P.DT_CORRECTION_CONST_VALUE
    prm.dt_dynamic_mode = 0
    P.fill_in_computing_params(prm)
    P.post_fill_in_model_params(prm)
    prm.dt = cfl * prm.hsml / prm.eos_c
    prm.simulation_time = 1e9
    v = np.zeros_like(r)
    rho = hydrostatic_rho(r[:, 1], surface, prm.rho0, prm.eos_c)
    itype = np.concatenate([np.full(nfluid, P.TYPE_WATER), np.full(n - nfluid, P.TYPE_BOUNDARY)]).astype(np.int32)
    return Layout(prm, r, v, rho, itype)


def c1_1m_delta() -> float:
    """Spacing that puts ~1M fluid particles in the 20 m x 0.5 m flume."""
    return 0.0031623
