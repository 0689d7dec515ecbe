"""ctypes binding of the FP64 CPU oracle (oracle/sph_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference results.  The product package
(sph2d_amd) never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
MAX_HARMONICS = 32

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_uint32_p = ctypes.POINTER(ctypes.c_uint32)


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32), ("ntotal", ctypes.c_uint32), ("nfluid", ctypes.c_uint32),
        ("max_neighbours", ctypes.c_uint32),
        ("x_mingeom", ctypes.c_double), ("x_maxgeom", ctypes.c_double),
        ("y_mingeom", ctypes.c_double), ("y_maxgeom", ctypes.c_double),
        ("z_mingeom", ctypes.c_double), ("z_maxgeom", ctypes.c_double),
        ("hsml", ctypes.c_double), ("delta", ctypes.c_double), ("mass", ctypes.c_double),
        ("rho0", ctypes.c_double), ("cell_scale_k", ctypes.c_double), ("depth", ctypes.c_double),
        ("g", ctypes.c_double), ("pi", ctypes.c_double),
        ("eos_c_sqr", ctypes.c_double), ("eos_c", ctypes.c_double), ("eos_sound_vel", ctypes.c_double),
        ("density_treatment", ctypes.c_int32), ("density_skf", ctypes.c_int32),
        ("density_delta_sph_coef", ctypes.c_double),
        ("intf_sph_approximation", ctypes.c_int32), ("intf_skf", ctypes.c_int32),
        ("artificial_pressure", ctypes.c_int32), ("artificial_pressure_skf", ctypes.c_int32),
        ("artificial_pressure_index", ctypes.c_double), ("artificial_pressure_coef", ctypes.c_double),
        ("artificial_viscosity", ctypes.c_int32), ("artificial_viscosity_skf", ctypes.c_int32),
        ("artificial_shear_visc", ctypes.c_double), ("artificial_bulk_visc", ctypes.c_double),
        ("average_velocity", ctypes.c_int32), ("average_velocity_skf", ctypes.c_int32),
        ("average_velocity_coef", ctypes.c_double),
        ("boundary_treatment", ctypes.c_int32),
        ("dt_correction_method", ctypes.c_int32), ("dt", ctypes.c_double), ("CFL_coef", ctypes.c_double),
        ("consistency_check", ctypes.c_int32), ("consistency_treatment", ctypes.c_int32),
        ("nwm", ctypes.c_int32), ("nwm_direction", ctypes.c_int32),
        ("nwm_particles_start", ctypes.c_uint32), ("nwm_particles_end", ctypes.c_uint32),
        ("nwm_time_start", ctypes.c_double), ("nwm_freq", ctypes.c_double),
        ("nwm_wave_number", ctypes.c_double), ("nwm_wave_magnitude", ctypes.c_double),
        ("nwm_phase", ctypes.c_double),
        ("piston_nharm", ctypes.c_int32), ("piston_omega", ctypes.c_double),
        ("piston_a", ctypes.c_double * MAX_HARMONICS), ("piston_b", ctypes.c_double * MAX_HARMONICS),
        ("piston_segments", ctypes.c_int32), ("piston_seg_zmin", ctypes.c_double),
        ("piston_seg_zmax", ctypes.c_double),
        ("boundary_motion", ctypes.c_int32), ("dt_dynamic_mode", ctypes.c_int32),
        ("num_threads", ctypes.c_int32),
        ("density_normalization", ctypes.c_int32), ("density_normalization_every", ctypes.c_uint32),
    ]


_lib: Optional[ctypes.CDLL] = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER(OracleParams)
        sig = {
            "oracle_kernel_w": (ctypes.c_double, [P, ctypes.c_double, ctypes.c_int32]),
            "oracle_kernel_dwdr": (None, [P, ctypes.c_double, c_double_p, ctypes.c_int32, c_double_p]),
            "oracle_eos_p": (ctypes.c_double, [P, ctypes.c_double]),
            "oracle_eos_rho": (ctypes.c_double, [P, ctypes.c_double]),
            "oracle_morton2": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32]),
            "oracle_morton2_decode": (None, [ctypes.c_uint32, c_uint32_p, c_uint32_p]),
            "oracle_morton3": (ctypes.c_uint32, [ctypes.c_uint32] * 3),
            "oracle_morton3_decode": (None, [ctypes.c_uint32, c_uint32_p, c_uint32_p, c_uint32_p]),
            "oracle_cell_coord": (ctypes.c_uint64, [P, ctypes.c_double, ctypes.c_double]),
            "oracle_cell_idx": (ctypes.c_uint64, [P, c_double_p]),
            "oracle_neighbouring_cells": (None, [P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
            "oracle_count_cells": (ctypes.c_uint64, [P, c_int32_p]),
            "oracle_grid_find": (ctypes.c_int32, [P, c_double_p, c_int32_p, c_uint32_p]),
            "oracle_probe": (ctypes.c_int32, [P, c_double_p, c_int32_p, c_double_p, ctypes.c_uint32, ctypes.c_uint32,
                                              c_double_p, ctypes.c_uint32, c_double_p]),
            "oracle_density_sum": (None, [P, c_double_p, c_uint32_p, c_double_p, c_double_p]),
            "oracle_density_sum_normalized": (None, [P, c_double_p, c_uint32_p, c_double_p, c_double_p,
                                                     c_double_p]),
            "oracle_density_con": (None, [P, c_double_p, c_double_p, c_uint32_p, c_double_p, c_double_p,
                                          c_double_p]),
            "oracle_term_external": (None, [P, c_double_p, c_int32_p, c_uint32_p, c_double_p]),
            "oracle_term_internal": (None, [P, c_double_p, c_uint32_p, c_double_p, c_double_p, c_double_p]),
            "oracle_term_art_visc": (None, [P, c_double_p, c_double_p, c_uint32_p, c_double_p, c_double_p,
                                            c_double_p]),
            "oracle_term_average_velocity": (None, [P, c_double_p, c_double_p, c_int32_p, c_uint32_p, c_double_p,
                                                    ctypes.c_uint32, c_double_p]),
            "oracle_update_acceleration": (None, [P, c_int32_p, c_double_p, c_double_p, c_uint32_p, c_double_p,
                                                  c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                                                  c_double_p]),
            "oracle_create": (ctypes.c_void_p, [P]),
            "oracle_destroy": (None, [ctypes.c_void_p]),
            "oracle_set_state": (ctypes.c_int32, [ctypes.c_void_p, c_double_p, c_double_p, c_double_p, c_double_p,
                                                  c_int32_p, ctypes.c_double, ctypes.c_double]),
            "oracle_step": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32]),
            "oracle_get_state": (None, [ctypes.c_void_p, c_double_p, c_double_p, c_double_p, c_double_p, c_int32_p,
                                        c_double_p, c_double_p, ctypes.POINTER(ctypes.c_uint64)]),
            "oracle_get_fields": (None, [ctypes.c_void_p, c_double_p, c_double_p, c_double_p, c_double_p]),
            "oracle_neighbours_ptr": (c_uint32_p, [ctypes.c_void_p]),
            "oracle_last_outside": (ctypes.c_uint32, [ctypes.c_void_p]),
            "oracle_piston_velocity": (ctypes.c_double, [P, ctypes.c_double]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a: Optional[np.ndarray], ctype=ctypes.c_double):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def from_params(p, num_threads: int = 0) -> OracleParams:
    """Build the oracle's struct from a :class:`sph2d_amd.params.ExperimentParams`."""
    o = OracleParams()
    for name, _t in OracleParams._fields_:
        if name in ("piston_a", "piston_b", "piston_nharm", "num_threads", "eos_c_sqr", "eos_c", "piston_segments",
                    "piston_seg_zmin", "piston_seg_zmax", "density_normalization_every"):
            continue
        val = getattr(p, name)
        setattr(o, name, type(getattr(o, name))(val))
    o.eos_c_sqr = p.eos_c_sqr
    o.eos_c = p.eos_c
    o.boundary_motion = int(p.boundary_motion)
    o.dt_dynamic_mode = int(p.dt_dynamic_mode)
    segs = max(int(getattr(p, "piston_segments", 1)), 1)
    o.piston_segments = segs
    o.piston_seg_zmin = float(getattr(p, "piston_seg_zmin", 0.0))
    o.piston_seg_zmax = float(getattr(p, "piston_seg_zmax", 0.0))
    o.piston_nharm = len(p.piston_a) // segs
    for k, (a, b) in enumerate(zip(p.piston_a, p.piston_b)):
        o.piston_a[k] = a
        o.piston_b[k] = b
    o.num_threads = num_threads
    o.density_normalization_every = max(int(getattr(p, "density_normalization_every", 1)), 1)
    return o


class Oracle:
    """Whole-loop oracle with persistent state (original particle order)."""

    STATUS = {0: "ok", 1: "Hit the maximum number of neighbours.", 2: "never ending simulation",
              3: "Encounter particle outside boundaries", 4: "cell_idx was >= params.max_cells"}

    def __init__(self, params, num_threads: int = 0):
        self.params = params
        self.op = from_params(params, num_threads)
        self.L = lib()
        self.ctx = self.L.oracle_create(ctypes.byref(self.op))
        if not self.ctx:
            raise RuntimeError("oracle_create failed")
        self.n = params.ntotal
        self.dim = params.dim

    def __del__(self):
        if getattr(self, "ctx", None):
            self.L.oracle_destroy(self.ctx)
            self.ctx = None

    def set_state(self, r, v, rho, itype, p=None, time=0.0, dt=None):
        r = np.ascontiguousarray(r, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        itype = np.ascontiguousarray(itype, dtype=np.int32)
        pp = None if p is None else np.ascontiguousarray(p, dtype=np.float64)
        self.L.oracle_set_state(self.ctx, ptr(r), ptr(v), ptr(rho), ptr(pp), ptr(itype, ctypes.c_int32),
                                float(time), float(self.params.dt if dt is None else dt))

    def step(self, n: int = 1) -> int:
        return int(self.L.oracle_step(self.ctx, n))

    def state(self):
        n, d = self.n, self.dim
        r = np.empty((n, d)); v = np.empty((n, d)); rho = np.empty(n); p = np.empty(n)
        it = np.empty(n, dtype=np.int32)
        t = ctypes.c_double(); dt = ctypes.c_double(); st = ctypes.c_uint64()
        self.L.oracle_get_state(self.ctx, ptr(r), ptr(v), ptr(rho), ptr(p), ptr(it, ctypes.c_int32),
                                ctypes.byref(t), ctypes.byref(dt), ctypes.byref(st))
        return {"r": r, "v": v, "rho": rho, "p": p, "itype": it, "time": t.value, "dt": dt.value, "step": st.value}

    def fields(self):
        n, d = self.n, self.dim
        a = np.empty((n, d)); drho = np.empty(n); av = np.empty((n, d)); mu = np.empty(n)
        self.L.oracle_get_fields(self.ctx, ptr(a), ptr(drho), ptr(av), ptr(mu))
        return {"a": a, "drho": drho, "av": av, "arvmu": mu}

    def neighbours(self) -> np.ndarray:
        m = self.params.max_neighbours
        p = self.L.oracle_neighbours_ptr(self.ctx)
        return np.ctypeslib.as_array(p, shape=(self.n * m,)).reshape(self.n, m).copy()

    def outside(self) -> int:
        return int(self.L.oracle_last_outside(self.ctx))


def grid_find(params, r, itype):
    L = lib()
    op = from_params(params)
    m = params.max_neighbours
    nb = np.zeros(params.ntotal * m, dtype=np.uint32)
    r = np.ascontiguousarray(r, dtype=np.float64)
    itype = np.ascontiguousarray(itype, dtype=np.int32)
    st = L.oracle_grid_find(ctypes.byref(op), ptr(r), ptr(itype, ctypes.c_int32), ptr(nb, ctypes.c_uint32))
    return st, nb.reshape(params.ntotal, m)


PROBE_FIELDS = {"vx": ("v", 0), "vy": ("v", 1), "vz": ("v", 2), "p": ("p", 0), "rho": ("rho", 0)}


def probe(params, state, points, field):
    """FuncAtPoint of `field` (vx, vy, vz, p, rho) of `state` (r, v, rho, p,
    itype) at `points` (npoints x dim): (status, values)."""
    L = lib()
    op = from_params(params)
    name, comp = PROBE_FIELDS[field]
    f = np.ascontiguousarray(state[name], dtype=np.float64)
    stride = f.shape[1] if f.ndim == 2 else 1
    pts = np.ascontiguousarray(points, dtype=np.float64)
    out = np.zeros(len(pts))
    st = L.oracle_probe(ctypes.byref(op), ptr(np.ascontiguousarray(state["r"], dtype=np.float64)),
                        ptr(np.ascontiguousarray(state["itype"], dtype=np.int32), ctypes.c_int32), ptr(f), stride, comp,
                        ptr(pts), len(pts), ptr(out))
    return st, out


def neighbour_sets(nb: np.ndarray, ntotal: int):
    """Neighbour list (reference layout, sentinel ntotal) -> list of sorted arrays."""
    out = []
    for row in nb:
        k = np.where(row == ntotal)[0]
        cnt = int(k[0]) if len(k) else len(row)
        out.append(np.sort(row[:cnt]))
    return out
