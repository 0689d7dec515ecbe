"""ctypes binding of the RRSPH_HIP C-ABI (include/rrsph_hip.h).

This is the Python side of the drop-in boundary: it builds the flat
``rrsph_params_c`` struct from :class:`sph2d_amd.params.ExperimentParams`
and wraps the context in :class:`HipStep`.  There is no CPU fallback: if
``librrsph_hip.so`` is missing or no HIP device is present, construction
raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional, Sequence

import numpy as np

from . import params as P

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librrsph_hip.so")
MAX_HARMONICS = 32

FP64 = 0
FP32 = 1

FIELD_R, FIELD_V, FIELD_RHO, FIELD_P, FIELD_ITYPE = 1, 2, 4, 8, 16

STATUS_NAMES = {
    0: "RRSPH_OK", 1: "RRSPH_ERR_INVALID_ARGUMENT", 2: "RRSPH_ERR_HIP", 3: "RRSPH_ERR_TOO_MANY_NEIGHBOURS",
    4: "RRSPH_ERR_NEVER_ENDING", 5: "RRSPH_ERR_OUTSIDE_BOUNDARIES", 6: "RRSPH_ERR_TOO_MANY_CELLS",
    7: "RRSPH_ERR_NOT_IMPLEMENTED", 8: "RRSPH_ERR_NO_DEVICE",
}

# exported symbols of the C-ABI (all of include/rrsph_hip.h)
EXPORTED = [
    "rrsph_hip_create", "rrsph_hip_destroy", "rrsph_hip_last_error", "rrsph_hip_upload", "rrsph_hip_step",
    "rrsph_hip_run_until", "rrsph_hip_download", "rrsph_hip_get_time", "rrsph_hip_set_export_fields",
    "rrsph_hip_fields", "rrsph_hip_cell_keys", "rrsph_hip_neighbours", "rrsph_hip_diagnostics", "rrsph_hip_gauges",
    "rrsph_hip_stream", "rrsph_hip_set_graph", "rrsph_hip_profile_steps", "rrsph_hip_kernel_class_count",
    "rrsph_hip_kernel_class_name", "rrsph_hip_time_steps", "rrsph_hip_sort_stats", "rrsph_hip_list_stats",
    "rrsph_hip_smooth_gauges", "rrsph_hip_smooth_gauges_noalloc", "rrsph_hip_set_piston", "rrsph_hip_adjoint_forward", "rrsph_hip_adjoint_backward",
    "rrsph_hip_probe",
    "RRSPH_HIP_GetSpecificVersionMajor",
    "RRSPH_HIP_GetSpecificVersionMinor", "RRSPH_HIP_GetSpecificVersionPatch", "RRSPH_HIP_GetSpecificVersionName",
]


class RrsphParams(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32), ("dim", ctypes.c_int32), ("ntotal", ctypes.c_uint32),
        ("nfluid", ctypes.c_uint32), ("max_neighbours", ctypes.c_uint32),
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
        ("boundary_motion", ctypes.c_int32), ("dt_dynamic_mode", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("start_time", ctypes.c_double),
        ("piston_segments", ctypes.c_int32), ("piston_seg_zmin", ctypes.c_double),
        ("piston_seg_zmax", ctypes.c_double),
        ("density_normalization", ctypes.c_int32), ("density_normalization_every", ctypes.c_uint32),
    ]


class Diagnostics(ctypes.Structure):
    _fields_ = [
        ("time", ctypes.c_double), ("dt", ctypes.c_double), ("step", ctypes.c_uint64),
        ("n_fluid_alive", ctypes.c_uint32), ("n_outside", ctypes.c_uint32),
        ("mass_fluid", ctypes.c_double), ("kinetic_energy", ctypes.c_double),
        ("potential_energy", ctypes.c_double), ("internal_energy", ctypes.c_double),
        ("max_speed", ctypes.c_double), ("rho_min", ctypes.c_double), ("rho_max", ctypes.c_double),
    ]


MAX_GAUGES = 64


class GaugeSpec(ctypes.Structure):
    """rrsph_gauge_spec: smooth gauge surrogate (window radius, softness tau);
    gauges at x (2D) or at (x, z) (3D)."""
    _fields_ = [("n", ctypes.c_uint32), ("ptype", ctypes.c_int32), ("radius", ctypes.c_double),
                ("tau", ctypes.c_double), ("y_ref", ctypes.c_double), ("x", ctypes.c_double * MAX_GAUGES),
                ("z", ctypes.c_double * MAX_GAUGES)]


def gauge_spec(xs: Sequence[float], radius: float, tau: float, y_ref: float, ptype: int = 2,
               zs: Optional[Sequence[float]] = None) -> GaugeSpec:
    if len(xs) > MAX_GAUGES:
        raise ValueError(f"at most {MAX_GAUGES} gauges")
    if zs is not None and len(zs) != len(xs):
        raise ValueError("xs and zs need the same length")
    g = GaugeSpec()
    g.n = len(xs)
    g.ptype = int(ptype)
    g.radius, g.tau, g.y_ref = float(radius), float(tau), float(y_ref)
    for i, x in enumerate(xs):
        g.x[i] = float(x)
        g.z[i] = float(zs[i]) if zs is not None else 0.0
    return g


class RrsphError(RuntimeError):
    def __init__(self, status: int, message: str):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")


_lib: Optional[ctypes.CDLL] = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the C-ABI library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"RRSPH_HIP library not built: {path} (run python -m sph2d_amd.build)")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    up = ctypes.POINTER(ctypes.c_uint32)
    st = ctypes.c_int32
    sig = {
        "rrsph_hip_create": (st, [ctypes.POINTER(RrsphParams), ctypes.POINTER(vp)]),
        "rrsph_hip_destroy": (None, [vp]),
        "rrsph_hip_last_error": (ctypes.c_char_p, [vp]),
        "rrsph_hip_upload": (st, [vp, dp, dp, dp, ip, ctypes.c_uint32]),
        "rrsph_hip_step": (st, [vp, ctypes.c_uint32]),
        "rrsph_hip_run_until": (st, [vp, ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
        "rrsph_hip_download": (st, [vp, ctypes.c_uint32, dp, dp, dp, dp, ip]),
        "rrsph_hip_get_time": (st, [vp, dp, dp, ctypes.POINTER(ctypes.c_uint64)]),
        "rrsph_hip_set_export_fields": (st, [vp, ctypes.c_int32]),
        "rrsph_hip_fields": (st, [vp, dp, dp, dp, dp]),
        "rrsph_hip_cell_keys": (st, [vp, up]),
        "rrsph_hip_neighbours": (st, [vp, up, ctypes.c_uint32]),
        "rrsph_hip_diagnostics": (st, [vp, ctypes.POINTER(Diagnostics)]),
        "rrsph_hip_gauges": (st, [vp, dp, ctypes.c_uint32, ctypes.c_double, ctypes.c_int32, dp]),
        "rrsph_hip_probe": (st, [vp, dp, ctypes.c_uint32, ctypes.c_uint32, dp]),
        "rrsph_hip_stream": (vp, [vp]),
        "rrsph_hip_set_graph": (st, [vp, ctypes.c_int32]),
        "rrsph_hip_profile_steps": (st, [vp, ctypes.c_uint32, dp, ctypes.c_uint32, dp]),
        "rrsph_hip_kernel_class_count": (ctypes.c_uint32, []),
        "rrsph_hip_kernel_class_name": (ctypes.c_char_p, [ctypes.c_uint32]),
        "rrsph_hip_time_steps": (st, [vp, ctypes.c_uint32, dp]),
        "rrsph_hip_sort_stats": (st, [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_uint64), up]),
        "rrsph_hip_list_stats": (st, [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_uint64), up]),
        "rrsph_hip_smooth_gauges": (st, [vp, ctypes.POINTER(GaugeSpec), dp]),
        "rrsph_hip_smooth_gauges_noalloc": (st, [vp, ctypes.POINTER(GaugeSpec), dp]),
        "rrsph_hip_set_piston": (st, [vp, ctypes.c_double, ctypes.c_uint32, dp, dp]),
        "rrsph_hip_adjoint_forward": (st, [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(GaugeSpec), dp]),
        "rrsph_hip_adjoint_backward": (st, [vp, dp, dp, dp, dp, dp, dp]),
        "RRSPH_HIP_GetSpecificVersionMajor": (ctypes.c_uint32, []),
        "RRSPH_HIP_GetSpecificVersionMinor": (ctypes.c_uint32, []),
        "RRSPH_HIP_GetSpecificVersionPatch": (ctypes.c_uint32, []),
        "RRSPH_HIP_GetSpecificVersionName": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def make_c_params(prm: "P.ExperimentParams", precision: int = FP64) -> RrsphParams:
    c = RrsphParams()
    c.struct_size = ctypes.sizeof(RrsphParams)
    skip = {"struct_size", "eos_c_sqr", "eos_c", "piston_a", "piston_b", "piston_nharm", "precision", "start_time",
            "piston_segments", "piston_seg_zmin", "piston_seg_zmax", "density_normalization_every"}
    for name, _t in RrsphParams._fields_:
        if name in skip:
            continue
        setattr(c, name, type(getattr(c, name))(getattr(prm, name)))
    c.eos_c_sqr = prm.eos_c_sqr
    c.eos_c = prm.eos_c
    c.boundary_motion = int(prm.boundary_motion)
    c.dt_dynamic_mode = int(prm.dt_dynamic_mode)
    segs = max(int(getattr(prm, "piston_segments", 1)), 1)
    if len(prm.piston_a) > MAX_HARMONICS or len(prm.piston_a) % segs:
        raise ValueError("at most 32 piston coefficients per array, a multiple of piston_segments")
    c.piston_nharm = len(prm.piston_a) // segs
    c.piston_segments = segs
    c.piston_seg_zmin = float(getattr(prm, "piston_seg_zmin", 0.0))
    c.piston_seg_zmax = float(getattr(prm, "piston_seg_zmax", 0.0))
    for k, (a, b) in enumerate(zip(prm.piston_a, prm.piston_b)):
        c.piston_a[k] = a
        c.piston_b[k] = b
    c.precision = int(precision)
    c.start_time = prm.start_simulation_time
    c.density_normalization_every = max(int(getattr(prm, "density_normalization_every", 1)), 1)
    return c


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class HipStep:
    """One RRSPH_HIP context: the device-resident WCSPH time step."""

    def __init__(self, prm: "P.ExperimentParams", precision: int = FP64, graph: bool = True):
        self.L = load_library()
        self.prm = prm
        self.n = prm.ntotal
        self.dim = prm.dim
        self.cparams = make_c_params(prm, precision)
        self.ctx = ctypes.c_void_p()
        st = self.L.rrsph_hip_create(ctypes.byref(self.cparams), ctypes.byref(self.ctx))
        if st != 0:
            msg = self.L.rrsph_hip_last_error(self.ctx).decode() if self.ctx else ""
            if self.ctx:
                self.L.rrsph_hip_destroy(self.ctx)
                self.ctx = ctypes.c_void_p()
            raise RrsphError(st, msg)
        if not graph:
            self.L.rrsph_hip_set_graph(self.ctx, 0)
        # counts the tapes adjoint_forward has recorded: a backward pass
        # belongs to the tape of the forward it follows
        self.tape_generation = 0

    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            self.L.rrsph_hip_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int):
        if st != 0:
            raise RrsphError(st, self.L.rrsph_hip_last_error(self.ctx).decode())

    def upload(self, r, v, rho, itype):
        r = np.ascontiguousarray(r, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        itype = np.ascontiguousarray(itype, dtype=np.int32)
        self._check(self.L.rrsph_hip_upload(self.ctx, _dp(r), _dp(v), _dp(rho),
                                            itype.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(rho)))

    def step(self, n: int = 1):
        self._check(self.L.rrsph_hip_step(self.ctx, int(n)))

    def run_until(self, t_end: float, max_steps: int = 1 << 62) -> int:
        taken = ctypes.c_uint64()
        self._check(self.L.rrsph_hip_run_until(self.ctx, float(t_end), int(max_steps), ctypes.byref(taken)))
        return taken.value

    def download(self) -> Dict[str, np.ndarray]:
        n, d = self.n, self.dim
        r = np.empty((n, d)); v = np.empty((n, d)); rho = np.empty(n); p = np.empty(n)
        it = np.empty(n, dtype=np.int32)
        self._check(self.L.rrsph_hip_download(self.ctx, 31, _dp(r), _dp(v), _dp(rho), _dp(p),
                                              it.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return {"r": r, "v": v, "rho": rho, "p": p, "itype": it}

    def time(self):
        t = ctypes.c_double(); dt = ctypes.c_double(); st = ctypes.c_uint64()
        self._check(self.L.rrsph_hip_get_time(self.ctx, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(st)))
        return t.value, dt.value, st.value

    def set_export_fields(self, on: bool = True):
        self._check(self.L.rrsph_hip_set_export_fields(self.ctx, int(on)))

    def fields(self) -> Dict[str, np.ndarray]:
        n, d = self.n, self.dim
        a = np.empty((n, d)); drho = np.empty(n); av = np.empty((n, d)); mu = np.empty(n)
        self._check(self.L.rrsph_hip_fields(self.ctx, _dp(a), _dp(drho), _dp(av), _dp(mu)))
        return {"a": a, "drho": drho, "av": av, "arvmu": mu}

    def cell_keys(self) -> np.ndarray:
        k = np.empty(self.n, dtype=np.uint32)
        self._check(self.L.rrsph_hip_cell_keys(self.ctx, k.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        return k

    def neighbours(self, max_nbr: Optional[int] = None) -> np.ndarray:
        m = int(max_nbr or self.prm.max_neighbours)
        out = np.empty(self.n * m, dtype=np.uint32)
        self._check(self.L.rrsph_hip_neighbours(self.ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), m))
        return out.reshape(self.n, m)

    def diagnostics(self) -> Dict[str, float]:
        d = Diagnostics()
        self._check(self.L.rrsph_hip_diagnostics(self.ctx, ctypes.byref(d)))
        return {name: getattr(d, name) for name, _ in Diagnostics._fields_}

    def gauges(self, xs: Sequence[float], search_n: float = 2.0, ptype: int = P.TYPE_WATER) -> np.ndarray:
        x = np.ascontiguousarray(xs, dtype=np.float64)
        out = np.zeros(len(x))
        self._check(self.L.rrsph_hip_gauges(self.ctx, _dp(x), len(x), float(search_n), int(ptype), _dp(out)))
        return out

    PROBE_FIELDS = {"vx": 0, "vy": 1, "vz": 2, "p": 3, "rho": 4}

    def probe(self, points, field: str) -> np.ndarray:
        """FuncAtPoint (tools/FuncAtPoint/lib/src/fatp.cpp:14-58) of `field`
        (vx, vy, vz, p, rho) at `points` (npoints x dim) on the current state."""
        pts = np.ascontiguousarray(points, dtype=np.float64)
        out = np.zeros(len(pts))
        code = self.PROBE_FIELDS.get(field, 99)
        self._check(self.L.rrsph_hip_probe(self.ctx, _dp(pts), len(pts), code, _dp(out)))
        return out

    def time_steps(self, n: int) -> float:
        ms = ctypes.c_double()
        self._check(self.L.rrsph_hip_time_steps(self.ctx, int(n), ctypes.byref(ms)))
        return ms.value

    def smooth_gauges(self, spec: GaugeSpec) -> np.ndarray:
        out = np.zeros(spec.n)
        self._check(self.L.rrsph_hip_smooth_gauges(self.ctx, ctypes.byref(spec), _dp(out)))
        return out

    def smooth_gauges_noalloc(self, spec: GaugeSpec) -> np.ndarray:
        """smooth_gauges through the entry point that allocates nothing."""
        out = np.zeros(spec.n)
        self._check(self.L.rrsph_hip_smooth_gauges_noalloc(self.ctx, ctypes.byref(spec), _dp(out)))
        return out

    def set_piston(self, omega: float, a: Sequence[float], b: Sequence[float]):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        if a.shape != b.shape:
            raise ValueError("a and b need the same length")
        segs = self.cparams.piston_segments
        if len(a) % segs:
            raise ValueError("the coefficients of every paddle: a multiple of piston_segments")
        self._check(self.L.rrsph_hip_set_piston(self.ctx, float(omega), len(a) // segs, _dp(a), _dp(b)))

    def adjoint_forward(self, nsteps: int, every: int, spec: GaugeSpec) -> np.ndarray:
        """Differentiable run from the uploaded state; gauges after each step,
        shape (nsteps, ngauges)."""
        eta = np.zeros((nsteps, spec.n))
        self.tape_generation += 1
        self._check(self.L.rrsph_hip_adjoint_forward(self.ctx, int(nsteps), int(every), ctypes.byref(spec),
                                                     _dp(eta)))
        return eta

    def adjoint_backward(self, eta_bar: np.ndarray, ncoef: int) -> Dict[str, np.ndarray]:
        """Gradients of the loss whose derivative with respect to the gauges
        is eta_bar: the ncoef piston coefficients of a and of b (every
        paddle's, paddle-major) and the initial state."""
        eta_bar = np.ascontiguousarray(eta_bar, dtype=np.float64)
        ga, gb = np.zeros(MAX_HARMONICS), np.zeros(MAX_HARMONICS)
        d = self.dim
        gr, gv, grho = np.zeros((self.n, d)), np.zeros((self.n, d)), np.zeros(self.n)
        self._check(self.L.rrsph_hip_adjoint_backward(self.ctx, _dp(eta_bar), _dp(ga), _dp(gb), _dp(gr), _dp(gv),
                                                      _dp(grho)))
        return {"a": ga[:ncoef], "b": gb[:ncoef], "r": gr, "v": gv, "rho": grho}

    def sort_stats(self) -> Dict[str, int]:
        """Steps sorted incrementally / fully, re-sort fallbacks, particles
        that changed cell in the last step."""
        inc, full, fb = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        last = ctypes.c_uint32()
        self._check(self.L.rrsph_hip_sort_stats(self.ctx, ctypes.byref(inc), ctypes.byref(full), ctypes.byref(fb),
                                                ctypes.byref(last)))
        return {"incremental": inc.value, "full": full.value, "fallbacks": fb.value, "last_changed": last.value}

    def list_stats(self) -> Dict[str, int]:
        """Steps that read the neighbour lists, build steps, halts on
        possibly stale lists, and the current build interval (0: no lists)."""
        ls, b, h = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        iv = ctypes.c_uint32()
        self._check(self.L.rrsph_hip_list_stats(self.ctx, ctypes.byref(ls), ctypes.byref(b), ctypes.byref(h),
                                                ctypes.byref(iv)))
        return {"list_steps": ls.value, "builds": b.value, "stale_halts": h.value, "interval": iv.value}

    def profile_steps(self, n: int):
        k = self.L.rrsph_hip_kernel_class_count()
        avg = np.zeros(k)
        tot = ctypes.c_double()
        self._check(self.L.rrsph_hip_profile_steps(self.ctx, int(n), _dp(avg), k, ctypes.byref(tot)))
        names = [self.L.rrsph_hip_kernel_class_name(i).decode() for i in range(k)]
        return dict(zip(names, avg.tolist())), tot.value
