"""Experiment parameters: the reference's ModelParams.json / ParticleParams.json
formats and the constants the time step derives from them.

Mirrors, in behaviour and error messages:
  * ``load_model_params``    -- RRSPH_PARAMS/src/ModelParams.cpp:185-311
    (mandatory keys, defaults, optional keys) and ``check_optional_params``
    (ModelParams.cpp:12-97);
  * ``load_particle_params`` -- RRSPH_PARAMS/src/ParticleParams.cpp:51-148;
  * ``apply_*``              -- ModelParams.cpp:110-176, ParticleParams.cpp:22-46
    (optional values default to zero when absent);
  * ``fill_in_computing_params`` / ``post_fill_in_model_params`` /
    ``find_depth`` / ``count_cells`` -- RRSPH_COMMON/src/Input.cpp:17-138.

Everything here is plain host logic; the numbers it derives are passed to the
HIP library through the C-ABI struct built in :mod:`sph2d_amd.hip`.
"""
from __future__ import annotations

import dataclasses
import json
import math
import os
from typing import Any, Dict, Optional

# ParamsEnumeration.h:5-60
DENSITY_SUMMATION = 0
DENSITY_CONTINUITY = 1
DENSITY_CONTINUITY_DELTA = 2
CONSISTENCY_PRINT = 0
CONSISTENCY_STOP = 1
CONSISTENCY_FIX = 2
NWM_NO_WAVES = 0
NWM_METHOD_DYNAMIC_1 = 2
NWM_METHOD_DYNAMIC_2 = 3
NWM_METHOD_WALL_DISAPPEAR = 4
NWM_METHOD_SOLITARY_RAYLEIGH = 5
NWM_METHOD_FOURIER = 6  # extension: piston x(t) = sum a_k sin(k w t) + b_k cos(k w t)
NWM_DIRECTION_UNDEFINED = 0
NWM_DIRECTION_AXIS_X = 1
NWM_DIRECTION_AXIS_Y = 2
NWM_DIRECTION_AXIS_Z = 3
NWM_DIRECTION_RADIAL_Y = 4
SKF_CUBIC = 1
SKF_GAUSS = 2
SKF_WENDLAND = 3
SKF_DESBRUN = 4
INTF_SPH_APPROXIMATION_1 = 1
INTF_SPH_APPROXIMATION_2 = 2
EOS_SOUND_VEL_DAM_BREAK = 0
EOS_SOUND_VEL_SPECIFIC = 1
DT_CORRECTION_CONST_VALUE = 0
DT_CORRECTION_CONST_CFL = 1
DT_CORRECTION_DYNAMIC = 2
SBT_DYNAMIC = 0
SBT_REPULSIVE = 1

TYPE_BOUNDARY = -2
TYPE_NON_EXISTENT = 0
TYPE_WATER = 2

PARAMS_VERSION = (3, 1, 4)  # RRSPH_PARAMS/CMakeLists.txt:2
PI = 3.14159265358979323846
G = 9.81

# ModelParams.h:11-79: name -> (kind, default). kind: "mandatory", "default", "optional".
_MODEL_KEYS = [
    ("density_treatment", "default", DENSITY_CONTINUITY),
    ("density_normalization", "default", 0),
    ("density_skf", "default", SKF_CUBIC),
    ("density_delta_sph_coef", "optional", None),
    ("eos_sound_vel_method", "mandatory", None),
    ("eos_sound_vel", "optional", None),
    ("eos_sound_vel_coef", "optional", None),
    ("intf_hsml_coef", "default", 1.0),
    ("intf_sph_approximation", "default", INTF_SPH_APPROXIMATION_2),
    ("intf_skf", "default", SKF_CUBIC),
    ("artificial_pressure", "default", False),
    ("artificial_pressure_skf", "optional", SKF_CUBIC),
    ("artificial_pressure_index", "optional", 4.0),
    ("artificial_pressure_coef", "optional", 0.2),
    ("visc", "default", True),
    ("visc_coef", "default", 0.001),
    ("artificial_viscosity", "default", False),
    ("artificial_shear_visc", "optional", None),
    ("artificial_bulk_visc", "optional", None),
    ("artificial_viscosity_skf", "default", SKF_CUBIC),
    ("average_velocity", "default", False),
    ("average_velocity_coef", "optional", None),
    ("average_velocity_skf", "default", SKF_CUBIC),
    ("save_every_step", "default", False),
    ("save_time", "mandatory", None),
    ("dump_time", "optional", None),
    ("step_time_estimate", "optional", None),
    ("use_dump", "default", False),
    ("use_crash_dump", "default", True),
    ("use_custom_time_estimate_step", "default", False),
    ("consistency_check", "default", True),
    ("consistency_treatment", "default", CONSISTENCY_STOP),
    ("boundary_treatment", "mandatory", None),
    ("nwm", "default", NWM_NO_WAVES),
    ("nwm_time_start", "default", 0.0),
    ("nwm_wave_length", "optional", None),
    ("nwm_wave_magnitude", "optional", None),
    ("nwm_direction", "optional", None),
    ("nwm_phase", "default", 0.0),
    ("simulation_time", "mandatory", None),
    ("dt_correction_method", "default", DT_CORRECTION_DYNAMIC),
    ("dt", "optional", None),
    ("CFL_coef", "optional", None),
    ("max_neighbours", "default", 64),
    ("local_threads", "optional", None),
    ("save_velocity", "default", True),
    ("save_pressure", "default", True),
    ("save_density", "default", True),
]

# Keys of the RRSPH_HIP extension (Fourier-series piston, optionally split
# into paddles across z); they are ignored by the reference loader and
# optional here.
_EXTENSION_KEYS = ("piston_omega", "piston_a", "piston_b", "piston_segments", "piston_seg_zmin", "piston_seg_zmax",
                   "density_normalization_every")


def _version_tuple(d: Dict[str, Any]):
    return (int(d.get("params_target_version_major", 0)),
            int(d.get("params_target_version_minor", 0)),
            int(d.get("params_target_version_patch", 0)))


def _read_json(path: str) -> Dict[str, Any]:
    if not os.path.exists(path):
        raise RuntimeError(f"No params file provided: '{path}' expected")
    with open(path) as f:
        return json.load(f)


def check_optional_params(mp: Dict[str, Any]) -> None:
    """ModelParams.cpp:12-97."""

    def need(name):
        if mp.get(name) is None:
            raise RuntimeError(f"Need optional param: {name}")

    if mp["visc"]:
        raise RuntimeError("Not implemented: visc")
    # the reference's loader rejects density_normalization as not implemented;
    # this backend implements it as its documentation states it
    # (docs/latex/ModelParams.tex:124-133: summation density only)
    if mp["density_normalization"] and mp["density_treatment"] != DENSITY_SUMMATION:
        raise RuntimeError("density_normalization needs density_treatment == 0")
    if mp["eos_sound_vel_method"] == EOS_SOUND_VEL_SPECIFIC:
        need("eos_sound_vel")
    elif mp["eos_sound_vel_method"] == EOS_SOUND_VEL_DAM_BREAK:
        need("eos_sound_vel_coef")
    else:
        raise RuntimeError("Invalid enum value: eos_sound_vel_method")
    if mp["density_treatment"] == DENSITY_CONTINUITY_DELTA:
        need("density_delta_sph_coef")
    if mp["artificial_viscosity"]:
        need("artificial_shear_visc")
        need("artificial_bulk_visc")
    if mp["average_velocity"]:
        need("average_velocity_coef")
    if mp["use_dump"]:
        need("dump_time")
    if mp["use_custom_time_estimate_step"]:
        need("step_time_estimate")
    nwm = mp["nwm"]
    if nwm in (NWM_METHOD_WALL_DISAPPEAR, NWM_NO_WAVES):
        pass
    elif nwm in (NWM_METHOD_DYNAMIC_1, NWM_METHOD_DYNAMIC_2):
        need("nwm_wave_length")
        need("nwm_wave_magnitude")
        need("nwm_direction")
    elif nwm == NWM_METHOD_SOLITARY_RAYLEIGH:
        need("nwm_wave_magnitude")
    elif nwm == NWM_METHOD_FOURIER:
        need("piston_omega")
        need("piston_a")
        need("piston_b")
    else:
        raise RuntimeError("Invalid enum value: nwm")
    dcm = mp["dt_correction_method"]
    if dcm == DT_CORRECTION_CONST_VALUE:
        need("dt")
    elif dcm in (DT_CORRECTION_CONST_CFL, DT_CORRECTION_DYNAMIC):
        need("CFL_coef")
    else:
        raise RuntimeError("Invalid enum value: dt_correction_method")


def parse_model_params(js: Dict[str, Any]) -> Dict[str, Any]:
    """ModelParams.cpp:185-311 applied to an already-parsed JSON object."""
    js = dict(js)
    if _version_tuple(js) < PARAMS_VERSION:
        js["nwm_direction"] = NWM_DIRECTION_AXIS_X  # ModelParams.cpp:178-183
    mp: Dict[str, Any] = {}
    for name, kind, default in _MODEL_KEYS:
        if name in js:
            mp[name] = js[name]
        elif kind == "mandatory":
            raise RuntimeError(f"Mandatory param not specified: {name}")
        else:
            mp[name] = default
    for name in _EXTENSION_KEYS:
        mp[name] = js.get(name)
    check_optional_params(mp)
    return mp


def load_model_params(experiment_directory: str) -> Dict[str, Any]:
    return parse_model_params(_read_json(os.path.join(experiment_directory, "ModelParams.json")))


def parse_particle_params(js: Dict[str, Any]) -> Dict[str, Any]:
    """ParticleParams.cpp:51-148 applied to an already-parsed JSON object."""
    pp: Dict[str, Any] = {}

    def load(name):
        if name not in js:
            raise RuntimeError(f"Mandatory param not specified: {name}")
        pp[name] = js[name]

    pp["dim"] = int(js.get("dim", 2))
    for name in ("x_mingeom", "x_maxgeom", "y_mingeom", "y_maxgeom"):
        load(name)
    if pp["dim"] == 3:
        load("z_mingeom")
        load("z_maxgeom")
    else:
        pp["z_mingeom"] = 0.0
        pp["z_maxgeom"] = 0.0
    for name in ("delta", "ntotal", "nfluid", "nvirt"):
        load(name)
    pp["rho0"] = js.get("rho0", 1000)
    pp["nwm_particles_start"] = js.get("nwm_particles_start", pp["ntotal"])
    pp["nwm_particles_end"] = js.get("nwm_particles_end", pp["ntotal"])
    pp["depth"] = js.get("depth")
    for name in ("boundary_delta", "boundary_separation", "x_boundary_left", "x_boundary_right",
                 "y_boundary_bottom", "y_boundary_top", "use_chess_order"):
        if name in js:
            pp[name] = js[name]
    return pp


def load_particle_params(experiment_directory: str) -> Dict[str, Any]:
    return parse_particle_params(_read_json(os.path.join(experiment_directory, "ParticleParams.json")))


@dataclasses.dataclass
class ExperimentParams:
    """The global ``params`` of the reference (Params.h:8-194), restricted to
    the fields the time step and its boundary use."""

    dim: int = 2
    maxn: int = 0
    max_neighbours: int = 64
    max_cells: int = 0
    x_mingeom: float = 0.0
    x_maxgeom: float = 0.0
    y_mingeom: float = 0.0
    y_maxgeom: float = 0.0
    z_mingeom: float = 0.0
    z_maxgeom: float = 0.0
    nfluid: int = 0
    nvirt: int = 0
    ntotal: int = 0
    nwm_wave_length: float = 0.0
    depth: float = 0.0
    nwm_freq: float = 0.0
    nwm_piston_magnitude: float = 0.0
    nwm_wave_magnitude: float = 0.0
    nwm_wave_number: float = 0.0
    nwm_direction: int = NWM_DIRECTION_UNDEFINED
    nwm_phase: float = 0.0
    nwm_particles_start: int = 0
    nwm_particles_end: int = 0
    nwm_time_start: float = 0.0
    CFL_coef: float = 0.0
    dt: float = 0.0
    simulation_time: float = 0.0
    dt_correction_method: int = DT_CORRECTION_CONST_CFL
    local_threads: int = 32
    eos_sound_vel_method: int = 0
    eos_sound_vel_coef: float = 1.0
    eos_sound_vel: float = 0.0
    intf_sph_approximation: int = INTF_SPH_APPROXIMATION_2
    artificial_pressure: bool = False
    artificial_pressure_skf: int = SKF_CUBIC
    artificial_pressure_index: float = 4.0
    artificial_pressure_coef: float = 0.2
    density_skf: int = SKF_CUBIC
    intf_skf: int = SKF_DESBRUN
    artificial_viscosity_skf: int = SKF_CUBIC
    average_velocity_skf: int = SKF_CUBIC
    cell_scale_k: float = 2.0
    nwm: int = NWM_NO_WAVES
    boundary_treatment: int = SBT_REPULSIVE
    hsml: float = 0.0
    intf_hsml_coef: float = 1.0
    delta: float = 0.0
    density_treatment: int = DENSITY_CONTINUITY
    density_normalization: int = 0
    # extension: the boundary-deficiency correction on every n-th step only
    density_normalization_every: int = 1
    density_delta_sph_coef: float = 0.1
    average_velocity: bool = True
    average_velocity_coef: float = 0.3
    visc: bool = True
    visc_coef: float = 1.0e-3
    artificial_viscosity: bool = True
    artificial_shear_visc: float = 1.0
    artificial_bulk_visc: float = 0.0
    rho0: float = 1000.0
    mass: float = 0.0
    consistency_treatment: int = CONSISTENCY_STOP
    consistency_check: bool = True
    start_simulation_time: float = 0.0
    save_time: float = 0.0
    save_every_step: bool = False
    save_velocity: bool = True
    save_pressure: bool = True
    save_density: bool = True
    use_crash_dump: bool = True
    use_dump: bool = False
    dump_time: float = 0.0
    # extension: Fourier-series piston; with piston_segments > 1 (3D) the
    # coefficients hold every paddle's harmonics, paddle-major, and the
    # paddles split [piston_seg_zmin, piston_seg_zmax) into equal parts
    piston_omega: float = 0.0
    piston_a: tuple = ()
    piston_b: tuple = ()
    piston_segments: int = 1
    piston_seg_zmin: float = 0.0
    piston_seg_zmax: float = 0.0
    # switches for RRSPH_CL-only behaviour (described in the oracle header)
    boundary_motion: bool = False
    dt_dynamic_mode: int = 0
    pi: float = PI
    g: float = G

    # ---- derived helpers ---------------------------------------------------
    @property
    def cell_size(self) -> float:
        return self.cell_scale_k * self.hsml

    @property
    def eos_c_sqr(self) -> float:
        """RRSPH_COMMON/src/EOS.cpp:3-20: method 1 uses eos_sound_vel itself as c^2."""
        if self.eos_sound_vel_method == EOS_SOUND_VEL_DAM_BREAK:
            return 200 * self.g * self.depth * self.eos_sound_vel_coef
        return self.eos_sound_vel

    @property
    def eos_c(self) -> float:
        return math.sqrt(self.eos_c_sqr)


def apply_model_params(params: ExperimentParams, mp: Dict[str, Any]) -> None:
    """ModelParams.cpp:110-176: optional values that are absent become zero."""
    for name, _kind, _default in _MODEL_KEYS:
        if not hasattr(params, name):
            continue
        val = mp[name]
        if val is None:
            val = 0
        setattr(params, name, type(getattr(params, name))(val))
    if mp.get("piston_omega") is not None:
        params.piston_omega = float(mp["piston_omega"])
        params.piston_a = tuple(float(x) for x in mp["piston_a"])
        params.piston_b = tuple(float(x) for x in mp["piston_b"])
    if mp.get("density_normalization_every") is not None:
        params.density_normalization_every = int(mp["density_normalization_every"])
    if mp.get("piston_segments") is not None:
        params.piston_segments = int(mp["piston_segments"])
        params.piston_seg_zmin = float(mp["piston_seg_zmin"])
        params.piston_seg_zmax = float(mp["piston_seg_zmax"])


def apply_particle_params(params: ExperimentParams, pp: Dict[str, Any]) -> None:
    """ParticleParams.cpp:22-46."""
    for name in ("dim", "x_mingeom", "x_maxgeom", "y_mingeom", "y_maxgeom", "z_mingeom", "z_maxgeom",
                 "delta", "ntotal", "nfluid", "nvirt", "rho0", "nwm_particles_start", "nwm_particles_end"):
        setattr(params, name, type(getattr(params, name))(pp[name]))
    params.depth = float(pp["depth"]) if pp.get("depth") is not None else 0.0


def intlog2(size: int) -> int:
    """CommonIncl.h:69-77."""
    passes = 0
    while size > 1:
        size >>= 1
        passes += 1
    return passes


def cell_scale_k_for(skfs) -> float:
    """Input.cpp:17-43: all kernels must share one cell scale."""
    table = {SKF_CUBIC: 2, SKF_GAUSS: 3, SKF_WENDLAND: 2, SKF_DESBRUN: 2}
    ks = []
    for skf in skfs:
        if skf not in table:
            raise RuntimeError(f"get_cell_scale_k: unknown skf {skf}")
        ks.append(table[skf])
    if min(ks) != max(ks):
        raise RuntimeError("selected incompatible skf")
    return float(ks[0])


def _split2(x: int) -> int:
    x &= 0xFFFF
    x = (x | (x << 8)) & 0x00FF00FF
    x = (x | (x << 4)) & 0x0F0F0F0F
    x = (x | (x << 2)) & 0x33333333
    x = (x | (x << 1)) & 0x55555555
    return x


def _split3(x: int) -> int:
    x &= 0x3FF
    x = (x | (x << 16)) & 0x030000FF
    x = (x | (x << 8)) & 0x0300F00F
    x = (x | (x << 4)) & 0x030C30C3
    x = (x | (x << 2)) & 0x09249249
    return x


def cell_coord(params: ExperimentParams, r: float, rmin: float) -> int:
    """GridUtils.h:135-137 (truncating conversion)."""
    return int((r - rmin) / params.cell_size)


def count_cells(params: ExperimentParams) -> int:
    """Input.cpp:45-75: size of the reference's dense Morton cell table."""
    x_id = cell_coord(params, params.x_maxgeom, params.x_mingeom)
    y_id = cell_coord(params, params.y_maxgeom, params.y_mingeom)
    if params.dim == 3:
        z_id = cell_coord(params, params.z_maxgeom, params.z_mingeom)
        max_id = 1 << 10
        key = _split3(x_id) | (_split3(y_id) << 1) | (_split3(z_id) << 2)
    else:
        z_id = 0
        max_id = 1 << 16
        key = _split2(x_id) | (_split2(y_id) << 1)
    if x_id > max_id or y_id > max_id or z_id >= max_id:
        raise RuntimeError(f"can't make grid with so many cells: ({x_id}/{max_id}; {y_id}/{max_id}; {z_id}/{max_id})")
    return 1 << (intlog2(key) + 1)


def fill_in_computing_params(params: ExperimentParams) -> None:
    """Input.cpp:93-111."""
    params.hsml = params.delta * params.intf_hsml_coef
    params.maxn = 1 << (1 + intlog2(params.ntotal))
    # the reference sizes the cell table before it sets cell_scale_k, so the
    # count uses the default scale of 2 (Input.cpp:100-110)
    params.max_cells = count_cells(params)
    m = 1.0  # powun(delta, dim): repeated multiplication
    for _ in range(params.dim):
        m *= params.delta
    params.mass = params.rho0 * m
    params.cell_scale_k = cell_scale_k_for([params.artificial_pressure_skf, params.artificial_viscosity_skf,
                                            params.average_velocity_skf, params.intf_skf, params.density_skf])


def find_depth(params: ExperimentParams, r) -> float:
    """Input.cpp:77-91: explicit depth, else the y-extent of the fluid."""
    if params.depth != 0:
        return params.depth
    import numpy as np

    y = np.asarray(r)[: params.nfluid, 1]
    return float(y.max() - y.min())


def post_fill_in_model_params(params: ExperimentParams) -> None:
    """Input.cpp:112-138."""
    if params.eos_sound_vel_method == EOS_SOUND_VEL_DAM_BREAK:
        params.eos_sound_vel = math.sqrt(200 * params.g * params.depth * params.eos_sound_vel_coef)
    if params.dt_correction_method == DT_CORRECTION_CONST_CFL:
        params.dt = params.CFL_coef * params.hsml / (params.eos_sound_vel * (1 + 1.2 * params.artificial_shear_visc))
    if params.nwm in (NWM_METHOD_DYNAMIC_1, NWM_METHOD_DYNAMIC_2):
        params.nwm_wave_number = 2.0 * params.pi / params.nwm_wave_length
        kd = params.nwm_wave_number * params.depth
        params.nwm_freq = math.sqrt(params.nwm_wave_number * params.g * math.tanh(kd))
        params.nwm_piston_magnitude = (params.nwm_wave_magnitude * 0.5 / math.sinh(kd) ** 2 *
                                       (math.sinh(kd) * math.cosh(kd) + kd))


def make_params(model: Dict[str, Any], particle: Dict[str, Any], start_time: float = 0.0) -> ExperimentParams:
    """apply_* + fillInComputingParams (Input.cpp:250-259); depth and the
    post-fill constants need the particle positions (see ``finalize``)."""
    p = ExperimentParams()
    apply_particle_params(p, particle)
    apply_model_params(p, model)
    p.start_simulation_time = start_time
    fill_in_computing_params(p)
    return p


def finalize(params: ExperimentParams, r) -> ExperimentParams:
    """find_depth + postFillInModelParams (Input.cpp:277-321)."""
    params.depth = find_depth(params, r)
    post_fill_in_model_params(params)
    return params


def start_time_from_dump_name(path: str) -> float:
    """Input.cpp:258: the dump file stem is the start time."""
    stem = os.path.splitext(os.path.basename(path))[0]
    return float(stem)


def params_to_json(params: ExperimentParams) -> Dict[str, Any]:
    return dataclasses.asdict(params)


# ---------------------------------------------------------------- writers
# The parameter files the reference writes into the experiment directory at
# the start of a run (Input.cpp:323-325).  nlohmann::json keeps object keys
# sorted and dump(4) indents by four spaces with ": " after each key, which
# json.dumps(indent=4, sort_keys=True) reproduces; both print a double as its
# shortest round-trip decimal.  A file ends with a newline (std::endl).
RRSPH_VERSION = (3, 1, 4)         # CMakeLists.txt:3
RRSPH_COMMON_VERSION = (1, 3, 4)  # RRSPH_COMMON/CMakeLists.txt:2

# ParticleParams.cpp, params_make_particles_json: every field but depth is
# printed, the ones a layout did not set as the struct's zero
# (ParticleParams.h:13-60)
_PARTICLE_PRINT = ("dim", "x_maxgeom", "x_mingeom", "y_maxgeom", "y_mingeom", "z_mingeom", "z_maxgeom", "delta",
                   "ntotal", "nfluid", "nvirt", "x_fluid_particles", "y_fluid_particles", "z_fluid_particles",
                   "x_fluid_min", "x_fluid_max", "y_fluid_min", "y_fluid_max", "z_fluid_min", "z_fluid_max",
                   "x_boundary_left", "x_boundary_right", "x_boundary_center", "y_boundary_bottom",
                   "y_boundary_top", "z_boundary_near", "z_boundary_far", "rho0", "nwm_particles_start",
                   "nwm_particles_end")
_PARTICLE_INTS = ("dim", "ntotal", "nfluid", "nvirt", "x_fluid_particles", "y_fluid_particles", "z_fluid_particles",
                  "nwm_particles_start", "nwm_particles_end")


def _versions(d: Dict[str, Any]) -> Dict[str, int]:
    v = PARAMS_VERSION
    return {"params_target_version_major": int(d.get("params_target_version_major", v[0])),
            "params_target_version_minor": int(d.get("params_target_version_minor", v[1])),
            "params_target_version_patch": int(d.get("params_target_version_patch", v[2]))}


def model_params_json(mp: Dict[str, Any], versions: Optional[Dict[str, Any]] = None) -> Dict[str, Any]:
    """ModelParams.cpp:313-399 (params_make_model_json): the parsed model
    parameters, optional ones only when set; the RRSPH_HIP Fourier piston
    keys when set."""
    out = _versions(versions or {})
    for name, kind, _default in _MODEL_KEYS:
        value = mp.get(name)
        if kind == "optional" and value is None:
            continue
        out[name] = value
    for name in _EXTENSION_KEYS:
        if mp.get(name) is not None:
            out[name] = list(mp[name]) if isinstance(mp[name], (list, tuple)) else mp[name]
    return out


def particle_params_json(pp: Dict[str, Any], versions: Optional[Dict[str, Any]] = None) -> Dict[str, Any]:
    """ParticleParams.cpp (params_make_particles_json)."""
    out = _versions(versions or {})
    for name in _PARTICLE_PRINT:
        value = pp.get(name)
        out[name] = (0 if name in _PARTICLE_INTS else 0.0) if value is None else value
    if pp.get("depth") is not None:
        out["depth"] = pp["depth"]
    return out


def computing_params_json(params: ExperimentParams, specific_version=(0, 0, 0, "RRSPH_HIP")) -> Dict[str, Any]:
    """Input.cpp:140-175 (make_ComputingParams) + ComputingParams.cpp:108-157
    (params_make_computing_json): derived constants, the wave numbers of the
    first-order piston when set, and the version stamps."""
    out: Dict[str, Any] = {
        "start_simulation_time": params.start_simulation_time, "pi": params.pi, "g": params.g,
        "TYPE_BOUNDARY": TYPE_BOUNDARY, "TYPE_NON_EXISTENT": TYPE_NON_EXISTENT, "TYPE_WATER": TYPE_WATER,
        "cell_scale_k": params.cell_scale_k, "max_cells": params.max_cells, "maxn": params.maxn,
        "hsml": params.hsml, "mass": params.mass,
    }
    for name in ("nwm_wave_number", "nwm_freq", "nwm_piston_magnitude"):
        if getattr(params, name) != 0:
            out[name] = getattr(params, name)
    for key, v in (("params_version", PARAMS_VERSION), ("RRSPH_common_version", RRSPH_COMMON_VERSION),
                   ("RRSPH_version", RRSPH_VERSION), ("RRSPH_specific_version", specific_version)):
        out[key + "_major"], out[key + "_minor"], out[key + "_patch"] = int(v[0]), int(v[1]), int(v[2])
    out["RRSPH_specific_version_name"] = str(specific_version[3])
    return out


def dump_params_json(obj: Dict[str, Any]) -> str:
    """The text nlohmann's json.dump(4) followed by std::endl produces."""
    return json.dumps(obj, indent=4, sort_keys=True) + "\n"


def write_experiment_params(experiment_directory: str, mp: Dict[str, Any], pp: Dict[str, Any],
                            params: ExperimentParams, specific_version=(0, 0, 0, "RRSPH_HIP"),
                            versions: Optional[Dict[str, Any]] = None) -> None:
    """Input.cpp:323-325: ComputingParams.json, ModelParams.json and
    ParticleParams.json of a run, loadable by load_model_params /
    load_particle_params here and by the reference's loaders."""
    os.makedirs(experiment_directory, exist_ok=True)
    files = (("ComputingParams.json", computing_params_json(params, specific_version)),
             ("ModelParams.json", model_params_json(mp, versions)),
             ("ParticleParams.json", particle_params_json(pp, versions)))
    for name, obj in files:
        with open(os.path.join(experiment_directory, name), "w") as f:
            f.write(dump_params_json(obj))
