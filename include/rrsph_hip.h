/*
 * RRSPH_HIP -- C-ABI of the MI355X-native WCSPH time step.
 *
 * The entry points replace, one for one, the step-module API of the
 * RackotRR/SPH2D backends; each declaration names the reference interface it
 * stands in for.  All types are fixed width; no C++ or torch types cross
 * this boundary.  Arrays passed in or out are in ORIGINAL particle order and
 * use the reference's interleaved layout (r[dim*i + k]).  One context per
 * host thread; a context owns one HIP stream (default priority).
 *
 * Status codes map one to one to the reference's exceptions; the message of
 * rrsph_hip_last_error() is the reference's text.
 */
#ifndef RRSPH_HIP_H
#define RRSPH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RRSPH_HIP_VERSION_MAJOR 1
#define RRSPH_HIP_VERSION_MINOR 0
#define RRSPH_HIP_VERSION_PATCH 0
#define RRSPH_MAX_HARMONICS 32

typedef enum rrsph_status {
    RRSPH_OK = 0,
    RRSPH_ERR_INVALID_ARGUMENT = 1,   /* bad parameter / size (std::runtime_error from Input.cpp) */
    RRSPH_ERR_HIP = 2,                /* HIP runtime failure */
    RRSPH_ERR_TOO_MANY_NEIGHBOURS = 3,/* "making neighbours grid: error occured" GridFind.cpp:186-188 */
    RRSPH_ERR_NEVER_ENDING = 4,       /* "never ending simulation" UpdateAcceleration.cpp:26-28 */
    RRSPH_ERR_OUTSIDE_BOUNDARIES = 5, /* "Encounter particle outside boundaries" ConsistencyCheck.cpp:184-195 */
    RRSPH_ERR_TOO_MANY_CELLS = 6,     /* "can't make grid with so many cells" Input.cpp:64-70 */
    RRSPH_ERR_NOT_IMPLEMENTED = 7,
    RRSPH_ERR_NO_DEVICE = 8
} rrsph_status;

typedef enum rrsph_precision {
    RRSPH_FP64 = 0,       /* all state and pair arithmetic in FP64 */
    RRSPH_FP32 = 1        /* FP32 pair arithmetic and state; FP64 positions, dt, time, reductions */
} rrsph_precision;

/* Everything the step needs, flattened.  Field names follow the reference's
 * global `params` (RRSPH_PARAMS/include/Params.h:8-194); the derived values
 * (hsml, mass, cell_scale_k, eos constants, dt, wave-maker constants) are the
 * ones Input.cpp:93-138 computes. */
typedef struct rrsph_params_c {
    uint32_t struct_size;          /* = sizeof(rrsph_params_c) */
    int32_t dim;                   /* 2 or 3 */
    uint32_t ntotal;
    uint32_t nfluid;
    uint32_t max_neighbours;

    double x_mingeom, x_maxgeom;
    double y_mingeom, y_maxgeom;
    double z_mingeom, z_maxgeom;

    double hsml, delta, mass, rho0, cell_scale_k, depth, g, pi;
    double eos_c_sqr;              /* c^2 of the EOS */
    double eos_c;                  /* c of the dt rule and delta-SPH */
    double eos_sound_vel;          /* c of artificial viscosity */

    int32_t density_treatment;     /* 0 summation, 1 continuity, 2 continuity + delta-SPH */
    int32_t density_skf;
    double density_delta_sph_coef;

    int32_t intf_sph_approximation;
    int32_t intf_skf;

    int32_t artificial_pressure;
    int32_t artificial_pressure_skf;
    double artificial_pressure_index;
    double artificial_pressure_coef;

    int32_t artificial_viscosity;
    int32_t artificial_viscosity_skf;
    double artificial_shear_visc;
    double artificial_bulk_visc;

    int32_t average_velocity;
    int32_t average_velocity_skf;
    double average_velocity_coef;

    int32_t boundary_treatment;    /* 0 dynamic, 1 repulsive (Lennard-Jones walls) */

    int32_t dt_correction_method;  /* 0 const value, 1 const CFL, 2 dynamic */
    double dt;
    double CFL_coef;

    int32_t consistency_check;
    int32_t consistency_treatment; /* 0 print, 1 stop, 2 fix */

    int32_t nwm;                   /* 0 none, 2/3 piston, 4 wall disappear, 5 solitary, 6 Fourier piston */
    int32_t nwm_direction;
    uint32_t nwm_particles_start;
    uint32_t nwm_particles_end;
    double nwm_time_start;
    double nwm_freq;
    double nwm_wave_number;
    double nwm_wave_magnitude;
    double nwm_phase;
    int32_t piston_nharm;
    double piston_omega;
    double piston_a[RRSPH_MAX_HARMONICS];
    double piston_b[RRSPH_MAX_HARMONICS];

    int32_t boundary_motion;       /* 1: non-fluid r += v dt and the wave maker run (RRSPH_CL behaviour) */
    int32_t dt_dynamic_mode;       /* 0: RRSPH_OMP OpenMP build (reductions off); 1: max |a|, max mu over all particles */
    int32_t precision;             /* rrsph_precision */
    double start_time;

    /* Segmented Fourier piston (nwm = 6, 3D): piston_segments paddles across
     * z in [piston_seg_zmin, piston_seg_zmax); a wave-maker particle at z moves
     * with paddle floor((z - zmin) * segments / (zmax - zmin)) (clamped).
     * piston_a / piston_b hold piston_nharm harmonics per paddle, paddle-major
     * (segments * nharm <= RRSPH_MAX_HARMONICS).  0 or 1: one piston. */
    int32_t piston_segments;
    double piston_seg_zmin, piston_seg_zmax;

    /* Boundary-deficiency correction of the summation density
     * (density_normalization = 1, DENSITY_NORMALIZATION_BASE of the
     * reference's ParamsEnumeration.h:9-10, documented in
     * docs/latex/ModelParams.tex:124-133, rejected by its loader): on every
     * density_normalization_every-th step (1: every step) the summation
     * density becomes rho_j = sum_i m W_ji / sum_i (m / rho_i) W_ji, the sums
     * over j and its neighbours, rho_i the density before the step.  Only
     * with density_treatment 0. */
    int32_t density_normalization;
    uint32_t density_normalization_every;
} rrsph_params_c;

typedef struct rrsph_ctx rrsph_ctx;

/* Field selection for rrsph_hip_download. */
#define RRSPH_FIELD_R     1u
#define RRSPH_FIELD_V     2u
#define RRSPH_FIELD_RHO   4u
#define RRSPH_FIELD_P     8u
#define RRSPH_FIELD_ITYPE 16u

typedef struct rrsph_diagnostics {
    double time, dt;
    uint64_t step;
    uint32_t n_fluid_alive;        /* itype > 0 */
    uint32_t n_outside;            /* particles outside geometry after the last step */
    double mass_fluid;             /* m * count(itype > 0) */
    double kinetic_energy;         /* sum_fluid 0.5 m |v|^2 */
    double potential_energy;       /* sum_fluid m g y */
    double internal_energy;        /* sum_fluid m * Tait elastic energy per unit mass */
    double max_speed;              /* max_fluid |v| */
    double rho_min, rho_max;       /* over fluid */
} rrsph_diagnostics;

/* ---- lifetime -------------------------------------------------------- */
/* Replaces the setup half of time_integration (RRSPH_OMP/src/TimeIntegration.cpp:21-58)
 * and cl_time_integration's buffer creation (RRSPH_CL/src/CLTimeIntegration.cpp:415-467). */
rrsph_status rrsph_hip_create(const rrsph_params_c* params, rrsph_ctx** out);
void rrsph_hip_destroy(rrsph_ctx* ctx);
const char* rrsph_hip_last_error(const rrsph_ctx* ctx);

/* Host -> device state.  Stands in for the makeBufferCopyHost uploads of
 * CLTimeIntegration.cpp:427-431 (r, v, rho, itype; p is recomputed). Resets
 * time to params.start_time, dt to params.dt and the step counter to 0. */
rrsph_status rrsph_hip_upload(rrsph_ctx* ctx, const double* r, const double* v, const double* rho,
                              const int32_t* itype, uint32_t ntotal);

/* Advance nsteps whole time steps: predict_half_step -> grid_find -> density
 * -> update_acceleration (+update_dt) -> wave maker -> whole_step, i.e. one
 * iteration of the loop body of TimeIntegration.cpp:60-102 per step. */
rrsph_status rrsph_hip_step(rrsph_ctx* ctx, uint32_t nsteps);

/* Steps while time <= t_end (the loop condition of TimeIntegration.cpp:60),
 * at most max_steps; returns the number taken. */
rrsph_status rrsph_hip_run_until(rrsph_ctx* ctx, double t_end, uint64_t max_steps, uint64_t* steps_taken);

/* Device -> host in original order: the RRSPHOutput loaders of
 * CLTimeIntegration.cpp:83-100 (load_floatn_array / load_array). p is the
 * pressure of the last density evaluation, as the reference's p array. */
rrsph_status rrsph_hip_download(rrsph_ctx* ctx, uint32_t fields, double* r, double* v, double* rho, double* p,
                                int32_t* itype);

rrsph_status rrsph_hip_get_time(rrsph_ctx* ctx, double* time, double* dt, uint64_t* step);

/* Fields of the last step in original order: a (dim per particle), drho,
 * av (XSPH, dim per particle) and per-particle max |mu| of artificial
 * viscosity (UpdateAcceleration.cpp:59-116 outputs).  Requires
 * rrsph_hip_set_export_fields(ctx, 1) before the step. */
rrsph_status rrsph_hip_set_export_fields(rrsph_ctx* ctx, int32_t enable);
rrsph_status rrsph_hip_fields(rrsph_ctx* ctx, double* a, double* drho, double* av, double* arvmu);

/* Morton cell key of every particle at the current positions (get_cell_idx2,
 * RRSPH_COMMON/include/GridUtils.h:139-144), original order. */
rrsph_status rrsph_hip_cell_keys(rrsph_ctx* ctx, uint32_t* keys);

/* Neighbour lists of the current positions in the reference layout
 * out[n + max_nbr * j] with sentinel ntotal (grid_find, GridFind.cpp:215-236;
 * HeapDArray.h:78-85).  Order within a list follows the cell walk. */
rrsph_status rrsph_hip_neighbours(rrsph_ctx* ctx, uint32_t* out, uint32_t max_nbr);

/* Mass / energy budget and fluid statistics of the current state. */
rrsph_status rrsph_hip_diagnostics(rrsph_ctx* ctx, rrsph_diagnostics* out);

/* Wave-gauge elevation: max y of particles of type `ptype` (0 = any) with
 * |x - x_g| < search_n * hsml, for each gauge (HeightTesting::maxInLayer,
 * tools/WaterProfile/include/HeightTesting.h:25-54). */
rrsph_status rrsph_hip_gauges(rrsph_ctx* ctx, const double* x_gauges, uint32_t ngauges, double search_n,
                              int32_t ptype, double* out);

/* SPH interpolation of a field at points, FuncAtPoint
 * (tools/FuncAtPoint/lib/src/fatp.cpp:14-58 findValue, :121-144 calculate):
 * for each point x, sum_i f_i W(|r_i - x|) m / rho_i over the particles of
 * the 3^dim cells around x's cell with |r_i - x|^2 < (k h)^2, W the cubic
 * spline, rho_i the summation density of the current positions
 * (Density.h:25-61), f the field `field` as rrsph_hip_download returns it.
 * points holds npoints dim-tuples inside the geometry box (inclusive), at
 * most RRSPH_MAX_PROBES; out gets npoints values.  Errors as check_input
 * (fatp.cpp:92-119): "wrong target value", "point is outside bounds"; a
 * neighbour overflow with consistency_check on is the reference's
 * RRSPH_ERR_TOO_MANY_NEIGHBOURS.  A time series is one call per saved
 * time, as calculate() runs one per layer. */
#define RRSPH_MAX_PROBES 2048
#define RRSPH_PROBE_VX  0u
#define RRSPH_PROBE_VY  1u
#define RRSPH_PROBE_VZ  2u
#define RRSPH_PROBE_P   3u
#define RRSPH_PROBE_RHO 4u
rrsph_status rrsph_hip_probe(rrsph_ctx* ctx, const double* points, uint32_t npoints, uint32_t field,
                             double* out);

/* ---- differentiation ------------------------------------------------- */
/* No reference counterpart: the reference has no derivative of its step.
 * The adjoint differentiates the step, 2D or 3D, in FP64 arithmetic on the
 * trajectory of the context's precision (an FP32 context tapes and replays
 * the FP32 step; in 3D its pair adjoint reads the neighbours' adjoint values
 * rounded to FP32 and their separations from FP32 records) (continuity density,
 * cubic kernels, pressure, artificial viscosity, Lennard-Jones walls, XSPH,
 * boundary motion with the Fourier piston, nwm = 6, one piston or paddles
 * in 3D) with dt independent of the state (dt_correction_method 0, or 2
 * with dt_dynamic_mode 0). */
#define RRSPH_MAX_GAUGES 64
typedef struct rrsph_gauge_spec {
    uint32_t n;                    /* gauges, at most RRSPH_MAX_GAUGES */
    int32_t ptype;                 /* particle type the gauges look at (fluid: 2) */
    double radius;                 /* window half-width R: weight (1 - (x - x_g)^2 / R^2)^3 */
    double tau;                    /* softness; eta -> max y in the window as tau -> 0 */
    double y_ref;                  /* shift of the exponent, keep (y - y_ref) / tau below ~700 */
    double x[RRSPH_MAX_GAUGES];
    double z[RRSPH_MAX_GAUGES];    /* 3D: gauge at (x, z), weight (1 - ((x - x_g)^2 + (z - z_g)^2) / R^2)^3 */
} rrsph_gauge_spec;

/* Smooth surrogate of the gauge elevation on the current state:
 * eta_g = y_ref + tau * log(sum_i w_i exp((y_i - y_ref) / tau) / sum_i w_i)
 * over particles of type ptype (0 where the window is empty). */
rrsph_status rrsph_hip_smooth_gauges(rrsph_ctx* ctx, const rrsph_gauge_spec* spec, double* eta);
/* rrsph_hip_smooth_gauges without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device or host memory (tight
 * control loops reading the gauges every step, code that forbids hipMalloc):
 * its block partial sums live in a fixed device array of the library
 * (1.03 MiB per device, kept until the process ends, its address looked up
 * once per device when a context is created) and the results pass through a
 * pinned buffer the context allocated at creation.  At most
 * RRSPH_MAX_GAUGES gauges (more: RRSPH_ERR_INVALID_ARGUMENT), any particle
 * count.  Results equal rrsph_hip_smooth_gauges' bit for bit.  All contexts
 * on one device share that array: calls that overlap in time on one device
 * (other host threads, other streams) race; calls in sequence do not. */
rrsph_status rrsph_hip_smooth_gauges_noalloc(rrsph_ctx* ctx, const rrsph_gauge_spec* spec, double* eta);
/* Replace the Fourier piston coefficients (nwm = 6):
 * x(t) = sum_k a_k sin(k w t) + b_k cos(k w t), k = 1 .. nharm; with
 * piston_segments paddles, a and b hold nharm harmonics per paddle,
 * paddle-major (segments * nharm entries). */
rrsph_status rrsph_hip_set_piston(rrsph_ctx* ctx, double omega, uint32_t nharm, const double* a, const double* b);
/* Differentiable run from the current state: nsteps steps, a checkpoint
 * every checkpoint_every steps (device memory), the smooth gauges after
 * every step into eta[step * spec->n + g] (host).  Steps that sort every
 * step keep the state before every step instead when all of them fit in
 * half the free device memory (then the reverse pass replays nothing). */
rrsph_status rrsph_hip_adjoint_forward(rrsph_ctx* ctx, uint32_t nsteps, uint32_t checkpoint_every,
                                       const rrsph_gauge_spec* spec, double* eta);
/* Reverse pass of the last adjoint_forward for eta_bar = dL/d eta (same
 * layout): gradients with respect to the piston coefficients (segments *
 * nharm each, the layout of set_piston) and, when the pointers are not
 * NULL, with respect to the initial state (original particle order; r, v as
 * dim-tuples).  Replays each checkpoint segment forward (unless the
 * forward kept every state) and runs the step adjoints backward; leaves the
 * state as adjoint_forward left it. */
rrsph_status rrsph_hip_adjoint_backward(rrsph_ctx* ctx, const double* eta_bar, double* grad_a, double* grad_b,
                                        double* grad_r0, double* grad_v0, double* grad_rho0);

/* ---- streams and timing ---------------------------------------------- */
void* rrsph_hip_stream(rrsph_ctx* ctx);           /* hipStream_t the context launches on */
rrsph_status rrsph_hip_set_graph(rrsph_ctx* ctx, int32_t enable);
/* Runs nsteps steps with HIP events around every launch on the context's
 * stream; fills average milliseconds per launch for each kernel class
 * (names via rrsph_hip_kernel_class_name) and the total wall time. */
rrsph_status rrsph_hip_profile_steps(rrsph_ctx* ctx, uint32_t nsteps, double* avg_ms, uint32_t n_classes,
                                     double* total_ms);
uint32_t rrsph_hip_kernel_class_count(void);
const char* rrsph_hip_kernel_class_name(uint32_t k);
/* Times nsteps steps between two events on the context's stream (no
 * per-launch events); returns milliseconds. */
rrsph_status rrsph_hip_time_steps(rrsph_ctx* ctx, uint32_t nsteps, double* total_ms);

/* Cell-sort statistics.  No reference counterpart (the reference re-sorts
 * from scratch every step, GridFind.cpp:47-88 / CLTimeIntegration.cpp:139-210):
 * steps whose order came from the incremental re-sort, from the full radix
 * sort, re-runs after a step changed more cells than the incremental path
 * takes, and the number of particles that changed cell in the last step. */
rrsph_status rrsph_hip_sort_stats(rrsph_ctx* ctx, uint64_t* incremental_steps, uint64_t* full_steps,
                                  uint64_t* fallbacks, uint32_t* last_changed);

/* Neighbour-list statistics.  No reference counterpart (the reference
 * builds its lists from scratch every step, GridFind.cpp:112-189): in the
 * cubic configuration (2D and 3D) the step sorts and walks the cells only on build
 * steps, which list every particle within k h + skin, and the steps between
 * test those lists with the exact cut-off.  Steps that read the lists,
 * build steps, steps that halted because the lists might be stale (re-run
 * as builds), and the current build interval (0: lists off). */
rrsph_status rrsph_hip_list_stats(rrsph_ctx* ctx, uint64_t* list_steps, uint64_t* builds, uint64_t* stale_halts,
                                  uint32_t* interval);

/* Version hooks of the backend (RRSPH_COMMON/include/CommonIncl.h:14-17). */
uint32_t RRSPH_HIP_GetSpecificVersionMajor(void);
uint32_t RRSPH_HIP_GetSpecificVersionMinor(void);
uint32_t RRSPH_HIP_GetSpecificVersionPatch(void);
const char* RRSPH_HIP_GetSpecificVersionName(void);

#ifdef __cplusplus
}
#endif
#endif
