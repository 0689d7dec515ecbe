/*
 * FP64 CPU oracle for the WCSPH time step -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C++/OpenMP, the RRSPH_OMP time step of
 * RackotRR/SPH2D (the .cpp files of sources/RRSPH_OMP/src and the per-pair
 * physics headers of sources/RRSPH_COMMON/include).  It is what the HIP
 * product path is compared against, and the CPU baseline that bench.py
 * times.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library never links or calls it.
 *
 * Semantics follow RRSPH_OMP (data flow: density from v and rho, forces from
 * v_predict and rho_predict, sec. "time_integration" TimeIntegration.cpp:60-102)
 * with two named switches for behaviour that only RRSPH_CL has:
 *   boundary_motion = 1 : non-fluid particles advance r += v*dt and the piston
 *                         wave maker sets their velocity
 *                         (cl/TimeIntegration.cl:190-193, :196-245);
 *   dt_dynamic_mode = 1 : DT_CORRECTION_DYNAMIC reduces max|a| and max mu over
 *                         all particles, as the serial (non-OpenMP) build of
 *                         UpdateAcceleration.cpp:107-124 does; mode 0 keeps
 *                         the OpenMP build's disabled reductions, which yields
 *                         dt = CFL*h/c.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_HARMONICS 32

typedef struct oracle_params {
    int32_t dim;                 /* 2 or 3 */
    uint32_t ntotal;
    uint32_t nfluid;
    uint32_t max_neighbours;

    double x_mingeom, x_maxgeom;
    double y_mingeom, y_maxgeom;
    double z_mingeom, z_maxgeom;

    double hsml;
    double delta;
    double mass;
    double rho0;
    double cell_scale_k;
    double depth;
    double g;
    double pi;

    double eos_c_sqr;            /* c^2 used by the EOS (EOS.cpp:3-20) */
    double eos_c;                /* c used by dt (EOS.cpp:17-20) */
    double eos_sound_vel;        /* params.eos_sound_vel used by AV */

    int32_t density_treatment;   /* 0 summation, 1 continuity, 2 continuity+delta */
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

    int32_t boundary_treatment;  /* 0 dynamic, 1 repulsive (LJ) */

    int32_t dt_correction_method;
    double dt;
    double CFL_coef;

    int32_t consistency_check;
    int32_t consistency_treatment; /* 0 print, 1 stop, 2 fix */

    /* wave maker */
    int32_t nwm;                 /* 0 none, 2/3 piston 1st/2nd order, 4 wall disappear, 5 solitary, 6 fourier piston */
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
    double piston_a[ORACLE_MAX_HARMONICS];
    double piston_b[ORACLE_MAX_HARMONICS];
    int32_t piston_segments;     /* > 1: paddles across z, harmonics paddle-major */
    double piston_seg_zmin, piston_seg_zmax;

    int32_t boundary_motion;
    int32_t dt_dynamic_mode;
    int32_t num_threads;         /* 0: OpenMP default */
    /* summation density renormalised (ModelParams.tex:124-133) on every
     * density_normalization_every-th step */
    int32_t density_normalization;
    uint32_t density_normalization_every;
} oracle_params;

typedef struct oracle_ctx oracle_ctx;

/* status codes: 0 ok, 1 too many neighbours, 2 never ending simulation,
 * 3 particle outside boundaries (STOP), 4 cell index overflow, 5 bad args */

/* ---- known-answer helpers ---- */
double oracle_kernel_w(const oracle_params* p, double dist, int32_t skf);
void oracle_kernel_dwdr(const oracle_params* p, double dist, const double* diff, int32_t skf, double* out);
double oracle_eos_p(const oracle_params* p, double rho);
double oracle_eos_rho(const oracle_params* p, double pressure);
uint32_t oracle_morton2(uint32_t x, uint32_t y);
void oracle_morton2_decode(uint32_t key, uint32_t* x, uint32_t* y);
uint32_t oracle_morton3(uint32_t x, uint32_t y, uint32_t z);
void oracle_morton3_decode(uint32_t key, uint32_t* x, uint32_t* y, uint32_t* z);
uint64_t oracle_cell_coord(const oracle_params* p, double r, double rmin);
uint64_t oracle_cell_idx(const oracle_params* p, const double* r);
void oracle_neighbouring_cells(const oracle_params* p, uint64_t idx, uint64_t* cells);
uint64_t oracle_count_cells(const oracle_params* p, int32_t* status);

/* ---- step modules (arrays are AoS like the reference: r[dim*i + k]) ---- */
/* FuncAtPoint (tools/FuncAtPoint/lib/src/fatp.cpp:14-58, :121-144): the SPH
 * interpolation sum_i f_i W_cubic(|r_i - x|) m / rho_i at each point, rho the
 * summation density of r; f = field[i * stride + comp].  Returns grid_find's
 * status (nonzero: no values). */
int32_t oracle_probe(const oracle_params* p, const double* r, const int32_t* itype, const double* field,
                     uint32_t stride, uint32_t comp, const double* points, uint32_t npoints, double* out);

int32_t oracle_grid_find(const oracle_params* p, const double* r, const int32_t* itype,
                         uint32_t* neighbours);
void oracle_density_sum(const oracle_params* p, const double* r, const uint32_t* neighbours,
                        double* rho, double* pres);
/* renormalised summation density: rho = sum m W / sum (m / rho_old) W
 * (docs/latex/ModelParams.tex:124-133), self term included in both sums */
void oracle_density_sum_normalized(const oracle_params* p, const double* r, const uint32_t* neighbours,
                                   const double* rho_old, double* rho, double* pres);
void oracle_density_con(const oracle_params* p, const double* r, const double* v,
                        const uint32_t* neighbours, const double* rho, double* drho, double* pres);
/* per-term sums over the neighbour list, as the reference's module tests do */
void oracle_term_external(const oracle_params* p, const double* r, const int32_t* itype,
                          const uint32_t* neighbours, double* out);
void oracle_term_internal(const oracle_params* p, const double* r, const uint32_t* neighbours,
                          const double* pres, const double* rho, double* out);
void oracle_term_art_visc(const oracle_params* p, const double* r, const double* v,
                          const uint32_t* neighbours, const double* rho, double* out, double* mu);
void oracle_term_average_velocity(const oracle_params* p, const double* r, const double* v,
                                  const int32_t* itype, const uint32_t* neighbours,
                                  const double* rho, uint32_t jmax, double* out);
void oracle_update_acceleration(const oracle_params* p, const int32_t* itype, const double* r,
                                const double* v, const uint32_t* neighbours, const double* rho,
                                const double* pres, double* a, double* av, double* arvmu,
                                double* amax, double* mumax);

/* ---- whole time loop with persistent state ---- */
oracle_ctx* oracle_create(const oracle_params* p);
void oracle_destroy(oracle_ctx* c);
int32_t oracle_set_state(oracle_ctx* c, const double* r, const double* v, const double* rho,
                         const double* pres, const int32_t* itype, double time, double dt);
int32_t oracle_step(oracle_ctx* c, uint32_t nsteps);
void oracle_get_state(const oracle_ctx* c, double* r, double* v, double* rho, double* pres,
                      int32_t* itype, double* time, double* dt, uint64_t* step);
/* fields of the last completed step: a, drho, av (XSPH), arvmu, outside count */
void oracle_get_fields(const oracle_ctx* c, double* a, double* drho, double* av, double* arvmu);
const uint32_t* oracle_neighbours_ptr(const oracle_ctx* c);
uint32_t oracle_last_outside(const oracle_ctx* c);
double oracle_piston_velocity(const oracle_params* p, double time);

#ifdef __cplusplus
}
#endif
#endif
