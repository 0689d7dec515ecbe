// Shared device-side types for the RRSPH_HIP kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rrsph {

constexpr int kBlock = 256;            // threads per block for particle kernels (4 waves)
constexpr int kWave = 64;
constexpr int kNumXcd = 8;             // MI355X: 8 XCDs, blocks dealt round-robin
constexpr uint32_t kRadixBits = 8;
constexpr uint32_t kRadix = 1u << kRadixBits;
constexpr int kSortItems = 8;          // keys per thread per sort tile
constexpr int kSortTile = kBlock * kSortItems;

template <class R> struct V4;
template <> struct V4<float> { using type = float4; };
template <> struct V4<double> { using type = double4; };
template <class R> using real4 = typename V4<R>::type;
template <class R> struct V2;
template <> struct V2<float> { using type = float2; };
template <> struct V2<double> { using type = double2; };
template <class R> using real2 = typename V2<R>::type;

// |r|^2 = dx*dx + dy*dy with every product and sum rounded separately (no
// fused multiply-add), the arithmetic of the reference's FP64 neighbour
// test, so that the neighbour sets agree bit for bit.
__device__ __forceinline__ double dist2_exact(double dx, double dy) {
#pragma clang fp contract(off)
    const double xx = dx * dx;
    const double yy = dy * dy;
    return xx + yy;
}

// Per-step control block in device memory.  The prologue (one thread, start
// of each step) commits the previous step and snapshots the values every
// other kernel of the step reads; kernels only read the snapshot.
struct Ctrl {
    // committed state
    double time;
    double dt;              // dt used by the next predict (params.dt of the reference)
    unsigned long long step;
    // produced by the integrator of the step in flight
    double dt_new;
    uint32_t pending;       // 1: dt_new / time advance not committed yet
    uint32_t halted;        // nonzero: every kernel returns at once
    uint32_t status;        // rrsph_status of the halt
    uint32_t nwm_done;      // wall-disappear wave maker already fired
    // reductions of the step in flight
    unsigned long long amax_bits;   // max |a| (non-negative double bits)
    unsigned long long mumax_bits;  // max |mu|
    uint32_t outside;       // particles outside the geometry after whole_step
    uint32_t nbr_overflow;  // a particle had more than max_neighbours - 1 neighbours
    uint32_t grid_overflow; // a live particle lay outside the cell grid
    uint32_t resort_count;  // slots that changed cell key since the last sort
    // snapshot of the step in flight (written by the prologue)
    double cur_time;
    double cur_dt;          // dt of predict_half_step
    double cur_dt_const;    // dt of whole_step when it is known before the step
    double piston_v;        // wave-maker velocity along its axis
    double t_end;           // the prologue stops (status 0) once time > t_end
    uint32_t cur_first;     // step == 0 -> v_dt = dt / 2
    uint32_t piston_active;
    uint32_t wall_disappear;
    uint32_t rollback;      // 1: the halted step was integrated by the fused pair pass; the host restores its input
    double piston_vseg[8];  // segmented piston: velocity of each paddle
    // neighbour-list steps
    double disp_acc;        // sum over the steps since the last list build of each step's largest displacement
    double disp_last;       // the largest displacement of the last committed list or build step
    uint32_t list_overflow; // the last build found more candidates for some particle than its list holds
    uint32_t list_valid;    // the lists describe the current slot order (a build ran; no sort since)
    uint32_t cur_renorm;    // this step's summation density is renormalised
    uint32_t renorm_pad;
};

constexpr int kMaxSegments = 8;  // paddles of a segmented piston

// Everything the kernels need, by value (kernel arguments live in SGPRs).
struct StepParams {
    // sizes
    uint32_t n;
    uint32_t nfluid;
    uint32_t max_nbr;
    int32_t dim;            // 2 or 3
    int32_t ncx, ncy, ncz;  // ncz = 1 in 2D
    uint32_t ncells;        // = ncx * ncy (* ncz); also the dead-particle key
    // grid
    double xmin, ymin, xmax, ymax, zmin, zmax, cell;
    double cut2;            // (k h)^2, compared in FP64
    // physics
    double h, mass, rho0, g;
    double f_cubic, f_gauss, f_wend, f_desb;
    int32_t skf_density, skf_intf, skf_av, skf_xsph, skf_ap;
    double eos_B;
    int32_t density_treatment;
    int32_t renorm;         // summation density renormalised on every renorm_every-th step
    uint32_t renorm_every;
    double dhc;             // delta-SPH: coef * h * c
    int32_t intf_approx;
    int32_t art_pressure;
    double ap_coef, ap_index, ap_wdelta;
    int32_t av_on;
    double av_alpha, av_beta, av_c, av_eta2;
    int32_t xsph_on;
    double xsph_eps;
    int32_t lj_on;
    double lj_r0, lj_D;
    // time stepping
    int32_t dt_method;
    int32_t dt_dynamic_mode;
    double dt0;
    double cfl, c0, never_ending;
    // consistency
    int32_t consistency_check;
    int32_t consistency_treatment;
    // wave maker
    int32_t boundary_motion;
    int32_t nwm;
    int32_t nwm_axis;       // 0 x, 1 y, 2 z, -1 none
    uint32_t nwm_start, nwm_end;
    double nwm_time_start;
    double nwm_freq, nwm_delta, nwm_S0, nwm_e2;
    double sol_k, sol_c, sol_Tf, sol_H, sol_h;
    int32_t piston_nharm;   // harmonics per paddle
    int32_t piston_segments;  // paddles across z (1: one piston); coefficients paddle-major
    double piston_omega;
    double piston_seg_z0, piston_seg_inv_w;  // paddle of a particle: floor((z - z0) * inv_w)
    double piston_a[32];
    double piston_b[32];
    // launch geometry
    uint32_t nblocks;       // logical blocks of kBlock particles
    uint32_t blocks_per_xcd;
    uint32_t ntiles;        // tiles of one wavefront (64 particles) for the tiled pair pass
    uint32_t tile_blocks_per_xcd;
    // incremental re-sort: the integrator stores the ballot of the slots of
    // [64 w, 64 w + 64) whose particle moved to another cell in chg_mask[w]
    unsigned long long* chg_mask;
    // neighbour lists: every particle within list_reach = k h + skin at the
    // build (squared, with a rounding margin: list_r2), at most list_cap per
    // particle; valid while the summed largest displacements stay below
    // skin_half = skin / 2 (less a rounding margin)
    double list_reach, list_r2, skin_half;
    uint32_t list_cap;
};

// XCD-aware block index: the hardware deals consecutive workgroups to the 8
// XCDs round-robin; this mapping gives each XCD one contiguous slice of the
// sorted particle range so neighbour data is shared inside one L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t blocks_per_xcd) {
    const uint32_t b = blockIdx.x;
    return (b % kNumXcd) * blocks_per_xcd + b / kNumXcd;
}

}  // namespace rrsph
