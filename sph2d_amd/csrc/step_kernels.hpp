// The WCSPH step as HIP kernels for gfx950 (2D).
//
// Device layout (all arrays in the particle order of the last cell sort):
//   pos   double2   position (FP64 in both precisions)
//   va    real4     (vx, vy, rho, itype)
//   ad    real4     (ax, ay, drho, 0) of the previous step (predictor input)
//   oid   uint32    original particle index
//   pred  real4     (vx^, vy^, rho^, p / rho^2): predicted velocity and density,
//                   pressure over predicted density squared
//   pres  real      p of the current density (the reference's p array)
// Two copies of (pos, va, oid) ping-pong: "A" holds the state between steps,
// "B" the freshly sorted copy a step reads from.
//
// Step = prologue -> radix sort of the cell keys -> reorder (gather A -> B,
// predict, pressure, cell-start table) -> [density summation pass] -> pair
// pass (continuity + forces, and whole_step fused in when dt is known before
// the step) -> [integrate, when dt comes from this step's reductions].
#pragma once

#include "device_common.hpp"
#include "physics.hpp"

namespace rrsph {

template <class R> struct StateA {
    double2* pos;
    real4<R>* va;
    real4<R>* ad;
    uint32_t* oid;
};
template <class R> struct StateB {
    double2* pos;
    real4<R>* va;
    real4<R>* pred;   // (vx^, vy^, rho^, p / rho^2)
    uint32_t* oid;
    R* pres;          // p, the reference's pressure array
    real4<R>* ad = nullptr;  // copy of A.ad, kept when a fused step may have to be undone
};

// ------------------------------------------------------------------ grid keys
// Cell coordinate by truncation of (r - r_min) / cell, as get_cell_idx2
// (GridUtils.h:135-144); the linear key is x-major (cx * ncy + cy) so the
// three cells of one grid column are contiguous after sorting and are walked
// in the reference's x-major neighbour-cell order (GridUtils.h:209-214).
__device__ __forceinline__ int cell_coord(double r, double rmin, double cell, int ncell, uint32_t& bad) {
    const double t = (r - rmin) / cell;
    if (t >= 0.0 && t < (double)ncell) return (int)t;
    bad = 1;
    return t >= (double)ncell ? ncell - 1 : 0;  // NaN and negative -> 0
}

__device__ __forceinline__ uint32_t cell_key(const StepParams& P, double x, double y, uint32_t& bad) {
    const int cx = cell_coord(x, P.xmin, P.cell, P.ncx, bad);
    const int cy = cell_coord(y, P.ymin, P.cell, P.ncy, bad);
    return (uint32_t)cx * (uint32_t)P.ncy + (uint32_t)cy;
}

__device__ __forceinline__ uint32_t split2(uint32_t x) {
    x &= 0x0000FFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

// ------------------------------------------------------------------ prologue
// Velocity of the wave maker at time t (paddle `seg` of a segmented piston).
__device__ double piston_velocity(const StepParams& P, double t, int seg = 0) {
    if (P.nwm == 6) {
        const double* a = P.piston_a + seg * P.piston_nharm;
        const double* b = P.piston_b + seg * P.piston_nharm;
        double v = 0;
        for (int k = 1; k <= P.piston_nharm; ++k) {
            const double kw = k * P.piston_omega;
            v += kw * (a[k - 1] * cos(kw * t) - b[k - 1] * sin(kw * t));
        }
        return v;
    }
    if (P.nwm == 5) {  // WaveMaker.cpp:105-128
        const double u = tanh(P.sol_k * P.sol_c * (t - P.sol_Tf));
        const double H = P.sol_H, h = P.sol_h, k = P.sol_k;
        const double top = k * H * (h + H - H * u * u) + 2 * k * (H * u) * (H * u);
        const double bottom = (k * (h + H - H * u * u)) * (k * (h + H - H * u * u));
        return top / bottom * (1 - u * u) * (k * P.sol_c);
    }
    // cl/TimeIntegration.cl:206-229 (first order + product-form second order)
    double vx = 0.5 * P.nwm_S0 * P.nwm_freq * cos(P.nwm_freq * t + P.nwm_delta);
    if (P.nwm == 3) vx += 2 * P.nwm_freq * P.nwm_e2 * cos(2 * P.nwm_freq * t + 2 * P.nwm_delta);
    return vx;
}

// Paddle of a segmented piston that a wave-maker particle at z belongs to.
__device__ __forceinline__ int piston_segment(const StepParams& P, double z) {
    const double u = floor((z - P.piston_seg_z0) * P.piston_seg_inv_w);
    return u < 0.0 ? 0 : (u >= (double)P.piston_segments ? P.piston_segments - 1 : (int)u);
}

__device__ __forceinline__ double dbits(unsigned long long b) { return __longlong_as_double((long long)b); }

// Block maximum of two non-negative values, folded into one of kMaxSlots
// slots per quantity with an atomic max on the bit pattern (spreading the
// atomics over many addresses keeps them from serialising on one).
constexpr uint32_t kMaxSlots = 256;
__device__ __forceinline__ void block_max2_to_slots(double a, double b, unsigned long long* __restrict__ slots,
                                                    uint32_t blk) {
    __shared__ double s_a[kBlock / kWave], s_b[kBlock / kWave];
    for (int off = 32; off > 0; off >>= 1) {
        a = fmax(a, __shfl_xor(a, off));
        b = fmax(b, __shfl_xor(b, off));
    }
    const uint32_t w = threadIdx.x / kWave;
    const uint32_t nw = blockDim.x / kWave;
    if (__lane_id() == 0) {
        s_a[w] = a;
        s_b[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t k = 1; k < nw; ++k) {
            a = fmax(a, s_a[k]);
            b = fmax(b, s_b[k]);
        }
        const uint32_t slot = blk % kMaxSlots;
        if (a > 0.0) atomicMax(&slots[slot], (unsigned long long)__double_as_longlong(a));
        if (b > 0.0) atomicMax(&slots[kMaxSlots + slot], (unsigned long long)__double_as_longlong(b));
    }
}

// Maximum over the slots, computed by every block (deterministic).
__device__ __forceinline__ void slots_max2(const unsigned long long* __restrict__ slots, double& a, double& b) {
    __shared__ double s_a[kBlock / kWave], s_b[kBlock / kWave];
    double x = dbits(slots[threadIdx.x]), y = dbits(slots[kMaxSlots + threadIdx.x]);
    for (int off = 32; off > 0; off >>= 1) {
        x = fmax(x, __shfl_xor(x, off));
        y = fmax(y, __shfl_xor(y, off));
    }
    const uint32_t w = threadIdx.x / kWave;
    if (__lane_id() == 0) {
        s_a[w] = x;
        s_b[w] = y;
    }
    __syncthreads();
    a = s_a[0];
    b = s_b[0];
    for (int k = 1; k < kBlock / kWave; ++k) {
        a = fmax(a, s_a[k]);
        b = fmax(b, s_b[k]);
    }
}

// One thread: commit the previous step, decide whether to halt, snapshot.
// The commit half of the prologue (one thread, on the control block c):
// the previous step's time advance, the halts, and this step's snapshot.
// Returns false when the step halts.
__device__ __forceinline__ bool prologue_commit(const StepParams& P, Ctrl& c) {
    if (c.halted) return false;
    c.list_valid = 0;  // a sorted step reorders the slots; a list build sets it again
    if (c.pending) {
        if (P.consistency_check && c.nbr_overflow) {
            // the reference throws in find_neighbours, before whole_step
            // (GridFind.cpp:186-188): the step that overflowed is not
            // committed, and the host puts back the state it started from
            // (the fused pair pass has already integrated it)
            c.pending = 0;
            c.rollback = 1;
            if (c.wall_disappear) c.nwm_done = 0;
            c.halted = 1;
            c.status = 3;  // RRSPH_ERR_TOO_MANY_NEIGHBOURS
            return false;
        }
        c.time += c.dt_new;
        c.dt = c.dt_new;
        c.step += 1;
        c.pending = 0;
        if (P.consistency_check && c.outside > 0 && P.consistency_treatment == 1) {
            c.halted = 1;
            c.status = 5;  // RRSPH_ERR_OUTSIDE_BOUNDARIES
            return false;
        }
    }
    if (c.time > c.t_end) {  // loop condition time <= simulation_time
        c.halted = 1;
        c.status = 0;
        return false;
    }
    c.cur_time = c.time;
    c.cur_dt = c.dt;
    c.cur_first = c.step == 0;
    // the boundary-deficiency correction of the summation density on every
    // renorm_every-th step (steps counted from 1)
    c.cur_renorm = P.renorm && (c.step + 1) % P.renorm_every == 0;
    // dt of whole_step when it does not depend on this step's reductions
    if (P.dt_method == 2 && P.dt_dynamic_mode == 0) {
        // RRSPH_OMP OpenMP build: max |a| and max mu stay 0 (UpdateAcceleration.cpp:107-124)
        c.cur_dt_const = P.cfl * fmin(sqrt(P.h / 0.0), P.h / (P.c0 + 0.0));
    } else {
        c.cur_dt_const = c.dt;
    }
    const bool nwm_on = P.boundary_motion && P.nwm != 0 && c.time >= P.nwm_time_start;
    c.piston_active = nwm_on && P.nwm != 4 && P.nwm_axis >= 0;
    c.wall_disappear = nwm_on && P.nwm == 4 && !c.nwm_done;
    if (c.wall_disappear) c.nwm_done = 1;
    c.amax_bits = 0;
    c.mumax_bits = 0;
    c.outside = 0;
    return true;
}

// Term k of paddle seg of the Fourier piston's velocity at time t (k from 1).
__device__ __forceinline__ double fourier_term(const StepParams& P, double t, int seg, int k) {
    const double kw = k * P.piston_omega;
    double sn, cs;
    sincos(kw * t, &sn, &cs);
    return kw * (P.piston_a[seg * P.piston_nharm + k - 1] * cs - P.piston_b[seg * P.piston_nharm + k - 1] * sn);
}

// The step prologue, run by the first wave of the block with the control
// block in scalar registers: each of the first kCtrlWords lanes loads one
// 8-byte word (with the kernel's other loads), every lane reads them all
// with readlane (the values are uniform), the wave commits the previous
// step and makes this step's snapshot on that copy, the Fourier piston's
// terms are evaluated one per lane and summed in harmonic order, and lane
// i writes word i back.  (Updating the block from one thread through
// memory, LDS or global, is a chain of dependent round trips.)  `before`
// and `after` act on the copy.  Every thread of the block calls it; the
// barrier at the end orders the write-back before the block's later
// accesses to the control block.
constexpr int kPistonTerms = 32;  // paddles x harmonics of the Fourier piston (at most one per lane)
static_assert(kPistonTerms <= kWave, "the piston's terms are evaluated one per lane");
constexpr uint32_t kCtrlWords = sizeof(Ctrl) / 8;
static_assert(sizeof(Ctrl) % 8 == 0 && kCtrlWords <= kWave, "the control block is copied in 8-byte words");

// Word threadIdx.x of the control block (the first kCtrlWords threads), to
// be handed to block_prologue; a kernel loads it first so that the load is
// in flight with its other loads.
__device__ __forceinline__ unsigned long long ctrl_word(const Ctrl* ctrl) {
    return threadIdx.x < kCtrlWords ? reinterpret_cast<const unsigned long long*>(ctrl)[threadIdx.x] : 0ull;
}

template <class B, class A>
__device__ __forceinline__ void block_prologue(const StepParams& P, Ctrl* ctrl, unsigned long long word, B&& before,
                                               A&& after) {
    if (threadIdx.x < kWave) {
        union {
            Ctrl c;
            unsigned long long w[kCtrlWords];
        } u;
        const int lo = (int)(uint32_t)word, hi = (int)(uint32_t)(word >> 32);
#pragma unroll
        for (uint32_t i = 0; i < kCtrlWords; ++i)
            u.w[i] = (unsigned long long)(uint32_t)__builtin_amdgcn_readlane(lo, i) |
                     ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane(hi, i) << 32);
        Ctrl& c = u.c;
        before(c);
        if (prologue_commit(P, c)) {
            double vseg[kMaxSegments];
#pragma unroll
            for (int g = 0; g < kMaxSegments; ++g) vseg[g] = 0.0;
            if (c.piston_active && P.nwm == 6) {
                const int nh = P.piston_nharm, nseg = max(P.piston_segments, 1);
                const int lane = (int)threadIdx.x;
                const double term = lane < nseg * nh ? fourier_term(P, c.time, lane / nh, lane % nh + 1) : 0.0;
#pragma unroll
                for (int g = 0; g < kMaxSegments; ++g) {
                    double v = 0;
                    if (g < nseg)
                        for (int k = 0; k < nh; ++k) v += __shfl(term, g * nh + k);
                    vseg[g] = v;
                }
            } else if (c.piston_active) {
                vseg[0] = piston_velocity(P, c.time);
            }
            c.piston_v = vseg[0];
#pragma unroll
            for (int g = 0; g < kMaxSegments; ++g) c.piston_vseg[g] = g < P.piston_segments ? vseg[g] : 0.0;
        }
        after(c);
        unsigned long long out = 0;
#pragma unroll
        for (uint32_t i = 0; i < kCtrlWords; ++i) out = threadIdx.x == i ? u.w[i] : out;
        if (threadIdx.x < kCtrlWords) reinterpret_cast<unsigned long long*>(ctrl)[threadIdx.x] = out;
    }
    __syncthreads();
}

// The prologue of a sorted step (one block of kWave threads).
__global__ void __launch_bounds__(kWave) step_prologue(StepParams P, Ctrl* ctrl) {
    block_prologue(P, ctrl, ctrl_word(ctrl), [](Ctrl&) {}, [](Ctrl&) {});
}

// Cell-start table over keys 0 .. ncells (+1 sentinel entry):
// LB[c] = first sorted index whose key >= c.  Lane s owns the cells
// (key[s-1], key[s]]; the last lane also owns the cells after its key.
// Short runs are written by the owning lane, long runs (empty columns above
// the water, the domain margins) by the whole wave.  Every lane of the wave
// calls it (idle lanes with live = false).
__device__ __forceinline__ void fill_cell_starts(const StepParams& P, uint32_t s, bool live, uint32_t k,
                                                 const uint32_t* __restrict__ skey, uint32_t* __restrict__ LB) {
    uint32_t lo = s == 0 ? 0u : skey[s - 1] + 1u;
    uint32_t hi = live ? k + 1 : lo;  // exclusive
    for (int pass = 0; pass < 2; ++pass) {
        const uint32_t val = pass == 0 ? s : P.n;
        if (pass == 1) {
            lo = live && s == P.n - 1 ? k + 1 : 0u;
            hi = live && s == P.n - 1 ? P.ncells + 2 : 0u;
        }
        const bool small = hi - lo <= 4;
        if (small)
            for (uint32_t c = lo; c < hi; ++c) LB[c] = val;
        unsigned long long big = __ballot(!small);
        while (big) {
            const int src = __ffsll((long long)big) - 1;
            big &= big - 1;
            const uint32_t blo = __shfl(lo, src), bhi = __shfl(hi, src), bval = __shfl(val, src);
            for (uint32_t c = blo + __lane_id(); c < bhi; c += kWave) LB[c] = bval;
        }
    }
}

// |r| of a neighbour pair in the pass precision.  FP64 takes the square root
// of the FP64 sum; FP32 rounds the differences to float and uses the
// hardware square root (1 ulp), the FP64 root being a long instruction
// sequence paid on every neighbour of the summation-density passes.
template <class R> __device__ __forceinline__ R pair_dist(double dx, double dy, double dz);
template <> __device__ __forceinline__ double pair_dist<double>(double dx, double dy, double dz) {
    return sqrt(dx * dx + dy * dy + dz * dz);
}
template <> __device__ __forceinline__ float pair_dist<float>(double dx, double dy, double dz) {
    const float x = (float)dx, y = (float)dy, z = (float)dz;
    return __builtin_amdgcn_sqrtf(x * x + y * y + z * z);
}

// ------------------------------------------------------------------ reorder
// Gather the state into sorted order, predict (PredictHalfStep.cpp:16-22),
// compute the pressure of the current density (Density.h:112) and fill the
// cell-start table: LB[c] = first sorted index whose key >= c.  With
// skey_keep, also keep the sorted keys (the next step's incremental re-sort
// compares them with the keys after integration).
template <class R>
__global__ void __launch_bounds__(kBlock) reorder_kernel(StepParams P, const Ctrl* __restrict__ ctrl,
                                                         const uint32_t* __restrict__ skey,
                                                         const uint32_t* __restrict__ perm, StateA<R> A, StateB<R> B,
                                                         uint32_t* __restrict__ LB,
                                                         unsigned long long* __restrict__ maxslots,
                                                         uint32_t* __restrict__ skey_keep) {
    if (ctrl->halted) return;
    if (blockIdx.x == 0) {  // clear the max slots of this step's reductions
        maxslots[threadIdx.x] = 0;
        maxslots[kMaxSlots + threadIdx.x] = 0;
    }
    const uint32_t s0 = blockIdx.x * kBlock + threadIdx.x;
    const bool live = s0 < P.n;
    const uint32_t s = live ? s0 : P.n - 1;  // idle lanes shadow the last particle (no writes)
    const uint32_t k = skey[s];
    const uint32_t src = min(perm[s], P.n - 1);  // a permutation of [0, n); the clamp only guards memory
    const double2 pos = A.pos[src];
    const real4<R> va = A.va[src];
    const real4<R> ad = A.ad[src];
    const uint32_t oid = A.oid[src];
    if (live) {
        B.pos[s] = pos;
        B.va[s] = va;
        B.oid[s] = oid;
        if (skey_keep) skey_keep[s] = k;  // the key slot s is sorted by, for the next re-sort
    }
    const R dt = (R)ctrl->cur_dt;
    const bool cont = P.density_treatment != 0;
    real4<R> pr;
    pr.x = va.x + ad.x * dt * R(0.5);
    pr.y = va.y + ad.y * dt * R(0.5);
    pr.z = cont ? va.z + ad.z * dt * R(0.5) : va.z;
    const R p = eos_p<R>(va.z, (R)P.rho0, (R)P.eos_B);
    pr.w = p / (pr.z * pr.z);
    if (live) {
        B.pred[s] = pr;
        B.pres[s] = p;
        if (B.ad) B.ad[s] = ad;
    }
    fill_cell_starts(P, s, live, k, skey, LB);
}

// ------------------------------------------------------------------ neighbour walk
// Calls f(i, dx, dy, d2) for every neighbour i != s of particle s with
// |r_i - r_s|^2 < (k h)^2 decided in FP64 without contraction, in the
// reference's cell order; stops adding after max_nbr - 1 neighbours and
// reports the overflow (GridFind.cpp:144-176).
// With kFilter (FP32 pair arithmetic) the distance test runs in FP32 on the
// FP64 differences and falls back to the exact FP64 test only inside a
// relative band of 1e-5 around the cut-off, so the neighbour sets are those
// of the FP64 test.
template <bool kFilter, class F>
__device__ __forceinline__ void for_each_neighbour(const StepParams& P, uint32_t cap, uint32_t s, uint32_t key,
                                                   double2 pj, const double2* __restrict__ pos,
                                                   const uint32_t* __restrict__ LB, bool& overflow, F&& f) {
    const int cx = (int)(key / (uint32_t)P.ncy);
    const int cy = (int)key - cx * P.ncy;
    const int y0 = cy > 0 ? cy - 1 : 0;
    const int y1 = cy + 1 < P.ncy ? cy + 1 : P.ncy - 1;
    const float cut_lo = (float)(P.cut2 * (1.0 - 1e-5)), cut_hi = (float)(P.cut2 * (1.0 + 1e-5));
    uint32_t count = 0;
#pragma unroll 1
    for (int dc = -1; dc <= 1; ++dc) {
        const int col = cx + dc;
        if (col < 0 || col >= P.ncx) continue;
        const uint32_t b = LB[col * P.ncy + y0];
        const uint32_t e = LB[col * P.ncy + y1 + 1];
        double2 nxt = b < e ? pos[b] : pj;
#pragma unroll 1
        for (uint32_t i = b; i < e; ++i) {
            const double2 pi = nxt;
            if (i + 1 < e) nxt = pos[i + 1];
            if (i == s) continue;
            const double dx = __dsub_rn(pi.x, pj.x);
            const double dy = __dsub_rn(pi.y, pj.y);
            if (kFilter) {
                const float fx = (float)dx, fy = (float)dy;
                const float f2 = fx * fx + fy * fy;
                if (f2 >= cut_hi) continue;
                if (f2 > cut_lo) {
                    const double d2 = dist2_exact(dx, dy);
                    if (!(d2 < P.cut2)) continue;
                }
            } else {
                const double d2 = dist2_exact(dx, dy);
                if (!(d2 < P.cut2)) continue;
            }
            if (count == cap) {
                overflow = true;
                continue;
            }
            ++count;
            f(i, dx, dy);
        }
    }
}

// Batched walk for the FP32/FP64 fast pair pass: candidate positions are
// loaded kBatchN at a time (independent loads in flight), tested into a bit
// mask, and the neighbour function runs over the set bits only.  Same
// neighbour set and order as for_each_neighbour.
template <bool kFilter, class F>
__device__ __forceinline__ void for_each_neighbour_batched(const StepParams& P, uint32_t cap, uint32_t s,
                                                           uint32_t key, double2 pj,
                                                           const double2* __restrict__ pos,
                                                           const uint32_t* __restrict__ LB, bool& overflow,
                                                           F&& f) {
    constexpr uint32_t kBatchN = 8;
    const int cx = (int)(key / (uint32_t)P.ncy);
    const int cy = (int)key - cx * P.ncy;
    const int y0 = cy > 0 ? cy - 1 : 0;
    const int y1 = cy + 1 < P.ncy ? cy + 1 : P.ncy - 1;
    const float cut_lo = (float)(P.cut2 * (1.0 - 1e-5)), cut_hi = (float)(P.cut2 * (1.0 + 1e-5));
    uint32_t count = 0;
#pragma unroll 1
    for (int dc = -1; dc <= 1; ++dc) {
        const int col = cx + dc;
        if (col < 0 || col >= P.ncx) continue;
        const uint32_t b = LB[col * P.ncy + y0];
        const uint32_t e = LB[col * P.ncy + y1 + 1];
#pragma unroll 1
        for (uint32_t base = b; base < e; base += kBatchN) {
            double2 q[kBatchN];
#pragma unroll
            for (uint32_t u = 0; u < kBatchN; ++u) q[u] = pos[min(base + u, e - 1)];
            uint32_t mask = 0;
#pragma unroll
            for (uint32_t u = 0; u < kBatchN; ++u) {
                const uint32_t i = base + u;
                const double dx = __dsub_rn(q[u].x, pj.x);
                const double dy = __dsub_rn(q[u].y, pj.y);
                bool in;
                if (kFilter) {
                    const float fx = (float)dx, fy = (float)dy;
                    const float f2 = fx * fx + fy * fy;
                    in = f2 < cut_hi;
                    if (in && f2 > cut_lo) in = dist2_exact(dx, dy) < P.cut2;
                } else {
                    in = dist2_exact(dx, dy) < P.cut2;
                }
                in = in && i < e && i != s;
                mask |= (in ? 1u : 0u) << u;
            }
#pragma unroll 1
            while (mask) {
                const uint32_t u = (uint32_t)__ffs((int)mask) - 1;
                mask &= mask - 1;
                if (count == cap) {
                    overflow = true;
                    continue;
                }
                ++count;
                const double2 qi = pos[base + u];
                f(base + u, __dsub_rn(qi.x, pj.x), __dsub_rn(qi.y, pj.y));
            }
        }
    }
}

// Two-phase walk: each lane tests its candidates kCollect at a time
// (independent position loads in flight, exact FP64 distance test without
// branches) and appends the hits to its own column of an LDS list (`hits`,
// stride kBlock, kChunk rows); whenever the list could overflow, and at the
// end, flush(nh) consumes hits[0 .. nh).  The wave thereby runs the pair
// body max_lanes(hits) times instead of once per candidate iteration in
// which any lane hits.  Same neighbour set, visiting order and overflow
// rule (first `cap` hits kept, the rest flagged) as for_each_neighbour.
// Eight candidates per batch: measured 3.5% faster than four at C1-1M FP32
// (two and twelve are slower, sixteen much slower).
constexpr uint32_t kCollect = 8;
template <uint32_t kChunk, class F>
__device__ __forceinline__ void collect_neighbours(const StepParams& P, uint32_t cap, uint32_t s, uint32_t key,
                                                   double2 pj, const double2* __restrict__ pos,
                                                   const uint32_t* __restrict__ LB, uint32_t* __restrict__ hits,
                                                   bool& overflow, F&& flush) {
    static_assert(kChunk >= 2 * kCollect, "hit list too short");
    const int cx = (int)(key / (uint32_t)P.ncy);
    const int cy = (int)key - cx * P.ncy;
    const int y0 = cy > 0 ? cy - 1 : 0;
    const int y1 = cy + 1 < P.ncy ? cy + 1 : P.ncy - 1;
    uint32_t nh = 0, done = 0;  // hits in the list, hits already flushed
#pragma unroll 1
    for (int c = 0; c < 3; ++c) {
        const int col = cx + c - 1;
        if (col < 0 || col >= P.ncx) continue;
        const uint32_t b = LB[col * P.ncy + y0];
        const uint32_t e = LB[col * P.ncy + y1 + 1];
#pragma unroll 1
        for (uint32_t i = b; i < e; i += kCollect) {
            double2 q[kCollect];
#pragma unroll
            for (uint32_t u = 0; u < kCollect; ++u) q[u] = pos[min(i + u, e - 1)];
            bool hit[kCollect];
#pragma unroll
            for (uint32_t u = 0; u < kCollect; ++u) {
                const double dx = __dsub_rn(q[u].x, pj.x);
                const double dy = __dsub_rn(q[u].y, pj.y);
                const double d2 = dist2_exact(dx, dy);
                hit[u] = (i + u < e) & (i + u != s) & (d2 < P.cut2);
            }
            // nh <= kChunk - kCollect here, so every slot written is in range
            if (done + nh + kCollect <= cap) {
#pragma unroll
                for (uint32_t u = 0; u < kCollect; ++u) {
                    hits[nh * kBlock] = i + u;  // kept only when it is a hit
                    nh += hit[u] ? 1u : 0u;
                }
            } else {
#pragma unroll
                for (uint32_t u = 0; u < kCollect; ++u) {
                    if (!hit[u]) continue;
                    if (done + nh < cap) {
                        hits[nh * kBlock] = i + u;
                        ++nh;
                    } else {
                        overflow = true;
                    }
                }
            }
            if (nh + kCollect > kChunk) {
                flush(nh);
                done += nh;
                nh = 0;
            }
        }
    }
    if (nh) flush(nh);
}

// ------------------------------------------------------------------ summation density
// Density.h:25-61: rho = m W(0) + sum m W_ij, p = EOS(rho); written into the
// predicted slots (forces use this rho) of pred.
template <class R>
__global__ void __launch_bounds__(kBlock) density_sum_kernel(StepParams P, Ctrl* __restrict__ ctrl, StateB<R> B,
                                                             const uint32_t* __restrict__ skey,
                                                             const uint32_t* __restrict__ LB) {
    if (ctrl->halted) return;
    const uint32_t lb = xcd_block(P.blocks_per_xcd);
    if (lb >= P.nblocks) return;
    const uint32_t s = lb * kBlock + threadIdx.x;
    if (s >= P.n) return;
    const KernelConsts<R> K = kernel_consts<R>(P);
    const R m = (R)P.mass;
    const real4<R> va = B.va[s];
    const R w0 = kernel_w<R>(K, P.skf_density, R(0));
    R rho = m * w0;
    // renormalised step: the denominator sum_i (m / rho_i) W_ji, rho_i the
    // density before the step (ModelParams.tex:124-133)
    const bool renorm = ctrl->cur_renorm != 0;
    R wsum = renorm ? m / va.z * w0 : R(0);
    bool overflow = false;
    if ((int)va.w != 0) {
        const double2 pj = B.pos[s];
        for_each_neighbour<false>(P, P.max_nbr - 1, s, skey[s], pj, B.pos, LB, overflow,
                                  [&](uint32_t i, double dx, double dy) {
                                      const R w = kernel_w<R>(K, P.skf_density, pair_dist<R>(dx, dy, 0.0));
                                      rho += m * w;
                                      if (renorm) wsum += m / B.va[i].z * w;
                                  });
    }
    if (renorm) rho = rho / wsum;
    if (overflow) atomicOr(&ctrl->nbr_overflow, 1u);
    real4<R> pr = B.pred[s];
    const R p = eos_p<R>(rho, (R)P.rho0, (R)P.eos_B);
    pr.z = rho;
    pr.w = p / (rho * rho);
    B.pred[s] = pr;
    B.pres[s] = p;
}

// ------------------------------------------------------------------ whole_step
// WholeStep.cpp:19-61 for one particle, with the RRSPH_CL boundary motion
// (cl/TimeIntegration.cl:190-193) and wave maker (:196-255) when enabled:
// the new position and (v, rho, itype); `outside` counts a particle that
// ends outside the geometry (ConsistencyCheck.cpp:162-189).
template <class R> struct Stepped2 {
    double2 pos;
    real4<R> va;
};
template <class R>
__device__ __forceinline__ Stepped2<R> whole_step_one(const StepParams& P, const Ctrl& c, double2 pos, real4<R> va,
                                                      const real4<R>& pred, uint32_t oid, R ax, R ay, R drho, R avx,
                                                      R avy, double dt, uint32_t& outside) {
    const double vdt_d = c.cur_first ? dt * 0.5 : dt;
    const R vdt = (R)vdt_d;
    int itype = (int)va.w;
    R rho = P.density_treatment != 0 ? va.z + drho * vdt : pred.z;
    R vx = va.x, vy = va.y;
    if (c.wall_disappear && oid >= P.nwm_start && oid < P.nwm_end && oid >= P.nfluid) itype = 0;
    if (itype > 0) {
        vx += ax * vdt;
        vy += ay * vdt;
        if (P.xsph_on) {
            vx += avx;
            vy += avy;
        }
        const double nx = pos.x + (double)vx * dt;
        const double ny = pos.y + (double)vy * dt;
        if (P.consistency_treatment == 2) {
            if (nx < P.xmax && nx > P.xmin && ny < P.ymax && ny > P.ymin) {
                pos.x = nx;
                pos.y = ny;
            } else {
                itype = 0;
            }
        } else {
            pos.x = nx;
            pos.y = ny;
        }
    } else if (P.boundary_motion && itype < 0) {
        if (c.piston_active && oid >= P.nwm_start && oid < P.nwm_end && oid >= P.nfluid) {
            vx = P.nwm_axis == 0 ? (R)c.piston_v : R(0);
            vy = P.nwm_axis == 1 ? (R)c.piston_v : R(0);
        }
        pos.x += (double)vx * dt;
        pos.y += (double)vy * dt;
    }
    const bool inside = pos.x < P.xmax && pos.x > P.xmin && pos.y < P.ymax && pos.y > P.ymin;
    if (!inside) outside += 1;  // NaN positions fail the comparisons too
    Stepped2<R> o;
    o.pos = pos;
    o.va.x = vx;
    o.va.y = vy;
    o.va.z = rho;
    o.va.w = (R)itype;
    return o;
}

// whole_step of slot s into state A, and the slot's key for the next sort.
template <class R>
__device__ __forceinline__ void integrate_one(const StepParams& P, const Ctrl& c, uint32_t s, double2 pos,
                                              real4<R> va, const real4<R>& pred, uint32_t oid, R ax, R ay, R drho,
                                              R avx, R avy, double dt, StateA<R> A, uint32_t* __restrict__ key_next,
                                              uint32_t key_old, uint32_t& outside, uint32_t& bad,
                                              uint32_t& changed) {
    const Stepped2<R> o = whole_step_one<R>(P, c, pos, va, pred, oid, ax, ay, drho, avx, avy, dt, outside);
    A.pos[s] = o.pos;
    A.va[s] = o.va;
    real4<R> nad;
    nad.x = ax;
    nad.y = ay;
    nad.z = drho;
    nad.w = R(0);
    A.ad[s] = nad;
    A.oid[s] = oid;
    const uint32_t key = (int)o.va.w == 0 ? P.ncells : cell_key(P, o.pos.x, o.pos.y, bad);
    key_next[s] = key;
    changed = key != key_old ? 1u : 0u;
}

// Store the ballot of the lanes whose particle changed cell, for the next
// step's incremental re-sort.  Every lane of the wave calls it, lane l with
// slot s = 64 w + l; every wave with slots below n stores its word.
__device__ __forceinline__ void store_changed_mask(const StepParams& P, uint32_t changed, uint32_t s) {
    const unsigned long long m = __ballot(changed != 0);
    if (__lane_id() == 0 && s < P.n) P.chg_mask[s / kWave] = m;
}

__device__ __forceinline__ void wave_count_add(uint32_t* dst, uint32_t v) {
    // one atomic per wave
    const unsigned long long any = __ballot(v != 0);
    if (any == 0) return;
    uint32_t sum = v;
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if (__lane_id() == 0) atomicAdd(dst, sum);
}


// Shared tail of the pair passes: dt reductions, optional export of the
// fields, and whole_step fused in when dt is known before the step.
template <class R, bool FUSED, bool EXPORT>
__device__ __forceinline__ void pair_epilogue(const StepParams& P, Ctrl* __restrict__ ctrl, uint32_t lb, uint32_t s,
                                              bool live, bool overflow, double2 pj, const real4<R>& vaj,
                                              const real4<R>& prj, uint32_t oid, R ax, R ay, R drho, R avx, R avy,
                                              R amu, StateA<R> A, uint32_t* __restrict__ key_next, uint32_t key_old,
                                              real4<R>* __restrict__ der, real4<R>* __restrict__ xs,
                                              unsigned long long* __restrict__ maxslots) {
    if (overflow) atomicOr(&ctrl->nbr_overflow, 1u);
    if (P.dt_method == 2 && P.dt_dynamic_mode == 1) {
        block_max2_to_slots(live ? sqrt((double)ax * ax + (double)ay * ay) : 0.0, live ? (double)amu : 0.0,
                            maxslots, lb);
    }
    if (EXPORT || !FUSED) {
        if (live) {
            real4<R> d;
            d.x = ax;
            d.y = ay;
            d.z = drho;
            d.w = amu;
            der[s] = d;
            real2<R> x;  // 2D: xs holds (avx, avy) pairs
            x.x = avx;
            x.y = avy;
            ((real2<R>*)xs)[s] = x;
        }
    }
    if (FUSED) {
        uint32_t outside = 0, bad = 0, changed = 0;
        if (lb == 0 && threadIdx.x == 0) {  // nobody reads these during the pass
            ctrl->dt_new = ctrl->cur_dt_const;
            ctrl->pending = 1;
        }
        if (live) {
            const Ctrl& c = *ctrl;
            integrate_one<R>(P, c, s, pj, vaj, prj, oid, ax, ay, drho, avx, avy, c.cur_dt_const, A, key_next,
                             key_old, outside, bad, changed);
        }
        wave_count_add(&ctrl->outside, outside);
        store_changed_mask(P, changed, s);  // lanes past n did not change
        if (bad) atomicOr(&ctrl->grid_overflow, 1u);
    }
}

// ------------------------------------------------------------------ pair pass
// Per particle j (sorted index s): continuity density rate (Density.h:88-113),
// acceleration = gravity + LJ + pressure + artificial viscosity
// (UpdateAcceleration.cpp:59-100) and XSPH, in one walk over the neighbours.
// FUSED: whole_step runs in the epilogue (dt known before the step).
// EXPORT: also store a, drho, av, mu in sorted order for download.
template <class R, bool FUSED, bool EXPORT>
__global__ void __launch_bounds__(kBlock) pair_kernel(StepParams P, Ctrl* __restrict__ ctrl, StateB<R> B,
                                                      const uint32_t* __restrict__ skey,
                                                      const uint32_t* __restrict__ LB, StateA<R> A,
                                                      uint32_t* __restrict__ key_next, real4<R>* __restrict__ der,
                                                      real4<R>* __restrict__ xs,
                                                      unsigned long long* __restrict__ maxslots) {
    if (ctrl->halted) return;
    const uint32_t lb = xcd_block(P.blocks_per_xcd);
    if (lb >= P.nblocks) return;
    const uint32_t s = lb * kBlock + threadIdx.x;
    const bool live = s < P.n;
    const KernelConsts<R> K = kernel_consts<R>(P);
    const R m = (R)P.mass;
    R ax = R(0), ay = -(R)P.g, drho = R(0), avx = R(0), avy = R(0), amu = R(0);
    bool overflow = false;
    double2 pj = make_double2(0, 0);
    real4<R> vaj{}, prj{};
    uint32_t oid = 0;
    if (live) {
        pj = B.pos[s];
        vaj = B.va[s];
        prj = B.pred[s];
        oid = B.oid[s];
    }
    const int tj = (int)vaj.w;
    if (live && tj != 0) {
        const bool cont = P.density_treatment != 0;
        const bool delta = P.density_treatment == 2;
        const R h = K.h;
        const R rhoj = vaj.z, rhohj = prj.z, pjres = B.pres[s];
        const R pj_rho2 = pjres / (rhohj * rhohj);
        const R alpha = (R)P.av_alpha, beta = (R)P.av_beta, cs = (R)P.av_c, eta2 = (R)P.av_eta2;
        const R two_eps = R(2) * (R)P.xsph_eps;
        const R lj_r0 = (R)P.lj_r0, lj_D = (R)P.lj_D;
        const R dhc2m = R(2) * (R)P.dhc * m;
        for_each_neighbour<false>(P, P.max_nbr - 1, s, skey[s], pj, B.pos, LB, overflow,
                           [&](uint32_t i, double dxd, double dyd) {
            const double d2d = dxd * dxd + dyd * dyd;
            const real4<R> vai = B.va[i];
            const real4<R> pri = B.pred[i];
            const R dx = (R)dxd, dy = (R)dyd;  // diff = r_i - r_j
            const R d2 = (R)d2d;
            const R dist = sqrt(d2);
            const int ti = (int)vai.w;
            // continuity density with v and rho of the current step
            if (cont) {
                const R F = kernel_gradf<R>(K, P.skf_density, dist);
                const R gx = dx * F, gy = dy * F;
                drho += m * ((vai.x - vaj.x) * gx + (vai.y - vaj.y) * gy);
                if (delta) {
                    const R rf = -(dx * gx + dy * gy) / d2;
                    const R pf = (vai.z - rhoj) / vai.z;
                    drho += dhc2m * (pf * rf);
                }
            }
            // Lennard-Jones wall force on fluid from boundary particles
            if (P.lj_on && tj > 0 && ti < 0 && dist < lj_r0) {
                const R rn = lj_r0 / dist;
                const R rn2 = rn * rn, rn4 = rn2 * rn2;
                const R f = lj_D * (rn4 * rn4 * rn4 - rn4) / d2;
                ax -= dx * f;
                ay -= dy * f;
            }
            // pressure (predicted density, pressure of the current density)
            {
                const R F = kernel_gradf<R>(K, P.skf_intf, dist);
                const R rhohi = pri.z, pi = B.pres[i];
                R pf = P.intf_approx == 1 ? (pjres + pi) / (rhohj * rhohi) : pi / (rhohi * rhohi) + pj_rho2;
                if (P.art_pressure && (pi < R(0) || pjres < R(0))) {
                    const R fij = kernel_w<R>(K, P.skf_ap, dist) / (R)P.ap_wdelta;
                    const R api = pi < R(0) ? (-(R)P.ap_coef * pi / (rhohi * rhohi)) : R(0);
                    const R apj = pjres < R(0) ? (-(R)P.ap_coef * pjres / (rhohj * rhohj)) : R(0);
                    pf += (api + apj) * pow(fij, (R)P.ap_index);
                }
                const R c = F * pf * m;
                ax += dx * c;
                ay += dy * c;
            }
            // artificial viscosity with the predicted velocity and density
            if (P.av_on) {
                const R dvx = prj.x - pri.x, dvy = prj.y - pri.y;
                const R vr = -(dvx * dx + dvy * dy);  // (v_j - v_i) . (r_j - r_i)
                if (vr < R(0)) {
                    const R mu = h * vr / (d2 + eta2);
                    amu = fmax(amu, fabs(mu));
                    const R rhoij = R(0.5) * (pri.z + rhohj);
                    const R piv = (beta * mu - alpha * cs) * mu / rhoij;
                    const R F = kernel_gradf<R>(K, P.skf_av, dist);
                    const R c = F * piv * m;
                    ax += dx * c;
                    ay += dy * c;
                }
            }
            // XSPH with fluid neighbours only
            if (P.xsph_on && ti > 0) {
                const R w = kernel_w<R>(K, P.skf_xsph, dist);
                const R c = m / (pri.z + rhohj) * w * two_eps;
                avx += (pri.x - prj.x) * c;
                avy += (pri.y - prj.y) * c;
            }
        });
    }
    pair_epilogue<R, FUSED, EXPORT>(P, ctrl, lb, s, live, overflow, pj, vaj, prj, oid, ax, ay, drho, avx, avy, amu, A,
                                    key_next, live ? skey[s] : 0u, der, xs, maxslots);
}

// Reciprocal and reciprocal square root.  FP32: the hardware v_rcp / v_rsq
// (1 ulp).  FP64: the hardware estimate refined by two Newton-Raphson steps
// (y' = y + y (1 - x y) and y' = y + y (1 - x y^2) / 2), which recovers the
// last bits of the double result.
template <class R> __device__ __forceinline__ R fast_rcp(R x);
template <> __device__ __forceinline__ float fast_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ __forceinline__ double fast_rcp<double>(double x) {
    double y = __builtin_amdgcn_rcp(x);
    y = fma(y, fma(-x, y, 1.0), y);
    y = fma(y, fma(-x, y, 1.0), y);
    return y;
}
template <class R> __device__ __forceinline__ R fast_rsqrt(R x);
template <> __device__ __forceinline__ float fast_rsqrt<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <> __device__ __forceinline__ double fast_rsqrt<double>(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    y = fma(y, fma(-hx * y, y, 0.5), y);
    y = fma(y, fma(-hx * y, y, 0.5), y);
    return y;
}

// Pair pass for the common configuration: cubic spline for every term,
// pressure in the p_i / rho_i^2 + p_j / rho_j^2 form without artificial
// pressure, summation or plain continuity density.  Same physics as
// pair_kernel, restructured for instruction count: one reciprocal square
// root gives 1/|r| and |r|; W and grad W share q and are selected without
// branches; pressure and artificial viscosity share one coefficient on
// r_i - r_j; 1 / (rho_i + rho_j) serves artificial viscosity and XSPH; the
// mass factors are applied once at the end.
template <class R> struct FastPair {
    R h, inv_h, c_q1, c_q2, w_q1, w_q2, alpha_c, beta, eta2, lj_r0, lj_D;
    bool cont, lj_on, av_on, xsph_on;
    __device__ explicit FastPair(const StepParams& P)
        : h((R)P.h), inv_h((R)(1.0 / P.h)),
          c_q1((R)(P.f_cubic / (P.h * P.h))),   // grad W = diff * c_q1 (1.5 q - 2), q <= 1
          c_q2((R)(0.5 * P.f_cubic / P.h)),     // grad W = -diff * c_q2 (2 - q)^2 / |r|, 1 < q <= 2
          w_q1((R)P.f_cubic), w_q2((R)(P.f_cubic / 6.0)), alpha_c((R)(P.av_alpha * P.av_c)), beta((R)P.av_beta),
          eta2((R)P.av_eta2), lj_r0((R)P.lj_r0), lj_D((R)P.lj_D), cont(P.density_treatment != 0),
          lj_on(P.lj_on), av_on(P.av_on), xsph_on(P.xsph_on) {}
};
template <class R> struct PairSums {
    R sx = 0, sy = 0, sd = 0, xx = 0, xy = 0, amu = 0, ljx = 0, ljy = 0;
};

// One neighbour i of particle j: dx, dy = r_i - r_j.
template <class R>
__device__ __forceinline__ void fast_pair_term(const FastPair<R>& K, PairSums<R>& S, R dx, R dy,
                                               const real4<R>& vai, const real4<R>& pri, const real4<R>& vaj,
                                               const real4<R>& prj, int tj) {
    const R d2 = dx * dx + dy * dy;
    // coincident particles (|r| = 0) contribute grad W = 0 as in the
    // reference, where q = 0 selects the inner branch
    const R rinv = d2 > R(0) ? fast_rsqrt<R>(d2) : R(0);
    const R dist = d2 * rinv;
    const R q = dist * K.inv_h;
    const R t = R(2) - q;
    const R t2 = t * t;
    const bool inner = q <= R(1);
    const R F = inner ? K.c_q1 * (R(1.5) * q - R(2)) : -K.c_q2 * t2 * rinv;
    const R W = inner ? K.w_q1 * (R(2) / R(3) - q * q + R(0.5) * q * q * q) : K.w_q2 * t2 * t;
    if (K.cont) S.sd += F * ((vai.x - vaj.x) * dx + (vai.y - vaj.y) * dy);
    const int ti = (int)vai.w;
    if (K.lj_on && tj > 0 && ti < 0 && dist < K.lj_r0) {
        const R rn = K.lj_r0 / dist;  // |r| = 0 gives a non-finite force, as the reference
        const R rn2 = rn * rn, rn4 = rn2 * rn2;
        const R f = K.lj_D * (rn4 * rn4 * rn4 - rn4) / d2;
        S.ljx -= dx * f;
        S.ljy -= dy * f;
    }
    const R inv_rs = fast_rcp<R>(pri.z + prj.z);  // 1 / (rho^_i + rho^_j)
    R coef = pri.w + prj.w;                         // p_i / rho_i^2 + p_j / rho_j^2
    const R dvx = pri.x - prj.x, dvy = pri.y - prj.y;
    if (K.av_on) {
        const R vr = dvx * dx + dvy * dy;  // (v_j - v_i) . (r_j - r_i)
        const R mu = K.h * vr * fast_rcp<R>(d2 + K.eta2);
        const bool app = vr < R(0);
        S.amu = app ? fmax(S.amu, -mu) : S.amu;
        coef += app ? (K.beta * mu - K.alpha_c) * mu * R(2) * inv_rs : R(0);
    }
    const R fc = F * coef;
    S.sx += fc * dx;
    S.sy += fc * dy;
    if (K.xsph_on) {
        const R cw = ti > 0 ? W * inv_rs : R(0);
        S.xx += dvx * cw;
        S.xy += dvy * cw;
    }
}

// LDS hit-list rows per lane of the two-phase walk (32 KiB per block)
template <class R> constexpr uint32_t kPairChunk = 32u;

template <class R, bool FUSED, bool EXPORT, bool COMPACT = true>
__global__ void __launch_bounds__(kBlock) pair_fast_kernel(StepParams P, Ctrl* __restrict__ ctrl, StateB<R> B,
                                                           const uint32_t* __restrict__ skey,
                                                           const uint32_t* __restrict__ LB, StateA<R> A,
                                                           uint32_t* __restrict__ key_next,
                                                           real4<R>* __restrict__ der, real4<R>* __restrict__ xs,
                                                           unsigned long long* __restrict__ maxslots) {
    __shared__ uint32_t hits[COMPACT ? kPairChunk<R> * kBlock : 1];
    if (ctrl->halted) return;
    const uint32_t lb = xcd_block(P.blocks_per_xcd);
    if (lb >= P.nblocks) return;
    const uint32_t s = lb * kBlock + threadIdx.x;
    const bool live = s < P.n;
    constexpr bool kF32 = sizeof(R) == 4;
    const FastPair<R> K(P);
    PairSums<R> S;
    bool overflow = false;
    double2 pj = make_double2(0, 0);
    real4<R> vaj{}, prj{};
    uint32_t oid = 0;
    if (live) {
        pj = B.pos[s];
        vaj = B.va[s];
        prj = B.pred[s];
        oid = B.oid[s];
    }
    const int tj = (int)vaj.w;
    if (live && tj != 0) {
        if (COMPACT) {
            uint32_t* my = hits + threadIdx.x;
            // consume hits[0 .. nh): the loads of hit u + 1 are issued before
            // the arithmetic of hit u, the list index of u + 2 before those
            auto flush = [&](uint32_t nh) {
                const uint32_t last = nh - 1;
                uint32_t i1 = my[min(1u, last) * kBlock];
                const uint32_t i0 = my[0];
                double2 qn = B.pos[i0];
                real4<R> van = B.va[i0], prn = B.pred[i0];
#pragma unroll 1
                for (uint32_t u = 0; u < nh; ++u) {
                    const double2 q = qn;
                    const real4<R> vai = van, pri = prn;
                    qn = B.pos[i1];
                    van = B.va[i1];
                    prn = B.pred[i1];
                    i1 = my[min(u + 2, last) * kBlock];
                    fast_pair_term<R>(K, S, (R)__dsub_rn(q.x, pj.x), (R)__dsub_rn(q.y, pj.y), vai, pri, vaj, prj,
                                      tj);
                }
            };
            collect_neighbours<kPairChunk<R>>(P, P.max_nbr - 1, s, skey[s], pj, B.pos, LB, my, overflow, flush);
        } else {
            auto body = [&](uint32_t i, double dxd, double dyd) {
                fast_pair_term<R>(K, S, (R)dxd, (R)dyd, B.va[i], B.pred[i], vaj, prj, tj);
            };
            if (kF32)
                for_each_neighbour<true>(P, P.max_nbr - 1, s, skey[s], pj, B.pos, LB, overflow, body);
            else
                for_each_neighbour_batched<false>(P, P.max_nbr - 1, s, skey[s], pj, B.pos, LB, overflow, body);
        }
    }
    const R m = (R)P.mass;
    const R ax = S.ljx + m * S.sx;
    const R ay = -(R)P.g + S.ljy + m * S.sy;
    const R drho = m * S.sd;
    const R xf = m * R(2) * (R)P.xsph_eps;
    pair_epilogue<R, FUSED, EXPORT>(P, ctrl, lb, s, live, overflow, pj, vaj, prj, oid, ax, ay, drho, xf * S.xx,
                                    xf * S.xy, S.amu, A, key_next, live ? skey[s] : 0u, der, xs, maxslots);
}

// ------------------------------------------------------------------ integrate (dynamic dt)
// dt = CFL min(sqrt(h / max|a|), h / (c + max mu)) (UpdateAcceleration.cpp:15-29)
template <class R>
__global__ void __launch_bounds__(kBlock) integrate_kernel(StepParams P, Ctrl* __restrict__ ctrl, StateB<R> B,
                                                           const real4<R>* __restrict__ der,
                                                           const real4<R>* __restrict__ xs, StateA<R> A,
                                                           uint32_t* __restrict__ key_next,
                                                           const unsigned long long* __restrict__ maxslots,
                                                           const uint32_t* __restrict__ skey) {
    Ctrl& c = *ctrl;
    if (c.halted) return;
    double dt;
    if (P.dt_method == 2 && P.dt_dynamic_mode == 1) {
        double amax, mumax;
        slots_max2(maxslots, amax, mumax);
        dt = P.cfl * fmin(sqrt(P.h / amax), P.h / (P.c0 + mumax));
    } else {
        dt = c.cur_dt_const;
    }
    const bool never = P.dt_method == 2 && dt <= P.never_ending;
    const bool nbr_err = P.consistency_check && c.nbr_overflow;
    if (never || nbr_err) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            c.halted = 1;
            c.status = nbr_err ? 3u : 4u;
        }
        return;
    }
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    uint32_t outside = 0, bad = 0, changed = 0;
    if (s < P.n) {
        const real4<R> d = der[s];
        const real2<R> x = ((const real2<R>*)xs)[s];
        // the predicted state is needed for the summed density only
        const real4<R> pr = P.density_treatment != 0 ? real4<R>{} : B.pred[s];
        integrate_one<R>(P, c, s, B.pos[s], B.va[s], pr, B.oid[s], d.x, d.y, d.z, x.x, x.y, dt, A, key_next,
                         skey[s], outside, bad, changed);
    }
    wave_count_add(&c.outside, outside);
    store_changed_mask(P, changed, s);
    if (bad) atomicOr(&c.grid_overflow, 1u);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c.dt_new = dt;
        c.pending = 1;
    }
}

// ------------------------------------------------------------------ utilities
template <class R>
__global__ void init_kernel(StepParams P, const double* __restrict__ r, const double* __restrict__ v,
                            const double* __restrict__ rho, const int32_t* __restrict__ itype, StateA<R> A,
                            uint32_t* __restrict__ key_next, Ctrl* __restrict__ ctrl) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.n) return;
    double2 p = make_double2(r[2 * s], r[2 * s + 1]);
    A.pos[s] = p;
    real4<R> va;
    va.x = (R)v[2 * s];
    va.y = (R)v[2 * s + 1];
    va.z = (R)rho[s];
    va.w = (R)itype[s];
    A.va[s] = va;
    real4<R> z{};
    A.ad[s] = z;
    A.oid[s] = s;
    uint32_t bad = 0;
    key_next[s] = itype[s] == 0 ? P.ncells : cell_key(P, p.x, p.y, bad);
    if (bad) atomicOr(&ctrl->grid_overflow, 1u);
}

// Un-permute the state into original order (AoS doubles).
template <class R>
__global__ void download_kernel(StepParams P, StateA<R> A, const R* __restrict__ pres, double* __restrict__ r,
                                double* __restrict__ v, double* __restrict__ rho, double* __restrict__ p,
                                int32_t* __restrict__ itype) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.n) return;
    const uint32_t o = A.oid[s];
    const double2 pos = A.pos[s];
    const real4<R> va = A.va[s];
    r[2 * o] = pos.x;
    r[2 * o + 1] = pos.y;
    v[2 * o] = (double)va.x;
    v[2 * o + 1] = (double)va.y;
    rho[o] = (double)va.z;
    itype[o] = (int32_t)va.w;
    p[o] = pres ? (double)pres[s] : (double)eos_p<R>(va.z, (R)P.rho0, (R)P.eos_B);
}

template <class R>
__global__ void fields_kernel(uint32_t n, const uint32_t* __restrict__ oid, const real4<R>* __restrict__ der,
                              const real4<R>* __restrict__ xs, double* __restrict__ a, double* __restrict__ drho,
                              double* __restrict__ av, double* __restrict__ mu) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t o = oid[s];
    const real4<R> d = der[s];
    const real2<R> x = ((const real2<R>*)xs)[s];  // 2D: (avx, avy) pairs
    a[2 * o] = (double)d.x;
    a[2 * o + 1] = (double)d.y;
    drho[o] = (double)d.z;
    mu[o] = (double)d.w;
    av[2 * o] = (double)x.x;
    av[2 * o + 1] = (double)x.y;
}

// Morton cell key of the reference (GridUtils.h:46-151) per particle.
__global__ void morton_kernel(StepParams P, const double2* __restrict__ pos, const uint32_t* __restrict__ oid,
                              uint32_t* __restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.n) return;
    const double2 p = pos[s];
    const unsigned long long cx = (unsigned long long)((p.x - P.xmin) / P.cell);
    const unsigned long long cy = (unsigned long long)((p.y - P.ymin) / P.cell);
    out[oid[s]] = split2((uint32_t)cx) | (split2((uint32_t)cy) << 1);
}

// Neighbour lists in the reference layout (original indices, sentinel n).
template <class R>
__global__ void neighbours_kernel(StepParams P, StateB<R> B, const uint32_t* __restrict__ skey,
                                  const uint32_t* __restrict__ LB, uint32_t* __restrict__ out, uint32_t max_nbr,
                                  Ctrl* __restrict__ ctrl) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.n) return;
    const uint32_t o = B.oid[s];
    uint32_t cnt = 0;
    bool overflow = false;
    if ((int)B.va[s].w != 0) {
        for_each_neighbour<false>(P, max_nbr - 1, s, skey[s], B.pos[s], B.pos, LB, overflow,
                                  [&](uint32_t i, double, double) {
            out[cnt + (size_t)max_nbr * o] = B.oid[i];
            ++cnt;
        });
    }
    out[cnt + (size_t)max_nbr * o] = P.n;
    if (overflow) atomicOr(&ctrl->nbr_overflow, 1u);
}

}  // namespace rrsph
