// Neighbour-list steps (2D): the cell sort and the cell walk once every
// several steps instead of every step.
//
// A build step sorts the particles by cell, reorders the state, and walks
// the cells within reach k h + skin of each particle (list_build_kernel):
// every particle i with |r_i - r_j| < k h + skin goes into j's list, in the
// walk order, k-major (entry k of slot j at nbl[k n + j], count nbc[j]).
// Until the next build the slot order stays as it is, and a step reads the
// lists instead of walking the cells: the pair pass decides each entry with
// the reference's FP64 cut-off test (GridFind.cpp:144-176, |r|^2 < (k h)^2
// with separate roundings), so the neighbour sets are those of the cell
// walk.  A pair closer than k h now was closer than k h + skin at the build
// as long as no particle has moved skin / 2 since: the integrator reduces
// each step's largest displacement, the prologue sums them, and a step
// whose lists may be stale halts before it reads them (status
// kStatusRebuild); the host re-runs it as a build step.  The host schedules
// builds ahead from the measured displacement rate, so such halts are rare.
//
// What the pair pass gathers per entry, one record of two halves per slot
// (32 bytes in FP32, 64 in FP64, so one cache line):
//   near  FP32: (xi_x, xi_y, vx, vy), xi the position relative to the
//         origin of the particle's home cell (its cell at the last build),
//         rounded from the FP64 position; FP64: (x, y, vx, vy), the FP64
//         position itself
//   pred  (v^x, v^y, s rho^, p / rho^2): predicted velocity and density,
//         s = +1 for fluid, -1 for boundary, 0 for a particle that no
//         longer exists
// An entry holds the neighbour's slot in its low 26 bits and the offset of
// the neighbour's home cell from the particle's, (dcx + 3) | (dcy + 3) << 3,
// in its high 6, so that in FP32 r_i - r_j = xi_i - xi_j + (dcx, dcy) cell
// to a few 1e-9 m.  FP32 decides an entry from that difference unless its
// squared distance lies within 1e-5 of the cut-off (or a position lies far
// from its home cell), where it takes the FP64 positions and the reference's
// test; the rounding of xi stays far below that band (2e-6 of the cut-off).
//
// Step kernels: list_prologue -> [density_sum_list] -> pair_list ->
// integrate_list (in place; it also writes the next step's records, so no
// reorder runs between builds).  A build step runs list_prologue -> radix
// sort -> reorder_list_kernel -> list_build_kernel -> the same step kernels
// on the reordered copy, integrating into A.
#pragma once

#include "device_common.hpp"
#include "step_kernels.hpp"

namespace rrsph {

constexpr uint32_t kStatusRebuild = 101;       // internal: re-run this step as a list build
constexpr uint32_t kStatusListOverflow = 102;  // internal: a list ran out of room; run without lists

// The state a list step reads (the reordered copy B on a build step, A after).
constexpr uint32_t kSlotBits = 26;
constexpr uint32_t kSlotMask = (1u << kSlotBits) - 1;  // list steps take at most 2^26 particles

template <class R> struct ListState {
    const double2* pos;
    const real4<R>* va;
    const uint32_t* oid;
};

// The record of a slot: near and pred halves.
template <class R> struct alignas(2 * sizeof(real4<R>)) ListRec {
    real4<R> nr;
    real4<R> pr;
};

// The sign-encoded predicted density of the pred record.
template <class R> __device__ __forceinline__ R typed_rho(R rho, int itype) {
    return itype > 0 ? rho : (itype < 0 ? -rho : R(0));
}

// The near record of a particle at pos with velocity (vx, vy) and home cell
// `home` (a cell key cx ncy + cy).
template <class R> __device__ __forceinline__ real4<R> near_record(const StepParams& P, double2 pos, R vx, R vy,
                                                                   uint32_t home);
template <> __device__ __forceinline__ float4 near_record<float>(const StepParams& P, double2 pos, float vx, float vy,
                                                                 uint32_t home) {
    const uint32_t cx = home / (uint32_t)P.ncy, cy = home - cx * (uint32_t)P.ncy;
    const double ox = P.xmin + (double)cx * P.cell, oy = P.ymin + (double)cy * P.cell;
    return make_float4((float)(pos.x - ox), (float)(pos.y - oy), vx, vy);
}
template <> __device__ __forceinline__ double4 near_record<double>(const StepParams&, double2 pos, double vx,
                                                                   double vy, uint32_t) {
    return make_double4(pos.x, pos.y, vx, vy);
}

// The prediction of PredictHalfStep.cpp:16-22 (v^ = v + a dt/2, rho^ = rho +
// drho dt/2 with continuity) and the pressure of the current density
// (Density.h:112) as reorder_kernel computes them; p / rho^2 in w.
template <class R> __device__ __forceinline__ real4<R> predict_record(const StepParams& P, const real4<R>& va,
                                                                      R ax, R ay, R drho, R dt, R& p) {
    const bool cont = P.density_treatment != 0;
    real4<R> pr;
    pr.x = va.x + ax * dt * R(0.5);
    pr.y = va.y + ay * dt * R(0.5);
    pr.z = cont ? va.z + drho * dt * R(0.5) : va.z;
    p = eos_p<R>(va.z, (R)P.rho0, (R)P.eos_B);
    pr.w = p / (pr.z * pr.z);
    pr.z = typed_rho<R>(pr.z, (int)va.w);
    return pr;
}

// The reorder of a build step: reorder_kernel's gather, prediction,
// pressure and cell-start table, plus the records of the list pass and each
// slot's home cell (its cell key).
template <class R>
__global__ void __launch_bounds__(kBlock) reorder_list_kernel(StepParams P, const Ctrl* __restrict__ ctrl,
                                                              const uint32_t* __restrict__ skey,
                                                              const uint32_t* __restrict__ perm, StateA<R> A,
                                                              StateB<R> B, ListRec<R>* __restrict__ rec,
                                                              uint32_t* __restrict__ home, uint32_t* __restrict__ LB,
                                                              unsigned long long* __restrict__ maxslots) {
    if (ctrl->halted) return;
    if (blockIdx.x == 0) {
        maxslots[threadIdx.x] = 0;
        maxslots[kMaxSlots + threadIdx.x] = 0;
    }
    const uint32_t s0 = blockIdx.x * kBlock + threadIdx.x;
    const bool live = s0 < P.n;
    const uint32_t s = live ? s0 : P.n - 1;
    const uint32_t k = skey[s];
    const uint32_t src = min(perm[s], P.n - 1);  // a permutation of [0, n); the clamp only guards memory
    const double2 pos = A.pos[src];
    const real4<R> va = A.va[src];
    const real4<R> ad = A.ad[src];
    const uint32_t oid = A.oid[src];
    R p;
    const real4<R> pr = predict_record<R>(P, va, ad.x, ad.y, ad.z, (R)ctrl->cur_dt, p);
    if (live) {
        B.pos[s] = pos;
        B.va[s] = va;
        B.oid[s] = oid;
        B.pres[s] = p;
        ListRec<R> r;
        r.nr = near_record<R>(P, pos, va.x, va.y, k);
        r.pr = pr;
        rec[s] = r;
        home[s] = k;
    }
    fill_cell_starts(P, s, live, k, skey, LB);
}

// Prologue of a list step or a build step (one block of kBlock threads):
// the previous step's largest displacement (reduced from its slots), the
// commit of the previous step, the staleness test of the lists (a list
// step also halts when no build has run since the last sort, e.g. a build
// launched while the run was stopped), and the clearing of this step's
// reduction slots.
__global__ void __launch_bounds__(kBlock) list_prologue(StepParams P, Ctrl* __restrict__ ctrl,
                                                        unsigned long long* __restrict__ dslots,
                                                        unsigned long long* __restrict__ maxslots, uint32_t build) {
    __shared__ double s_m[kBlock / kWave];
    const unsigned long long word = ctrl_word(ctrl);
    double d = dbits(dslots[threadIdx.x]);
    for (int off = 32; off > 0; off >>= 1) d = fmax(d, __shfl_xor(d, off));
    if (__lane_id() == 0) s_m[threadIdx.x / kWave] = d;
    __syncthreads();
    uint32_t valid = 0;
    block_prologue(
        P, ctrl, word,
        [&](Ctrl& c) {
            double dmax = s_m[0];
            for (int w = 1; w < kBlock / kWave; ++w) dmax = fmax(dmax, s_m[w]);
            if (!c.halted && c.pending) {  // the step being committed moved no particle further than dmax
                c.disp_acc += dmax;
                c.disp_last = dmax;
            }
            valid = c.list_valid;
        },
        [&](Ctrl& c) {
            if (c.halted) return;
            c.list_valid = build ? 1u : valid;
            if (build) {
                c.disp_acc = 0.0;
                c.list_overflow = 0;
            } else if (!valid || !(c.disp_acc < P.skin_half)) {
                c.halted = 1;
                c.status = kStatusRebuild;
                if (c.wall_disappear) c.nwm_done = 0;  // the re-run prologue fires it again
            }
        });
    maxslots[threadIdx.x] = 0;
    maxslots[kMaxSlots + threadIdx.x] = 0;
    dslots[threadIdx.x] = 0;
    dslots[kMaxSlots + threadIdx.x] = 0;
}

// Cell index (clamped to the grid) of coordinate r.
__device__ __forceinline__ int clamped_cell(double r, double rmin, double cell, int ncell) {
    const double t = floor((r - rmin) / cell);
    return t < 0.0 ? 0 : (t >= (double)ncell ? ncell - 1 : (int)t);
}

// The lists of a build step, on the reordered copy (pos, va) with its
// cell-start table LB: for each live particle, every live particle i != s
// within list_reach, in the order of the cell walk (columns, then rows
// within a column).  The cells visited are those within list_reach of the
// particle (3 or 4 per axis, the reach being under two cells).
template <class R>
__global__ void __launch_bounds__(kBlock) list_build_kernel(StepParams P, Ctrl* __restrict__ ctrl,
                                                            const double2* __restrict__ pos,
                                                            const real4<R>* __restrict__ va,
                                                            const uint32_t* __restrict__ skey,
                                                            const uint32_t* __restrict__ LB, uint32_t* __restrict__ nbl,
                                                            uint32_t* __restrict__ nbc) {
    if (ctrl->halted) return;
    const uint32_t lb = xcd_block(P.blocks_per_xcd);
    if (lb >= P.nblocks) return;
    const uint32_t s = lb * kBlock + threadIdx.x;
    if (s >= P.n) return;
    uint32_t cnt = 0;
    bool bad = false;
    if ((int)va[s].w != 0) {
        const double2 pj = pos[s];
        const uint32_t kj = skey[s];
        const int hx = (int)(kj / (uint32_t)P.ncy), hy = (int)kj - hx * P.ncy;
        // the list pass takes the FP32 positions relative to the home cell
        // for within 2.1 cells of its origin: a particle outside the cell
        // grid (keyed to the nearest edge cell) runs the step without lists
        const double ox = pj.x - (P.xmin + hx * P.cell), oy = pj.y - (P.ymin + hy * P.cell);
        bad = !(ox > -P.cell && ox < 2.0 * P.cell && oy > -P.cell && oy < 2.0 * P.cell);
        const double reach = P.list_reach * (1.0 + 1e-9);
        const int x0 = clamped_cell(pj.x - reach, P.xmin, P.cell, P.ncx);
        const int x1 = clamped_cell(pj.x + reach, P.xmin, P.cell, P.ncx);
        const int y0 = clamped_cell(pj.y - reach, P.ymin, P.cell, P.ncy);
        const int y1 = clamped_cell(pj.y + reach, P.ymin, P.cell, P.ncy);
#pragma unroll 1
        for (int col = x0; col <= x1; ++col) {
            const uint32_t b = LB[col * P.ncy + y0];
            const uint32_t e = LB[col * P.ncy + y1 + 1];
#pragma unroll 1
            for (uint32_t i = b; i < e; i += kCollect) {
                double2 q[kCollect];
#pragma unroll
                for (uint32_t u = 0; u < kCollect; ++u) q[u] = pos[min(i + u, e - 1)];
#pragma unroll
                for (uint32_t u = 0; u < kCollect; ++u) {
                    const double dx = q[u].x - pj.x, dy = q[u].y - pj.y;
                    const bool in = (i + u < e) & (i + u != s) & (dx * dx + dy * dy < P.list_r2);
                    if (in) {
                        // the offset of the neighbour's home cell: column col, and the
                        // row of its key, from the same position by the same rule
                        // (cell_key), without loading the key
                        uint32_t ybad = 0;
                        const int dcx = col - hx, dcy = cell_coord(q[u].y, P.ymin, P.cell, P.ncy, ybad) - hy;
                        bad |= dcx < -3 || dcx > 4 || dcy < -3 || dcy > 4;
                        const uint32_t code = (uint32_t)(dcx + 3) | ((uint32_t)(dcy + 3) << 3);
                        if (cnt < P.list_cap) nbl[(size_t)cnt * P.n + s] = (i + u) | (code << kSlotBits);
                        ++cnt;
                    }
                }
            }
        }
    }
    if (cnt > P.list_cap || bad) atomicOr(&ctrl->list_overflow, 1u);
    nbc[s] = min(cnt, P.list_cap);
}

// One list entry as the pair pass sees it.
template <class R> struct ListEntry {
    bool take;        // a neighbour: within the cut-off, alive, among the first `cap`
    R dx, dy;         // r_i - r_j
    real4<R> nr;      // near record of i: (xi or x, xi or y, vx, vy)
    real4<R> pr;      // pred record of i: (v^x, v^y, s rho^, p / rho^2)
};

// r_i - r_j of an entry and whether it is within the cut-off.  FP32: from
// the near records and the offset of the two home cells (a table in LDS
// indexed by the entry's code), the reference's FP64 test on the FP64
// positions only inside the band around the cut-off.  The build guarantees
// that every listed position lies within a cell of its home cell, and the
// lists go stale before any particle moves half the skin, so xi stays below
// 2.1 cells and its rounding far inside the band.  FP64: the reference's test
// on the FP64 positions of the near records.
struct CutFP32 {
    float lo, hi;
};
__device__ __forceinline__ CutFP32 cut_band(const StepParams& P) {
    CutFP32 C;
    C.lo = (float)(P.cut2 * (1.0 - 1e-5));
    C.hi = (float)(P.cut2 * (1.0 + 1e-5));
    return C;
}
__device__ __forceinline__ bool list_cut(const StepParams& P, const CutFP32& C, const float2* s_off, uint32_t e,
                                         const float4& ni, const float4& nj, double2 pj,
                                         const double2* __restrict__ pos, float& dx, float& dy) {
    const float2 o = s_off[e >> kSlotBits];
    dx = (ni.x - nj.x) + o.x;
    dy = (ni.y - nj.y) + o.y;
    const float d2 = dx * dx + dy * dy;
    if (!(d2 > C.lo && d2 < C.hi)) return d2 <= C.lo;  // on the band's edge: within a rounding of lo < cut2
    const double2 q = pos[e & kSlotMask];  // rare: the band around the cut-off
    const double ddx = __dsub_rn(q.x, pj.x), ddy = __dsub_rn(q.y, pj.y);
    dx = (float)ddx;
    dy = (float)ddy;
    return dist2_exact(ddx, ddy) < P.cut2;
}
__device__ __forceinline__ bool list_cut(const StepParams& P, const CutFP32&, const float2*, uint32_t,
                                         const double4& ni, const double4& nj, double2, const double2* __restrict__,
                                         double& dx, double& dy) {
    dx = __dsub_rn(ni.x, nj.x);
    dy = __dsub_rn(ni.y, nj.y);
    return dist2_exact(dx, dy) < P.cut2;
}

// The table of home-cell offsets of the entry codes, code = (dcx + 3) |
// (dcy + 3) << 3 -> (dcx cell, dcy cell) rounded to FP32, filled by the
// first 64 threads of the block (a barrier must follow).
__device__ __forceinline__ void fill_cell_offsets(const StepParams& P, float2* s_off) {
    if (threadIdx.x < 64) {
        const int c = (int)threadIdx.x;
        s_off[c] = make_float2((float)(((c & 7) - 3) * P.cell), (float)(((c >> 3) - 3) * P.cell));
    }
}

// Entry k of slot s's list: the lists hold list_cap + kListPad rows, zero
// where no build wrote (so the loads of the entries after a list's last one
// stay in bounds and name valid slots), and the byte offsets fit 32 bits.
constexpr uint32_t kListPad = 8;
__device__ __forceinline__ uint32_t list_entry(const uint32_t* __restrict__ nbl, uint32_t byte_off) {
    return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(nbl) + byte_off);
}

// Visits the entries of slot s's list: f(ListEntry) for every entry (f must
// add nothing when take is false: every entry runs the same instructions,
// most entries being neighbours).  kCap: count the neighbours, and entries
// beyond the first `cap` set `overflow` (GridFind.cpp:158-188); a wave
// whose lists are all no longer than cap cannot overflow and skips the
// count.  Entries go in groups of G: the group's records are gathered
// together and the next group's entries loaded behind them, so that no
// gathered value lives across iterations (the loop waits only on its own
// loads).  Larger groups keep more gathers in flight per wave but take
// registers: the pass is bound by load latency, and 2 entries per group
// (74 VGPRs, 6 waves per SIMD) gave the least time.
template <class R, bool kCap, class F>
__device__ __forceinline__ void for_each_listed(const StepParams& P, const float2* s_off, uint32_t cap, uint32_t s,
                                                const real4<R>& nj, double2 pj, const double2* __restrict__ pos,
                                                const ListRec<R>* __restrict__ rec, const uint32_t* __restrict__ nbl,
                                                uint32_t cnt, bool& overflow,
                                                F&& f) {
    constexpr uint32_t G = 2;  // measured at C1-1M FP32: 57.7 us per pass against 58.8 (3), 61.4 (4), 65.1 (1)
    static_assert(2 * G - 1 <= kListPad, "the entries loaded past a list's end lie in its padding rows");
    if (cnt == 0) return;
    const uint32_t row = P.n * 4u;
    uint32_t off = s * 4u;
    const CutFP32 C = cut_band(P);
    uint32_t hits = 0;
    bool over = false;
    uint32_t e[G];
#pragma unroll
    for (uint32_t u = 0; u < G; ++u) e[u] = list_entry(nbl, off + u * row);
    off += G * row;
#pragma unroll 1
    for (uint32_t k = 0; k < cnt; k += G) {
        ListRec<R> r[G];
        uint32_t en[G];
#pragma unroll
        for (uint32_t u = 0; u < G; ++u) r[u] = rec[e[u] & kSlotMask];
#pragma unroll
        for (uint32_t u = 0; u < G; ++u) en[u] = list_entry(nbl, off + u * row);
        off += G * row;
#pragma unroll
        for (uint32_t u = 0; u < G; ++u) {
            ListEntry<R> E;
            const bool in = list_cut(P, C, s_off, e[u], r[u].nr, nj, pj, pos, E.dx, E.dy);
            const bool hit = in & (r[u].pr.z != R(0)) & (k + u < cnt);
            if (kCap) {
                E.take = hit & (hits < cap);
                over |= hit & !E.take;
                hits += E.take ? 1u : 0u;
            } else {
                E.take = hit;
            }
            E.nr = r[u].nr;
            E.pr = r[u].pr;
            f(E);
        }
#pragma unroll
        for (uint32_t u = 0; u < G; ++u) e[u] = en[u];
    }
    overflow |= over;
}

// fast_pair_term (the cubic configuration) for one list entry, without
// branches except the rare Lennard-Jones case: both pieces of the cubic
// spline are evaluated and selected, and nothing is added when take is
// false.  The Lennard-Jones ratios use 1/|r| from the reciprocal square
// root instead of two divisions.  j's own records: vaj (v, rho, itype) and
// prj (pred record).
template <class R>
__device__ __forceinline__ void list_pair_term(const FastPair<R>& K, PairSums<R>& S, const ListEntry<R>& E,
                                               const real4<R>& vaj, const real4<R>& prj, int tj) {
    const R dx = E.dx, dy = E.dy;
    const bool take = E.take;
    const R d2 = dx * dx + dy * dy;
    const R rinv = d2 > R(0) ? fast_rsqrt<R>(d2) : R(0);
    const R dist = d2 * rinv;
    const R q = dist * K.inv_h;
    const R t = R(2) - q;
    const R t2 = t * t;
    const bool inner = q <= R(1);
    const R F1 = K.c_q1 * (R(1.5) * q - R(2));
    const R F2 = -K.c_q2 * t2 * rinv;
    const R W1 = K.w_q1 * (R(2) / R(3) - q * q + R(0.5) * q * q * q);
    const R W2 = K.w_q2 * t2 * t;
    const R F = take ? (inner ? F1 : F2) : R(0);
    const R W = inner ? W1 : W2;
    if (K.cont) S.sd += F * ((E.nr.z - vaj.x) * dx + (E.nr.w - vaj.y) * dy);
    const R rho_i = fabs(E.pr.z), rho_j = fabs(prj.z);
    if (K.lj_on && take && tj > 0 && E.pr.z < R(0) && dist < K.lj_r0) {
        const R rn = K.lj_r0 * rinv;
        const R rn2 = rn * rn, rn4 = rn2 * rn2;
        const R f = K.lj_D * (rn4 * rn4 * rn4 - rn4) * (rinv * rinv);
        S.ljx -= dx * f;
        S.ljy -= dy * f;
    }
    const R inv_rs = fast_rcp<R>(rho_i + rho_j);
    R coef = E.pr.w + prj.w;
    const R dvx = E.pr.x - prj.x, dvy = E.pr.y - prj.y;
    if (K.av_on) {
        const R vr = dvx * dx + dvy * dy;
        const R mu = K.h * vr * fast_rcp<R>(d2 + K.eta2);
        const bool app = take & (vr < R(0));
        S.amu = app ? fmax(S.amu, -mu) : S.amu;
        coef += app ? (K.beta * mu - K.alpha_c) * mu * R(2) * inv_rs : R(0);
    }
    const R fc = F * coef;
    S.sx += fc * dx;
    S.sy += fc * dy;
    if (K.xsph_on) {
        const R cw = (take & (E.pr.z > R(0))) ? W * inv_rs : R(0);
        S.xx += dvx * cw;
        S.xy += dvy * cw;
    }
}

// Summation density from the lists (Density.h:25-61): rho = m W(0) +
// sum m W_ij, p = EOS(rho), into the pred records (the forces use this rho)
// and the pressure array; renormalised on the steps that ask for it
// (ModelParams.tex:124-133, rho_i the density before the step).  Cubic
// spline.  Aliveness and the old densities come from va (pred is being
// rewritten by this pass).
template <class R>
__global__ void __launch_bounds__(kBlock) density_sum_list(StepParams P, Ctrl* __restrict__ ctrl, ListState<R> X,
                                                           ListRec<R>* __restrict__ rec, R* __restrict__ pres,
                                                           const uint32_t* __restrict__ nbl,
                                                           const uint32_t* __restrict__ nbc) {
    if (ctrl->halted) return;
    const uint32_t lb = xcd_block(P.blocks_per_xcd);
    if (lb >= P.nblocks) return;
    __shared__ float2 s_off[64];
    fill_cell_offsets(P, s_off);
    __syncthreads();
    const uint32_t s = lb * kBlock + threadIdx.x;
    if (s >= P.n) return;
    const KernelConsts<R> K = kernel_consts<R>(P);
    const R m = (R)P.mass;
    const real4<R> vaj = X.va[s];
    const R w0 = kernel_w<R>(K, P.skf_density, R(0));
    R rho = m * w0;
    const bool renorm = ctrl->cur_renorm != 0;
    R wsum = renorm ? m / vaj.z * w0 : R(0);
    bool overflow = false;
    const uint32_t cnt = (int)vaj.w != 0 ? nbc[s] : 0u;
    if (cnt) {
        const CutFP32 C = cut_band(P);
        const real4<R> nj = rec[s].nr;
        const double2 pj = X.pos[s];
        const uint32_t cap = P.max_nbr - 1;
        const uint32_t row = P.n * 4u;
        uint32_t off = s * 4u;
        uint32_t hits = 0;
#pragma unroll 2
        for (uint32_t k = 0; k < cnt; ++k, off += row) {
            const uint32_t e = list_entry(nbl, off);
            const uint32_t i = e & kSlotMask;
            const real4<R> vai = X.va[i];
            R dx, dy;
            const bool hit = list_cut(P, C, s_off, e, rec[i].nr, nj, pj, X.pos, dx, dy) && (int)vai.w != 0;
            const bool take = hit & (hits < cap);
            overflow |= hit & !take;
            hits += take ? 1u : 0u;
            const R w = take ? kernel_w<R>(K, P.skf_density, pair_dist<R>(dx, dy, 0.0)) : R(0);
            rho += m * w;
            if (renorm) wsum += m / vai.z * w;
        }
    }
    if (renorm) rho = rho / wsum;
    if (overflow) atomicOr(&ctrl->nbr_overflow, 1u);
    real4<R> pr = rec[s].pr;
    const R p = eos_p<R>(rho, (R)P.rho0, (R)P.eos_B);
    pr.z = typed_rho<R>(rho, (int)vaj.w);
    pr.w = p / (rho * rho);
    rec[s].pr = pr;
    pres[s] = p;
}

// Pair pass from the lists: continuity density rate, pressure, artificial
// viscosity, Lennard-Jones walls and XSPH (the cubic configuration of
// pair_fast_kernel), the dt maxima, and the fields for the integrator.  With
// continuity it also writes the pressure of the current density (the
// reference's p array).
template <class R>
__global__ void __launch_bounds__(kBlock) pair_list_kernel(StepParams P, Ctrl* __restrict__ ctrl, ListState<R> X,
                                                           const ListRec<R>* __restrict__ rec, R* __restrict__ pres,
                                                           const uint32_t* __restrict__ nbl,
                                                           const uint32_t* __restrict__ nbc,
                                                           real4<R>* __restrict__ der, real2<R>* __restrict__ xs,
                                                           unsigned long long* __restrict__ maxslots) {
    if (ctrl->halted) return;
    if (ctrl->list_overflow) {  // the build of this step ran out of room: run it without lists
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            ctrl->halted = 1;
            ctrl->status = kStatusListOverflow;
        }
        return;
    }
    const uint32_t lb = xcd_block(P.blocks_per_xcd);
    if (lb >= P.nblocks) return;
    __shared__ float2 s_off[64];
    fill_cell_offsets(P, s_off);
    __syncthreads();
    const uint32_t s = lb * kBlock + threadIdx.x;
    const bool live = s < P.n;
    const FastPair<R> K(P);
    PairSums<R> S;
    bool overflow = false;
    real4<R> vaj{}, prj{};
    if (live) {
        vaj = X.va[s];
        prj = rec[s].pr;
    }
    const int tj = (int)vaj.w;
    const uint32_t cnt = live && tj != 0 ? nbc[s] : 0u;
    const uint32_t cap = P.max_nbr - 1;
    if (cnt) {
        auto term = [&](const ListEntry<R>& E) { list_pair_term<R>(K, S, E, vaj, prj, tj); };
        const real4<R> nj = rec[s].nr;
        const double2 pj = X.pos[s];
        if (__ballot(cnt > cap))  // some list of the wave may hold more than cap neighbours
            for_each_listed<R, true>(P, s_off, cap, s, nj, pj, X.pos, rec, nbl, cnt, overflow, term);
        else
            for_each_listed<R, false>(P, s_off, cap, s, nj, pj, X.pos, rec, nbl, cnt, overflow, term);
    }
    if (overflow) atomicOr(&ctrl->nbr_overflow, 1u);
    const R m = (R)P.mass;
    const R ax = S.ljx + m * S.sx;
    const R ay = -(R)P.g + S.ljy + m * S.sy;
    const R drho = m * S.sd;
    const R xf = m * R(2) * (R)P.xsph_eps;
    if (P.dt_method == 2 && P.dt_dynamic_mode == 1)
        block_max2_to_slots(live ? sqrt((double)ax * ax + (double)ay * ay) : 0.0, live ? (double)S.amu : 0.0,
                            maxslots, lb);
    if (live) {
        real4<R> d;
        d.x = ax;
        d.y = ay;
        d.z = drho;
        d.w = S.amu;
        der[s] = d;
        real2<R> x;
        x.x = xf * S.xx;
        x.y = xf * S.xy;
        xs[s] = x;
        if (P.density_treatment != 0) pres[s] = eos_p<R>(vaj.z, (R)P.rho0, (R)P.eos_B);
    }
}

// whole_step of every slot from X into A (in place on list steps), the cell
// key for the next build, the step's largest displacement, and the records
// of the next step's pair pass: near (the new position relative to the home
// cell, the new velocity) and pred (PredictHalfStep.cpp:16-22 with this
// step's dt, which is the next step's).
template <class R>
__global__ void __launch_bounds__(kBlock) integrate_list_kernel(StepParams P, Ctrl* __restrict__ ctrl, ListState<R> X,
                                                                const real4<R>* __restrict__ der,
                                                                const real2<R>* __restrict__ xs,
                                                                ListRec<R>* __restrict__ rec,
                                                                const uint32_t* __restrict__ home, StateA<R> A,
                                                                uint32_t* __restrict__ key_next,
                                                                const unsigned long long* __restrict__ maxslots,
                                                                unsigned long long* __restrict__ dslots) {
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
    uint32_t outside = 0, bad = 0;
    double moved = 0.0;
    if (s < P.n) {
        const real4<R> d = der[s];
        const real2<R> x = xs[s];
        const double2 p0 = X.pos[s];
        real4<R> pr{};
        if (P.density_treatment == 0) {  // the summed density (sign-encoded)
            pr = rec[s].pr;
            pr.z = fabs(pr.z);
        }
        const uint32_t oid = X.oid[s];
        const Stepped2<R> o = whole_step_one<R>(P, c, p0, X.va[s], pr, oid, d.x, d.y, d.z, x.x, x.y, dt, outside);
        A.pos[s] = o.pos;
        A.va[s] = o.va;
        real4<R> nad;
        nad.x = d.x;
        nad.y = d.y;
        nad.z = d.z;
        nad.w = R(0);
        A.ad[s] = nad;
        if (A.oid != X.oid) A.oid[s] = oid;
        const int it = (int)o.va.w;
        key_next[s] = it == 0 ? P.ncells : cell_key(P, o.pos.x, o.pos.y, bad);
        const double mx = o.pos.x - p0.x, my = o.pos.y - p0.y;
        moved = sqrt(mx * mx + my * my);
        R p;
        ListRec<R> r;
        r.pr = predict_record<R>(P, o.va, d.x, d.y, d.z, (R)dt, p);
        r.nr = near_record<R>(P, o.pos, o.va.x, o.va.y, home[s]);
        rec[s] = r;
    }
    wave_count_add(&c.outside, outside);
    if (bad) atomicOr(&c.grid_overflow, 1u);
    block_max2_to_slots(moved, 0.0, dslots, blockIdx.x);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c.dt_new = dt;
        c.pending = 1;
    }
}

}  // namespace rrsph
