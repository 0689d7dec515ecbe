// Neighbour-list steps in 3D: the design of list_step.hpp (builds that sort
// the cells and list every particle within k h + skin, list steps that
// decide each entry with the reference's cut-off test, a device-side
// staleness halt) with three-dimensional records.
//
// Per slot, in slot order:
//   rec.nr  FP32: (xi_x, xi_y, xi_z, p / rho^2), xi the position relative to
//           the origin of the home cell; FP64: (x, y, z, p / rho^2)
//   rec.pr  (v^x, v^y, v^z, s rho^), s = +1 fluid, -1 boundary, 0 dead
//   dens    (xi or x, s rho), rho the density before the step: what the
//           summation density pass gathers (its own record, so that the
//           pass can write the predicted density and p / rho^2 of its slot
//           into rec while others read dens)
// An entry holds the neighbour's slot in its low 25 bits and the offset of
// the two home cells, (dcx + 2) + 5 (dcy + 2) + 25 (dcz + 2), in its high 7.
// The FP32 cut-off test takes r_i - r_j = xi_i - xi_j + dc cell, the
// offsets from a table in LDS, and the FP64 positions and the reference's
// test (GridFind.cpp:144-176, dist2_exact3) within 1e-5 of the cut-off.
//
// Step kernels: list_prologue -> [density_sum_list3] -> pair_list3_kernel
// -> integrate_list3_kernel (in place).  A build step runs list_prologue ->
// radix sort -> reorder_list3_kernel -> list_build3_kernel -> the same step
// kernels on the reordered copy, integrating into A.
#pragma once

#include "list_step.hpp"
#include "step3d.hpp"

namespace rrsph {

constexpr uint32_t kSlotBits3 = 25;
constexpr uint32_t kSlotMask3 = (1u << kSlotBits3) - 1;  // 3D list steps take fewer than 2^25 particles
constexpr uint32_t kCodes3 = 125;

template <class R> struct ListState3 {
    const double4* pos;
    const real4<R>* va;
    const uint32_t* oid;
};

// Origin of home cell `home` (a key (cx ncy + cy) ncz + cz).
__device__ __forceinline__ double3 cell_origin3(const StepParams& P, uint32_t home) {
    const uint32_t cz = home % (uint32_t)P.ncz;
    const uint32_t cxy = home / (uint32_t)P.ncz;
    const uint32_t cx = cxy / (uint32_t)P.ncy, cy = cxy - cx * (uint32_t)P.ncy;
    return make_double3(P.xmin + (double)cx * P.cell, P.ymin + (double)cy * P.cell, P.zmin + (double)cz * P.cell);
}

// A record's position half with w in the fourth component.
template <class R> __device__ __forceinline__ real4<R> near3(const StepParams& P, double4 pos, R w, uint32_t home);
template <> __device__ __forceinline__ float4 near3<float>(const StepParams& P, double4 pos, float w, uint32_t home) {
    const double3 o = cell_origin3(P, home);
    return make_float4((float)(pos.x - o.x), (float)(pos.y - o.y), (float)(pos.z - o.z), w);
}
template <> __device__ __forceinline__ double4 near3<double>(const StepParams&, double4 pos, double w, uint32_t) {
    return make_double4(pos.x, pos.y, pos.z, w);
}

// The prediction of PredictHalfStep.cpp:16-22 as reorder3_kernel makes it:
// (v^, s rho^) and p / rho^2 with p the pressure of the current density.
template <class R>
__device__ __forceinline__ real4<R> predict3(const StepParams& P, const real4<R>& va, int itype, const real4<R>& ad,
                                              R dt, R& pw, R& p) {
    real4<R> pr;
    pr.x = va.x + ad.x * dt * R(0.5);
    pr.y = va.y + ad.y * dt * R(0.5);
    pr.z = va.z + ad.z * dt * R(0.5);
    const R rho_hat = P.density_treatment != 0 ? va.w + ad.w * dt * R(0.5) : va.w;
    p = eos_p<R>(va.w, (R)P.rho0, (R)P.eos_B);
    pw = p / (rho_hat * rho_hat);
    pr.w = typed_rho<R>(rho_hat, itype);
    return pr;
}

// The reorder of a 3D build step: reorder3_kernel's gather, pressure and
// cell-start table, plus the records, each slot's home cell and its
// candidate record for the build's walk, cand = (position relative to the
// home cell's origin in FP32, the home cell's z index).
template <class R>
__global__ void __launch_bounds__(kBlock) reorder_list3_kernel(StepParams P, const Ctrl* __restrict__ ctrl,
                                                               const uint32_t* __restrict__ skey,
                                                               const uint32_t* __restrict__ perm, StateA3<R> A,
                                                               StateB3<R> B, ListRec<R>* __restrict__ rec,
                                                               real4<R>* __restrict__ dens, uint32_t* __restrict__ home,
                                                               float4* __restrict__ cand, uint32_t* __restrict__ LB,
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
    const double4 pos = A.pos[src];
    const real4<R> va = A.va[src];
    const real4<R> ad = A.ad[src];
    const uint32_t oid = A.oid[src];
    const int itype = (int)pos.w;
    R pw, p;
    const real4<R> pr = predict3<R>(P, va, itype, ad, (R)ctrl->cur_dt, pw, p);
    if (live) {
        B.pos[s] = pos;
        B.va[s] = va;
        B.oid[s] = oid;
        B.pres[s] = p;
        ListRec<R> r;
        r.nr = near3<R>(P, pos, pw, k);
        r.pr = pr;
        rec[s] = r;
        dens[s] = near3<R>(P, pos, typed_rho<R>(va.w, itype), k);
        home[s] = k;
        const double3 o = cell_origin3(P, k);
        cand[s] = make_float4((float)(pos.x - o.x), (float)(pos.y - o.y), (float)(pos.z - o.z),
                              __int_as_float((int)(k % (uint32_t)P.ncz)));
    }
    fill_cell_starts(P, s, live, k, skey, LB);
}

// The lists of a 3D build step on the reordered copy: for each live
// particle, every live particle i != s within list_reach, in the order of
// the reference's cell walk (x-major columns, then y, each column's z range
// contiguous).  The cells visited are those within list_reach (3 or 4 per
// axis).  The walk is wave-uniform: the slots of a wave share one (x, y)
// column of cells or a few, so the wave loops over the union of its lanes'
// columns and, in each, over its longest z range, lanes masked outside
// their own ranges (same entries and order as a per-lane walk).  Entries
// go to the lane's column of an LDS buffer first; when a lane's column
// might overflow, the wave writes the buffered rows out together, so each
// row's stores from the lanes of a wave are contiguous (written straight
// from the walk, they left lines partly written in L2).
constexpr uint32_t kCollectL3 = 4;
constexpr uint32_t kBuildChunk = 32;  // LDS rows per lane
__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    return v;
}
constexpr uint32_t kBuildL3 = 8;  // candidates in flight per lane (6.6M dam break: 4.06 ms per build, 4.22 at 4)
__global__ void __launch_bounds__(kBlock) list_build3_kernel(StepParams P, Ctrl* __restrict__ ctrl,
                                                             const double4* __restrict__ pos,
                                                             const float4* __restrict__ cand,
                                                             const uint32_t* __restrict__ skey,
                                                             const uint32_t* __restrict__ LB,
                                                             uint32_t* __restrict__ nbl, uint32_t* __restrict__ nbc) {
    __shared__ uint32_t buf[kBuildChunk * kBlock];
    if (ctrl->halted) return;
    const uint32_t lb = xcd_block(P.blocks_per_xcd);
    if (lb >= P.nblocks) return;
    const uint32_t s0 = lb * kBlock + threadIdx.x;
    const bool inb = s0 < P.n;
    const uint32_t s = inb ? s0 : P.n - 1;
    uint32_t cnt = 0;
    bool bad = false;
    const double4 pj = pos[s];
    const bool alive = inb && (int)pj.w != 0;
    const uint32_t kj = skey[s];
    const int hz = (int)(kj % (uint32_t)P.ncz);
    const uint32_t hxy = kj / (uint32_t)P.ncz;
    const int hx = (int)(hxy / (uint32_t)P.ncy), hy = (int)hxy - hx * P.ncy;
    const double reach = P.list_reach * (1.0 + 1e-9);
    // empty ranges (x0 > x1) for lanes with nothing to list
    int x0 = 1, x1 = 0, y0 = 1, y1 = 0, z0 = 0, z1 = 0;
    if (alive) {
        // the list pass takes the FP32 positions relative to the home cell
        // for within 2.1 cells of its origin: a particle outside the cell
        // grid (keyed to the nearest edge cell) runs the step without lists
        const double3 o = cell_origin3(P, kj);
        const double ox = pj.x - o.x, oy = pj.y - o.y, oz = pj.z - o.z;
        bad = !(ox > -P.cell && ox < 2.0 * P.cell && oy > -P.cell && oy < 2.0 * P.cell && oz > -P.cell &&
                oz < 2.0 * P.cell);
        x0 = clamped_cell(pj.x - reach, P.xmin, P.cell, P.ncx);
        x1 = clamped_cell(pj.x + reach, P.xmin, P.cell, P.ncx);
        y0 = clamped_cell(pj.y - reach, P.ymin, P.cell, P.ncy);
        y1 = clamped_cell(pj.y + reach, P.ymin, P.cell, P.ncy);
        z0 = clamped_cell(pj.z - reach, P.zmin, P.cell, P.ncz);
        z1 = clamped_cell(pj.z + reach, P.zmin, P.cell, P.ncz);
        bad |= x0 < hx - 2 || x1 > hx + 2 || y0 < hy - 2 || y1 > hy + 2 || z0 < hz - 2 || z1 > hz + 2;
    }
    uint32_t* my = buf + threadIdx.x;
    uint32_t nh = 0;
    auto flush = [&]() {
#pragma unroll 1
        for (uint32_t r = 0; r < nh; ++r)
            if (cnt + r < P.list_cap) nbl[(size_t)(cnt + r) * P.n + s] = my[r * kBlock];
        cnt += nh;
        nh = 0;
    };
    const int wx0 = wave_min_i32(alive ? x0 : INT_MAX), wx1 = wave_max_i32(alive ? x1 : INT_MIN);
    const int wy0 = wave_min_i32(alive ? y0 : INT_MAX), wy1 = wave_max_i32(alive ? y1 : INT_MIN);
    const double reach2 = reach * reach;
    // candidates are tested in FP32 from the candidate records: r_i - r_s =
    // (xi_i - xi_s) + (home-cell offset) cell, rounded far below the 1e-5
    // by which the FP32 reach exceeds list_reach, so the lists hold every
    // pair within list_reach (and at most a few just beyond, which every
    // list step tests again)
    const float4 cj = cand[s];
    const float cellf = (float)P.cell;
    const float r2f = (float)(P.list_r2 * (1.0 + 1e-5));
#pragma unroll 1
    for (int x = wx0; x <= wx1; ++x) {
        // distance from the particle to the column's x slab
        const double xl = P.xmin + x * P.cell;
        const double dxm = fmax(0.0, fmax(xl - pj.x, pj.x - (xl + P.cell)));
        const float ox = (float)(x - hx) * cellf;
#pragma unroll 1
        for (int y = wy0; y <= wy1; ++y) {
            const double yl = P.ymin + y * P.cell;
            const double dym = fmax(0.0, fmax(yl - pj.y, pj.y - (yl + P.cell)));
            const float oy = (float)(y - hy) * cellf;
            // the z cells of this column within reach: a column off the
            // particle's own needs a shorter range (corner columns most)
            const double rem = reach2 - dxm * dxm - dym * dym;
            bool need = x >= x0 && x <= x1 && y >= y0 && y <= y1 && rem >= 0.0;
            int za = z0, zb = z1;
            if (need) {
                const double zr = sqrt(rem) * (1.0 + 1e-9);
                za = max(z0, clamped_cell(pj.z - zr, P.zmin, P.cell, P.ncz));
                zb = min(z1, clamped_cell(pj.z + zr, P.zmin, P.cell, P.ncz));
            }
            const uint32_t col = need ? ((uint32_t)x * (uint32_t)P.ncy + (uint32_t)y) * (uint32_t)P.ncz : 0u;
            const uint32_t b = need ? LB[col + za] : 0u;
            const uint32_t len = need ? LB[col + zb + 1] - b : 0u;
            const uint32_t code_xy = (uint32_t)(x - hx + 2) + 5u * (uint32_t)(y - hy + 2);
            const uint32_t lmax = wave_max_u32(len);
#pragma unroll 1
            for (uint32_t i = 0; i < lmax; i += kBuildL3) {
                float4 q[kBuildL3];
#pragma unroll
                for (uint32_t u = 0; u < kBuildL3; ++u) q[u] = cand[len ? b + min(i + u, len - 1) : s];
#pragma unroll
                for (uint32_t u = 0; u < kBuildL3; ++u) {
                    const int dcz = __float_as_int(q[u].w) - hz;  // the neighbour's z cell, from its key
                    const float dx = (q[u].x - cj.x) + ox, dy = (q[u].y - cj.y) + oy;
                    const float dz = (q[u].z - cj.z) + (float)dcz * cellf;
                    const bool in = (i + u < len) & (b + i + u != s) & (dx * dx + dy * dy + dz * dz < r2f);
                    if (in) {
                        const uint32_t code = code_xy + 25u * (uint32_t)(dcz + 2);
                        my[nh * kBlock] = (b + i + u) | (code << kSlotBits3);
                        ++nh;
                    }
                }
                if (__any(nh + kBuildL3 > kBuildChunk)) flush();  // wave-uniform
            }
        }
    }
    flush();
    if (!inb) return;
    if (cnt > P.list_cap || bad) atomicOr(&ctrl->list_overflow, 1u);
    nbc[s] = min(cnt, P.list_cap);
}

// collect_neighbours3 over a particle's list instead of the cell walk: the
// entries within the cut-off (the reference's FP64 test on pos), alive, in
// list order, first `cap` kept, into the lane's column of the LDS hit list,
// flushed wave-uniformly.  Every lane of the wave calls it; lanes with
// active = false have no entries.
template <uint32_t kChunk, class F>
__device__ __forceinline__ void collect_listed3(const StepParams& P, bool active, uint32_t cap, uint32_t s,
                                                double4 pj, const double4* __restrict__ pos,
                                                const uint32_t* __restrict__ nbl, const uint32_t* __restrict__ nbc,
                                                uint32_t* __restrict__ hits, bool& overflow, F&& flush) {
    static_assert(kChunk >= 2 * kCollectL3, "hit list too short");
    const uint32_t cnt = active ? nbc[s] : 0u;
    const uint32_t cmax = wave_max_u32(cnt);
    const uint32_t row = P.n * 4u;
    uint32_t nh = 0, done = 0;
#pragma unroll 1
    for (uint32_t k = 0; k < cmax; k += kCollectL3) {
        uint32_t i[kCollectL3];
        double4 q[kCollectL3];
#pragma unroll
        for (uint32_t u = 0; u < kCollectL3; ++u) i[u] = list_entry(nbl, (k + u) * row + s * 4u) & kSlotMask3;
#pragma unroll
        for (uint32_t u = 0; u < kCollectL3; ++u) q[u] = pos[i[u]];
#pragma unroll
        for (uint32_t u = 0; u < kCollectL3; ++u) {
            const double dx = __dsub_rn(q[u].x, pj.x), dy = __dsub_rn(q[u].y, pj.y), dz = __dsub_rn(q[u].z, pj.z);
            const bool hit = (k + u < cnt) & ((int)q[u].w != 0) & (dist2_exact3(dx, dy, dz) < P.cut2);
            if (hit) {
                if (done + nh < cap) {
                    hits[nh * kBlock] = i[u];
                    ++nh;
                } else {
                    overflow = true;
                }
            }
        }
        if (__any(nh + kCollectL3 > kChunk)) {  // wave-uniform
            flush(nh);
            done += nh;
            nh = 0;
        }
    }
    flush(nh);
}

// The table of home-cell offsets of the 3D entry codes (first kCodes3
// threads; a barrier must follow).
__device__ __forceinline__ void fill_cell_offsets3(const StepParams& P, float4* s_off) {
    if (threadIdx.x < kCodes3) {
        const int c = (int)threadIdx.x;
        const int dx = c % 5 - 2, dy = (c / 5) % 5 - 2, dz = c / 25 - 2;
        s_off[c] = make_float4((float)(dx * P.cell), (float)(dy * P.cell), (float)(dz * P.cell), 0.f);
    }
}

// r_i - r_j of an entry and whether it is within the cut-off (see the
// header): FP32 from the records, FP64 on the band; FP64 from the FP64
// record positions.
__device__ __forceinline__ bool list_cut3(const StepParams& P, const CutFP32& C, const float4* s_off, uint32_t e,
                                          const float4& ni, const float4& nj, const double4& pj,
                                          const double4* __restrict__ pos, float& dx, float& dy, float& dz) {
    const float4 o = s_off[e >> kSlotBits3];
    dx = (ni.x - nj.x) + o.x;
    dy = (ni.y - nj.y) + o.y;
    dz = (ni.z - nj.z) + o.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    if (!(d2 > C.lo && d2 < C.hi)) return d2 <= C.lo;  // on the band's edge: within a rounding of lo < cut2
    const double4 q = pos[e & kSlotMask3];  // rare: the band around the cut-off
    const double ddx = __dsub_rn(q.x, pj.x), ddy = __dsub_rn(q.y, pj.y), ddz = __dsub_rn(q.z, pj.z);
    dx = (float)ddx;
    dy = (float)ddy;
    dz = (float)ddz;
    return dist2_exact3(ddx, ddy, ddz) < P.cut2;
}
__device__ __forceinline__ bool list_cut3(const StepParams& P, const CutFP32&, const float4*, uint32_t,
                                          const double4& ni, const double4& nj, const double4&,
                                          const double4* __restrict__, double& dx, double& dy, double& dz) {
    dx = __dsub_rn(ni.x, nj.x);
    dy = __dsub_rn(ni.y, nj.y);
    dz = __dsub_rn(ni.z, nj.z);
    return dist2_exact3(dx, dy, dz) < P.cut2;
}

// collect_listed3 from the near records of the list steps (position
// relative to the home cell of the build, and type, in nearf): each entry
// decided as the FP32 list pass decides it (list_cut3: FP32, the FP64 test
// on pos inside the band), the same hits; the hit list keeps the whole
// entry (slot and home-cell offset code).
template <uint32_t kChunk, class F>
__device__ __forceinline__ void collect_near3(const StepParams& P, bool active, uint32_t cap, uint32_t s,
                                              const float4* s_off, const float4& nk, const float4* __restrict__ nearf,
                                              const double4& pj, const double4* __restrict__ pos,
                                              const uint32_t* __restrict__ nbl, const uint32_t* __restrict__ nbc,
                                              uint32_t* __restrict__ hits, bool& overflow, F&& flush) {
    static_assert(kChunk >= 2 * kCollectL3, "hit list too short");
    const CutFP32 C = cut_band(P);
    const uint32_t cnt = active ? nbc[s] : 0u;
    const uint32_t cmax = wave_max_u32(cnt);
    const uint32_t row = P.n * 4u;
    uint32_t nh = 0, done = 0;
#pragma unroll 1
    for (uint32_t k = 0; k < cmax; k += kCollectL3) {
        uint32_t e[kCollectL3];
        float4 q[kCollectL3];
#pragma unroll
        for (uint32_t u = 0; u < kCollectL3; ++u) e[u] = list_entry(nbl, (k + u) * row + s * 4u);
#pragma unroll
        for (uint32_t u = 0; u < kCollectL3; ++u) q[u] = nearf[e[u] & kSlotMask3];
#pragma unroll
        for (uint32_t u = 0; u < kCollectL3; ++u) {
            float dx, dy, dz;
            const bool in = list_cut3(P, C, s_off, e[u], q[u], nk, pj, pos, dx, dy, dz);
            const bool hit = (k + u < cnt) & (q[u].w != 0.f) & in;
            if (hit) {
                if (done + nh < cap) {
                    hits[nh * kBlock] = e[u];
                    ++nh;
                } else {
                    overflow = true;
                }
            }
        }
        if (__any(nh + kCollectL3 > kChunk)) {  // wave-uniform
            flush(nh);
            done += nh;
            nh = 0;
        }
    }
    flush(nh);
}

// Summation density from the lists (Density.h:25-61) with the
// renormalisation of the steps that ask for it (ModelParams.tex:124-133):
// rho into rec.pr.w (sign-typed), p / rho^2 into rec.nr.w, p into pres.
template <class R>
__global__ void __launch_bounds__(kBlock) density_sum_list3(StepParams P, Ctrl* __restrict__ ctrl, ListState3<R> X,
                                                            const real4<R>* __restrict__ dens,
                                                            ListRec<R>* __restrict__ rec, R* __restrict__ pres,
                                                            const uint32_t* __restrict__ nbl,
                                                            const uint32_t* __restrict__ nbc) {
    if (ctrl->halted) return;
    const uint32_t lb = xcd_block(P.blocks_per_xcd);
    if (lb >= P.nblocks) return;
    __shared__ float4 s_off[kCodes3];
    fill_cell_offsets3(P, s_off);
    __syncthreads();
    const uint32_t s = lb * kBlock + threadIdx.x;
    if (s >= P.n) return;
    const R m = (R)P.mass, inv_h = (R)(1.0 / P.h), w_q1 = (R)P.f_cubic, w_q2 = (R)(P.f_cubic / 6.0);
    const real4<R> dj = dens[s];
    const R w0 = w_q1 * (R(2) / R(3));
    R rho = m * w0;
    const bool renorm = ctrl->cur_renorm != 0;
    R wsum = renorm && dj.w != R(0) ? m / fabs(dj.w) * w0 : R(0);
    bool overflow = false;
    const uint32_t cnt = dj.w != R(0) ? nbc[s] : 0u;
    if (cnt) {
        const CutFP32 C = cut_band(P);
        const double4 pj = X.pos[s];
        const uint32_t cap = P.max_nbr - 1;
        const uint32_t row = P.n * 4u;
        uint32_t hits = 0;
        // groups of G entries: the group's records gathered together, the
        // next group's entries loaded behind them (as the pair pass does);
        // the sums run in list order.  Rows past a list's end name valid
        // slots (padding, or entries of an earlier build); the loads past
        // the last allocated row read that row instead.
        // 6.6M dam break: FP32 1.13 ms per pass at G = 8 (1.45 at G = 4),
        // FP64 2.06 ms at G = 4 with incremental offsets (2.21 clamped)
        constexpr bool kWide = sizeof(R) == 4;
        constexpr uint32_t G = kWide ? 8 : 4;
        static_assert(kWide || 2 * G - 1 <= kListPad, "the entries loaded past a list's end lie in its padding rows");
        const uint32_t last = P.list_cap + kListPad - 1;
        auto entry = [&](uint32_t k) {
            return list_entry(nbl, (kWide ? min(k, last) : k) * row + s * 4u);
        };
        uint32_t e[G];
#pragma unroll
        for (uint32_t u = 0; u < G; ++u) e[u] = entry(u);
#pragma unroll 1
        for (uint32_t k = 0; k < cnt; k += G) {
            real4<R> di[G];
            uint32_t en[G];
#pragma unroll
            for (uint32_t u = 0; u < G; ++u) di[u] = dens[e[u] & kSlotMask3];
#pragma unroll
            for (uint32_t u = 0; u < G; ++u) en[u] = entry(k + G + u);
#pragma unroll
            for (uint32_t u = 0; u < G; ++u) {
                R dx, dy, dz;
                const bool hit =
                    list_cut3(P, C, s_off, e[u], di[u], dj, pj, X.pos, dx, dy, dz) & (di[u].w != R(0)) & (k + u < cnt);
                const bool take = hit & (hits < cap);
                overflow |= hit & !take;
                hits += take ? 1u : 0u;
                const R q = pair_dist<R>(dx, dy, dz) * inv_h;
                const R t = R(2) - q;
                const R w = q <= R(1) ? w_q1 * (R(2) / R(3) - q * q + R(0.5) * q * q * q) : w_q2 * t * t * t;
                const R wt = take ? w : R(0);
                rho += m * wt;
                // (a neighbour that died since the build has density 0 in its
                // record: only taken entries enter the sum)
                if (renorm && take) wsum += m / fabs(di[u].w) * wt;
            }
#pragma unroll
            for (uint32_t u = 0; u < G; ++u) e[u] = en[u];
        }
    }
    // a particle that died since the build has no sums to renormalise (its
    // record's density is 0): it keeps the plain sum, so that its records,
    // which the lists of its former neighbours still name, stay finite
    if (renorm && wsum > R(0)) rho = rho / wsum;
    if (overflow) atomicOr(&ctrl->nbr_overflow, 1u);
    const int itype = (int)X.pos[s].w;
    const R p = eos_p<R>(rho, (R)P.rho0, (R)P.eos_B);
    ListRec<R> r = rec[s];
    r.pr.w = typed_rho<R>(rho, itype);
    r.nr.w = p / (rho * rho);
    rec[s] = r;
    pres[s] = p;
}

// fast_pair_term3 (the cubic configuration) for one list entry, without
// branches except the rare Lennard-Jones case; nothing is added when take
// is false.  i's records: nr (position half, p / rho^2), pr (v^, s rho^),
// vai (its current velocity; continuity only).
template <class R>
__device__ __forceinline__ void list_pair_term3(const FastPair<R>& K, R (&S)[12], bool take, R dx, R dy, R dz,
                                                const real4<R>& nri, const real4<R>& pri, const real4<R>& vai,
                                                const real4<R>& vaj, const real4<R>& prj, R pwj, int tj) {
    // S: sx, sy, sz, sd, xx, xy, xz, amu, ljx, ljy, ljz, (unused)
    const R d2 = dx * dx + dy * dy + dz * dz;
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
    if (K.cont) S[3] += F * ((vai.x - vaj.x) * dx + (vai.y - vaj.y) * dy + (vai.z - vaj.z) * dz);
    if (K.lj_on && take && tj > 0 && pri.w < R(0) && dist < K.lj_r0) {
        const R rn = K.lj_r0 * rinv;
        const R rn2 = rn * rn, rn4 = rn2 * rn2;
        const R f = K.lj_D * (rn4 * rn4 * rn4 - rn4) * (rinv * rinv);
        S[8] -= dx * f;
        S[9] -= dy * f;
        S[10] -= dz * f;
    }
    const R inv_rs = fast_rcp<R>(fabs(pri.w) + fabs(prj.w));
    R coef = nri.w + pwj;
    const R dvx = pri.x - prj.x, dvy = pri.y - prj.y, dvz = pri.z - prj.z;
    if (K.av_on) {
        const R vr = dvx * dx + dvy * dy + dvz * dz;
        const R mu = K.h * vr * fast_rcp<R>(d2 + K.eta2);
        const bool app = take & (vr < R(0));
        S[7] = app ? fmax(S[7], -mu) : S[7];
        coef += app ? (K.beta * mu - K.alpha_c) * mu * R(2) * inv_rs : R(0);
    }
    const R fc = F * coef;
    S[0] += fc * dx;
    S[1] += fc * dy;
    S[2] += fc * dz;
    if (K.xsph_on) {
        const R cw = (take & (pri.w > R(0))) ? W * inv_rs : R(0);
        S[4] += dvx * cw;
        S[5] += dvy * cw;
        S[6] += dvz * cw;
    }
}

// Pair pass from the 3D lists: the cubic configuration of pair3_kernel
// (continuity, pressure, artificial viscosity, Lennard-Jones walls, XSPH),
// the dt maxima and the fields for the integrator.  Entries in groups of
// G, as for_each_listed.
template <class R>
__global__ void __launch_bounds__(kBlock) pair_list3_kernel(StepParams P, Ctrl* __restrict__ ctrl, ListState3<R> X,
                                                            const ListRec<R>* __restrict__ rec, R* __restrict__ pres,
                                                            const uint32_t* __restrict__ nbl,
                                                            const uint32_t* __restrict__ nbc,
                                                            real4<R>* __restrict__ der, real4<R>* __restrict__ xs,
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
    __shared__ float4 s_off[kCodes3];
    fill_cell_offsets3(P, s_off);
    __syncthreads();
    const uint32_t s = lb * kBlock + threadIdx.x;
    const bool live = s < P.n;
    const FastPair<R> K(P);
    R S[12] = {};
    bool overflow = false;
    double4 pj = make_double4(0, 0, 0, 0);
    real4<R> vaj{};
    ListRec<R> rj{};
    if (live) {
        pj = X.pos[s];
        vaj = X.va[s];
        rj = rec[s];
    }
    const int tj = (int)pj.w;
    const uint32_t cnt = live && tj != 0 ? nbc[s] : 0u;
    const uint32_t cap = P.max_nbr - 1;
    if (cnt) {
        // FP32 6.6M dam break: 1.87 ms per pass at G = 8 (168 VGPRs, 3 waves
        // per SIMD), 2.00 at 4, 2.39 at 2.  Rows past a list's end name
        // valid slots (padding, or entries of an earlier build); the loads
        // past the last allocated row read that row instead.
        constexpr uint32_t G = sizeof(R) == 4 ? 8 : 2;
        const uint32_t lastrow = P.list_cap + kListPad - 1;
        const bool capped = __ballot(cnt > cap) != 0;  // some list of the wave may hold more than cap neighbours
        const CutFP32 C = cut_band(P);
        const uint32_t row = P.n * 4u;
        uint32_t hits = 0;
        bool over = false;
        uint32_t e[G];
#pragma unroll
        for (uint32_t u = 0; u < G; ++u) e[u] = list_entry(nbl, min(u, lastrow) * row + s * 4u);
#pragma unroll 1
        for (uint32_t k = 0; k < cnt; k += G) {
            ListRec<R> r[G];
            real4<R> va[G]{};
            uint32_t en[G];
#pragma unroll
            for (uint32_t u = 0; u < G; ++u) {
                r[u] = rec[e[u] & kSlotMask3];
                if (K.cont) va[u] = X.va[e[u] & kSlotMask3];
            }
#pragma unroll
            for (uint32_t u = 0; u < G; ++u) en[u] = list_entry(nbl, min(k + G + u, lastrow) * row + s * 4u);
#pragma unroll
            for (uint32_t u = 0; u < G; ++u) {
                R dx, dy, dz;
                const bool in = list_cut3(P, C, s_off, e[u], r[u].nr, rj.nr, pj, X.pos, dx, dy, dz);
                const bool hit = in & (r[u].pr.w != R(0)) & (k + u < cnt);
                bool take = hit;
                if (capped) {
                    take = hit & (hits < cap);
                    over |= hit & !take;
                    hits += take ? 1u : 0u;
                }
                list_pair_term3<R>(K, S, take, dx, dy, dz, r[u].nr, r[u].pr, va[u], vaj, rj.pr, rj.nr.w, tj);
            }
#pragma unroll
            for (uint32_t u = 0; u < G; ++u) e[u] = en[u];
        }
        overflow |= over;
    }
    if (overflow) atomicOr(&ctrl->nbr_overflow, 1u);
    const R m = (R)P.mass;
    const R ax = S[8] + m * S[0];
    const R ay = -(R)P.g + S[9] + m * S[1];
    const R az = S[10] + m * S[2];
    const R drho = m * S[3];
    const R xf = m * R(2) * (R)P.xsph_eps;
    if (P.dt_method == 2 && P.dt_dynamic_mode == 1)
        block_max2_to_slots(live ? sqrt((double)ax * ax + (double)ay * ay + (double)az * az) : 0.0,
                            live ? (double)S[7] : 0.0, maxslots, lb);
    if (live) {
        real4<R> d;
        d.x = ax;
        d.y = ay;
        d.z = az;
        d.w = drho;
        der[s] = d;
        real4<R> x;
        x.x = xf * S[4];
        x.y = xf * S[5];
        x.z = xf * S[6];
        x.w = S[7];
        xs[s] = x;
        if (P.density_treatment != 0) pres[s] = eos_p<R>(vaj.w, (R)P.rho0, (R)P.eos_B);
    }
}

// whole_step of every slot from X into A (in place on list steps), the
// cell key for the next build, the step's largest displacement, and the
// next step's records.
template <class R>
__global__ void __launch_bounds__(kBlock) integrate_list3_kernel(StepParams P, Ctrl* __restrict__ ctrl,
                                                                 ListState3<R> X, const real4<R>* __restrict__ der,
                                                                 const real4<R>* __restrict__ xs,
                                                                 ListRec<R>* __restrict__ rec,
                                                                 real4<R>* __restrict__ dens,
                                                                 const uint32_t* __restrict__ home, StateA3<R> A,
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
    uint32_t outside = 0, bad = 0, changed = 0;
    double moved = 0.0;
    if (s < P.n) {
        const real4<R> d = der[s];
        const real4<R> x = xs[s];
        const double4 p0 = X.pos[s];
        real4<R> pr{};
        if (P.density_treatment == 0) pr.w = fabs(rec[s].pr.w);  // the summed density
        const uint32_t h = home[s];
        const Stepped3<R> o = integrate_one3<R>(P, c, s, p0, X.va[s], pr, X.oid[s], d.x, d.y, d.z, d.w, x.x, x.y,
                                                x.z, dt, A, key_next, h, outside, bad, changed);
        const double mx = o.pos.x - p0.x, my = o.pos.y - p0.y, mz = o.pos.z - p0.z;
        moved = sqrt(mx * mx + my * my + mz * mz);
        const int itype = (int)o.pos.w;
        R pw, p;
        ListRec<R> r;
        r.pr = predict3<R>(P, o.va, itype, d, (R)dt, pw, p);
        r.nr = near3<R>(P, o.pos, pw, h);
        rec[s] = r;
        dens[s] = near3<R>(P, o.pos, typed_rho<R>(o.va.w, itype), h);
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
