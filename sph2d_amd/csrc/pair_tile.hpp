// LDS-tiled pair pass (the hot kernel).
//
// One wavefront = one tile of 64 consecutive particles in cell order.  The
// tile's particles lie in at most two adjacent grid columns ("groups"); for
// each group the wave stages the three neighbour columns' row ranges into LDS
// with coalesced loads (positions relative to a tile origin, velocity /
// density records), then every lane walks its own three column ranges out
// of LDS.  This turns ~70 dependent global-memory round trips per particle
// into a handful of independent ones and keeps the neighbour data traffic in
// LDS instead of L2.  Tiles whose staged set does not fit, or that span
// more than two columns, walk global memory instead (same arithmetic).
//
// Physics and cut-off decision are those of pair_fast_kernel (cubic spline,
// p_i / rho_i^2 + p_j / rho_j^2, artificial viscosity, Lennard-Jones walls,
// XSPH, summation or continuity density); the neighbour decision is made on
// the relative positions and re-checked exactly in FP64 inside a 1e-5 band
// around the cut-off, so the neighbour sets are the FP64 ones.
#pragma once

#include "step_kernels.hpp"

namespace rrsph {

constexpr int kTile = kWave;          // particles per tile (one wavefront)
constexpr int kTilesPerBlock = 4;     // independent tiles (wavefronts) per workgroup

// candidates staged per tile (LDS: 48 B each in FP32, 80 B in FP64)
template <class R> constexpr int tile_cap() { return sizeof(R) == 4 ? 320 : 256; }

__device__ __forceinline__ int wave_min_i(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    return v;
}

// Per-neighbour physics shared by the LDS and global walks.
template <class R> struct PairAcc {
    R sx = R(0), sy = R(0), sd = R(0), xx = R(0), xy = R(0), amu = R(0), ljx = R(0), ljy = R(0);
};

template <class R> struct PairConst {
    R h, inv_h, c_q1, c_q2, w_q1, w_q2, alpha_c, beta, eta2, lj_r0, lj_D;
    bool cont, av_on, xsph_on, lj_on;
};

template <class R> __device__ __forceinline__ PairConst<R> pair_const(const StepParams& P) {
    PairConst<R> C;
    C.h = (R)P.h;
    C.inv_h = (R)(1.0 / P.h);
    C.c_q1 = (R)(P.f_cubic / (P.h * P.h));
    C.c_q2 = (R)(0.5 * P.f_cubic / P.h);
    C.w_q1 = (R)P.f_cubic;
    C.w_q2 = (R)(P.f_cubic / 6.0);
    C.alpha_c = (R)(P.av_alpha * P.av_c);
    C.beta = (R)P.av_beta;
    C.eta2 = (R)P.av_eta2;
    C.lj_r0 = (R)P.lj_r0;
    C.lj_D = (R)P.lj_D;
    C.cont = P.density_treatment != 0;
    C.av_on = P.av_on;
    C.xsph_on = P.xsph_on;
    C.lj_on = P.lj_on;
    return C;
}

// j = self (vaj, prj, tj), i = neighbour; dx, dy = r_i - r_j, d2 = |r_i - r_j|^2
template <class R>
__device__ __forceinline__ void pair_body(const PairConst<R>& C, PairAcc<R>& A, R dx, R dy, R d2,
                                          const real4<R>& vai, const real4<R>& pri, const real4<R>& vaj,
                                          const real4<R>& prj, int tj) {
    // coincident particles (|r| = 0) get grad W = 0 as in the reference (q = 0)
    const R rinv = d2 > R(0) ? fast_rsqrt<R>(d2) : R(0);
    const R dist = d2 * rinv;
    const R q = dist * C.inv_h;
    const R t = R(2) - q;
    R F, W;
    if (q <= R(1)) {
        F = C.c_q1 * (R(1.5) * q - R(2));
        W = C.w_q1 * (R(2) / R(3) - q * q + R(0.5) * q * q * q);
    } else {
        F = -C.c_q2 * t * t * rinv;
        W = C.w_q2 * t * t * t;
    }
    if (C.cont) A.sd += F * ((vai.x - vaj.x) * dx + (vai.y - vaj.y) * dy);
    const int ti = (int)vai.w;
    if (C.lj_on && tj > 0 && ti < 0 && dist < C.lj_r0) {
        const R rn = C.lj_r0 / dist;  // |r| = 0 gives a non-finite force, as the reference
        const R rn2 = rn * rn, rn4 = rn2 * rn2;
        const R f = C.lj_D * (rn4 * rn4 * rn4 - rn4) / d2;
        A.ljx -= dx * f;
        A.ljy -= dy * f;
    }
    const R inv_rs = fast_rcp<R>(pri.z + prj.z);  // 1 / (rho^_i + rho^_j)
    R coef = pri.w + prj.w;                         // p_i / rho_i^2 + p_j / rho_j^2
    if (C.av_on) {
        const R vr = (pri.x - prj.x) * dx + (pri.y - prj.y) * dy;  // (v_j - v_i) . (r_j - r_i)
        if (vr < R(0)) {
            const R mu = C.h * vr * fast_rcp<R>(d2 + C.eta2);
            A.amu = fmax(A.amu, -mu);
            coef += (C.beta * mu - C.alpha_c) * mu * R(2) * inv_rs;
        }
    }
    const R fc = F * coef;
    A.sx += fc * dx;
    A.sy += fc * dy;
    if (C.xsph_on && ti > 0) {
        const R cw = W * inv_rs;
        A.xx += (pri.x - prj.x) * cw;
        A.xy += (pri.y - prj.y) * cw;
    }
}

// Exact FP64 cut-off decision (no contraction), as GridFind.cpp:153-156.
__device__ __forceinline__ bool exact_inside(double2 pi, double2 pj, double cut2) {
    const double dx = __dsub_rn(pi.x, pj.x);
    const double dy = __dsub_rn(pi.y, pj.y);
    return dist2_exact(dx, dy) < cut2;
}

template <class R, bool FUSED, bool EXPORT>
__global__ void __launch_bounds__(kTile * kTilesPerBlock) pair_tile_kernel(StepParams P, Ctrl* __restrict__ ctrl, StateB<R> B,
                                                          const uint32_t* __restrict__ skey,
                                                          const uint32_t* __restrict__ LB, StateA<R> A,
                                                          uint32_t* __restrict__ key_next,
                                                          real4<R>* __restrict__ der, real4<R>* __restrict__ xs,
                                                          unsigned long long* __restrict__ maxslots) {
    constexpr int kCap = tile_cap<R>();
    __shared__ real2<R> s_pos_all[kTilesPerBlock][kCap];
    __shared__ real4<R> s_va_all[kTilesPerBlock][kCap];
    __shared__ real4<R> s_pr_all[kTilesPerBlock][kCap];
    if (ctrl->halted) return;
    const uint32_t lb = xcd_block(P.tile_blocks_per_xcd);  // block of kTilesPerBlock tiles
    if (lb * kTilesPerBlock >= P.ntiles) return;
    const uint32_t wave = threadIdx.x / kTile;
    const uint32_t lane = threadIdx.x % kTile;
    real2<R>* s_pos = s_pos_all[wave];
    real4<R>* s_va = s_va_all[wave];
    real4<R>* s_pr = s_pr_all[wave];
    const uint32_t tile = lb * kTilesPerBlock + wave;
    const uint32_t s = tile * kTile + lane;
    const bool live = s < P.n;
    const PairConst<R> C = pair_const<R>(P);
    PairAcc<R> acc;
    bool overflow = false;
    double2 pj = make_double2(0, 0);
    real4<R> vaj{}, prj{};
    uint32_t oid = 0, key = 0;
    if (live) {
        pj = B.pos[s];
        vaj = B.va[s];
        prj = B.pred[s];
        oid = B.oid[s];
        key = skey[s];
    }
    const int tj = (int)vaj.w;
    const bool active = live && tj != 0;
    const int cx = active ? (int)(key / (uint32_t)P.ncy) : 0;
    const int cy = active ? (int)key - cx * P.ncy : 0;
    const int cxa = wave_min_i(active ? cx : INT_MAX);
    const int cxb = wave_max_i(active ? cx : -1);
    const uint32_t cap = P.max_nbr - 1;
    const float band_lo = (float)(P.cut2 * (1.0 - 1e-5)), band_hi = (float)(P.cut2 * (1.0 + 1e-5));
    if (cxb >= 0) {  // at least one particle with neighbours to find
        const int g = cx - cxa;  // group of this lane (0 or 1 when the tile fits)
        const int rmin0 = wave_min_i(active && g == 0 ? cy : INT_MAX);
        const int rmax0 = wave_max_i(active && g == 0 ? cy : -1);
        const int rmin1 = wave_min_i(active && g == 1 ? cy : INT_MAX);
        const int rmax1 = wave_max_i(active && g == 1 ? cy : -1);
        // segment t (lanes 0..5): group t / 3, column offset t % 3 - 1
        uint32_t seg_b = 0, seg_n = 0;
        if (lane < 6) {
            const int tg = lane / 3, col = (tg == 0 ? cxa : cxa + 1) + (int)(lane % 3) - 1;
            const int r0 = tg == 0 ? rmin0 : rmin1, r1 = tg == 0 ? rmax0 : rmax1;
            if (r1 >= 0 && col >= 0 && col < P.ncx) {
                const int y0 = r0 > 0 ? r0 - 1 : 0;
                const int y1 = r1 + 1 < P.ncy ? r1 + 1 : P.ncy - 1;
                seg_b = LB[col * P.ncy + y0];
                seg_n = LB[col * P.ncy + y1 + 1] - seg_b;
            }
        }
        // exclusive prefix of the 6 segment lengths -> LDS offsets
        uint32_t seg_o = seg_n;
        for (int off = 1; off < 8; off <<= 1) {
            const uint32_t y = __shfl_up(seg_o, off);
            if ((int)lane >= off) seg_o += y;
        }
        const uint32_t total = __shfl(seg_o, 5);
        seg_o -= seg_n;
        const bool use_lds = cxb - cxa <= 1 && total <= (uint32_t)kCap;
        // tile origin: first active lane's position
        const int first = __ffsll((long long)__ballot(active)) - 1;
        const double ox = __shfl(pj.x, first), oy = __shfl(pj.y, first);
        if (use_lds) {
            // flat staging: element f of the concatenated segments; every lane
            // issues all of its loads before the first LDS store
            uint32_t sb[6], so[6];
#pragma unroll
            for (int t = 0; t < 6; ++t) {
                sb[t] = __shfl(seg_b, t);
                so[t] = __shfl(seg_o, t);
            }
            constexpr int kPer = (tile_cap<R>() + kTile - 1) / kTile;
            constexpr int kBatch = sizeof(R) == 4 ? 3 : 2;  // loads in flight per lane per round
#pragma unroll
            for (int u0 = 0; u0 < kPer; u0 += kBatch) {
                double2 lp[kBatch];
                real4<R> lv[kBatch], lr[kBatch];
#pragma unroll
                for (int u = 0; u < kBatch; ++u) {
                    const uint32_t f = lane + (u0 + u) * kTile;
                    uint32_t gi = s < P.n ? s : 0;  // any valid index for lanes past the end (not stored)
                    if (u0 + u < kPer && f < total) {
#pragma unroll
                        for (int q = 0; q < 6; ++q)
                            if (f >= so[q]) gi = sb[q] + (f - so[q]);  // offsets ascend; last match wins
                    }
                    lp[u] = B.pos[gi];
                    lv[u] = B.va[gi];
                    lr[u] = B.pred[gi];
                }
#pragma unroll
                for (int u = 0; u < kBatch; ++u) {
                    const uint32_t f = lane + (u0 + u) * kTile;
                    if (u0 + u < kPer && f < total) {
                        real2<R> rp;
                        rp.x = (R)(lp[u].x - ox);
                        rp.y = (R)(lp[u].y - oy);
                        s_pos[f] = rp;
                        s_va[f] = lv[u];
                        s_pr[f] = lr[u];
                    }
                }
            }
        }
        // the tile belongs to this wavefront alone: order its LDS writes before its reads
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (active) {
            const R myx = (R)(pj.x - ox), myy = (R)(pj.y - oy);
            const int y0 = cy > 0 ? cy - 1 : 0;
            const int y1 = cy + 1 < P.ncy ? cy + 1 : P.ncy - 1;
            uint32_t count = 0;
            if (use_lds) {
                // this lane's three ranges as (LDS start, length, global start)
                uint32_t rl[3], rn[3], rg[3];
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const int col = cx + r - 1;
                    rl[r] = 0;
                    rn[r] = 0;
                    rg[r] = 0;
                    if (col >= 0 && col < P.ncx) {
                        const uint32_t b = LB[col * P.ncy + y0];
                        const uint32_t e = LB[col * P.ncy + y1 + 1];
                        const int t = g * 3 + r;
                        rl[r] = __shfl(seg_o, t) + (b - __shfl(seg_b, t));
                        rn[r] = e - b;
                        rg[r] = b;
                    }
                }
                // per range: test pass -> bit mask, then the neighbour pass over
                // the set bits only (keeps the lanes of a wavefront on the same
                // amount of neighbour work); ranges longer than 64 go in chunks
#pragma unroll 1
                for (int r = 0; r < 3; ++r) {
                    const uint32_t lo = r == 0 ? rl[0] : (r == 1 ? rl[1] : rl[2]);
                    const uint32_t len = r == 0 ? rn[0] : (r == 1 ? rn[1] : rn[2]);
                    const uint32_t gb = r == 0 ? rg[0] : (r == 1 ? rg[1] : rg[2]);
                    const uint32_t self = s - gb;  // == k - lo for the particle itself (may be out of range)
#pragma unroll 1
                    for (uint32_t base = 0; base < len; base += 32) {
                        const uint32_t lim = min(32u, len - base);
                        uint32_t mask = 0, band = 0;
#pragma unroll 4
                        for (uint32_t f = 0; f < lim; ++f) {
                            const real2<R> rp = s_pos[lo + base + f];
                            const R dx = rp.x - myx, dy = rp.y - myy;
                            const float d2f = (float)(dx * dx + dy * dy);
                            mask |= (d2f < band_hi ? 1u : 0u) << f;
                            band |= (d2f > band_lo ? 1u : 0u) << f;
                        }
                        const uint32_t so_ = self - base;  // the particle itself, if in this chunk
                        if (so_ < lim) mask &= ~(1u << so_);
                        // exact FP64 re-check of the candidates inside the band
                        uint32_t chk = mask & band;
                        while (chk) {
                            const uint32_t f = (uint32_t)__ffs((int)chk) - 1;
                            chk &= chk - 1;
                            if (!exact_inside(B.pos[gb + base + f], pj, P.cut2)) mask &= ~(1u << f);
                        }
#pragma unroll 1
                        while (mask) {
                            const uint32_t f = (uint32_t)__ffs((int)mask) - 1;
                            mask &= mask - 1;
                            if (count == cap) {
                                overflow = true;
                                continue;
                            }
                            ++count;
                            const uint32_t k = lo + base + f;
                            const real2<R> rp = s_pos[k];
                            const R dx = rp.x - myx, dy = rp.y - myy;
                            pair_body<R>(C, acc, dx, dy, dx * dx + dy * dy, s_va[k], s_pr[k], vaj, prj, tj);
                        }
                    }
                }
            } else {
#pragma unroll 1
                for (int dc = -1; dc <= 1; ++dc) {
                    const int col = cx + dc;
                    if (col < 0 || col >= P.ncx) continue;
                    const uint32_t b = LB[col * P.ncy + y0];
                    const uint32_t e = LB[col * P.ncy + y1 + 1];
#pragma unroll 1
                    for (uint32_t i = b; i < e; ++i) {
                        if (i == s) continue;
                        const double2 pi = B.pos[i];
                        const R dx = (R)(pi.x - ox) - myx, dy = (R)(pi.y - oy) - myy;
                        const R d2 = dx * dx + dy * dy;
                        const float d2f = (float)d2;
                        if (d2f >= band_hi) continue;
                        if (d2f > band_lo && !exact_inside(pi, pj, P.cut2)) continue;
                        if (count == cap) {
                            overflow = true;
                            continue;
                        }
                        ++count;
                        pair_body<R>(C, acc, dx, dy, d2, B.va[i], B.pred[i], vaj, prj, tj);
                    }
                }
            }
        }
    }
    const R m = (R)P.mass;
    const R ax = acc.ljx + m * acc.sx;
    const R ay = -(R)P.g + acc.ljy + m * acc.sy;
    const R drho = m * acc.sd;
    const R xf = m * R(2) * (R)P.xsph_eps;
    pair_epilogue<R, FUSED, EXPORT>(P, ctrl, lb, s, live, overflow, pj, vaj, prj, oid, ax, ay, drho, xf * acc.xx,
                                    xf * acc.xy, acc.amu, A, key_next, live ? skey[s] : 0u, der, xs, maxslots);
}

}  // namespace rrsph
