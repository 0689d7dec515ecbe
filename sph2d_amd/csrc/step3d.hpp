// The WCSPH step in 3D (dim = 3): the same pipeline as the 2D step
// (prologue -> cell sort -> reorder -> [density summation] -> pair pass ->
// [integrate]) with
//   pos   double4   (x, y, z, itype)  FP64 positions; the type rides along so
//                   the candidate loads of the neighbour walk carry it
//   va    real4     (vx, vy, vz, rho)
//   ad    real4     (ax, ay, az, drho) of the previous step
//   pred  real4     (vx^, vy^, vz^, rho^);  pw = p / rho^2;  pres = p
// Cell key (cx * ncy + cy) * ncz + cz: the three z-cells of one (x, y)
// column are contiguous, so a neighbour walk is nine contiguous ranges in
// the reference's x, y, z cell order (GridUtils.h:215-250).  y is up
// (gravity along -y, UpdateAcceleration.cpp:63) in 3D as in 2D.
// The 3D step covers the cubic-spline configuration (every term with the
// cubic kernel, pressure as p_i/rho_i^2 + p_j/rho_j^2, no artificial
// pressure, summation or continuity density).
#pragma once

#include "device_common.hpp"
#include "physics.hpp"
#include "step_kernels.hpp"

namespace rrsph {

template <class R> struct StateA3 {
    double4* pos;
    real4<R>* va;
    real4<R>* ad;
    uint32_t* oid;
};
template <class R> struct StateB3 {
    double4* pos;
    real4<R>* va;
    real4<R>* pred;
    R* pw;
    R* pres;
    uint32_t* oid;
    real4<R>* ad = nullptr;  // copy of A.ad, kept when a fused step may have to be undone
};

__device__ __forceinline__ uint32_t cell_key3(const StepParams& P, double x, double y, double z, uint32_t& bad) {
    const int cx = cell_coord(x, P.xmin, P.cell, P.ncx, bad);
    const int cy = cell_coord(y, P.ymin, P.cell, P.ncy, bad);
    const int cz = cell_coord(z, P.zmin, P.cell, P.ncz, bad);
    return ((uint32_t)cx * (uint32_t)P.ncy + (uint32_t)cy) * (uint32_t)P.ncz + (uint32_t)cz;
}

// |r|^2 of the reference's FP64 test, products and sums rounded separately.
__device__ __forceinline__ double dist2_exact3(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    const double xx = dx * dx;
    const double yy = dy * dy;
    const double zz = dz * dz;
    return (xx + yy) + zz;
}

// ------------------------------------------------------------------ reorder
template <class R>
__global__ void __launch_bounds__(kBlock) reorder3_kernel(StepParams P, const Ctrl* __restrict__ ctrl,
                                                          const uint32_t* __restrict__ skey,
                                                          const uint32_t* __restrict__ perm, StateA3<R> A,
                                                          StateB3<R> B, uint32_t* __restrict__ LB,
                                                          unsigned long long* __restrict__ maxslots,
                                                          uint32_t* __restrict__ skey_keep) {
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
    const R dt = (R)ctrl->cur_dt;
    const bool cont = P.density_treatment != 0;
    real4<R> pr;
    pr.x = va.x + ad.x * dt * R(0.5);
    pr.y = va.y + ad.y * dt * R(0.5);
    pr.z = va.z + ad.z * dt * R(0.5);
    pr.w = cont ? va.w + ad.w * dt * R(0.5) : va.w;
    const R p = eos_p<R>(va.w, (R)P.rho0, (R)P.eos_B);
    if (live) {
        B.pos[s] = pos;
        B.va[s] = va;
        B.oid[s] = oid;
        B.pred[s] = pr;
        B.pw[s] = p / (pr.w * pr.w);
        B.pres[s] = p;
        if (B.ad) B.ad[s] = ad;
        if (skey_keep) skey_keep[s] = k;
    }
    fill_cell_starts(P, s, live, k, skey, LB);
}

// ------------------------------------------------------------------ neighbour walks
// Calls f(i, dx, dy, dz) for every neighbour i != s, nine column ranges in
// the reference's cell order, first `cap` hits kept (GridFind.cpp:144-188).
template <class F>
__device__ __forceinline__ void for_each_neighbour3(const StepParams& P, uint32_t cap, uint32_t s, uint32_t key,
                                                    double4 pj, const double4* __restrict__ pos,
                                                    const uint32_t* __restrict__ LB, bool& overflow, F&& f) {
    const uint32_t cz = key % (uint32_t)P.ncz;
    const uint32_t cxy = key / (uint32_t)P.ncz;
    const int cx = (int)(cxy / (uint32_t)P.ncy);
    const int cy = (int)cxy - cx * P.ncy;
    const int z0 = cz > 0 ? (int)cz - 1 : 0;
    const int z1 = (int)cz + 1 < P.ncz ? (int)cz + 1 : P.ncz - 1;
    uint32_t count = 0;
#pragma unroll 1
    for (int dxc = -1; dxc <= 1; ++dxc) {
        const int x = cx + dxc;
        if (x < 0 || x >= P.ncx) continue;
#pragma unroll 1
        for (int dyc = -1; dyc <= 1; ++dyc) {
            const int y = cy + dyc;
            if (y < 0 || y >= P.ncy) continue;
            const uint32_t col = ((uint32_t)x * (uint32_t)P.ncy + (uint32_t)y) * (uint32_t)P.ncz;
            const uint32_t b = LB[col + z0], e = LB[col + z1 + 1];
#pragma unroll 1
            for (uint32_t i = b; i < e; ++i) {
                if (i == s) continue;
                const double4 q = pos[i];
                const double dx = __dsub_rn(q.x, pj.x), dy = __dsub_rn(q.y, pj.y), dz = __dsub_rn(q.z, pj.z);
                if (!(dist2_exact3(dx, dy, dz) < P.cut2)) continue;
                if (count == cap) {
                    overflow = true;
                    continue;
                }
                ++count;
                f(i, dx, dy, dz);
            }
        }
    }
}

// Two-phase walk of the 3D pair pass: hits go to the lane's column of the
// LDS list, flush(nh) consumes them.  A particle has ~75 neighbours in 3D,
// so lists fill during the walk; to keep the flushes of a wave together the
// walk is wave-uniform: a uniform loop over the nine columns and, in each,
// over the wave's longest range (shorter ranges masked), with the flush
// decided by a ballot.  Every lane of the wave calls it; lanes with
// active = false walk empty ranges.  Same neighbour set, visiting order and
// overflow rule as for_each_neighbour3.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}

constexpr uint32_t kCollect3 = 4;  // candidates per batch of the 3D walk
template <uint32_t kChunk, class F>
__device__ __forceinline__ void collect_neighbours3(const StepParams& P, bool active, uint32_t cap, uint32_t s,
                                                    uint32_t key, double4 pj, const double4* __restrict__ pos,
                                                    const uint32_t* __restrict__ LB, uint32_t* __restrict__ hits,
                                                    bool& overflow, F&& flush) {
    static_assert(kChunk >= 2 * kCollect3, "hit list too short");
    const uint32_t cz = key % (uint32_t)P.ncz;
    const uint32_t cxy = key / (uint32_t)P.ncz;
    const int cx = (int)(cxy / (uint32_t)P.ncy);
    const int cy = (int)cxy - cx * P.ncy;
    const int z0 = cz > 0 ? (int)cz - 1 : 0;
    const int z1 = (int)cz + 1 < P.ncz ? (int)cz + 1 : P.ncz - 1;
    uint32_t nh = 0, done = 0;
#pragma unroll 1
    for (int c = 0; c < 9; ++c) {
        const int x = cx + c / 3 - 1, y = cy + c % 3 - 1;
        const bool ok = active && x >= 0 && x < P.ncx && y >= 0 && y < P.ncy;
        const uint32_t col = ok ? ((uint32_t)x * (uint32_t)P.ncy + (uint32_t)y) * (uint32_t)P.ncz : 0u;
        const uint32_t b = ok ? LB[col + z0] : 0u;
        const uint32_t len = ok ? LB[col + z1 + 1] - b : 0u;
        const uint32_t lmax = wave_max_u32(len);
#pragma unroll 1
        for (uint32_t i = 0; i < lmax; i += kCollect3) {
            double4 q[kCollect3];
#pragma unroll
            for (uint32_t u = 0; u < kCollect3; ++u) q[u] = pos[len ? b + min(i + u, len - 1) : s];
            bool hit[kCollect3];
#pragma unroll
            for (uint32_t u = 0; u < kCollect3; ++u) {
                const double dx = __dsub_rn(q[u].x, pj.x), dy = __dsub_rn(q[u].y, pj.y), dz = __dsub_rn(q[u].z, pj.z);
                hit[u] = (i + u < len) & (b + i + u != s) & (dist2_exact3(dx, dy, dz) < P.cut2);
            }
            if (done + nh + kCollect3 <= cap) {
#pragma unroll
                for (uint32_t u = 0; u < kCollect3; ++u) {
                    hits[nh * kBlock] = b + i + u;
                    nh += hit[u] ? 1u : 0u;
                }
            } else {
#pragma unroll
                for (uint32_t u = 0; u < kCollect3; ++u) {
                    if (!hit[u]) continue;
                    if (done + nh < cap) {
                        hits[nh * kBlock] = b + i + u;
                        ++nh;
                    } else {
                        overflow = true;
                    }
                }
            }
            if (__any(nh + kCollect3 > kChunk)) {  // wave-uniform
                flush(nh);
                done += nh;
                nh = 0;
            }
        }
    }
    flush(nh);
}

// ------------------------------------------------------------------ summation density
template <class R>
__global__ void __launch_bounds__(kBlock) density_sum3_kernel(StepParams P, Ctrl* __restrict__ ctrl, StateB3<R> B,
                                                              const uint32_t* __restrict__ skey,
                                                              const uint32_t* __restrict__ LB) {
    if (ctrl->halted) return;
    const uint32_t lb = xcd_block(P.blocks_per_xcd);
    if (lb >= P.nblocks) return;
    const uint32_t s = lb * kBlock + threadIdx.x;
    if (s >= P.n) return;
    const R m = (R)P.mass, inv_h = (R)(1.0 / P.h), w_q1 = (R)P.f_cubic, w_q2 = (R)(P.f_cubic / 6.0);
    const double4 pj = B.pos[s];
    const R w0 = w_q1 * (R(2) / R(3));
    R rho = m * w0;
    // renormalised step: the denominator sum_i (m / rho_i) W_ji, rho_i the
    // density before the step (ModelParams.tex:124-133)
    const bool renorm = ctrl->cur_renorm != 0;
    R wsum = renorm ? m / B.va[s].w * w0 : R(0);
    bool overflow = false;
    if ((int)pj.w != 0) {
        for_each_neighbour3(P, P.max_nbr - 1, s, skey[s], pj, B.pos, LB, overflow,
                            [&](uint32_t i, double dx, double dy, double dz) {
                                const R q = pair_dist<R>(dx, dy, dz) * inv_h;
                                const R t = R(2) - q;
                                const R w = q <= R(1) ? w_q1 * (R(2) / R(3) - q * q + R(0.5) * q * q * q)
                                                      : w_q2 * t * t * t;
                                rho += m * w;
                                if (renorm) wsum += m / B.va[i].w * w;
                            });
    }
    if (renorm) rho = rho / wsum;
    if (overflow) atomicOr(&ctrl->nbr_overflow, 1u);
    real4<R> pr = B.pred[s];
    const R p = eos_p<R>(rho, (R)P.rho0, (R)P.eos_B);
    pr.w = rho;
    B.pred[s] = pr;
    B.pw[s] = p / (rho * rho);
    B.pres[s] = p;
}

// ------------------------------------------------------------------ whole_step
// whole_step of slot s into A (WholeStep.cpp:19-61 with boundary motion);
// returns the new position (with the type in w) and velocity/density.
template <class R> struct Stepped3 {
    double4 pos;
    real4<R> va;
};
template <class R>
__device__ __forceinline__ Stepped3<R> integrate_one3(const StepParams& P, const Ctrl& c, uint32_t s, double4 pos,
                                               real4<R> va, const real4<R>& pred, uint32_t oid, R ax, R ay, R az,
                                               R drho, R avx, R avy, R avz, double dt, StateA3<R> A,
                                               uint32_t* __restrict__ key_next, uint32_t key_old, uint32_t& outside,
                                               uint32_t& bad, uint32_t& changed) {
    const double vdt_d = c.cur_first ? dt * 0.5 : dt;
    const R vdt = (R)vdt_d;
    int itype = (int)pos.w;
    const R rho = P.density_treatment != 0 ? va.w + drho * vdt : pred.w;
    R vx = va.x, vy = va.y, vz = va.z;
    if (c.wall_disappear && oid >= P.nwm_start && oid < P.nwm_end && oid >= P.nfluid) itype = 0;
    if (itype > 0) {
        vx += ax * vdt;
        vy += ay * vdt;
        vz += az * vdt;
        if (P.xsph_on) {
            vx += avx;
            vy += avy;
            vz += avz;
        }
        const double nx = pos.x + (double)vx * dt, ny = pos.y + (double)vy * dt, nz = pos.z + (double)vz * dt;
        if (P.consistency_treatment == 2) {
            if (nx < P.xmax && nx > P.xmin && ny < P.ymax && ny > P.ymin && nz < P.zmax && nz > P.zmin) {
                pos.x = nx;
                pos.y = ny;
                pos.z = nz;
            } else {
                itype = 0;
            }
        } else {
            pos.x = nx;
            pos.y = ny;
            pos.z = nz;
        }
    } else if (P.boundary_motion && itype < 0) {
        if (c.piston_active && oid >= P.nwm_start && oid < P.nwm_end && oid >= P.nfluid) {
            const double pv = P.piston_segments > 1 ? c.piston_vseg[piston_segment(P, pos.z)] : c.piston_v;
            vx = P.nwm_axis == 0 ? (R)pv : R(0);
            vy = P.nwm_axis == 1 ? (R)pv : R(0);
            vz = P.nwm_axis == 2 ? (R)pv : R(0);
        }
        pos.x += (double)vx * dt;
        pos.y += (double)vy * dt;
        pos.z += (double)vz * dt;
    }
    const bool inside =
        pos.x < P.xmax && pos.x > P.xmin && pos.y < P.ymax && pos.y > P.ymin && pos.z < P.zmax && pos.z > P.zmin;
    if (!inside) outside += 1;
    pos.w = (double)itype;
    A.pos[s] = pos;
    real4<R> nva;
    nva.x = vx;
    nva.y = vy;
    nva.z = vz;
    nva.w = rho;
    A.va[s] = nva;
    real4<R> nad;
    nad.x = ax;
    nad.y = ay;
    nad.z = az;
    nad.w = drho;
    A.ad[s] = nad;
    A.oid[s] = oid;
    const uint32_t key = itype == 0 ? P.ncells : cell_key3(P, pos.x, pos.y, pos.z, bad);
    key_next[s] = key;
    changed = key != key_old ? 1u : 0u;
    return Stepped3<R>{pos, nva};
}

// ------------------------------------------------------------------ pair pass
// der = (ax, ay, az, drho), xs = (avx, avy, avz, max mu) when dt comes from
// this step's reductions or the fields are exported.
template <class R>
__device__ __forceinline__ void fast_pair_term3(const FastPair<R>& K, R& sx, R& sy, R& sz, R& sd, R& xx, R& xy,
                                                R& xz, R& amu, R& ljx, R& ljy, R& ljz, R dx, R dy, R dz,
                                                const real4<R>& vai, const real4<R>& pri, R pwi, int ti,
                                                const real4<R>& vaj, const real4<R>& prj, R pwj, int tj) {
    const R d2 = dx * dx + dy * dy + dz * dz;
    const R rinv = d2 > R(0) ? fast_rsqrt<R>(d2) : R(0);
    const R dist = d2 * rinv;
    const R q = dist * K.inv_h;
    const R t = R(2) - q;
    const R t2 = t * t;
    const bool inner = q <= R(1);
    const R F = inner ? K.c_q1 * (R(1.5) * q - R(2)) : -K.c_q2 * t2 * rinv;
    const R W = inner ? K.w_q1 * (R(2) / R(3) - q * q + R(0.5) * q * q * q) : K.w_q2 * t2 * t;
    if (K.cont) sd += F * ((vai.x - vaj.x) * dx + (vai.y - vaj.y) * dy + (vai.z - vaj.z) * dz);
    if (K.lj_on && tj > 0 && ti < 0 && dist < K.lj_r0) {
        const R rn = K.lj_r0 / dist;
        const R rn2 = rn * rn, rn4 = rn2 * rn2;
        const R f = K.lj_D * (rn4 * rn4 * rn4 - rn4) / d2;
        ljx -= dx * f;
        ljy -= dy * f;
        ljz -= dz * f;
    }
    const R inv_rs = fast_rcp<R>(pri.w + prj.w);
    R coef = pwi + pwj;
    const R dvx = pri.x - prj.x, dvy = pri.y - prj.y, dvz = pri.z - prj.z;
    if (K.av_on) {
        const R vr = dvx * dx + dvy * dy + dvz * dz;
        const R mu = K.h * vr * fast_rcp<R>(d2 + K.eta2);
        const bool app = vr < R(0);
        amu = app ? fmax(amu, -mu) : amu;
        coef += app ? (K.beta * mu - K.alpha_c) * mu * R(2) * inv_rs : R(0);
    }
    const R fc = F * coef;
    sx += fc * dx;
    sy += fc * dy;
    sz += fc * dz;
    if (K.xsph_on) {
        const R cw = ti > 0 ? W * inv_rs : R(0);
        xx += dvx * cw;
        xy += dvy * cw;
        xz += dvz * cw;
    }
}

template <class R, bool FUSED, bool EXPORT>
__global__ void __launch_bounds__(kBlock) pair3_kernel(StepParams P, Ctrl* __restrict__ ctrl, StateB3<R> B,
                                                       const uint32_t* __restrict__ skey,
                                                       const uint32_t* __restrict__ LB, StateA3<R> A,
                                                       uint32_t* __restrict__ key_next, real4<R>* __restrict__ der,
                                                       real4<R>* __restrict__ xs,
                                                       unsigned long long* __restrict__ maxslots) {
    __shared__ uint32_t hits[kPairChunk<R> * kBlock];
    if (ctrl->halted) return;
    const uint32_t lb = xcd_block(P.blocks_per_xcd);
    if (lb >= P.nblocks) return;
    const uint32_t s = lb * kBlock + threadIdx.x;
    const bool live = s < P.n;
    const FastPair<R> K(P);
    R sx = 0, sy = 0, sz = 0, sd = 0, xx = 0, xy = 0, xz = 0, amu = 0, ljx = 0, ljy = 0, ljz = 0;
    bool overflow = false;
    double4 pj = make_double4(0, 0, 0, 0);
    real4<R> vaj{}, prj{};
    R pwj = 0;
    uint32_t oid = 0;
    if (live) {
        pj = B.pos[s];
        vaj = B.va[s];
        prj = B.pred[s];
        pwj = B.pw[s];
        oid = B.oid[s];
    }
    const int tj = (int)pj.w;
    {
        uint32_t* my = hits + threadIdx.x;
        auto flush = [&](uint32_t nh) {
            if (nh == 0) return;
            const uint32_t last = nh - 1;
            uint32_t i1 = my[min(1u, last) * kBlock];
            const uint32_t i0 = my[0];
            // the current velocity enters the continuity term only; with
            // summation density its gather is skipped (a uniform branch)
            double4 qn = B.pos[i0];
            real4<R> van{}, prn = B.pred[i0];
            if (K.cont) van = B.va[i0];
            R pwn = B.pw[i0];
#pragma unroll 1
            for (uint32_t u = 0; u < nh; ++u) {
                const double4 q = qn;
                const real4<R> vai = van, pri = prn;
                const R pwi = pwn;
                qn = B.pos[i1];
                if (K.cont) van = B.va[i1];
                prn = B.pred[i1];
                pwn = B.pw[i1];
                i1 = my[min(u + 2, last) * kBlock];
                fast_pair_term3<R>(K, sx, sy, sz, sd, xx, xy, xz, amu, ljx, ljy, ljz, (R)__dsub_rn(q.x, pj.x),
                                   (R)__dsub_rn(q.y, pj.y), (R)__dsub_rn(q.z, pj.z), vai, pri, pwi, (int)q.w, vaj,
                                   prj, pwj, tj);
            }
        };
        collect_neighbours3<kPairChunk<R>>(P, live && tj != 0, P.max_nbr - 1, live ? s : 0u, live ? skey[s] : 0u,
                                           pj, B.pos, LB, my, overflow, flush);
    }
    const R m = (R)P.mass;
    const R ax = ljx + m * sx;
    const R ay = -(R)P.g + ljy + m * sy;
    const R az = ljz + m * sz;
    const R drho = m * sd;
    const R xf = m * R(2) * (R)P.xsph_eps;
    const R avx = xf * xx, avy = xf * xy, avz = xf * xz;
    if (overflow) atomicOr(&ctrl->nbr_overflow, 1u);
    if (P.dt_method == 2 && P.dt_dynamic_mode == 1) {
        block_max2_to_slots(live ? sqrt((double)ax * ax + (double)ay * ay + (double)az * az) : 0.0,
                            live ? (double)amu : 0.0, maxslots, lb);
    }
    if ((EXPORT || !FUSED) && live) {
        real4<R> d;
        d.x = ax;
        d.y = ay;
        d.z = az;
        d.w = drho;
        der[s] = d;
        real4<R> x;
        x.x = avx;
        x.y = avy;
        x.z = avz;
        x.w = amu;
        xs[s] = x;
    }
    if (FUSED) {
        uint32_t outside = 0, bad = 0, changed = 0;
        if (lb == 0 && threadIdx.x == 0) {
            ctrl->dt_new = ctrl->cur_dt_const;
            ctrl->pending = 1;
        }
        if (live) {
            const Ctrl& c = *ctrl;
            integrate_one3<R>(P, c, s, pj, vaj, prj, oid, ax, ay, az, drho, avx, avy, avz, c.cur_dt_const, A,
                              key_next, skey[s], outside, bad, changed);
        }
        wave_count_add(&ctrl->outside, outside);
        store_changed_mask(P, changed, s);
        if (bad) atomicOr(&ctrl->grid_overflow, 1u);
    }
}

template <class R>
__global__ void __launch_bounds__(kBlock) integrate3_kernel(StepParams P, Ctrl* __restrict__ ctrl, StateB3<R> B,
                                                            const real4<R>* __restrict__ der,
                                                            const real4<R>* __restrict__ xs, StateA3<R> A,
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
        const real4<R> x = xs[s];
        // the predicted state is needed for the summed density only
        const real4<R> pr = P.density_treatment != 0 ? real4<R>{} : B.pred[s];
        integrate_one3<R>(P, c, s, B.pos[s], B.va[s], pr, B.oid[s], d.x, d.y, d.z, d.w, x.x, x.y, x.z, dt, A,
                          key_next, skey[s], outside, bad, changed);
    }
    wave_count_add(&c.outside, outside);
    store_changed_mask(P, changed, s);
    if (bad) atomicOr(&c.grid_overflow, 1u);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c.dt_new = dt;
        c.pending = 1;
    }
}

// ------------------------------------------------------------------ host-facing utilities
template <class R>
__global__ void init3_kernel(StepParams P, const double* __restrict__ r, const double* __restrict__ v,
                             const double* __restrict__ rho, const int32_t* __restrict__ itype, StateA3<R> A,
                             uint32_t* __restrict__ key_next, Ctrl* __restrict__ ctrl) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.n) return;
    const double4 p = make_double4(r[3 * s], r[3 * s + 1], r[3 * s + 2], (double)itype[s]);
    A.pos[s] = p;
    real4<R> va;
    va.x = (R)v[3 * s];
    va.y = (R)v[3 * s + 1];
    va.z = (R)v[3 * s + 2];
    va.w = (R)rho[s];
    A.va[s] = va;
    real4<R> z{};
    A.ad[s] = z;
    A.oid[s] = s;
    uint32_t bad = 0;
    key_next[s] = itype[s] == 0 ? P.ncells : cell_key3(P, p.x, p.y, p.z, bad);
    if (bad) atomicOr(&ctrl->grid_overflow, 1u);
}

template <class R>
__global__ void download3_kernel(StepParams P, StateA3<R> A, const R* __restrict__ pres, double* __restrict__ r,
                                 double* __restrict__ v, double* __restrict__ rho, double* __restrict__ p,
                                 int32_t* __restrict__ itype) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.n) return;
    const uint32_t o = A.oid[s];
    const double4 pos = A.pos[s];
    const real4<R> va = A.va[s];
    r[3 * o] = pos.x;
    r[3 * o + 1] = pos.y;
    r[3 * o + 2] = pos.z;
    v[3 * o] = (double)va.x;
    v[3 * o + 1] = (double)va.y;
    v[3 * o + 2] = (double)va.z;
    rho[o] = (double)va.w;
    itype[o] = (int32_t)pos.w;
    p[o] = pres ? (double)pres[s] : (double)eos_p<R>(va.w, (R)P.rho0, (R)P.eos_B);
}

template <class R>
__global__ void fields3_kernel(uint32_t n, const uint32_t* __restrict__ oid, const real4<R>* __restrict__ der,
                               const real4<R>* __restrict__ xs, double* __restrict__ a, double* __restrict__ drho,
                               double* __restrict__ av, double* __restrict__ mu) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t o = oid[s];
    const real4<R> d = der[s];
    const real4<R> x = xs[s];
    a[3 * o] = (double)d.x;
    a[3 * o + 1] = (double)d.y;
    a[3 * o + 2] = (double)d.z;
    drho[o] = (double)d.w;
    av[3 * o] = (double)x.x;
    av[3 * o + 1] = (double)x.y;
    av[3 * o + 2] = (double)x.z;
    mu[o] = (double)x.w;
}

__device__ __forceinline__ uint32_t split3(uint32_t x) {
    x &= 0x3FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// 3D Morton cell key of the reference (GridUtils.h:84-105, 145-151).
__global__ void morton3_kernel(StepParams P, const double4* __restrict__ pos, const uint32_t* __restrict__ oid,
                               uint32_t* __restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.n) return;
    const double4 p = pos[s];
    const unsigned long long cx = (unsigned long long)((p.x - P.xmin) / P.cell);
    const unsigned long long cy = (unsigned long long)((p.y - P.ymin) / P.cell);
    const unsigned long long cz = (unsigned long long)((p.z - P.zmin) / P.cell);
    out[oid[s]] = split3((uint32_t)cx) | (split3((uint32_t)cy) << 1) | (split3((uint32_t)cz) << 2);
}

template <class R>
__global__ void neighbours3_kernel(StepParams P, StateB3<R> B, const uint32_t* __restrict__ skey,
                                   const uint32_t* __restrict__ LB, uint32_t* __restrict__ out, uint32_t max_nbr,
                                   Ctrl* __restrict__ ctrl) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.n) return;
    const uint32_t o = B.oid[s];
    uint32_t cnt = 0;
    bool overflow = false;
    const double4 pj = B.pos[s];
    if ((int)pj.w != 0) {
        for_each_neighbour3(P, max_nbr - 1, s, skey[s], pj, B.pos, LB, overflow,
                            [&](uint32_t i, double, double, double) {
                                out[cnt + (size_t)max_nbr * o] = B.oid[i];
                                ++cnt;
                            });
    }
    out[cnt + (size_t)max_nbr * o] = P.n;
    if (overflow) atomicOr(&ctrl->nbr_overflow, 1u);
}

}  // namespace rrsph
