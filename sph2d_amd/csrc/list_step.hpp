// Neighbour-list steps (2D): the cell sort and the cell walk once every
// several steps instead of every step.
//
// A build step sorts the particles by cell, reorders the state, and walks
// the cells within reach k h + skin of each particle (list_build_kernel):
// every particle i with |r_i - r_j| < k h + skin goes into j's list, in the
// walk order, k-major (entry k of slot j at nbl[k n + j], count nbc[j]).
// Until the next build the slot order stays as it is, and a step reads the
// lists instead of walking the cells: the pair pass applies the reference's
// exact FP64 cut-off test (GridFind.cpp:144-176, |r|^2 < (k h)^2 with
// separate roundings) to each entry, so the neighbour sets are those of the
// cell walk.  A pair closer than k h now was closer than k h + skin at the
// build as long as no particle has moved skin / 2 since: the integrator
// reduces each step's largest displacement, the prologue sums them, and a
// step whose lists may be stale halts before it reads them (status
// kStatusRebuild); the host re-runs it as a build step.  The host schedules
// builds ahead from the measured displacement rate, so such halts are rare.
//
// Step kernels: list_prologue -> [density_sum_list] -> pair_list ->
// integrate_list (in place; it also writes the next step's predicted
// velocity and density, so no reorder runs between builds).  A build step
// runs list_prologue -> radix sort -> reorder_kernel -> list_build_kernel ->
// the same step kernels on the reordered copy, integrating into A.
#pragma once

#include "device_common.hpp"
#include "step_kernels.hpp"

namespace rrsph {

constexpr uint32_t kStatusRebuild = 101;       // internal: re-run this step as a list build
constexpr uint32_t kStatusListOverflow = 102;  // internal: a list ran out of room; run without lists

// The state a list step reads (the reordered copy B on a build step, A after).
template <class R> struct ListState {
    const double2* pos;
    const real4<R>* va;
    const uint32_t* oid;
};

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
    double d = dbits(dslots[threadIdx.x]);
    for (int off = 32; off > 0; off >>= 1) d = fmax(d, __shfl_xor(d, off));
    if (__lane_id() == 0) s_m[threadIdx.x / kWave] = d;
    __syncthreads();
    if (threadIdx.x == 0) {
        double dmax = s_m[0];
        for (int w = 1; w < kBlock / kWave; ++w) dmax = fmax(dmax, s_m[w]);
        Ctrl& c = *ctrl;
        if (!c.halted && c.pending) {  // the step being committed moved no particle further than dmax
            c.disp_acc += dmax;
            c.disp_last = dmax;
        }
        const uint32_t valid = c.list_valid;
        step_prologue_body(P, ctrl);
        if (!c.halted) {
            c.list_valid = build ? 1u : valid;
            if (build) {
                c.disp_acc = 0.0;
                c.list_overflow = 0;
            } else if (!valid || !(c.disp_acc < P.skin_half)) {
                c.halted = 1;
                c.status = kStatusRebuild;
                if (c.wall_disappear) c.nwm_done = 0;  // the re-run prologue fires it again
            }
        }
    }
    __syncthreads();
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
                                                            const uint32_t* __restrict__ LB, uint32_t* __restrict__ nbl,
                                                            uint32_t* __restrict__ nbc) {
    if (ctrl->halted) return;
    const uint32_t lb = xcd_block(P.blocks_per_xcd);
    if (lb >= P.nblocks) return;
    const uint32_t s = lb * kBlock + threadIdx.x;
    if (s >= P.n) return;
    uint32_t cnt = 0;
    if ((int)va[s].w != 0) {
        const double2 pj = pos[s];
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
                        if (cnt < P.list_cap) nbl[(size_t)cnt * P.n + s] = i + u;
                        ++cnt;
                    }
                }
            }
        }
    }
    if (cnt > P.list_cap) atomicOr(&ctrl->list_overflow, 1u);
    nbc[s] = min(cnt, P.list_cap);
}

// Visits the entries of slot s's list: f(take, dx, dy, va_i, pred_i) for
// every entry, take = the entry passes the reference's exact test, is alive
// and is among the first `cap` such (further ones set `overflow`,
// GridFind.cpp:158-188); f must add nothing when take is false.  Every
// entry runs the same instructions (most entries are neighbours, so
// skipping the others would not shorten a wave's work).  The list index
// two entries ahead and the positions, velocities and predictions one
// entry ahead are in flight while f runs.
template <class R, class F>
__device__ __forceinline__ void for_each_listed(const StepParams& P, uint32_t cap, uint32_t s, double2 pj,
                                                const double2* __restrict__ pos, const real4<R>* __restrict__ va,
                                                const real4<R>* __restrict__ pred, const uint32_t* __restrict__ nbl,
                                                uint32_t cnt, bool& overflow, F&& f) {
    if (cnt == 0) return;
    const size_t n = P.n;
    const uint32_t last = cnt - 1;
    uint32_t hits = 0;
    bool over = false;
    auto visit = [&](const double2& q, const real4<R>& a, const real4<R>& b) {
        const double dx = __dsub_rn(q.x, pj.x), dy = __dsub_rn(q.y, pj.y);
        const bool hit = (dist2_exact(dx, dy) < P.cut2) & ((int)a.w != 0);
        const bool take = hit & (hits < cap);
        over |= hit & !take;
        hits += take ? 1u : 0u;
        f(take, dx, dy, a, b);
    };
    // two register sets in turn (no copies between iterations): while one
    // entry is processed the next one's gathers and the index after it are
    // in flight; each index is loaded before the gathers issued with it
    uint32_t iA = nbl[s];
    uint32_t iB = nbl[min(1u, last) * n + s];
    double2 qA = pos[iA], qB;
    real4<R> aA = va[iA], bA = pred[iA], aB, bB;
    uint32_t k = 0;
#pragma unroll 1
    for (; k + 1 < cnt; k += 2) {
        iA = nbl[min(k + 2, last) * n + s];
        qB = pos[iB];
        aB = va[iB];
        bB = pred[iB];
        visit(qA, aA, bA);
        iB = nbl[min(k + 3, last) * n + s];
        qA = pos[iA];
        aA = va[iA];
        bA = pred[iA];
        visit(qB, aB, bB);
    }
    if (k < cnt) visit(qA, aA, bA);
    overflow |= over;
}

// fast_pair_term for one list entry without branches except the rare
// Lennard-Jones case: both pieces of the cubic spline are evaluated and
// selected, and nothing is added when take is false.  The Lennard-Jones
// ratios use 1/|r| from the reciprocal square root instead of two divisions.
template <class R>
__device__ __forceinline__ void list_pair_term(const FastPair<R>& K, PairSums<R>& S, bool take, R dx, R dy,
                                               const real4<R>& vai, const real4<R>& pri, const real4<R>& vaj,
                                               const real4<R>& prj, int tj) {
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
    if (K.cont) S.sd += F * ((vai.x - vaj.x) * dx + (vai.y - vaj.y) * dy);
    const int ti = (int)vai.w;
    if (K.lj_on && take && tj > 0 && ti < 0 && dist < K.lj_r0) {
        const R rn = K.lj_r0 * rinv;
        const R rn2 = rn * rn, rn4 = rn2 * rn2;
        const R f = K.lj_D * (rn4 * rn4 * rn4 - rn4) * (rinv * rinv);
        S.ljx -= dx * f;
        S.ljy -= dy * f;
    }
    const R inv_rs = fast_rcp<R>(pri.z + prj.z);
    R coef = pri.w + prj.w;
    const R dvx = pri.x - prj.x, dvy = pri.y - prj.y;
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
        const R cw = take & (ti > 0) ? W * inv_rs : R(0);
        S.xx += dvx * cw;
        S.xy += dvy * cw;
    }
}

// Summation density from the lists (Density.h:25-61): rho = m W(0) +
// sum m W_ij, p = EOS(rho), into the predicted slots (the forces use this
// rho) and the pressure array.  Cubic spline (list steps run the cubic
// configuration).
template <class R>
__global__ void __launch_bounds__(kBlock) density_sum_list(StepParams P, Ctrl* __restrict__ ctrl, ListState<R> X,
                                                           real4<R>* __restrict__ pred, R* __restrict__ pres,
                                                           const uint32_t* __restrict__ nbl,
                                                           const uint32_t* __restrict__ nbc) {
    if (ctrl->halted) return;
    const uint32_t lb = xcd_block(P.blocks_per_xcd);
    if (lb >= P.nblocks) return;
    const uint32_t s = lb * kBlock + threadIdx.x;
    if (s >= P.n) return;
    const KernelConsts<R> K = kernel_consts<R>(P);
    const R m = (R)P.mass;
    const real4<R> vaj = X.va[s];
    const R w0 = kernel_w<R>(K, P.skf_density, R(0));
    R rho = m * w0;
    const bool renorm = ctrl->cur_renorm != 0;  // ModelParams.tex:124-133, rho_i before the step
    R wsum = renorm ? m / vaj.z * w0 : R(0);
    bool overflow = false;
    if ((int)vaj.w != 0) {
        for_each_listed<R>(P, P.max_nbr - 1, s, X.pos[s], X.pos, X.va, pred, nbl, nbc[s], overflow,
                           [&](bool take, double dx, double dy, const real4<R>& vai, const real4<R>&) {
                               const R w = take ? kernel_w<R>(K, P.skf_density, pair_dist<R>(dx, dy, 0.0)) : R(0);
                               rho += m * w;
                               if (renorm) wsum += m / vai.z * w;
                           });
    }
    if (renorm) rho = rho / wsum;
    if (overflow) atomicOr(&ctrl->nbr_overflow, 1u);
    real4<R> pr = pred[s];
    const R p = eos_p<R>(rho, (R)P.rho0, (R)P.eos_B);
    pr.z = rho;
    pr.w = p / (rho * rho);
    pred[s] = pr;
    pres[s] = p;
}

// Pair pass from the lists: continuity density rate, pressure, artificial
// viscosity, Lennard-Jones walls and XSPH (the cubic fast path of
// pair_fast_kernel, the same per-pair arithmetic), the dt maxima, and the
// fields for the integrator.  With continuity it also writes the pressure
// of the current density (the reference's p array).
template <class R>
__global__ void __launch_bounds__(kBlock) pair_list_kernel(StepParams P, Ctrl* __restrict__ ctrl, ListState<R> X,
                                                           const real4<R>* __restrict__ pred, R* __restrict__ pres,
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
    const uint32_t s = lb * kBlock + threadIdx.x;
    const bool live = s < P.n;
    const FastPair<R> K(P);
    PairSums<R> S;
    bool overflow = false;
    double2 pj = make_double2(0, 0);
    real4<R> vaj{}, prj{};
    if (live) {
        pj = X.pos[s];
        vaj = X.va[s];
        prj = pred[s];
    }
    const int tj = (int)vaj.w;
    if (live && tj != 0) {
        for_each_listed<R>(P, P.max_nbr - 1, s, pj, X.pos, X.va, pred, nbl, nbc[s], overflow,
                           [&](bool take, double dx, double dy, const real4<R>& vai, const real4<R>& pri) {
                               list_pair_term<R>(K, S, take, (R)dx, (R)dy, vai, pri, vaj, prj, tj);
                           });
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
// key for the next build, the step's largest displacement, and the
// prediction of the next step (PredictHalfStep.cpp:16-22 with this step's
// dt, which is the next step's: (v^, rho^, p / rho^2)).
template <class R>
__global__ void __launch_bounds__(kBlock) integrate_list_kernel(StepParams P, Ctrl* __restrict__ ctrl, ListState<R> X,
                                                                const real4<R>* __restrict__ der,
                                                                const real2<R>* __restrict__ xs, real4<R>* pred,
                                                                StateA<R> A, uint32_t* __restrict__ key_next,
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
        const real4<R> pr = P.density_treatment != 0 ? real4<R>{} : pred[s];
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
        // the next step's prediction, as reorder_kernel computes it
        const R dtr = (R)dt;
        const bool cont = P.density_treatment != 0;
        real4<R> np;
        np.x = o.va.x + d.x * dtr * R(0.5);
        np.y = o.va.y + d.y * dtr * R(0.5);
        np.z = cont ? o.va.z + d.z * dtr * R(0.5) : o.va.z;
        const R p = eos_p<R>(o.va.z, (R)P.rho0, (R)P.eos_B);
        np.w = p / (np.z * np.z);
        pred[s] = np;
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
