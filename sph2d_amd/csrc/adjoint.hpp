// Discrete adjoint of the FP64 WCSPH step (continuity density, cubic
// spline, pressure p_i/rho_i^2 + p_j/rho_j^2, artificial viscosity, Lennard-
// Jones walls, XSPH, boundary motion with the Fourier piston), and the
// smooth gauge surrogate it is driven by.
//
// The step maps S_n = (r, v, rho, a_prev, drho_prev) to S_{n+1}:
//   u = v + a_prev dt/2, R = rho + drho_prev dt/2, p = EOS(rho), P = p / R^2
//   drho_j = m sum_i F_ij (v_i - v_j).e        (e = r_i - r_j, grad W = F e)
//   a_j    = -g y + LJ_j + m sum_i F_ij coef_ij e,  coef = P_i + P_j + AV_ij
//   av_j   = 2 eps m sum_{i fluid} (u_i - u_j) W_ij / (R_i + R_j)
//   whole_step (dt/2 for v and rho on the first step), piston velocity on
//   the wave-maker particles, a_prev' = a, drho_prev' = drho.
// Neighbour sets are those of the step (the kernel, its gradient, the LJ
// force and the AV switch are continuous at their thresholds, so the sets
// are constant on the pieces the derivative is taken on).
//
// Kernels (sorted order of the replayed step, adjoints stored in original
// particle order, lam = (r, v, rho, a_prev, drho_prev) adjoints, 8 doubles):
//   adj_prep  whole_step reverse: output adjoints (a, av, drho) per particle,
//             the direct input adjoints, and the piston sum
//             s_n = sum over wave-maker particles of (v' + dt r')-adjoints
//   adj_pair  gather over neighbours: every pair is differentiated in both
//             orientations (the particle as owner of the sum and as
//             neighbour), then predict and EOS reverse; writes lam_n
//   smooth_gauge_* the surrogate eta_g = y_ref + tau log(sum w e / sum w)
//             with w = (1 - (x - x_g)^2 / R^2)^3 inside |x - x_g| < R and
//             e = exp((y - y_ref) / tau); tau -> 0 gives max y in the window,
//             the definition of HeightTesting::maxInLayer
//             (tools/WaterProfile/include/HeightTesting.h:25-54)
#pragma once

#include "device_common.hpp"
#include "step_kernels.hpp"

namespace rrsph {

constexpr uint32_t kMaxGauges = 64;
constexpr uint32_t kAdjChunk = 32;

// The adjoint runs in FP64 on the state of either precision (an FP32 state
// is the FP32 step's trajectory; its Jacobian is taken in FP64).
template <class R> __device__ __forceinline__ double4 to_d4(const real4<R>& v) {
    return make_double4((double)v.x, (double)v.y, (double)v.z, (double)v.w);
}

struct GaugeSpec {
    uint32_t n;          // gauges
    int32_t ptype;       // particle type the gauge looks at
    double radius, tau, y_ref;
    double x[kMaxGauges];
};

// ------------------------------------------------------------------ smooth gauges
// Block partial sums (sum w e, sum w) of every gauge, gauge-major:
// partials[(2 g) nb + b] = sum w e, partials[(2 g + 1) nb + b] = sum w of
// block b (nb = gridDim.x).  The grid is at most kGaugeBlocks blocks; block
// b sums the particle chunks b, b + nb, b + 2 nb, ... of kBlock particles.
// A chunk with no particle inside any window adds nothing and skips the
// per-gauge reductions (its sums are exactly zero).
constexpr uint32_t kGaugeBlocks = 1024;

struct GaugeLds {
    double z[kBlock / kWave][kMaxGauges], w[kBlock / kWave][kMaxGauges];
};

__device__ __forceinline__ void gauge_lds_clear(const GaugeSpec& G, GaugeLds& L) {
    for (uint32_t i = threadIdx.x; i < (kBlock / kWave) * G.n; i += kBlock) {
        L.z[i / G.n][i % G.n] = 0.0;
        L.w[i / G.n][i % G.n] = 0.0;
    }
    __syncthreads();
}

// One chunk: every thread of the block calls it (uniform control flow).
template <class D2>
__device__ __forceinline__ void gauge_chunk_sums(const GaugeSpec& G, bool in, double y, D2&& d2, double inv_r2,
                                                 GaugeLds& L) {
    const uint32_t wave = threadIdx.x / kWave;
    bool near = false;
    for (uint32_t g = 0; g < G.n; ++g) near |= in && d2(g) * inv_r2 < 1.0;
    if (!__syncthreads_or(near)) return;
    for (uint32_t g = 0; g < G.n; ++g) {
        const double u = 1.0 - d2(g) * inv_r2;
        const double w = in && u > 0.0 ? u * u * u : 0.0;
        double z = w > 0.0 ? w * exp((y - G.y_ref) / G.tau) : 0.0;
        double ww = w;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            z += __shfl_xor(z, off);
            ww += __shfl_xor(ww, off);
        }
        if (__lane_id() == 0) {  // each wave owns its row
            L.z[wave][g] += z;
            L.w[wave][g] += ww;
        }
    }
}

__device__ __forceinline__ void gauge_block_store(const GaugeSpec& G, GaugeLds& L, double* __restrict__ partials) {
    __syncthreads();
    const uint32_t nb = gridDim.x;
    for (uint32_t g = threadIdx.x; g < G.n; g += kBlock) {
        double z = 0, ww = 0;
        for (uint32_t w = 0; w < kBlock / kWave; ++w) {
            z += L.z[w][g];
            ww += L.w[w][g];
        }
        partials[(size_t)(2 * g) * nb + blockIdx.x] = z;
        partials[(size_t)(2 * g + 1) * nb + blockIdx.x] = ww;
    }
}

__host__ __device__ inline uint32_t gauge_blocks(uint32_t n) {
    const uint32_t chunks = (n + kBlock - 1) / kBlock;
    return chunks < kGaugeBlocks ? (chunks ? chunks : 1u) : kGaugeBlocks;
}

template <class R>
__global__ void __launch_bounds__(kBlock) smooth_gauge_partials(uint32_t n, const double2* __restrict__ pos,
                                                                const real4<R>* __restrict__ va, GaugeSpec G,
                                                                double* __restrict__ partials) {
    __shared__ GaugeLds L;
    gauge_lds_clear(G, L);
    const double inv_r2 = 1.0 / (G.radius * G.radius);
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const uint32_t s = base + threadIdx.x;
        bool in = false;
        double x = 0, y = 0;
        if (s < n) {
            const double2 p = pos[s];
            in = (int)va[s].w == G.ptype;
            x = p.x;
            y = p.y;
        }
        auto d2 = [&](uint32_t g) { return (x - G.x[g]) * (x - G.x[g]); };
        gauge_chunk_sums(G, in, y, d2, inv_r2, L);
    }
    gauge_block_store(G, L, partials);
}

// One block per gauge: each thread sums a fixed stride of the block
// partials, then a fixed shuffle tree and wave order (deterministic);
// writes eta[g], Z[g] = sum w e and Wsum[g] = sum w.  With ctrl, the three
// arrays are a tape of nsteps rows of kMaxGauges and the row is the step
// index of the step just launched - base, as adj_record takes it.
__global__ void __launch_bounds__(kBlock) smooth_gauge_final(uint32_t nblocks, GaugeSpec G,
                                                             const double* __restrict__ partials,
                                                             double* __restrict__ eta, double* __restrict__ zsum,
                                                             double* __restrict__ wsum, const Ctrl* __restrict__ ctrl,
                                                             uint64_t base, uint32_t nsteps) {
    __shared__ double s_z[kBlock / kWave], s_w[kBlock / kWave];
    const uint32_t g = blockIdx.x;
    if (g >= G.n) return;
    if (ctrl) {
        const size_t row = (size_t)min(ctrl->step - base, (uint64_t)nsteps - 1) * kMaxGauges;
        eta += row;
        zsum += row;
        wsum += row;
    }
    const double* pz = partials + (size_t)(2 * g) * nblocks;
    const double* pw = partials + (size_t)(2 * g + 1) * nblocks;
    double z = 0, ww = 0;
    for (uint32_t b = threadIdx.x; b < nblocks; b += kBlock) {
        z += pz[b];
        ww += pw[b];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        z += __shfl_xor(z, off);
        ww += __shfl_xor(ww, off);
    }
    if (__lane_id() == 0) {
        s_z[threadIdx.x / kWave] = z;
        s_w[threadIdx.x / kWave] = ww;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        z = 0;
        ww = 0;
        for (uint32_t w = 0; w < kBlock / kWave; ++w) {
            z += s_z[w];
            ww += s_w[w];
        }
        eta[g] = ww > 0.0 ? G.y_ref + G.tau * log(z / ww) : 0.0;
        zsum[g] = z;
        wsum[g] = ww;
    }
}

// lam[oid].r += sum_g eta_bar[g] d eta_g / d r (state in sorted order).
template <class R>
__global__ void __launch_bounds__(kBlock) smooth_gauge_grad(uint32_t n, const double2* __restrict__ pos,
                                                            const real4<R>* __restrict__ va,
                                                            const uint32_t* __restrict__ oid, GaugeSpec G,
                                                            const double* __restrict__ eta_bar,
                                                            const double* __restrict__ zsum,
                                                            const double* __restrict__ wsum,
                                                            double* __restrict__ lam) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    if ((int)va[s].w != G.ptype) return;
    const double2 p = pos[s];
    const double inv_r2 = 1.0 / (G.radius * G.radius);
    double gx = 0, gy = 0;
    bool seen = false;  // inside a window: only those particles update lam
    for (uint32_t g = 0; g < G.n; ++g) {
        const double dx = p.x - G.x[g];
        const double u = 1.0 - dx * dx * inv_r2;
        if (!(u > 0.0) || !(wsum[g] > 0.0) || eta_bar[g] == 0.0) continue;
        seen = true;
        const double w = u * u * u;
        const double wp = -6.0 * dx * inv_r2 * u * u;  // dw/dx
        const double e = exp((p.y - G.y_ref) / G.tau);
        const double z = zsum[g], ws = wsum[g];
        gy += eta_bar[g] * w * e / z;
        gx += eta_bar[g] * G.tau * wp * (e / z - 1.0 / ws);
    }
    if (!seen) return;
    double* l = lam + 8 * (size_t)oid[s];
    l[0] += gx;
    l[1] += gy;
}

// Cell keys of a restored FP64 state (dead particles get the dead key).
template <class R>
__global__ void __launch_bounds__(kBlock) state_keys_kernel(StepParams P, const double2* __restrict__ pos,
                                                            const real4<R>* __restrict__ va,
                                                            uint32_t* __restrict__ key_next) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= P.n) return;
    uint32_t bad = 0;
    const double2 p = pos[s];
    key_next[s] = (int)va[s].w == 0 ? P.ncells : cell_key(P, p.x, p.y, bad);
}

// ------------------------------------------------------------------ step adjoint
struct AdjStep {
    double t;         // time at the start of the step
    double dt;        // dt of whole_step and of the time advance
    double dt_pred;   // dt of predict_half_step
    uint32_t first;   // first step: v and rho advance with dt / 2
    uint32_t piston;  // the wave maker moves its particles this step
    uint32_t halted;  // the step did not run
    uint32_t pad;
};

// After a step: what its adjoint needs from the control block.
// Record of the step just launched, at rec[k], k = its step index (the
// control block's step count, committed at the next prologue) - base,
// clamped to the tape's nsteps records: a step after a halt does not run,
// and its record lands on the halted step's, which the resumed range
// writes again.
__global__ void adj_record(const Ctrl* __restrict__ ctrl, AdjStep* __restrict__ rec, uint64_t base, uint32_t nsteps) {
    const Ctrl& c = *ctrl;
    const uint64_t k = min(c.step - base, (uint64_t)nsteps - 1);
    AdjStep r;
    r.t = c.cur_time;
    r.dt = c.dt_new;
    r.dt_pred = c.cur_dt;
    r.first = c.cur_first;
    r.piston = c.piston_active;
    r.halted = c.halted;
    r.pad = 0;
    rec[k] = r;
}

// whole_step reverse.  out_adj[s] = (a_x, a_y, av_x, av_y) adjoints,
// drho_adj[s]; direct[s] = (r_x, r_y, v_x, v_y, rho) adjoints that do not
// pass through the pair sums; piston_sum += wave-maker velocity adjoint.
template <class R>
__global__ void __launch_bounds__(kBlock) adj_prep(StepParams P, AdjStep T, StateB<R> B,
                                                   const double* __restrict__ lam_in,
                                                   double4* __restrict__ out_adj, double* __restrict__ drho_adj,
                                                   double* __restrict__ direct, double* __restrict__ piston_sum) {
    __shared__ double s_p[kBlock / kWave];
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    double pist = 0.0;
    if (s < P.n) {
        const uint32_t o = B.oid[s];
        const double* l = lam_in + 8 * (size_t)o;
        const double rx = l[0], ry = l[1], vx = l[2], vy = l[3], rb = l[4], ax = l[5], ay = l[6], db = l[7];
        const int t = (int)B.va[s].w;
        const double vdt = T.first ? T.dt * 0.5 : T.dt;
        double4 oa;
        double dvx = 0, dvy = 0;
        if (t > 0) {
            const double tx = vx + rx * T.dt, ty = vy + ry * T.dt;  // adjoint of the updated v
            oa.x = tx * vdt + ax;
            oa.y = ty * vdt + ay;
            oa.z = P.xsph_on ? tx : 0.0;
            oa.w = P.xsph_on ? ty : 0.0;
            dvx = tx;
            dvy = ty;
        } else {
            oa.x = ax;
            oa.y = ay;
            oa.z = 0.0;
            oa.w = 0.0;
            const bool moving = P.boundary_motion && t < 0;
            const bool piston = moving && T.piston && o >= P.nwm_start && o < P.nwm_end && o >= P.nfluid;
            if (piston) {
                pist = P.nwm_axis == 0 ? vx + rx * T.dt : vy + ry * T.dt;
            } else if (moving) {
                dvx = vx + rx * T.dt;
                dvy = vy + ry * T.dt;
            } else {
                dvx = vx;
                dvy = vy;
            }
        }
        if (t == 0) {  // no neighbours, no acceleration dependence
            oa.x = oa.y = 0.0;
        }
        out_adj[s] = oa;
        drho_adj[s] = P.density_treatment != 0 ? rb * vdt + db : 0.0;
        double* d = direct + 5 * (size_t)s;
        d[0] = rx;
        d[1] = ry;
        d[2] = dvx;
        d[3] = dvy;
        d[4] = rb;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) pist += __shfl_xor(pist, off);
    if (__lane_id() == 0) s_p[threadIdx.x / kWave] = pist;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0;
        for (int w = 0; w < kBlock / kWave; ++w) b += s_p[w];
        if (b != 0.0) atomicAdd(piston_sum, b);
    }
}

struct AdjConsts {
    double h, inv_h, c1, c2, w1, w2, alpha_c, beta, eta2, r0, D, m, xf;
    bool av_on, xsph_on, lj_on;
};

// Shared scalars of one pair (e = r_i - r_k, symmetric in i and k).
struct AdjPair {
    double d2, rinv, d, F, Fp, W, IR, mu, den, vr, A;
    bool app;
};

// Contributions of one orientation: the owner's sums at owner adjoints
// (Sb = m a-adjoint, s1 = m drho-adjoint, Xb = 2 eps m av-adjoint, Lb =
// a-adjoint) give adjoints of e, dv = v_nb - v_own, du = u_nb - u_own, the
// reciprocal IR and the coefficient.
__device__ __forceinline__ void adj_orient(const AdjConsts& K, const AdjPair& Q, double ex, double ey,
                                           double dvx, double dvy, double dux, double duy, double Sbx, double Sby,
                                           double s1, double Xbx, double Xby, double Lbx, double Lby, bool lj,
                                           bool xsph, double coef, double& ebx, double& eby, double& dvbx,
                                           double& dvby, double& dubx, double& duby, double& irb,
                                           double& coefb) {
    const double Fpe = Q.Fp * Q.rinv;  // grad_e F = Fpe * e
    // continuity: T1 = F (dv . e)
    const double dve = dvx * ex + dvy * ey;
    dvbx = s1 * Q.F * ex;
    dvby = s1 * Q.F * ey;
    ebx = s1 * (Q.F * dvx + dve * Fpe * ex);
    eby = s1 * (Q.F * dvy + dve * Fpe * ey);
    // acceleration: T2 = F coef e
    const double Se = Sbx * ex + Sby * ey;
    ebx += Q.F * coef * Sbx + coef * Se * Fpe * ex;
    eby += Q.F * coef * Sby + coef * Se * Fpe * ey;
    coefb = Q.F * Se;
    dubx = 0;
    duby = 0;
    irb = 0;
    // artificial viscosity: A = 2 IR (beta mu^2 - alpha c mu), mu = h vr / (d^2 + eta^2), vr = du . e
    if (Q.app) {
        const double Ab = coefb;
        const double mub = Ab * 2.0 * Q.IR * (2.0 * K.beta * Q.mu - K.alpha_c);
        irb += Ab * 2.0 * (K.beta * Q.mu * Q.mu - K.alpha_c * Q.mu);
        const double vrb = mub * K.h * Q.den;
        const double em = -2.0 * mub * Q.mu * Q.den;
        ebx += em * ex + vrb * dux;
        eby += em * ey + vrb * duy;
        dubx += vrb * ex;
        duby += vrb * ey;
    }
    // Lennard-Jones on a fluid owner from a wall neighbour: T3 = -e f(d)
    if (lj) {
        const double rn = K.r0 * Q.rinv;
        const double rn2 = rn * rn, rn4 = rn2 * rn2, rn12 = rn4 * rn4 * rn4;
        const double f = K.D * (rn12 - rn4) * Q.rinv * Q.rinv;
        const double fp = K.D * (-14.0 * rn12 + 6.0 * rn4) * Q.rinv * Q.rinv * Q.rinv;  // df/dd
        const double Le = Lbx * ex + Lby * ey;
        ebx += -f * Lbx - Le * fp * Q.rinv * ex;
        eby += -f * Lby - Le * fp * Q.rinv * ey;
    }
    // XSPH from a fluid neighbour: T4 = du W IR
    if (xsph) {
        const double xdu = Xbx * dux + Xby * duy;
        dubx += Xbx * Q.W * Q.IR;
        duby += Xby * Q.W * Q.IR;
        const double wb = Q.IR * xdu;
        ebx += wb * Q.F * ex;  // grad_e W = F e
        eby += wb * Q.F * ey;
        irb += Q.W * xdu;
    }
}

template <class R>
__global__ void __launch_bounds__(kBlock) adj_pair(StepParams P, AdjStep T, StateB<R> B,
                                                   const uint32_t* __restrict__ skey,
                                                   const uint32_t* __restrict__ LB,
                                                   const double4* __restrict__ out_adj,
                                                   const double* __restrict__ drho_adj,
                                                   const double* __restrict__ direct, double* __restrict__ lam_out,
                                                   uint32_t* __restrict__ overflow_flag) {
    __shared__ uint32_t hits[kAdjChunk * kBlock];
    // XCD-aware: each XCD walks one contiguous slice of the slots
    const uint32_t lb = xcd_block(P.blocks_per_xcd);
    const uint32_t s = lb * kBlock + threadIdx.x;
    if (lb >= P.nblocks || s >= P.n) return;
    AdjConsts K;
    K.h = P.h;
    K.inv_h = 1.0 / P.h;
    K.c1 = P.f_cubic / (P.h * P.h);
    K.c2 = 0.5 * P.f_cubic / P.h;
    K.w1 = P.f_cubic;
    K.w2 = P.f_cubic / 6.0;
    K.alpha_c = P.av_alpha * P.av_c;
    K.beta = P.av_beta;
    K.eta2 = P.av_eta2;
    K.r0 = P.lj_r0;
    K.D = P.lj_D;
    K.m = P.mass;
    K.xf = 2.0 * P.mass * P.xsph_eps;
    K.av_on = P.av_on;
    K.xsph_on = P.xsph_on;
    K.lj_on = P.lj_on;
    const double2 pk = B.pos[s];
    const double4 vak = to_d4<R>(B.va[s]);
    const double4 prk = to_d4<R>(B.pred[s]);
    const double pres_k = B.pres[s];
    const double4 oak = out_adj[s];
    const double dak = drho_adj[s];
    const int tk = (int)vak.w;
    double rbx = 0, rby = 0, vbx = 0, vby = 0, ubx = 0, uby = 0, Rb = 0, Pb = 0;
    bool overflow = false;
    if (tk != 0) {
        uint32_t* my = hits + threadIdx.x;
        auto flush = [&](uint32_t nh) {
#pragma unroll 1
            for (uint32_t u = 0; u < nh; ++u) {
                const uint32_t i = my[u * kBlock];
                const double2 pi = B.pos[i];
                const double4 vai = to_d4<R>(B.va[i]);
                const double4 pri = to_d4<R>(B.pred[i]);
                const double4 oai = out_adj[i];
                const double dai = drho_adj[i];
                const int ti = (int)vai.w;
                const double ex = pi.x - pk.x, ey = pi.y - pk.y;  // r_i - r_k
                AdjPair Q;
                Q.d2 = ex * ex + ey * ey;
                Q.rinv = Q.d2 > 0.0 ? fast_rsqrt<double>(Q.d2) : 0.0;
                Q.d = Q.d2 * Q.rinv;
                const double q = Q.d * K.inv_h;
                const double t = 2.0 - q;
                if (q <= 1.0) {
                    Q.F = K.c1 * (1.5 * q - 2.0);
                    Q.Fp = 1.5 * K.c1 * K.inv_h;
                    Q.W = K.w1 * (2.0 / 3.0 - q * q + 0.5 * q * q * q);
                } else {
                    Q.F = -K.c2 * t * t * Q.rinv;
                    Q.Fp = K.c2 * t * (2.0 * Q.rinv * K.inv_h + t * Q.rinv * Q.rinv);
                    Q.W = K.w2 * t * t * t;
                }
                Q.IR = fast_rcp<double>(pri.z + prk.z);
                const double dux = pri.x - prk.x, duy = pri.y - prk.y;  // u_i - u_k
                Q.vr = dux * ex + duy * ey;
                Q.den = fast_rcp<double>(Q.d2 + K.eta2);
                Q.mu = K.h * Q.vr * Q.den;
                Q.app = K.av_on && Q.vr < 0.0;
                Q.A = Q.app ? 2.0 * Q.IR * (K.beta * Q.mu * Q.mu - K.alpha_c * Q.mu) : 0.0;
                const double coef = pri.w + prk.w + Q.A;
                const double dvx = vai.x - vak.x, dvy = vai.y - vak.y;  // v_i - v_k
                double ebx, eby, dvbx, dvby, dubx, duby, irb, coefb;
                // Both orientations of the pair in one evaluation.  k's sum at
                // k's adjoints is adj_orient at (e, dv, du); i's sum at i's
                // adjoints is adj_orient at (-e, -dv, -du), entering rb, vb, ub
                // with the opposite sign and Rb, Pb with the same.  Every
                // output is linear in the owner's adjoints, and even (the a-
                // and av-adjoint terms, the wall force) or odd (the
                // drho-adjoint terms) in (e, dv, du), so the two make one
                // evaluation at (e, dv, du) with the owner adjoints
                // m (oa_k - oa_i) and m (drho_k + drho_i), the XSPH adjoint of
                // whichever owner has a fluid neighbour (i's negated), and the
                // wall force of whichever owner is the fluid particle (i's
                // negated).
                const bool xk = K.xsph_on && ti > 0, xi = K.xsph_on && tk > 0;
                const bool ljk = K.lj_on && tk > 0 && ti < 0 && Q.d < K.r0;
                const bool lji = K.lj_on && ti > 0 && tk < 0 && Q.d < K.r0;
                const double xdx = K.xf * ((xk ? oak.z : 0.0) - (xi ? oai.z : 0.0));
                const double xdy = K.xf * ((xk ? oak.w : 0.0) - (xi ? oai.w : 0.0));
                adj_orient(K, Q, ex, ey, dvx, dvy, dux, duy, K.m * (oak.x - oai.x), K.m * (oak.y - oai.y),
                           K.m * (dak + dai), xdx, xdy, ljk ? oak.x : -oai.x, ljk ? oak.y : -oai.y, ljk || lji,
                           xk || xi, coef, ebx, eby, dvbx, dvby, dubx, duby, irb, coefb);
                rbx -= ebx;
                rby -= eby;
                vbx -= dvbx;
                vby -= dvby;
                ubx -= dubx;
                uby -= duby;
                Rb -= Q.IR * Q.IR * irb;
                Pb += coefb;
            }
        };
        collect_neighbours<kAdjChunk>(P, P.max_nbr - 1, s, skey[s], pk, B.pos, LB, my, overflow, flush);
    }
    if (overflow) atomicOr(overflow_flag, 1u);
    // P = p / R^2, p = EOS(rho) = B ((rho / rho0)^7 - 1), u = v + a_prev dt/2, R = rho + drho_prev dt/2
    const double Rk = prk.z;
    const double pb = Pb / (Rk * Rk);
    Rb += -2.0 * Pb * pres_k / (Rk * Rk * Rk);
    const double x = vak.z / P.rho0;
    const double x2 = x * x, x3 = x2 * x;
    const double dpdrho = 7.0 * P.eos_B * x3 * x3 / P.rho0;
    const double* d = direct + 5 * (size_t)s;
    const double h2 = T.dt_pred * 0.5;
    double* l = lam_out + 8 * (size_t)B.oid[s];
    l[0] = d[0] + rbx;
    l[1] = d[1] + rby;
    l[2] = d[2] + vbx + ubx;
    l[3] = d[3] + vby + uby;
    l[4] = d[4] + (P.density_treatment != 0 ? Rb : 0.0) + pb * dpdrho;
    l[5] = ubx * h2;
    l[6] = uby * h2;
    l[7] = P.density_treatment != 0 ? Rb * h2 : 0.0;
}

}  // namespace rrsph
