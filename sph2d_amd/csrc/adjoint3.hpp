// Discrete adjoint of the FP64 3D step: the 2D adjoint (adjoint.hpp) with a
// z component, the state of step3d.hpp (positions carry the type), the
// wave-uniform neighbour walk, and a segmented piston whose paddles each
// collect their own velocity adjoint.
//
// Adjoint state per particle (original order, 12 doubles):
//   lam = (r x y z, v x y z, rho, a_prev x y z, drho_prev, 0)
// Output adjoints of the step (sorted order), packed by adj_prep3 with the
// state a neighbour's pair adjoint reads into one record per slot (AdjRec3):
//   oa = (a x y z, drho),  ox = (av x y z)
// Gauges: gauge g sits at (x_g, z_g) and weighs fluid particles by
// w = (1 - ((x - x_g)^2 + (z - z_g)^2) / R^2)^3 inside the disc of radius R;
// eta_g = y_ref + tau log(sum w e / sum w), e = exp((y - y_ref) / tau).
#pragma once

#include "adjoint.hpp"
#include "list_step3.hpp"
#include "step3d.hpp"

namespace rrsph {

constexpr uint32_t kLam3 = 12;
constexpr uint32_t kDirect3 = 8;

// What the pair adjoint of a neighbour reads of slot s, in one record:
// the output adjoints oa = (a x y z, drho) and ox = (av x y z), p / R^2,
// the velocity v and the predicted (u, R).  The FP64 step's record holds
// its position too, all in FP64.  The FP32 step's record is 64 bytes, two
// slots per cache line (a hit's gather costs by the lines it touches, and
// the neighbours of a particle lie in runs of slots): the state in its
// FP32 and the adjoints rounded to FP32 (the arithmetic stays FP64); the
// FP64 position comes from the packed position array, whose lines the
// neighbour walk has just read.
template <class R> struct AdjRec3;
template <> struct alignas(32) AdjRec3<double> {
    double4 pos;
    double4 oa;
    double ox, oy, oz, pw;
    double4 v;  // (v x y z, unused)
    double4 u;  // (u x y z, R)
};
template <> struct alignas(64) AdjRec3<float> {
    float4 vp;  // (v x y z, p / R^2)
    float4 u;   // (u x y z, R)
    float4 oa;
    float4 ox;  // (av x y z, unused)
};
static_assert(sizeof(AdjRec3<float>) == 64, "two records per cache line");

// The fields of a record, widened to FP64 (pos: the FP64 record's own,
// else the position array's entry).
struct AdjHit3 {
    double4 pos, v, u, oa, ox;
    double pw;
};
__device__ __forceinline__ AdjHit3 adj_hit3(const AdjRec3<double>& r, const double4&) {
    return AdjHit3{r.pos, r.v, r.u, r.oa, make_double4(r.ox, r.oy, r.oz, 0.0), r.pw};
}
__device__ __forceinline__ AdjHit3 adj_hit3(const AdjRec3<float>& r, const double4& pos) {
    return AdjHit3{pos,
                   make_double4(r.vp.x, r.vp.y, r.vp.z, 0.0),
                   make_double4(r.u.x, r.u.y, r.u.z, r.u.w),
                   make_double4(r.oa.x, r.oa.y, r.oa.z, r.oa.w),
                   make_double4(r.ox.x, r.ox.y, r.ox.z, 0.0),
                   (double)r.vp.w};
}

struct GaugeSpec3 {
    GaugeSpec g;
    double z[kMaxGauges];
};

__device__ __forceinline__ double gauge_weight3(const GaugeSpec3& G, uint32_t g, double x, double z, double inv_r2,
                                                double& dx, double& dz) {
    dx = x - G.g.x[g];
    dz = z - G.z[g];
    const double u = 1.0 - (dx * dx + dz * dz) * inv_r2;
    return u > 0.0 ? u : 0.0;
}

// Block partial sums (sum w e, sum w) of every gauge (layout of
// gauge_block_sums).
__global__ void __launch_bounds__(kBlock) smooth_gauge_partials3(uint32_t n, const double4* __restrict__ pos,
                                                                 GaugeSpec3 G, double* __restrict__ partials) {
    __shared__ GaugeLds L;
    gauge_lds_clear(G.g, L);
    const double inv_r2 = 1.0 / (G.g.radius * G.g.radius);
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const uint32_t s = base + threadIdx.x;
        bool in = false;
        double4 p = make_double4(0, 0, 0, 0);
        if (s < n) {
            p = pos[s];
            in = (int)p.w == G.g.ptype;
        }
        auto d2 = [&](uint32_t g) {
            const double dx = p.x - G.g.x[g], dz = p.z - G.z[g];
            return dx * dx + dz * dz;
        };
        gauge_chunk_sums(G.g, in, p.y, d2, inv_r2, L);
    }
    gauge_block_store(G.g, L, partials);
}

// lam[oid].r += sum_g eta_bar[g] d eta_g / d r (state in sorted order).
__global__ void __launch_bounds__(kBlock) smooth_gauge_grad3(uint32_t n, const double4* __restrict__ pos,
                                                             const uint32_t* __restrict__ oid, GaugeSpec3 G,
                                                             const double* __restrict__ eta_bar,
                                                             const double* __restrict__ zsum,
                                                             const double* __restrict__ wsum,
                                                             double* __restrict__ lam) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const double4 p = pos[s];
    if ((int)p.w != G.g.ptype) return;
    const double inv_r2 = 1.0 / (G.g.radius * G.g.radius);
    double gx = 0, gy = 0, gz = 0;
    bool seen = false;  // inside a window: only those particles update lam
    for (uint32_t g = 0; g < G.g.n; ++g) {
        double dx, dz;
        const double u = gauge_weight3(G, g, p.x, p.z, inv_r2, dx, dz);
        if (!(u > 0.0) || !(wsum[g] > 0.0) || eta_bar[g] == 0.0) continue;
        seen = true;
        const double w = u * u * u;
        const double wr = -6.0 * inv_r2 * u * u;  // dw/dx = wr dx, dw/dz = wr dz
        const double e = exp((p.y - G.g.y_ref) / G.g.tau);
        const double z = zsum[g], ws = wsum[g];
        const double c = eta_bar[g] * G.g.tau * wr * (e / z - 1.0 / ws);
        gy += eta_bar[g] * w * e / z;
        gx += c * dx;
        gz += c * dz;
    }
    if (!seen) return;
    double* l = lam + kLam3 * (size_t)oid[s];
    l[0] += gx;
    l[1] += gy;
    l[2] += gz;
}

// Cell keys of a restored FP64 3D state (dead particles get the dead key).
__global__ void __launch_bounds__(kBlock) state_keys3_kernel(StepParams P, const double4* __restrict__ pos,
                                                             uint32_t* __restrict__ key_next) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= P.n) return;
    uint32_t bad = 0;
    const double4 p = pos[s];
    key_next[s] = (int)p.w == 0 ? P.ncells : cell_key3(P, p.x, p.y, p.z, bad);
}

// whole_step reverse (WholeStep.cpp:19-61 with the boundary motion of
// cl/TimeIntegration.cl:190-245).  piston_sum[g] += adjoint of paddle g's
// velocity (one entry per paddle of this step).
template <class R>
__global__ void __launch_bounds__(kBlock) adj_prep3(StepParams P, AdjStep T, StateB3<R> B,
                                                    const double* __restrict__ lam_in, AdjRec3<R>* __restrict__ rec,
                                                    double* __restrict__ direct, double* __restrict__ piston_sum,
                                                    const uint32_t* __restrict__ home, float4* __restrict__ nearf) {
    __shared__ double s_p[kMaxSegments][kBlock / kWave];
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    double pist = 0.0;
    int seg = 0;
    if (s < P.n) {
        const uint32_t o = B.oid[s];
        const double4 ps = B.pos[s];
        const double* l = lam_in + kLam3 * (size_t)o;
        const double rx = l[0], ry = l[1], rz = l[2], vx = l[3], vy = l[4], vz = l[5], rb = l[6];
        const double ax = l[7], ay = l[8], az = l[9], db = l[10];
        const int t = (int)ps.w;
        const double vdt = T.first ? T.dt * 0.5 : T.dt;
        double4 oa, ox;
        double dvx = 0, dvy = 0, dvz = 0;
        if (t > 0) {
            const double tx = vx + rx * T.dt, ty = vy + ry * T.dt, tz = vz + rz * T.dt;  // updated-v adjoint
            oa = make_double4(tx * vdt + ax, ty * vdt + ay, tz * vdt + az, 0.0);
            ox = P.xsph_on ? make_double4(tx, ty, tz, 0.0) : make_double4(0, 0, 0, 0);
            dvx = tx;
            dvy = ty;
            dvz = tz;
        } else {
            oa = make_double4(ax, ay, az, 0.0);
            ox = make_double4(0, 0, 0, 0);
            const bool moving = P.boundary_motion && t < 0;
            const bool piston = moving && T.piston && o >= P.nwm_start && o < P.nwm_end && o >= P.nfluid;
            if (piston) {
                const double tv[3] = {vx + rx * T.dt, vy + ry * T.dt, vz + rz * T.dt};
                pist = tv[P.nwm_axis];
                seg = P.piston_segments > 1 ? piston_segment(P, ps.z) : 0;
            } else if (moving) {
                dvx = vx + rx * T.dt;
                dvy = vy + ry * T.dt;
                dvz = vz + rz * T.dt;
            } else {
                dvx = vx;
                dvy = vy;
                dvz = vz;
            }
        }
        if (t == 0) oa.x = oa.y = oa.z = 0.0;  // no neighbours, no acceleration dependence
        oa.w = P.density_treatment != 0 ? rb * vdt + db : 0.0;
        AdjRec3<R> r;
        if constexpr (sizeof(R) == 4) {
            const real4<R> va = B.va[s];
            r.vp = make_float4(va.x, va.y, va.z, B.pw[s]);
            r.u = B.pred[s];
            r.oa = make_float4((float)oa.x, (float)oa.y, (float)oa.z, (float)oa.w);
            r.ox = make_float4((float)ox.x, (float)ox.y, (float)ox.z, 0.f);
        } else {
            r.pos = ps;
            r.oa = oa;
            r.ox = ox.x;
            r.oy = ox.y;
            r.oz = ox.z;
            r.pw = B.pw[s];
            r.v = B.va[s];
            r.u = B.pred[s];
        }
        rec[s] = r;
        if (nearf) nearf[s] = near3<float>(P, ps, (float)t, home[s]);  // type in w
        double* d = direct + kDirect3 * (size_t)s;
        d[0] = rx;
        d[1] = ry;
        d[2] = rz;
        d[3] = dvx;
        d[4] = dvy;
        d[5] = dvz;
        d[6] = rb;
        d[7] = 0.0;
    }
    for (int g = 0; g < P.piston_segments; ++g) {
        double v = seg == g ? pist : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (__lane_id() == 0) s_p[g][threadIdx.x / kWave] = v;
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)P.piston_segments) {
        double b = 0;
        for (int w = 0; w < kBlock / kWave; ++w) b += s_p[threadIdx.x][w];
        if (b != 0.0) atomicAdd(piston_sum + threadIdx.x, b);
    }
}

// One orientation of a pair in 3D (see adj_orient): the owner's sums at the
// owner's adjoints (Sb = m a-adjoint, s1 = m drho-adjoint, Xb = 2 eps m
// av-adjoint, Lb = a-adjoint) give the adjoints of e, dv = v_nb - v_own,
// du = u_nb - u_own, the reciprocal IR and the coefficient.
struct V3 {
    double x, y, z;
};
__device__ __forceinline__ double dot3(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

__device__ __forceinline__ void adj_orient3(const AdjConsts& K, const AdjPair& Q, const V3& e, const V3& dv,
                                            const V3& du, const V3& Sb, double s1, const V3& Xb, const V3& Lb,
                                            bool lj, bool xsph, double coef, V3& eb, V3& dvb, V3& dub,
                                            double& irb, double& coefb) {
    const double Fpe = Q.Fp * Q.rinv;  // grad_e F = Fpe * e
    // continuity: T1 = F (dv . e)
    const double dve = dot3(dv, e);
    dvb = {s1 * Q.F * e.x, s1 * Q.F * e.y, s1 * Q.F * e.z};
    eb = {s1 * (Q.F * dv.x + dve * Fpe * e.x), s1 * (Q.F * dv.y + dve * Fpe * e.y),
          s1 * (Q.F * dv.z + dve * Fpe * e.z)};
    // acceleration: T2 = F coef e
    const double Se = dot3(Sb, e);
    eb.x += Q.F * coef * Sb.x + coef * Se * Fpe * e.x;
    eb.y += Q.F * coef * Sb.y + coef * Se * Fpe * e.y;
    eb.z += Q.F * coef * Sb.z + coef * Se * Fpe * e.z;
    coefb = Q.F * Se;
    dub = {0, 0, 0};
    irb = 0;
    // artificial viscosity: A = 2 IR (beta mu^2 - alpha c mu), mu = h vr / (d^2 + eta^2), vr = du . e
    if (Q.app) {
        const double Ab = coefb;
        const double mub = Ab * 2.0 * Q.IR * (2.0 * K.beta * Q.mu - K.alpha_c);
        irb += Ab * 2.0 * (K.beta * Q.mu * Q.mu - K.alpha_c * Q.mu);
        const double vrb = mub * K.h * Q.den;
        const double em = -2.0 * mub * Q.mu * Q.den;
        eb.x += em * e.x + vrb * du.x;
        eb.y += em * e.y + vrb * du.y;
        eb.z += em * e.z + vrb * du.z;
        dub.x += vrb * e.x;
        dub.y += vrb * e.y;
        dub.z += vrb * e.z;
    }
    // Lennard-Jones on a fluid owner from a wall neighbour: T3 = -e f(d)
    if (lj) {
        const double rn = K.r0 * Q.rinv;
        const double rn2 = rn * rn, rn4 = rn2 * rn2, rn12 = rn4 * rn4 * rn4;
        const double f = K.D * (rn12 - rn4) * Q.rinv * Q.rinv;
        const double fp = K.D * (-14.0 * rn12 + 6.0 * rn4) * Q.rinv * Q.rinv * Q.rinv;  // df/dd
        const double Le = dot3(Lb, e);
        eb.x += -f * Lb.x - Le * fp * Q.rinv * e.x;
        eb.y += -f * Lb.y - Le * fp * Q.rinv * e.y;
        eb.z += -f * Lb.z - Le * fp * Q.rinv * e.z;
    }
    // XSPH from a fluid neighbour: T4 = du W IR
    if (xsph) {
        const double xdu = dot3(Xb, du);
        dub.x += Xb.x * Q.W * Q.IR;
        dub.y += Xb.y * Q.W * Q.IR;
        dub.z += Xb.z * Q.W * Q.IR;
        const double wb = Q.IR * xdu;
        eb.x += wb * Q.F * e.x;  // grad_e W = F e
        eb.y += wb * Q.F * e.y;
        eb.z += wb * Q.F * e.z;
        irb += Q.W * xdu;
    }
}

// Gather over neighbours, both orientations of every pair, then the
// predict and EOS reverse; writes lam_out.  Every lane of a block runs the
// wave-uniform walk (lanes past n walk empty ranges).  LISTS: the
// neighbours come from the lists of the build the step read (nbl, nbc, in
// B's slot order) instead of the cell walk.
template <class R, bool LISTS>
__global__ void __launch_bounds__(kBlock) adj_pair3(StepParams P, AdjStep T, StateB3<R> B,
                                                    const uint32_t* __restrict__ skey,
                                                    const uint32_t* __restrict__ LB,
                                                    const uint32_t* __restrict__ nbl,
                                                    const uint32_t* __restrict__ nbc,
                                                    const AdjRec3<R>* __restrict__ rec,
                                                    const double* __restrict__ direct, double* __restrict__ lam_out,
                                                    uint32_t* __restrict__ overflow_flag,
                                                    const float4* __restrict__ nearf) {
    // FP32 step over lists: entries decided and hits' r_i - r_k taken from
    // the near records (16 bytes, as the list pass does) instead of the FP64
    // positions (32 bytes)
    constexpr bool kNear = LISTS && sizeof(R) == 4;
    __shared__ uint32_t hits[kAdjChunk * kBlock];
    __shared__ float4 s_off[kNear ? kCodes3 : 1];
    // XCD-aware: each XCD walks one contiguous slice of the slots
    const uint32_t lb = xcd_block(P.blocks_per_xcd);
    if (lb >= P.nblocks) return;
    if constexpr (kNear) {
        fill_cell_offsets3(P, s_off);
        __syncthreads();
    }
    const uint32_t s = lb * kBlock + threadIdx.x;
    const bool live = s < P.n;
    const uint32_t sc = live ? s : P.n - 1;
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
    const double4 pk = B.pos[sc];
    const double4 vak = to_d4<R>(B.va[sc]);    // (v, rho)
    const double4 prk = to_d4<R>(B.pred[sc]);  // (u, R)
    const double Pk = B.pw[sc];                // p / R^2
    const double pres_k = B.pres[sc];
    const AdjHit3 hk = adj_hit3(rec[sc], pk);
    float4 nk{};
    if constexpr (kNear) nk = nearf[sc];
    const double4 oak = hk.oa;
    const double4 oxk = hk.ox;
    const int tk = (int)pk.w;
    V3 rb{0, 0, 0}, vb{0, 0, 0}, ub{0, 0, 0};
    double Rb = 0, Pb = 0;
    bool overflow = false;
    uint32_t* my = hits + threadIdx.x;
    // the loads of hit u + 1 are issued before the arithmetic of hit u (the
    // kernel runs at 2 waves per SIMD either way, so the registers are free;
    // a second record in flight needs more registers than 2 waves leave)
    auto flush = [&](uint32_t nh) {
        if (nh == 0) return;
        const uint32_t last = nh - 1;
        // kNear: the hit list holds whole entries (slot and home-cell code)
        uint32_t in = my[0];
        AdjRec3<R> rn = rec[kNear ? in & kSlotMask3 : in];
        double4 pn{};
        float4 qn{};
        if constexpr (kNear) qn = nearf[in & kSlotMask3];
        else if constexpr (sizeof(R) == 4) pn = B.pos[in];
#pragma unroll 1
        for (uint32_t u = 0; u < nh; ++u) {
            const AdjHit3 h = adj_hit3(rn, pn);
            const double4 vai = h.v;
            const double4 pri = h.u;
            const double Pi = h.pw;
            const double4 oai = h.oa;
            const double4 oxi = h.ox;
            int ti;
            V3 e;  // r_i - r_k
            if constexpr (kNear) {
                const float4 o = s_off[in >> kSlotBits3];
                e = V3{(double)((qn.x - nk.x) + o.x), (double)((qn.y - nk.y) + o.y), (double)((qn.z - nk.z) + o.z)};
                ti = (int)qn.w;
            } else {
                e = V3{h.pos.x - pk.x, h.pos.y - pk.y, h.pos.z - pk.z};
                ti = (int)h.pos.w;
            }
            in = my[min(u + 1, last) * kBlock];
            rn = rec[kNear ? in & kSlotMask3 : in];
            if constexpr (kNear) qn = nearf[in & kSlotMask3];
            else if constexpr (sizeof(R) == 4) pn = B.pos[in];
            AdjPair Q;
            Q.d2 = dot3(e, e);
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
            Q.IR = fast_rcp<double>(pri.w + prk.w);
            const V3 du{pri.x - prk.x, pri.y - prk.y, pri.z - prk.z};  // u_i - u_k
            Q.vr = dot3(du, e);
            Q.den = fast_rcp<double>(Q.d2 + K.eta2);
            Q.mu = K.h * Q.vr * Q.den;
            Q.app = K.av_on && Q.vr < 0.0;
            Q.A = Q.app ? 2.0 * Q.IR * (K.beta * Q.mu * Q.mu - K.alpha_c * Q.mu) : 0.0;
            const double coef = Pi + Pk + Q.A;
            const V3 dv{vai.x - vak.x, vai.y - vak.y, vai.z - vak.z};  // v_i - v_k
            V3 eb, dvb, dub;
            double irb, coefb;
            // Both orientations of the pair in one evaluation.  k's sum at
            // k's adjoints is adj_orient3 at (e, dv, du); i's sum at i's
            // adjoints is adj_orient3 at (-e, -dv, -du), entering rb, vb, ub
            // with the opposite sign and Rb, Pb with the same.  Every output is
            // linear in the owner's adjoints, and even (the a- and
            // av-adjoint terms, the wall force) or odd (the drho-adjoint
            // terms) in (e, dv, du), so the two make one evaluation at
            // (e, dv, du) with the owner adjoints m (oa_k - oa_i) and
            // m (drho_k + drho_i), the XSPH adjoint of whichever owner has a
            // fluid neighbour (i's negated), and the wall force of whichever
            // owner is the fluid particle (i's negated).
            const bool xk = K.xsph_on && ti > 0, xi = K.xsph_on && tk > 0;
            const bool ljk = K.lj_on && tk > 0 && ti < 0 && Q.d < K.r0;
            const bool lji = K.lj_on && ti > 0 && tk < 0 && Q.d < K.r0;
            const V3 Xd{K.xf * ((xk ? oxk.x : 0.0) - (xi ? oxi.x : 0.0)),
                        K.xf * ((xk ? oxk.y : 0.0) - (xi ? oxi.y : 0.0)),
                        K.xf * ((xk ? oxk.z : 0.0) - (xi ? oxi.z : 0.0))};
            const V3 L = ljk ? V3{oak.x, oak.y, oak.z} : V3{-oai.x, -oai.y, -oai.z};
            adj_orient3(K, Q, e, dv, du, V3{K.m * (oak.x - oai.x), K.m * (oak.y - oai.y), K.m * (oak.z - oai.z)},
                        K.m * (oak.w + oai.w), Xd, L, ljk || lji, xk || xi, coef, eb, dvb, dub, irb, coefb);
            rb.x -= eb.x;
            rb.y -= eb.y;
            rb.z -= eb.z;
            vb.x -= dvb.x;
            vb.y -= dvb.y;
            vb.z -= dvb.z;
            ub.x -= dub.x;
            ub.y -= dub.y;
            ub.z -= dub.z;
            Rb -= Q.IR * Q.IR * irb;
            Pb += coefb;
        }
    };
    if constexpr (kNear)
        collect_near3<kAdjChunk>(P, live && tk != 0, P.max_nbr - 1, sc, s_off, nk, nearf, pk, B.pos, nbl, nbc, my,
                                 overflow, flush);
    else if (LISTS)
        collect_listed3<kAdjChunk>(P, live && tk != 0, P.max_nbr - 1, sc, pk, B.pos, nbl, nbc, my, overflow, flush);
    else
        collect_neighbours3<kAdjChunk>(P, live && tk != 0, P.max_nbr - 1, sc, skey[sc], pk, B.pos, LB, my,
                                       overflow, flush);
    if (overflow) atomicOr(overflow_flag, 1u);
    if (!live) return;
    // P = p / R^2, p = EOS(rho) = B ((rho / rho0)^7 - 1), u = v + a_prev dt/2, R = rho + drho_prev dt/2
    const double Rk = prk.w;
    const double pb = Pb / (Rk * Rk);
    Rb += -2.0 * Pb * pres_k / (Rk * Rk * Rk);
    const double x = vak.w / P.rho0;
    const double x2 = x * x, x3 = x2 * x;
    const double dpdrho = 7.0 * P.eos_B * x3 * x3 / P.rho0;
    const double* d = direct + kDirect3 * (size_t)s;
    const double h2 = T.dt_pred * 0.5;
    double* l = lam_out + kLam3 * (size_t)B.oid[s];
    l[0] = d[0] + rb.x;
    l[1] = d[1] + rb.y;
    l[2] = d[2] + rb.z;
    l[3] = d[3] + vb.x + ub.x;
    l[4] = d[4] + vb.y + ub.y;
    l[5] = d[5] + vb.z + ub.z;
    l[6] = d[6] + (P.density_treatment != 0 ? Rb : 0.0) + pb * dpdrho;
    l[7] = ub.x * h2;
    l[8] = ub.y * h2;
    l[9] = ub.z * h2;
    l[10] = P.density_treatment != 0 ? Rb * h2 : 0.0;
    l[11] = 0.0;
}

}  // namespace rrsph
