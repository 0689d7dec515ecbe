// Per-pair WCSPH physics, templated on the arithmetic type R (float or double).
//
// The formulas are those of the reference's per-pair "part" functions:
//   smoothing kernels   RRSPH_COMMON/include/SmoothingKernel.h:14-283
//   Tait EOS            RRSPH_COMMON/src/EOS.cpp:23-34
//   continuity density  RRSPH_COMMON/include/Density.h:88-113
//   pressure force      RRSPH_COMMON/include/InternalForce.h:24-141
//   Lennard-Jones walls RRSPH_COMMON/include/ExtForce.h:15-61
//   artificial visc.    RRSPH_COMMON/include/ArtificialViscosity.h:29-66
//   XSPH                RRSPH_COMMON/include/AverageVelocity.h:20-37
// Gradients are returned as a scalar F with grad W = diff * F, where
// diff = r_i - r_j, so one F serves every term that uses the same kernel.
#pragma once

#include "device_common.hpp"

namespace rrsph {

template <class R> __device__ __forceinline__ R sq(R v) { return v * v; }
template <class R> __device__ __forceinline__ R cb(R v) { return v * v * v; }
template <class R> __device__ __forceinline__ R pow7(R x) {
    // repeated multiplication as powun (CommonIncl.h:36-42)
    R r = x * x;
    r = r * x;
    r = r * x;
    r = r * x;
    r = r * x;
    return r * x;
}

template <class R> struct KernelConsts {
    R h, inv_h, f_cubic, f_gauss, f_wend, f_desb;
};

template <class R>
__device__ __forceinline__ KernelConsts<R> kernel_consts(const StepParams& P) {
    KernelConsts<R> K;
    K.h = (R)P.h;
    K.inv_h = (R)(1.0 / P.h);
    K.f_cubic = (R)P.f_cubic;
    K.f_gauss = (R)P.f_gauss;
    K.f_wend = (R)P.f_wend;
    K.f_desb = (R)P.f_desb;
    return K;
}

template <class R>
__device__ __forceinline__ R kernel_w(const KernelConsts<R>& K, int skf, R dist) {
    const R q = dist / K.h;
    switch (skf) {
    case 2: return K.f_gauss * exp(-q * q);
    case 3: {
        if (q > R(2)) return R(0);
        const R t = R(1) - R(0.5) * q;
        return K.f_wend * (t * t * t * t) * (R(2) * q + R(1));
    }
    case 4: return q <= R(2) ? K.f_desb * cb(R(2) - q) : R(0);
    default:
        if (q <= R(1)) return K.f_cubic * (R(2) / R(3) - q * q + q * q * q * R(0.5));
        if (q <= R(2)) return K.f_cubic * (R(1) / R(6) * cb(R(2) - q));
        return R(0);
    }
}

// grad W = diff * kernel_gradf(...)
template <class R>
__device__ __forceinline__ R kernel_gradf(const KernelConsts<R>& K, int skf, R dist) {
    const R q = dist / K.h;
    switch (skf) {
    case 2: return R(-2) / (K.h * K.h) * kernel_w(K, 2, dist);
    case 3: {
        if (q < R(1.e-6f) || q >= R(2)) return R(0);
        const R t = R(1) - R(0.5) * q;
        const R t3 = t * t * t;
        return R(2) * K.f_wend / K.h / dist * (t3 * t - (R(2) * q + R(1)) * t3);
    }
    case 4: return q < R(2) ? K.f_desb / dist / K.h * (R(-3) * sq(R(2) - q)) : R(0);
    default:
        if (q <= R(1)) return K.f_cubic / (K.h * K.h) * (R(-2) + R(1.5) * q);
        if (q <= R(2)) return -K.f_cubic / dist / K.h * (sq(R(2) - q) * R(0.5));
        return R(0);
    }
}

template <class R> __device__ __forceinline__ R eos_p(R rho, R rho0, R B) {
    return B * (pow7(rho / rho0) - R(1));
}

}  // namespace rrsph
