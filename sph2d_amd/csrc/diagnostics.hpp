// On-device diagnostics: mass / energy budget of the fluid and wave gauges.
//
// Gauges follow HeightTesting::maxInLayer (tools/WaterProfile/include/
// HeightTesting.h:25-54): the maximum y of particles of the chosen type within
// |x - x_g| < search_n * h, starting from 0 as the reference's zero-filled
// per-thread maxima do.
#pragma once

#include "device_common.hpp"
#include "step_kernels.hpp"
#include "step3d.hpp"

namespace rrsph {

constexpr int kDiagSlots = 8;  // 0 n_fluid, 1 n_outside, 2 KE, 3 PE, 4 IE (sums); 5 max|v|, 6 -rho_min, 7 rho_max (max)
constexpr int kDiagMaxBlocks = 1024;

template <class T> __device__ T block_reduce_sum(T v, T* sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int w = threadIdx.x / kWave;
    __syncthreads();
    if (__lane_id() == 0) sh[w] = v;
    __syncthreads();
    T r = 0;
    for (int k = 0; k < kBlock / kWave; ++k) r += sh[k];
    return r;
}

template <class T> __device__ T block_reduce_max(T v, T* sh) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    const int w = threadIdx.x / kWave;
    __syncthreads();
    if (__lane_id() == 0) sh[w] = v;
    __syncthreads();
    T r = sh[0];
    for (int k = 1; k < kBlock / kWave; ++k) r = fmax(r, sh[k]);
    return r;
}

// Tait elastic energy per unit mass, e(rho) = int_{rho0}^{rho} p / rho'^2 drho'
__device__ __forceinline__ double tait_energy(double rho, double rho0, double B) {
    const double x = rho / rho0;
    const double x6 = x * x * x * x * x * x;
    return B / rho0 * ((x6 - 1) / 6 + 1 / x - 1);
}

template <class R> __global__ void __launch_bounds__(kBlock) diag_partials(StepParams P, StateA<R> A, double* red) {
    __shared__ double sh[kBlock / kWave];
    double acc[kDiagSlots] = {0, 0, 0, 0, 0, 0, -INFINITY, -INFINITY};
    for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < P.n; s += gridDim.x * kBlock) {
        const double2 p = A.pos[s];
        const real4<R> va = A.va[s];
        const bool inside = p.x < P.xmax && p.x > P.xmin && p.y < P.ymax && p.y > P.ymin;
        acc[1] += inside ? 0.0 : 1.0;
        if ((int)va.w > 0) {
            const double vx = va.x, vy = va.y, rho = va.z;
            const double v2 = vx * vx + vy * vy;
            acc[0] += 1.0;
            acc[2] += 0.5 * P.mass * v2;
            acc[3] += P.mass * P.g * p.y;
            acc[4] += P.mass * tait_energy(rho, P.rho0, P.eos_B);
            acc[5] = fmax(acc[5], sqrt(v2));
            acc[6] = fmax(acc[6], -rho);
            acc[7] = fmax(acc[7], rho);
        }
    }
    for (int k = 0; k < kDiagSlots; ++k) {
        const double r = k < 5 ? block_reduce_sum(acc[k], sh) : block_reduce_max(acc[k], sh);
        if (threadIdx.x == 0) red[k * kDiagMaxBlocks + blockIdx.x] = r;
    }
}

template <class R> __global__ void __launch_bounds__(kBlock) diag_partials3(StepParams P, StateA3<R> A, double* red) {
    __shared__ double sh[kBlock / kWave];
    double acc[kDiagSlots] = {0, 0, 0, 0, 0, 0, -INFINITY, -INFINITY};
    for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < P.n; s += gridDim.x * kBlock) {
        const double4 p = A.pos[s];
        const real4<R> va = A.va[s];
        const bool inside =
            p.x < P.xmax && p.x > P.xmin && p.y < P.ymax && p.y > P.ymin && p.z < P.zmax && p.z > P.zmin;
        acc[1] += inside ? 0.0 : 1.0;
        if ((int)p.w > 0) {
            const double vx = va.x, vy = va.y, vz = va.z, rho = va.w;
            const double v2 = vx * vx + vy * vy + vz * vz;
            acc[0] += 1.0;
            acc[2] += 0.5 * P.mass * v2;
            acc[3] += P.mass * P.g * p.y;
            acc[4] += P.mass * tait_energy(rho, P.rho0, P.eos_B);
            acc[5] = fmax(acc[5], sqrt(v2));
            acc[6] = fmax(acc[6], -rho);
            acc[7] = fmax(acc[7], rho);
        }
    }
    for (int k = 0; k < kDiagSlots; ++k) {
        const double r = k < 5 ? block_reduce_sum(acc[k], sh) : block_reduce_max(acc[k], sh);
        if (threadIdx.x == 0) red[k * kDiagMaxBlocks + blockIdx.x] = r;
    }
}

__global__ void __launch_bounds__(kBlock) diag_final(double* red, uint32_t nblk) {
    __shared__ double sh[kBlock / kWave];
    for (int k = 0; k < kDiagSlots; ++k) {
        double v = k < 5 ? 0.0 : -INFINITY;
        for (uint32_t b = threadIdx.x; b < nblk; b += kBlock) {
            const double x = red[k * kDiagMaxBlocks + b];
            v = k < 5 ? v + x : fmax(v, x);
        }
        const double r = k < 5 ? block_reduce_sum(v, sh) : block_reduce_max(v, sh);
        if (threadIdx.x == 0) red[(size_t)kDiagSlots * kDiagMaxBlocks + k] = r;
    }
}

template <class R>
__global__ void __launch_bounds__(kBlock) gauge_kernel(uint32_t n, StateA<R> A, const double* __restrict__ xg,
                                                       double radius, int32_t ptype,
                                                       unsigned long long* __restrict__ out_bits) {
    __shared__ double sh[kBlock / kWave];
    const double x0 = xg[blockIdx.y];
    double m = 0.0;
    for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock) {
        const real4<R> va = A.va[s];
        if (ptype != 0 && (int)va.w != ptype) continue;
        const double2 p = A.pos[s];
        if (fabs(p.x - x0) < radius) m = fmax(m, p.y);
    }
    m = block_reduce_max(m, sh);
    if (threadIdx.x == 0 && m > 0.0) atomicMax(&out_bits[blockIdx.y], (unsigned long long)__double_as_longlong(m));
}

}  // namespace rrsph
