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

// ------------------------------------------------------------------ probes
// SPH interpolation of a field at points, as FuncAtPoint does
// (tools/FuncAtPoint/lib/src/fatp.cpp:14-58): at point x, the sum over the
// particles of the 3^dim cells around x's cell with |r_i - x|^2 < cell^2 of
// f_i W(|r_i - x|) m / rho_i, W the cubic spline whatever the step's kernels
// (SmoothingKernel.h:20-45), rho_i the summation density of the current
// positions (Density.h:25-61, recomputed per layer at fatp.cpp:134-139).
// The reference's grid holds every particle, dead ones (itype 0, rho = m W(0),
// no neighbours) included; here they are keyed past the last cell, so each
// point also scans that tail [LB[ncells], n).  One thread per point, FP64;
// fld is the field in original particle order with `stride` values per
// particle, read at `comp`.
template <class R>
__global__ void probe_kernel(StepParams P, StateB<R> B, const uint32_t* __restrict__ LB,
                             const double* __restrict__ pts, uint32_t np, const double* __restrict__ fld,
                             uint32_t stride, uint32_t comp, double* __restrict__ out) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= np) return;
    const double x = pts[2 * g], y = pts[2 * g + 1];
    uint32_t bad = 0;  // points lie in the geometry box (checked on the host)
    const int cx = cell_coord(x, P.xmin, P.cell, P.ncx, bad);
    const int cy = cell_coord(y, P.ymin, P.cell, P.ncy, bad);
    const KernelConsts<double> K = kernel_consts<double>(P);
    double val = 0.0;
    auto add = [&](uint32_t i) {
        const double2 q = B.pos[i];
        const double d2 = dist2_exact(__dsub_rn(q.x, x), __dsub_rn(q.y, y));
        if (d2 < P.cut2)
            val += fld[(size_t)B.oid[i] * stride + comp] * kernel_w<double>(K, 1, sqrt(d2)) * P.mass /
                   (double)B.pred[i].z;
    };
    const int y0 = cy > 0 ? cy - 1 : 0;
    const int y1 = cy + 1 < P.ncy ? cy + 1 : P.ncy - 1;
    for (int dc = -1; dc <= 1; ++dc) {
        const int col = cx + dc;
        if (col < 0 || col >= P.ncx) continue;
        for (uint32_t i = LB[col * P.ncy + y0]; i < LB[col * P.ncy + y1 + 1]; ++i) add(i);
    }
    for (uint32_t i = LB[P.ncells]; i < P.n; ++i) add(i);
    out[g] = val;
}

// The same in 3D (27 cells, rho_i in pred.w).
template <class R>
__global__ void probe3_kernel(StepParams P, StateB3<R> B, const uint32_t* __restrict__ LB,
                              const double* __restrict__ pts, uint32_t np, const double* __restrict__ fld,
                              uint32_t stride, uint32_t comp, double* __restrict__ out) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= np) return;
    const double x = pts[3 * g], y = pts[3 * g + 1], z = pts[3 * g + 2];
    uint32_t bad = 0;
    const int cx = cell_coord(x, P.xmin, P.cell, P.ncx, bad);
    const int cy = cell_coord(y, P.ymin, P.cell, P.ncy, bad);
    const int cz = cell_coord(z, P.zmin, P.cell, P.ncz, bad);
    const KernelConsts<double> K = kernel_consts<double>(P);
    double val = 0.0;
    auto add = [&](uint32_t i) {
        const double4 q = B.pos[i];
        const double d2 = dist2_exact3(__dsub_rn(q.x, x), __dsub_rn(q.y, y), __dsub_rn(q.z, z));
        if (d2 < P.cut2)
            val += fld[(size_t)B.oid[i] * stride + comp] * kernel_w<double>(K, 1, sqrt(d2)) * P.mass /
                   (double)B.pred[i].w;
    };
    const int z0 = cz > 0 ? cz - 1 : 0;
    const int z1 = cz + 1 < P.ncz ? cz + 1 : P.ncz - 1;
    for (int dx = -1; dx <= 1; ++dx) {
        const int xc = cx + dx;
        if (xc < 0 || xc >= P.ncx) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int yc = cy + dy;
            if (yc < 0 || yc >= P.ncy) continue;
            const uint32_t col = ((uint32_t)xc * (uint32_t)P.ncy + (uint32_t)yc) * (uint32_t)P.ncz;
            for (uint32_t i = LB[col + z0]; i < LB[col + z1 + 1]; ++i) add(i);
        }
    }
    for (uint32_t i = LB[P.ncells]; i < P.n; ++i) add(i);
    out[g] = val;
}

}  // namespace rrsph
