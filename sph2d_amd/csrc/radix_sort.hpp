// Stable LSD radix sort of (cell key, particle index) pairs, 8 bits per pass,
// any N (no power-of-two padding).  Replaces the reference's bitonic network
// (RRSPH_CL/cl/GridFind.cl:73-93, CLTimeIntegration.cpp:170-181) and the
// serial counting sort of RRSPH_OMP (GridFind.cpp:47-88).
//
// One pass = upsweep (per-tile digit histograms) -> per-digit scan across
// tiles (one block per digit) -> downsweep (stable in-tile ranking with wave
// ballots, scatter).  Tiles are kSortTile keys; within a tile the order of
// equal digits is preserved, and tiles are placed in block order, so the
// sort is stable and deterministic.
#pragma once

#include "device_common.hpp"

namespace rrsph {

// Lanes of this wave whose digit equals mine (8 ballots on the digit bits),
// restricted to `active` lanes.
__device__ __forceinline__ unsigned long long match_digit(uint32_t d, unsigned long long active) {
    unsigned long long m = active;
#pragma unroll
    for (uint32_t b = 0; b < kRadixBits; ++b) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

__device__ __forceinline__ unsigned long long lanemask_lt() {
    const uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Per-tile digit histogram; hist is digit-major: hist[d * ntiles + tile].
__global__ void __launch_bounds__(kBlock) radix_upsweep(const uint32_t* __restrict__ keys, uint32_t n, uint32_t shift,
                                                        uint32_t* __restrict__ hist, uint32_t ntiles,
                                                        const Ctrl* __restrict__ ctrl) {
    if (ctrl->halted) return;
    __shared__ uint32_t s_hist[kRadix];
    const uint32_t tid = threadIdx.x;
    s_hist[tid] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    for (int it = 0; it < kSortItems; ++it) {
        const uint32_t idx = base + it * kBlock + tid;
        const bool valid = idx < n;
        const uint32_t d = valid ? (keys[idx] >> shift) & (kRadix - 1) : 0;
        const unsigned long long act = __ballot(valid);
        const unsigned long long m = match_digit(d, act);
        if (valid && (m & lanemask_lt()) == 0) atomicAdd(&s_hist[d], (uint32_t)__popcll(m));
    }
    __syncthreads();
    hist[tid * ntiles + blockIdx.x] = s_hist[tid];
}

// Per-digit exclusive scan across tiles: block d scans hist[d * ntiles + t]
// over t in place and stores the digit's total in digit_total[d].
constexpr int kScanBlock = 512;
__global__ void __launch_bounds__(kScanBlock) radix_scan_digits(uint32_t* __restrict__ hist, uint32_t ntiles,
                                                                uint32_t* __restrict__ digit_total,
                                                                const Ctrl* __restrict__ ctrl) {
    if (ctrl->halted) return;
    __shared__ uint32_t s_w[kScanBlock / kWave];
    const uint32_t tid = threadIdx.x, lane = __lane_id(), wave = tid / kWave;
    uint32_t* row = hist + (size_t)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += kScanBlock) {
        const uint32_t t = base + tid;
        const uint32_t v = t < ntiles ? row[t] : 0;
        // inclusive wave scan
        uint32_t x = v;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if ((int)lane >= off) x += y;
        }
        if (lane == kWave - 1) s_w[wave] = x;
        __syncthreads();
        uint32_t wbase = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kScanBlock / kWave; ++w) {
            const uint32_t sw = s_w[w];
            if (w < (int)wave) wbase += sw;
            total += sw;
        }
        if (t < ntiles) row[t] = carry + wbase + x - v;
        carry += total;
        __syncthreads();
    }
    if (tid == 0) digit_total[blockIdx.x] = carry;
}

// Stable scatter of one pass.  vals_in == nullptr means vals = index.
__global__ void __launch_bounds__(kBlock) radix_downsweep(const uint32_t* __restrict__ keys_in,
                                                          const uint32_t* __restrict__ vals_in,
                                                          uint32_t* __restrict__ keys_out,
                                                          uint32_t* __restrict__ vals_out, uint32_t n, uint32_t shift,
                                                          const uint32_t* __restrict__ hist_scanned, uint32_t ntiles,
                                                          const uint32_t* __restrict__ digit_total,
                                                          const Ctrl* __restrict__ ctrl) {
    if (ctrl->halted) return;
    constexpr int kWaves = kBlock / kWave;
    __shared__ uint32_t s_tot[kWaves];
    __shared__ uint32_t s_base[kRadix];          // global start of digit d for this tile + digits seen so far
    __shared__ uint32_t s_wcnt[kWaves][kRadix];  // per-wave digit counts of the current iteration
    __shared__ uint32_t s_wofs[kWaves][kRadix];  // per-wave digit offsets of the current iteration
    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid / kWave;
    {
        // exclusive scan of the 256 digit totals (one digit per thread)
        const uint32_t lane = __lane_id();
        const uint32_t v = digit_total[tid];
        uint32_t x = v;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if ((int)lane >= off) x += y;
        }
        if (lane == kWave - 1) s_tot[wave] = x;
        __syncthreads();
        uint32_t wbase = 0;
        for (uint32_t w = 0; w < wave; ++w) wbase += s_tot[w];
        s_base[tid] = wbase + x - v + hist_scanned[tid * ntiles + blockIdx.x];
    }
    const uint32_t base = blockIdx.x * kSortTile;
    for (int it = 0; it < kSortItems; ++it) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) s_wcnt[w][tid] = 0;
        __syncthreads();
        const uint32_t idx = base + it * kBlock + tid;
        const bool valid = idx < n;
        const uint32_t key = valid ? keys_in[idx] : 0;
        const uint32_t val = valid ? (vals_in ? vals_in[idx] : idx) : 0;
        const uint32_t d = (key >> shift) & (kRadix - 1);
        const unsigned long long act = __ballot(valid);
        const unsigned long long m = match_digit(d, act);
        const uint32_t rank = (uint32_t)__popcll(m & lanemask_lt());
        if (valid && rank == 0) s_wcnt[wave][d] = (uint32_t)__popcll(m);
        __syncthreads();
        {
            uint32_t run = s_base[tid];
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                s_wofs[w][tid] = run;
                run += s_wcnt[w][tid];
            }
            s_base[tid] = run;
        }
        __syncthreads();
        if (valid) {
            const uint32_t pos = s_wofs[wave][d] + rank;
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
    }
}

}  // namespace rrsph
