// Stable LSD radix sort of (cell key, particle index) pairs, 8 bits per pass,
// any N (no power-of-two padding).  Replaces the reference's bitonic network
// (RRSPH_CL/cl/GridFind.cl:73-93, CLTimeIntegration.cpp:170-181) and the
// serial counting sort of RRSPH_OMP (GridFind.cpp:47-88).
//
// One pass = upsweep (per-tile digit histograms) -> one-block scan over the
// digit-major histogram table -> downsweep (stable in-tile ranking with wave
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

// Exclusive scan of `total` entries in place, one block of 1024 threads.
__global__ void __launch_bounds__(1024) radix_scan(uint32_t* __restrict__ hist, uint32_t total,
                                                   const Ctrl* __restrict__ ctrl) {
    if (ctrl->halted) return;
    __shared__ uint32_t s_part[1024];
    const uint32_t tid = threadIdx.x;
    const uint32_t chunk = (total + 1023) / 1024;
    const uint32_t b0 = tid * chunk;
    const uint32_t b1 = min(b0 + chunk, total);
    uint32_t sum = 0;
    for (uint32_t k = b0; k < b1; ++k) sum += hist[k];
    s_part[tid] = sum;
    __syncthreads();
    // Hillis-Steele inclusive scan over 1024 partial sums
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = tid >= off ? s_part[tid - off] : 0;
        __syncthreads();
        s_part[tid] += v;
        __syncthreads();
    }
    uint32_t run = tid ? s_part[tid - 1] : 0;
    for (uint32_t k = b0; k < b1; ++k) {
        const uint32_t v = hist[k];
        hist[k] = run;
        run += v;
    }
}

// Stable scatter of one pass.  vals_in == nullptr means vals = index.
__global__ void __launch_bounds__(kBlock) radix_downsweep(const uint32_t* __restrict__ keys_in,
                                                          const uint32_t* __restrict__ vals_in,
                                                          uint32_t* __restrict__ keys_out,
                                                          uint32_t* __restrict__ vals_out, uint32_t n, uint32_t shift,
                                                          const uint32_t* __restrict__ hist_scanned, uint32_t ntiles,
                                                          const Ctrl* __restrict__ ctrl) {
    if (ctrl->halted) return;
    constexpr int kWaves = kBlock / kWave;
    __shared__ uint32_t s_base[kRadix];          // global start of digit d for this tile + digits seen so far
    __shared__ uint32_t s_wcnt[kWaves][kRadix];  // per-wave digit counts of the current iteration
    __shared__ uint32_t s_wofs[kWaves][kRadix];  // per-wave digit offsets of the current iteration
    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid / kWave;
    s_base[tid] = hist_scanned[tid * ntiles + blockIdx.x];
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
