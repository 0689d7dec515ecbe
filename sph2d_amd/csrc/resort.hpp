// Incremental re-sort of the cell keys between time steps.
//
// After a step, slot s of the state holds the particle that was at sorted
// position s; skeyA[s] is the key it was sorted by and key_next[s] its new
// key.  Between two steps only the particles that crossed a cell boundary
// (a few per mille) change key, so the new order is a merge of
//   * the unchanged slots, already in (key, slot) order, and
//   * the changed slots, sorted by (key, slot) here,
// which is exactly the permutation the stable LSD radix sort of the keys in
// slot order produces (a stable sort orders by the composite (key, slot)).
//
//   (integrator)       counts the changed slots per tile of kResortTile
//   resort_compact     block 0 first runs the step prologue; every block
//                      writes chg_before[s] = changed slots before s and the
//                      changed slots of its tile as 64-bit composites
//                      (key << 32 | slot)
//   resort_sort_small  total C; more than `cap` halts the step with
//                      kStatusResort (the host re-runs it with the full
//                      radix sort), else a bitonic sort of the composites in
//                      LDS (one block).  (Folding it into the compaction as
//                      a last-block-done step needs a device-scope fence in
//                      every block, which on the multi-XCD chip writes back
//                      the L2 and cost 0.23 ms per step.)
//   resort_scatter     places every slot: unchanged s at s - chg_before[s] +
//                      the number of changed composites below (key, s);
//                      changed composite g at g + p - chg_before[p], p = the
//                      number of slots whose (skeyA[t], t) is below it
// Every position written is in [0, n) whatever the keys hold: s - chg_before[s]
// counts unchanged slots and g counts changed ones.
#pragma once

#include "device_common.hpp"
#include "step_kernels.hpp"

namespace rrsph {

constexpr uint32_t kResortBlock = 1024;
constexpr uint32_t kResortItems = kResortTile / kResortBlock;
constexpr uint32_t kResortCap = 8192;           // changed slots sorted in LDS (64 KiB)
constexpr uint32_t kStatusResort = 100;         // internal: re-run this step with the full sort

__device__ __forceinline__ unsigned long long resort_comp(uint32_t key, uint32_t slot) {
    return ((unsigned long long)key << 32) | slot;
}

// Sum of v over a block of kResortBlock threads (result in every thread).
__device__ __forceinline__ uint32_t resort_block_sum(uint32_t v, uint32_t* s_w) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (__lane_id() == 0) s_w[threadIdx.x / kWave] = v;
    __syncthreads();
    uint32_t t = 0;
#pragma unroll
    for (uint32_t w = 0; w < kResortBlock / kWave; ++w) t += s_w[w];
    __syncthreads();
    return t;
}

// Bitonic sort of chg[0, C) by one block of kResortBlock threads in LDS (C <= kResortCap).
__device__ __forceinline__ void resort_sort_block(uint32_t C, unsigned long long* __restrict__ chg,
                                                  unsigned long long* s) {
    uint32_t P2 = 1;
    while (P2 < C) P2 <<= 1;
    for (uint32_t i = threadIdx.x; i < P2; i += kResortBlock) s[i] = i < C ? chg[i] : ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= P2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < P2; i += kResortBlock) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const unsigned long long a = s[i], b = s[ixj];
                    const bool asc = (i & k) == 0;
                    if ((a > b) == asc) {
                        s[i] = b;
                        s[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = threadIdx.x; i < C; i += kResortBlock) chg[i] = s[i];
}

__global__ void __launch_bounds__(kResortBlock) resort_compact(StepParams P, const uint32_t* __restrict__ key_next,
                                                               const uint32_t* __restrict__ skeyA, uint32_t n,
                                                               const uint32_t* __restrict__ tile_cnt, uint32_t cap,
                                                               uint32_t* __restrict__ chg_before,
                                                               unsigned long long* __restrict__ chg,
                                                               Ctrl* __restrict__ ctrl) {
    __shared__ uint32_t s_w[kResortBlock / kWave];
    // the prologue's control-block fields are read by the later kernels of
    // the step only; the compaction below (scratch arrays) runs whether or
    // not the step halts
    if (blockIdx.x == 0 && threadIdx.x == 0) step_prologue_body(P, ctrl);
    const uint32_t tid = threadIdx.x, wave = tid / kWave;
    const uint32_t lane = __lane_id();
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    // changed slots in the tiles before this one
    uint32_t before = 0;
    for (uint32_t t = tid; t < blockIdx.x; t += kResortBlock) before += tile_cnt[t];
    uint32_t run = resort_block_sum(before, s_w);
    const uint32_t base = blockIdx.x * kResortTile;
    for (uint32_t it = 0; it < kResortItems; ++it) {
        const uint32_t s = base + it * kResortBlock + tid;
        const bool live = s < n;
        const uint32_t k = live ? key_next[s] : 0u;
        const bool changed = live && k != skeyA[s];
        const unsigned long long m = __ballot(changed);
        if (lane == 0) s_w[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t wbase = 0, total = 0;
#pragma unroll
        for (uint32_t w = 0; w < kResortBlock / kWave; ++w) {
            const uint32_t sw = s_w[w];
            if (w < wave) wbase += sw;
            total += sw;
        }
        const uint32_t c = run + wbase + (uint32_t)__popcll(m & lt);
        if (live) chg_before[s] = c;
        if (changed && c < cap) chg[c] = resort_comp(k, s);
        run += total;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kResortBlock) resort_sort_small(const uint32_t* __restrict__ tile_cnt,
                                                                  uint32_t ntiles, uint32_t* __restrict__ chg_before,
                                                                  uint32_t n, uint32_t cap,
                                                                  unsigned long long* __restrict__ chg,
                                                                  Ctrl* __restrict__ ctrl) {
    if (ctrl->halted) return;
    __shared__ uint32_t s_w[kResortBlock / kWave];
    __shared__ unsigned long long s[kResortCap];
    uint32_t part = 0;
    for (uint32_t t = threadIdx.x; t < ntiles; t += kResortBlock) part += tile_cnt[t];
    const uint32_t C = resort_block_sum(part, s_w);
    if (C > cap) {
        if (threadIdx.x == 0) {
            Ctrl& c = *ctrl;
            c.resort_count = C;
            c.halted = 1;
            c.status = kStatusResort;
            if (c.wall_disappear) c.nwm_done = 0;  // the re-run prologue fires it again
        }
        return;
    }
    if (threadIdx.x == 0) {
        ctrl->resort_count = C;
        chg_before[n] = C;
    }
    resort_sort_block(C, chg, s);
}

// Number of t in [0, len) with f(t) < x for a strictly increasing f, found by
// one wave in rounds of 64 probes (log64 rounds of dependent loads instead of
// log2).  Every lane of the wave calls it with the same x and len.
template <class F>
__device__ __forceinline__ uint32_t wave_count_below(F&& f, uint32_t len, unsigned long long x) {
    uint32_t lo = 0, hi = len;  // the answer lies in [lo, hi]
    const uint32_t lane = __lane_id();
    while (hi > lo) {
        const uint32_t step = (hi - lo + kWave - 1) / kWave;
        const uint32_t t = lo + lane * step;
        const bool below = t < hi && f(t) < x;
        const uint32_t k = (uint32_t)__popcll(__ballot(below));  // the probes below x are a prefix
        const uint32_t nlo = k == 0 ? lo : lo + (k - 1) * step + 1;
        hi = min(lo + k * step, hi);
        lo = nlo;
    }
    return lo;
}

__global__ void __launch_bounds__(kBlock) resort_scatter(const uint32_t* __restrict__ key_next,
                                                         const uint32_t* __restrict__ skeyA,
                                                         const uint32_t* __restrict__ chg_before,
                                                         const unsigned long long* __restrict__ chg, uint32_t n,
                                                         uint32_t* __restrict__ skeyB, uint32_t* __restrict__ perm,
                                                         const Ctrl* __restrict__ ctrl) {
    if (ctrl->halted) return;
    __shared__ uint32_t s_lo, s_hi;
    __shared__ unsigned long long s_c[kBlock];
    const uint32_t C = ctrl->resort_count;
    const uint32_t tid = threadIdx.x;
    const uint32_t s0 = blockIdx.x * kBlock;
    const uint32_t s = s0 + tid;
    auto changed_at = [&](uint32_t t) { return chg[t]; };
    // the changed composites that can fall between this block's slots
    if (tid < kWave) {
        const uint32_t last = min(s0 + kBlock, n) - 1;
        const uint32_t lo = wave_count_below(changed_at, C, resort_comp(skeyA[s0], s0));
        const uint32_t hi = wave_count_below(changed_at, C, resort_comp(skeyA[last], last) + 1);
        if (tid == 0) {
            s_lo = lo;
            s_hi = hi;
        }
    }
    __syncthreads();
    const uint32_t lo = s_lo, cnt = s_hi - s_lo;
    const bool staged = cnt <= kBlock;
    if (staged && tid < cnt) s_c[tid] = chg[lo + tid];
    __syncthreads();
    if (s < n) {
        const uint32_t k = skeyA[s];
        if (key_next[s] == k) {
            const unsigned long long x = resort_comp(k, s);
            uint32_t r = lo;
            if (staged) {
                for (uint32_t q = 0; q < cnt; ++q) r += s_c[q] < x ? 1u : 0u;
            } else {
                uint32_t a = 0, b = cnt;  // rare: a block spanning many changed slots
                while (a < b) {
                    const uint32_t mid = (a + b) >> 1;
                    if (chg[lo + mid] < x) a = mid + 1;
                    else b = mid;
                }
                r += a;
            }
            const uint32_t pos = s - chg_before[s] + r;
            skeyB[pos] = k;
            perm[pos] = s;
        }
    }
    // changed composites, one wave each: p = number of slots whose
    // (skeyA[t], t) is below it (strictly increasing in t)
    auto slot_at = [&](uint32_t t) { return resort_comp(skeyA[t], t); };
    const uint32_t waves = gridDim.x * (kBlock / kWave);
    for (uint32_t g = blockIdx.x * (kBlock / kWave) + tid / kWave; g < C; g += waves) {
        const unsigned long long x = chg[g];
        const uint32_t p = wave_count_below(slot_at, n, x);
        if (__lane_id() == 0) {
            const uint32_t pos = g + p - chg_before[p];
            skeyB[pos] = (uint32_t)(x >> 32);
            perm[pos] = (uint32_t)x;
        }
    }
}

}  // namespace rrsph
