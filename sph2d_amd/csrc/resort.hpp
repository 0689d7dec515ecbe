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
//   (integrator)       stores each wavefront's ballot of changed slots
//   resort_sort_small  one block: the step prologue; in passes of 16384
//                      wavefronts, the exclusive prefix wpre[w] of the
//                      changed counts of the wavefronts before w and the
//                      changed slots gathered as 64-bit composites
//                      (key << 32 | slot) in slot order; a total C above `cap`
//                      halts the step with kStatusResort (the host re-runs it
//                      with the full radix sort), else a bitonic sort of the
//                      composites in LDS
//   resort_scatter     places every slot: unchanged s at s - cb(s) + the
//                      number of changed composites below (skeyA[s], s);
//                      changed composite g at g + p - cb(p), p = the number
//                      of slots whose (skeyA[t], t) is below it; cb(s) = the
//                      changed slots below s = wpre[s / 64] + the set bits of
//                      the ballot below lane s % 64
// Every position written is in [0, n) whatever the keys hold: s - cb(s)
// counts unchanged slots and g changed ones, both from the same ballots.
#pragma once

#include "device_common.hpp"
#include "step_kernels.hpp"

namespace rrsph {

constexpr uint32_t kResortBlock = 1024;
constexpr uint32_t kResortCap = 8192;           // changed slots sorted in LDS (64 KiB)
constexpr uint32_t kStatusResort = 100;         // internal: re-run this step with the full sort

__device__ __forceinline__ unsigned long long resort_comp(uint32_t key, uint32_t slot) {
    return ((unsigned long long)key << 32) | slot;
}

__device__ __forceinline__ unsigned long long lanes_below(uint32_t lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Exclusive prefix of v over a block of kResortBlock threads; *total gets
// the sum (in every thread).
__device__ __forceinline__ uint32_t resort_block_scan(uint32_t v, uint32_t* s_w, uint32_t* total) {
    const uint32_t lane = __lane_id(), wave = threadIdx.x / kWave;
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t u = __shfl_up(incl, off);
        if (lane >= (uint32_t)off) incl += u;
    }
    if (lane == kWave - 1) s_w[wave] = incl;
    __syncthreads();
    uint32_t before = 0, t = 0;
#pragma unroll
    for (uint32_t w = 0; w < kResortBlock / kWave; ++w) {
        const uint32_t sw = s_w[w];
        if (w < wave) before += sw;
        t += sw;
    }
    __syncthreads();
    *total = t;
    return before + incl - v;
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

constexpr uint32_t kResortChunk = 16;  // ballots per thread and pass of the scan
// counts of one pass in LDS, one padding word after every 64 so that both the
// coalesced (j * kResortBlock + t) and the per-thread (t * kResortChunk + j)
// accesses of a wave hit 64 different banks
constexpr uint32_t kResortPass = kResortBlock * kResortChunk;
__device__ __forceinline__ uint32_t resort_pad(uint32_t i) { return i + (i >> 6); }
constexpr uint32_t kResortLds = kResortCap + kResortPass / 128;  // 64-bit words: sort buffer or padded counts
static_assert(kResortPass + kResortPass / 64 <= 2 * kResortLds, "a pass of counts fits the sort buffer");

// One block.  Each pass loads kResortChunk ballots per thread (coalesced),
// puts their counts in LDS, scans them (thread t owns the kResortChunk
// consecutive counts from t * kResortChunk), writes wpre coalesced and
// gathers the changed slots' composites to chg[wpre[w] + j] (j-th set bit).
__global__ void __launch_bounds__(kResortBlock) resort_sort_small(StepParams P, const uint32_t* __restrict__ key_next,
                                                                  uint32_t cap, unsigned long long* __restrict__ chg,
                                                                  uint32_t* __restrict__ wpre,
                                                                  Ctrl* __restrict__ ctrl) {
    __shared__ uint32_t s_w[kResortBlock / kWave];
    __shared__ unsigned long long s[kResortLds];
    uint32_t* s32 = (uint32_t*)s;
    const uint32_t tid = threadIdx.x;
    const uint32_t nw = (P.n + kWave - 1) / kWave;
    const unsigned long long* __restrict__ mask = P.chg_mask;
    unsigned long long m[kResortChunk];
    auto load_pass = [&](uint32_t base) {
#pragma unroll
        for (uint32_t j = 0; j < kResortChunk; ++j) {
            const uint32_t w = base + j * kResortBlock + tid;
            m[j] = w < nw ? mask[w] : 0ull;
        }
    };
    const unsigned long long word = ctrl_word(ctrl);
    load_pass(0);  // in flight during the prologue
    block_prologue(P, ctrl, word, [](Ctrl&) {}, [](Ctrl&) {});
    uint32_t run = 0;
    for (uint32_t base = 0; base < nw; base += kResortPass) {
        if (base != 0) load_pass(base);
#pragma unroll
        for (uint32_t j = 0; j < kResortChunk; ++j) s32[resort_pad(j * kResortBlock + tid)] = (uint32_t)__popcll(m[j]);
        __syncthreads();
        uint32_t part = 0;
#pragma unroll
        for (uint32_t j = 0; j < kResortChunk; ++j) part += s32[resort_pad(tid * kResortChunk + j)];
        uint32_t total;
        uint32_t pre = run + resort_block_scan(part, s_w, &total);
#pragma unroll
        for (uint32_t j = 0; j < kResortChunk; ++j) {  // this thread's own entries
            const uint32_t c = s32[resort_pad(tid * kResortChunk + j)];
            s32[resort_pad(tid * kResortChunk + j)] = pre;
            pre += c;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < kResortChunk; ++j) {
            const uint32_t w = base + j * kResortBlock + tid;
            if (w < nw) {
                uint32_t p = s32[resort_pad(j * kResortBlock + tid)];
                wpre[w] = p;
                // past cap the step overflows anyway: gather no further
                for (unsigned long long mm = m[j]; mm != 0 && p < cap; mm &= mm - 1, ++p) {
                    const uint32_t sl = w * kWave + (uint32_t)(__ffsll((long long)mm) - 1);
                    chg[p] = resort_comp(key_next[sl], sl);
                }
            }
        }
        run += total;
        __syncthreads();  // s32 is refilled by the next pass; chg is read by the sort
    }
    __syncthreads();
    if (ctrl->halted) return;
    const uint32_t C = run;
    if (C > cap) {
        if (tid == 0) {
            Ctrl& c = *ctrl;
            c.resort_count = C;
            c.halted = 1;
            c.status = kStatusResort;
            if (c.wall_disappear) c.nwm_done = 0;  // the re-run prologue fires it again
        }
        return;
    }
    if (tid == 0) ctrl->resort_count = C;
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

__global__ void __launch_bounds__(kBlock) resort_scatter(const uint32_t* __restrict__ skeyA,
                                                         const unsigned long long* __restrict__ chg,
                                                         const unsigned long long* __restrict__ mask,
                                                         const uint32_t* __restrict__ wpre, uint32_t n,
                                                         uint32_t* __restrict__ skeyB, uint32_t* __restrict__ perm,
                                                         const Ctrl* __restrict__ ctrl, uint32_t stage) {
    if (ctrl->halted) return;
    __shared__ uint32_t s_lo, s_hi;
    __shared__ unsigned long long s_c[kBlock];
    const uint32_t C = ctrl->resort_count;
    const uint32_t tid = threadIdx.x, lane = __lane_id();
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
    const bool staged = cnt <= stage;  // stage <= kBlock: the composites fit s_c
    if (staged && tid < cnt) s_c[tid] = chg[lo + tid];
    __syncthreads();
    if (s < n) {
        const unsigned long long m = mask[s / kWave];
        if (!((m >> lane) & 1ull)) {
            const uint32_t k = skeyA[s];
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
            const uint32_t cb = wpre[s / kWave] + (uint32_t)__popcll(m & lanes_below(lane));
            const uint32_t pos = s - cb + r;
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
        if (lane == 0) {
            const uint32_t cb = p < n ? wpre[p / kWave] + (uint32_t)__popcll(mask[p / kWave] & lanes_below(p % kWave))
                                      : C;
            const uint32_t pos = g + p - cb;
            skeyB[pos] = (uint32_t)(x >> 32);
            perm[pos] = (uint32_t)x;
        }
    }
}

}  // namespace rrsph
