// RRSPH_HIP: C-ABI implementation (host side) of the MI355X WCSPH step.
//
// The context owns the device buffers, one HIP stream at default priority
// and (optionally) a captured hipGraph of one step.  Every entry point is
// declared in include/rrsph_hip.h with the reference interface it replaces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rrsph_hip.h"
#include "device_common.hpp"
#include "diagnostics.hpp"
#include "radix_sort.hpp"
#include "resort.hpp"
#include "adjoint.hpp"
#include "adjoint3.hpp"
#include "step_kernels.hpp"
#include "step3d.hpp"
#include "pair_tile.hpp"
#include "list_step.hpp"
#include "list_step3.hpp"

using namespace rrsph;

namespace {

enum KernelClass : uint32_t {
    K_PROLOGUE = 0,
    K_UPSWEEP,
    K_SCAN,
    K_DOWNSWEEP,
    K_REORDER,
    K_DENSITY_SUM,
    K_PAIR,
    K_INTEGRATE,
    K_RESORT_SORT,
    K_RESORT_SCATTER,
    K_LIST_BUILD,
    K_NUM_CLASSES
};
const char* const kClassNames[K_NUM_CLASSES] = {"step_prologue", "radix_upsweep", "radix_scan", "radix_downsweep",
                                                "reorder_kernel", "density_sum_kernel", "pair_kernel",
                                                "integrate_kernel", "resort_sort_small", "resort_scatter",
                                                "list_build_kernel"};

struct Profiler {
    std::vector<std::pair<uint32_t, std::pair<hipEvent_t, hipEvent_t>>> rec;
};

template <class T> T* dalloc(size_t n, hipError_t& err) {
    void* p = nullptr;
    if (err == hipSuccess) err = hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T));
    return (T*)p;
}

// Host copies of the smoothing-kernel W (SmoothingKernel.h:14-242) for constants.
double host_w(const rrsph_params_c& p, double dist, int skf) {
    const double h = p.hsml, pi = p.pi, q = dist / h;
    switch (skf) {
    case 2: return 1 / (pi * h * h) * std::exp(-q * q);
    case 3: {
        if (q > 2) return 0;
        double t = 1 - 0.5 * q;
        return 7 / (4 * pi * h * h) * (t * t * t * t) * (2 * q + 1);
    }
    case 4: return q <= 2 ? 5 / (16 * pi * h * h) * (2 - q) * (2 - q) * (2 - q) : 0;
    default: {
        double f = double(15) / (double(7) * pi * (h * h));
        if (q <= 1) return f * (double(2) / double(3) - q * q + q * q * q * 0.5);
        if (q <= 2) return f * (double(1) / double(6) * ((2 - q) * (2 - q) * (2 - q)));
        return 0;
    }
    }
}

}  // namespace

// Device scratch of rrsph_hip_smooth_gauges_noalloc: the block partial sums
// of every gauge (2 kMaxGauges kGaugeBlocks doubles), then eta, sum w e and
// sum w of every gauge (3 kMaxGauges).  1.03 MiB per device, held until the
// process ends.  Host code never uses this name as an address: the device's
// own address comes from hipGetSymbolAddress (scratch_address below).
constexpr size_t kScratchDoubles = 2 * (size_t)kMaxGauges * kGaugeBlocks + 3 * (size_t)kMaxGauges;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::mutex g_scratch_mu;
std::vector<double*> g_scratch;  // per device ordinal: the address of `scratch` on that device

// Address of `scratch` on the current device, looked up once per device (at
// context creation, with the context's device current).
hipError_t scratch_address(int device, double** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0) return hipErrorInvalidDevice;
    if ((size_t)device >= g_scratch.size()) g_scratch.resize(device + 1, nullptr);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = (double*)p;
    }
    *out = g_scratch[device];
    return hipSuccess;
}
}  // namespace

// Checkpoints and per-step records of a differentiable run (FP64 only).
struct AdjTape {
    bool valid = false;
    uint32_t nsteps = 0, every = 0;
    GaugeSpec3 spec{};
    size_t blob = 0;                     // bytes of one saved state
    std::vector<void*> ck;               // checkpoint before step k * every; last entry: final state
    void* seg = nullptr;                 // states of one replayed segment (every blobs), or of every step
    bool seg_last = false;               // seg holds the last segment's states (kept by the forward)
    // seg holds the state before every step of the run (and the sorts):
    // the forward keeps them all when they fit in half the free device
    // memory (tapes of sorted steps), and the reverse pass replays nothing
    bool keep_all = false;
    // the cell sort each step of seg ran with (keys, permutation): the reverse
    // step reorders its state with them instead of sorting it again
    uint32_t *seg_keys = nullptr, *seg_perm = nullptr;
    std::vector<uint8_t> seg_sorted;     // per seg slot: keys / perm hold that step's sort
    AdjStep* rec = nullptr;              // per step (device)
    Ctrl* uctrl = nullptr;               // per step: the control block its reverse step sorts and reorders with
                                         // (cur_dt = the step's dt_pred, nothing halted)
    std::vector<AdjStep> hrec;           // per step (host copy)
    double *eta = nullptr, *zs = nullptr, *ws = nullptr;  // nsteps x kMaxGauges
    double *eta_bar = nullptr, *partials = nullptr, *piston_sum = nullptr;
    double *lam[2] = {nullptr, nullptr}, *direct = nullptr, *drho_adj = nullptr;
    double4* out_adj = nullptr;          // 2D: out_a (n); 3D: the AdjRec3 records (n)
    uint32_t* flag = nullptr;
    // 3D with neighbour lists: the tape runs list steps.  The forward builds
    // at every checkpoint step and records which steps built; a replay
    // repeats that schedule, so it repeats its steps exactly.  The segment
    // in hand keeps the lists (and sorted cell keys) of each of its builds;
    // the reverse step of a list step reads its state in the slot order of
    // its build (identity reorder) and that build's lists.
    bool lists = false;
    std::vector<uint8_t> builds;         // per step: the step built lists
    struct Lists {
        uint32_t *nbl = nullptr, *nbc = nullptr, *keys = nullptr;
    };
    std::vector<Lists> seg_lists;        // lists of the segment's builds (allocated on demand, reused)
    uint32_t nseg_lists = 0;             // builds of the segment in hand
    std::vector<int32_t> seg_build;      // per seg slot: the build (index in seg_lists) its step read
    uint32_t* iota = nullptr;            // identity permutation
    float4* nearf = nullptr;             // FP32 step: near records of the reverse step's state (n)
    uint64_t force_build = ~0ull;        // tape step that must build (after a stale-list halt)
    // the forward's sorted step with the incremental re-sort, its record and
    // its gauges in one graph (rows from the control block's step count
    // - tbase); captured on the forward's first such step
    hipGraph_t tgraph = nullptr;
    hipGraphExec_t texec = nullptr;
    uint64_t tbase = 0;
};

struct rrsph_ctx {
    rrsph_params_c prm{};
    StepParams P{};
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool fp32 = true;
    bool fused = true;
    bool export_fields = false;
    bool use_graph = true;
    bool uploaded = false;
    bool stepped = false;
    bool fields_valid = false;  // der / xs hold the last step's fields in A's slot order
    uint32_t passes = 0;
    uint32_t ntiles = 0;
    // buffers
    double2 *posA = nullptr, *posB = nullptr;
    void *vaA = nullptr, *vaB = nullptr, *adA = nullptr, *pred = nullptr, *der = nullptr, *xs = nullptr,
         *pred_scratch = nullptr, *pres = nullptr, *pres_scratch = nullptr;
    void *pw = nullptr, *pw_scratch = nullptr;  // 3D: p / rho^2 (2D keeps it in pred.w)
    void* adB = nullptr;  // A.ad of the step's input, kept only if a fused step may be undone
    bool fast_pair = false;
    int pair_impl = 1;  // 0 general, 1 two-phase walk with LDS hit lists, 2 LDS tiles, 3 single-phase walk
    uint32_t *oidA = nullptr, *oidB = nullptr;
    uint32_t *key_next = nullptr, *kbuf[2] = {nullptr, nullptr}, *vbuf[2] = {nullptr, nullptr};
    uint32_t* hist = nullptr;
    uint32_t* digit_total = nullptr;
    unsigned long long* maxslots = nullptr;
    uint32_t* LB = nullptr;
    Ctrl* ctrl = nullptr;
    Ctrl* util_ctrl = nullptr;
    Ctrl* host_ctrl = nullptr;  // pinned mirror
    double* scratch = nullptr;  // this device's `scratch` (rrsph_hip_smooth_gauges_noalloc)
    double* host_eta = nullptr; // pinned, kMaxGauges doubles: gauge results on their way to the caller
    double* red = nullptr;      // diagnostics partials
    // staging for up/download (device)
    double *st_r = nullptr, *st_v = nullptr, *st_rho = nullptr, *st_p = nullptr, *st_a = nullptr,
           *st_drho = nullptr, *st_av = nullptr, *st_mu = nullptr;
    int32_t* st_itype = nullptr;
    uint32_t* st_u32 = nullptr;
    hipGraph_t graph = nullptr;        // step with the incremental re-sort
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph_full = nullptr;   // step with the full radix sort
    hipGraphExec_t gexec_full = nullptr;
    const uint32_t* sorted_keys = nullptr;
    const uint32_t* sorted_perm = nullptr;
    const uint32_t* last_keys = nullptr;  // the cell sort of the last launched step
    const uint32_t* last_perm = nullptr;
    // incremental re-sort
    uint32_t* skeyA = nullptr;          // key each slot of A was sorted by
    unsigned long long* chg = nullptr;       // changed slots as (key << 32 | slot)
    unsigned long long* chg_mask = nullptr;  // per wavefront of slots: ballot of the changed ones
    uint32_t* wpre = nullptr;                // per wavefront: changed slots in the wavefronts before
    uint32_t resort_cap = kResortCap;
    uint32_t resort_stage = kBlock;     // changed composites a scatter block stages in LDS (more: binary search)
    bool incremental = true;            // false: every step runs the full radix sort
    bool need_full = true;              // skeyA does not describe A (after upload)
    uint64_t launched = 0;              // steps launched by this context
    uint64_t full_until = 0;            // after repeated re-sort overflows: full sorts until this launch count
    uint64_t last_fallback = ~0ull;     // step index of the last re-sort overflow
    uint64_t full_run_end = 0;          // step index where the last run of full sorts ended
    uint64_t full_run = 0;              // length of the last run of full sorts
    uint64_t n_incremental = 0, n_full = 0, n_fallback = 0;
    // neighbour-list steps (2D cubic configuration): lists built every
    // list_interval steps (host schedule, adapted to the measured
    // displacement rate); the device halts a step whose lists may be stale
    bool list_mode = false;
    bool list_suspended = false;        // the adjoint's tapes and replays run without lists
    uint32_t *nbl = nullptr, *nbc = nullptr;
    void* rec = nullptr;                // records of the list pass (ListRec per slot)
    void* dens = nullptr;               // 3D: records of the summation density pass (real4 per slot)
    uint32_t* home = nullptr;           // each slot's cell at the last build
    unsigned long long* dslots = nullptr;  // per-step largest displacement (block maxima)
    float4* cand = nullptr;                 // 3D builds: candidate records (FP32 home-relative position, z cell)
    hipGraph_t graph_list = nullptr, graph_build = nullptr;
    hipGraphExec_t gexec_list = nullptr, gexec_build = nullptr;
    uint64_t build_at = 0;              // step index of the last build launched
    uint64_t list_interval = 8;         // steps between builds
    uint64_t n_list = 0, n_build = 0, n_list_halts = 0;
    AdjTape adj;
};

namespace {

rrsph_status fail(rrsph_ctx* c, rrsph_status st, const std::string& msg) {
    if (c) c->err = msg;
    return st;
}

#define HIP_TRY(expr)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return fail(ctx, RRSPH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
    } while (0)

template <class R> StateA<R> stateA(rrsph_ctx* c) {
    return StateA<R>{c->posA, (real4<R>*)c->vaA, (real4<R>*)c->adA, c->oidA};
}
template <class R> StateB<R> stateB(rrsph_ctx* c, void* pred, void* pres) {
    return StateB<R>{c->posB, (real4<R>*)c->vaB, (real4<R>*)pred, c->oidB, (R*)pres, nullptr};
}
// the step's B copy; keeps A.ad too when the fused pass may have to be undone
template <class R> StateB<R> stepB(rrsph_ctx* c) {
    return StateB<R>{c->posB, (real4<R>*)c->vaB, (real4<R>*)c->pred, c->oidB, (R*)c->pres, (real4<R>*)c->adB};
}
// 3D: the position buffers hold double4 (two double2 slots per particle)
template <class R> StateA3<R> stateA3(rrsph_ctx* c) {
    return StateA3<R>{(double4*)c->posA, (real4<R>*)c->vaA, (real4<R>*)c->adA, c->oidA};
}
template <class R> StateB3<R> stateB3(rrsph_ctx* c, void* pred, void* pw, void* pres) {
    return StateB3<R>{(double4*)c->posB, (real4<R>*)c->vaB, (real4<R>*)pred, (R*)pw, (R*)pres, c->oidB, nullptr};
}
template <class R> StateB3<R> stepB3(rrsph_ctx* c) {
    return StateB3<R>{(double4*)c->posB, (real4<R>*)c->vaB, (real4<R>*)c->pred, (R*)c->pw, (R*)c->pres, c->oidB,
                      (real4<R>*)c->adB};
}

uint32_t grid_blocks(uint32_t n, uint32_t bs) { return (n + bs - 1) / bs; }

// After two incremental re-sort overflows within kFallbackWindow steps of
// each other, the next kFullSortRun steps use the full radix sort; an
// overflow soon after such a run ends doubles the run (up to kFullSortRunMax).
constexpr uint64_t kFullSortRun = 64;
constexpr uint64_t kFullSortRunMax = 4096;
constexpr uint64_t kFallbackWindow = 8;

// Launch helper that optionally records HIP events around the launch.
struct Launcher {
    rrsph_ctx* c;
    Profiler* prof;
    template <class F> void operator()(uint32_t cls, F&& launch) {
        if (prof) {
            hipEvent_t a, b;
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, c->stream);
            launch();
            (void)hipEventRecord(b, c->stream);
            prof->rec.push_back({cls, {a, b}});
        } else {
            launch();
        }
    }
};

// Sort key_next (values = index) into kbuf/vbuf; returns final buffers.
void enqueue_sort(rrsph_ctx* c, Ctrl* ctrl, Launcher& L) {
    const uint32_t n = c->P.n;
    const uint32_t* src_k = c->key_next;
    const uint32_t* src_v = nullptr;
    for (uint32_t p = 0; p < c->passes; ++p) {
        uint32_t* dk = c->kbuf[p & 1];
        uint32_t* dv = c->vbuf[p & 1];
        const uint32_t shift = p * kRadixBits;
        const uint32_t nt = c->ntiles;
        L(K_UPSWEEP, [&] { radix_upsweep<<<nt, kBlock, 0, c->stream>>>(src_k, n, shift, c->hist, nt, ctrl); });
        L(K_SCAN, [&] { radix_scan_digits<<<kRadix, kScanBlock, 0, c->stream>>>(c->hist, nt, c->digit_total, ctrl); });
        L(K_DOWNSWEEP, [&] {
            radix_downsweep<<<nt, kBlock, 0, c->stream>>>(src_k, src_v, dk, dv, n, shift, c->hist, nt, c->digit_total,
                                                          ctrl);
        });
        src_k = dk;
        src_v = dv;
    }
    c->sorted_keys = src_k;
    c->sorted_perm = src_v;
}

// Incremental re-sort of key_next against skeyA into kbuf[0] / vbuf[0]; its
// first kernel also runs the step prologue.
void enqueue_resort(rrsph_ctx* c, Launcher& L) {
    const uint32_t n = c->P.n;
    const uint32_t nb = grid_blocks(n, kBlock);
    L(K_RESORT_SORT, [&] {
        resort_sort_small<<<1, kResortBlock, 0, c->stream>>>(c->P, c->key_next, c->resort_cap, c->chg, c->wpre,
                                                             c->ctrl);
    });
    L(K_RESORT_SCATTER, [&] {
        resort_scatter<<<nb, kBlock, 0, c->stream>>>(c->skeyA, c->chg, c->chg_mask, c->wpre, n, c->kbuf[0],
                                                     c->vbuf[0], c->ctrl, c->resort_stage);
    });
    c->sorted_keys = c->kbuf[0];
    c->sorted_perm = c->vbuf[0];
}

template <class R> void enqueue_step_t(rrsph_ctx* c, Profiler* prof, bool full_sort) {
    Launcher L{c, prof};
    const StepParams& P = c->P;
    const uint32_t nb = grid_blocks(P.n, kBlock);
    const uint32_t nb_xcd = P.blocks_per_xcd * kNumXcd;
    if (full_sort) {
        L(K_PROLOGUE, [&] { step_prologue<<<1, kWave, 0, c->stream>>>(P, c->ctrl); });
        enqueue_sort(c, c->ctrl, L);
    } else {
        enqueue_resort(c, L);
    }
    StateA<R> A = stateA<R>(c);
    StateB<R> B = stepB<R>(c);
    L(K_REORDER, [&] {
        reorder_kernel<R><<<nb, kBlock, 0, c->stream>>>(P, c->ctrl, c->sorted_keys, c->sorted_perm, A, B, c->LB,
                                                        c->maxslots, c->skeyA);
    });
    if (P.density_treatment == 0) {
        L(K_DENSITY_SUM, [&] {
            density_sum_kernel<R><<<nb_xcd, kBlock, 0, c->stream>>>(P, c->ctrl, B, c->sorted_keys, c->LB);
        });
    }
    real4<R>* der = (real4<R>*)c->der;
    real4<R>* xs = (real4<R>*)c->xs;
    const uint32_t nt_xcd = P.tile_blocks_per_xcd * kNumXcd;
    const int impl = c->fast_pair ? c->pair_impl : 0;
    auto pair = [&](auto kernel) {
        const uint32_t grid = impl == 2 ? nt_xcd : nb_xcd;
        const uint32_t block = impl == 2 ? kTile * kTilesPerBlock : kBlock;
        L(K_PAIR, [&] {
            kernel<<<grid, block, 0, c->stream>>>(P, c->ctrl, B, c->sorted_keys, c->LB, A, c->key_next, der, xs,
                                                  c->maxslots);
        });
    };
    auto choose = [&](auto general, auto fast, auto tile, auto walk) {
        if (impl == 3) pair(walk);
        else if (impl == 2) pair(tile);
        else if (impl == 1) pair(fast);
        else pair(general);
    };
    if (c->fused) {
        if (c->export_fields)
            choose(pair_kernel<R, true, true>, pair_fast_kernel<R, true, true>, pair_tile_kernel<R, true, true>,
                   pair_fast_kernel<R, true, true, false>);
        else
            choose(pair_kernel<R, true, false>, pair_fast_kernel<R, true, false>, pair_tile_kernel<R, true, false>,
                   pair_fast_kernel<R, true, false, false>);
    } else {
        choose(pair_kernel<R, false, true>, pair_fast_kernel<R, false, true>, pair_tile_kernel<R, false, true>,
               pair_fast_kernel<R, false, true, false>);
        L(K_INTEGRATE, [&] {
            integrate_kernel<R><<<nb, kBlock, 0, c->stream>>>(P, c->ctrl, B, der, xs, A, c->key_next, c->maxslots,
                                                              c->sorted_keys);
        });
    }
}

template <class R> void enqueue_step3_t(rrsph_ctx* c, Profiler* prof, bool full_sort) {
    Launcher L{c, prof};
    const StepParams& P = c->P;
    const uint32_t nb = grid_blocks(P.n, kBlock);
    const uint32_t nb_xcd = P.blocks_per_xcd * kNumXcd;
    if (full_sort) {
        L(K_PROLOGUE, [&] { step_prologue<<<1, kWave, 0, c->stream>>>(P, c->ctrl); });
        enqueue_sort(c, c->ctrl, L);
    } else {
        enqueue_resort(c, L);
    }
    StateA3<R> A = stateA3<R>(c);
    StateB3<R> B = stepB3<R>(c);
    L(K_REORDER, [&] {
        reorder3_kernel<R><<<nb, kBlock, 0, c->stream>>>(P, c->ctrl, c->sorted_keys, c->sorted_perm, A, B, c->LB,
                                                         c->maxslots, c->skeyA);
    });
    if (P.density_treatment == 0) {
        L(K_DENSITY_SUM, [&] {
            density_sum3_kernel<R><<<nb_xcd, kBlock, 0, c->stream>>>(P, c->ctrl, B, c->sorted_keys, c->LB);
        });
    }
    real4<R>* der = (real4<R>*)c->der;
    real4<R>* xs = (real4<R>*)c->xs;
    auto pair = [&](auto kernel) {
        L(K_PAIR, [&] {
            kernel<<<nb_xcd, kBlock, 0, c->stream>>>(P, c->ctrl, B, c->sorted_keys, c->LB, A, c->key_next, der, xs,
                                                     c->maxslots);
        });
    };
    if (c->fused) {
        if (c->export_fields) pair(pair3_kernel<R, true, true>);
        else pair(pair3_kernel<R, true, false>);
    } else {
        pair(pair3_kernel<R, false, true>);
        L(K_INTEGRATE, [&] {
            integrate3_kernel<R><<<nb, kBlock, 0, c->stream>>>(P, c->ctrl, B, der, xs, A, c->key_next, c->maxslots,
                                                               c->sorted_keys);
        });
    }
}

// One neighbour-list step (build: sort, reorder and new lists first).
template <class R> void enqueue_list_step_t(rrsph_ctx* c, Profiler* prof, bool build) {
    Launcher L{c, prof};
    const StepParams& P = c->P;
    const uint32_t nb = grid_blocks(P.n, kBlock);
    const uint32_t nb_xcd = P.blocks_per_xcd * kNumXcd;
    L(K_PROLOGUE, [&] {
        list_prologue<<<1, kBlock, 0, c->stream>>>(P, c->ctrl, c->dslots, c->maxslots, build ? 1u : 0u);
    });
    ListState<R> X{c->posA, (const real4<R>*)c->vaA, c->oidA};
    ListRec<R>* rec = (ListRec<R>*)c->rec;
    if (build) {
        enqueue_sort(c, c->ctrl, L);
        StateA<R> A = stateA<R>(c);
        StateB<R> B = stateB<R>(c, c->pred, c->pres);
        L(K_REORDER, [&] {
            reorder_list_kernel<R><<<nb, kBlock, 0, c->stream>>>(P, c->ctrl, c->sorted_keys, c->sorted_perm, A, B,
                                                                 rec, c->home, c->LB, c->maxslots);
        });
        L(K_LIST_BUILD, [&] {
            list_build_kernel<R><<<nb_xcd, kBlock, 0, c->stream>>>(P, c->ctrl, c->posB, (const real4<R>*)c->vaB,
                                                                   c->sorted_keys, c->LB, c->nbl, c->nbc);
        });
        X = ListState<R>{c->posB, (const real4<R>*)c->vaB, c->oidB};
    }
    if (P.density_treatment == 0) {
        L(K_DENSITY_SUM, [&] {
            density_sum_list<R><<<nb_xcd, kBlock, 0, c->stream>>>(P, c->ctrl, X, rec, (R*)c->pres, c->nbl, c->nbc);
        });
    }
    L(K_PAIR, [&] {
        pair_list_kernel<R><<<nb_xcd, kBlock, 0, c->stream>>>(P, c->ctrl, X, rec, (R*)c->pres, c->nbl, c->nbc,
                                                              (real4<R>*)c->der, (real2<R>*)c->xs, c->maxslots);
    });
    L(K_INTEGRATE, [&] {
        integrate_list_kernel<R><<<nb, kBlock, 0, c->stream>>>(P, c->ctrl, X, (const real4<R>*)c->der,
                                                               (const real2<R>*)c->xs, rec, c->home, stateA<R>(c),
                                                               c->key_next, c->maxslots, c->dslots);
    });
}

// One 3D neighbour-list step (build: sort, reorder and new lists first).
template <class R> void enqueue_list_step3_t(rrsph_ctx* c, Profiler* prof, bool build) {
    Launcher L{c, prof};
    const StepParams& P = c->P;
    const uint32_t nb = grid_blocks(P.n, kBlock);
    const uint32_t nb_xcd = P.blocks_per_xcd * kNumXcd;
    L(K_PROLOGUE, [&] {
        list_prologue<<<1, kBlock, 0, c->stream>>>(P, c->ctrl, c->dslots, c->maxslots, build ? 1u : 0u);
    });
    ListState3<R> X{(const double4*)c->posA, (const real4<R>*)c->vaA, c->oidA};
    ListRec<R>* rec = (ListRec<R>*)c->rec;
    real4<R>* dens = (real4<R>*)c->dens;
    if (build) {
        enqueue_sort(c, c->ctrl, L);
        StateA3<R> A = stateA3<R>(c);
        StateB3<R> B = stateB3<R>(c, c->pred, c->pw, c->pres);
        L(K_REORDER, [&] {
            reorder_list3_kernel<R><<<nb, kBlock, 0, c->stream>>>(P, c->ctrl, c->sorted_keys, c->sorted_perm, A, B,
                                                                  rec, dens, c->home, c->cand, c->LB, c->maxslots);
        });
        L(K_LIST_BUILD, [&] {
            list_build3_kernel<<<nb_xcd, kBlock, 0, c->stream>>>(P, c->ctrl, (const double4*)c->posB, c->cand,
                                                                 c->sorted_keys, c->LB, c->nbl, c->nbc);
        });
        X = ListState3<R>{(const double4*)c->posB, (const real4<R>*)c->vaB, c->oidB};
    }
    if (P.density_treatment == 0) {
        L(K_DENSITY_SUM, [&] {
            density_sum_list3<R><<<nb_xcd, kBlock, 0, c->stream>>>(P, c->ctrl, X, dens, rec, (R*)c->pres, c->nbl,
                                                                   c->nbc);
        });
    }
    L(K_PAIR, [&] {
        pair_list3_kernel<R><<<nb_xcd, kBlock, 0, c->stream>>>(P, c->ctrl, X, rec, (R*)c->pres, c->nbl, c->nbc,
                                                               (real4<R>*)c->der, (real4<R>*)c->xs, c->maxslots);
    });
    L(K_INTEGRATE, [&] {
        integrate_list3_kernel<R><<<nb, kBlock, 0, c->stream>>>(P, c->ctrl, X, (const real4<R>*)c->der,
                                                                (const real4<R>*)c->xs, rec, dens, c->home,
                                                                stateA3<R>(c), c->key_next, c->maxslots, c->dslots);
    });
}

void enqueue_list_step(rrsph_ctx* c, Profiler* prof, bool build) {
    if (c->P.dim == 3) {
        if (c->fp32) enqueue_list_step3_t<float>(c, prof, build);
        else enqueue_list_step3_t<double>(c, prof, build);
        return;
    }
    if (c->fp32) enqueue_list_step_t<float>(c, prof, build);
    else enqueue_list_step_t<double>(c, prof, build);
}

bool list_active(const rrsph_ctx* c) { return c->list_mode && !c->list_suspended; }

void enqueue_step(rrsph_ctx* c, Profiler* prof, bool full_sort) {
    if (c->P.dim == 3) {
        if (c->fp32) enqueue_step3_t<float>(c, prof, full_sort);
        else enqueue_step3_t<double>(c, prof, full_sort);
        return;
    }
    if (c->fp32) enqueue_step_t<float>(c, prof, full_sort);
    else enqueue_step_t<double>(c, prof, full_sort);
}

void drop_graph(rrsph_ctx* c) {
    hipGraphExec_t* ex[] = {&c->gexec, &c->gexec_full, &c->gexec_list, &c->gexec_build};
    hipGraph_t* gr[] = {&c->graph, &c->graph_full, &c->graph_list, &c->graph_build};
    for (auto* e : ex) {
        if (*e) (void)hipGraphExecDestroy(*e);
        *e = nullptr;
    }
    for (auto* g : gr) {
        if (*g) (void)hipGraphDestroy(*g);
        *g = nullptr;
    }
}

// Capture one step with the incremental re-sort and one with the full sort.
rrsph_status build_graph(rrsph_ctx* ctx) {
    drop_graph(ctx);
    HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(ctx, nullptr, false);
    HIP_TRY(hipStreamEndCapture(ctx->stream, &ctx->graph));
    HIP_TRY(hipGraphInstantiate(&ctx->gexec, ctx->graph, nullptr, nullptr, 0));
    HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(ctx, nullptr, true);
    HIP_TRY(hipStreamEndCapture(ctx->stream, &ctx->graph_full));
    HIP_TRY(hipGraphInstantiate(&ctx->gexec_full, ctx->graph_full, nullptr, nullptr, 0));
    if (ctx->list_mode) {
        HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        enqueue_list_step(ctx, nullptr, false);
        HIP_TRY(hipStreamEndCapture(ctx->stream, &ctx->graph_list));
        HIP_TRY(hipGraphInstantiate(&ctx->gexec_list, ctx->graph_list, nullptr, nullptr, 0));
        HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        enqueue_list_step(ctx, nullptr, true);
        HIP_TRY(hipStreamEndCapture(ctx->stream, &ctx->graph_build));
        HIP_TRY(hipGraphInstantiate(&ctx->gexec_build, ctx->graph_build, nullptr, nullptr, 0));
    }
    return RRSPH_OK;
}

// Sort + reorder of the current state into B (with a scratch pred) for the
// utility queries, using a control block that is never halted.
// The arrays of a state saved in an adjoint blob (pos | va | ad | oid | ...),
// read in place by the reverse steps.
struct StateView {
    void *pos, *va, *ad;
    uint32_t* oid;
};
template <class R> void enqueue_util_reorder(rrsph_ctx* c, const StateView* src = nullptr, Ctrl* uc = nullptr);
template <class R> void enqueue_util_sort(rrsph_ctx* c, Ctrl* uc = nullptr) {
    Launcher L{c, nullptr};
    enqueue_sort(c, uc ? uc : c->util_ctrl, L);
    enqueue_util_reorder<R>(c, nullptr, uc);
}

// Reorder of the current state (or of the saved state src) by sorted_keys /
// sorted_perm into B (scratch pred) and the cell-start table, for the
// utility queries; with the control block uc (its cur_dt predicts) instead
// of the context's utility one.
template <class R> void enqueue_util_reorder(rrsph_ctx* c, const StateView* src, Ctrl* uc) {
    const uint32_t nb = grid_blocks(c->P.n, kBlock);
    const Ctrl* ctrl = uc ? uc : c->util_ctrl;
    if (c->P.dim == 3) {
        StateA3<R> A = stateA3<R>(c);
        if (src) A = StateA3<R>{(double4*)src->pos, (real4<R>*)src->va, (real4<R>*)src->ad, src->oid};
        reorder3_kernel<R><<<nb, kBlock, 0, c->stream>>>(c->P, ctrl, c->sorted_keys, c->sorted_perm, A,
                                                         stateB3<R>(c, c->pred_scratch, c->pw_scratch, c->pres_scratch),
                                                         c->LB, c->maxslots + 2 * kMaxSlots, nullptr);
        return;
    }
    StateA<R> A = stateA<R>(c);
    if (src) A = StateA<R>{(double2*)src->pos, (real4<R>*)src->va, (real4<R>*)src->ad, src->oid};
    StateB<R> B = stateB<R>(c, c->pred_scratch, c->pres_scratch);
    reorder_kernel<R><<<nb, kBlock, 0, c->stream>>>(c->P, ctrl, c->sorted_keys, c->sorted_perm, A, B, c->LB,
                                                    c->maxslots + 2 * kMaxSlots, nullptr);
}

// Up to kCopyRanges device-to-device copies in one launch (the saved-state
// blobs of the adjoint move five arrays at a time; separate copies cost a
// launch each, which dominates at tens of thousands of particles).  Lengths
// are in 4-byte words; a range moves 16 bytes per access when both ends are
// 16-byte aligned.
constexpr int kCopyRanges = 8;
struct CopyRange {
    const uint32_t* src;
    uint32_t* dst;
    size_t words;
};
struct CopyPlan {
    CopyRange r[kCopyRanges];
    int n = 0;
    void add(const void* src, void* dst, size_t bytes) { r[n++] = {(const uint32_t*)src, (uint32_t*)dst, bytes / 4}; }
};
__global__ void __launch_bounds__(kBlock) copy_ranges_kernel(CopyPlan plan) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
    for (int k = 0; k < plan.n; ++k) {
        const CopyRange c = plan.r[k];
        size_t done = 0;
        if (((reinterpret_cast<uintptr_t>(c.src) | reinterpret_cast<uintptr_t>(c.dst)) & 15u) == 0) {
            const size_t n4 = c.words / 4;
            const uint4* __restrict__ s = (const uint4*)c.src;
            uint4* __restrict__ d = (uint4*)c.dst;
            for (size_t i = t; i < n4; i += stride) d[i] = s[i];
            done = n4 * 4;
        }
        for (size_t i = done + t; i < c.words; i += stride) c.dst[i] = c.src[i];
    }
}
rrsph_status enqueue_copies(rrsph_ctx* ctx, const CopyPlan& plan) {
    size_t units = 0;
    for (int k = 0; k < plan.n; ++k) units = std::max(units, plan.r[k].words / 4 + 1);
    const uint32_t blocks = (uint32_t)std::min<size_t>(2048, (units + kBlock - 1) / kBlock);
    copy_ranges_kernel<<<std::max(blocks, 1u), kBlock, 0, ctx->stream>>>(plan);
    HIP_TRY(hipGetLastError());
    return RRSPH_OK;
}

std::string status_message(uint32_t st) {
    switch (st) {
    case RRSPH_ERR_TOO_MANY_NEIGHBOURS: return "making neighbours grid: error occured";
    case RRSPH_ERR_NEVER_ENDING: return "never ending simulation";
    case RRSPH_ERR_OUTSIDE_BOUNDARIES: return "Encounter particle outside boundaries";
    default: return "";
    }
}

void enqueue_state_keys(rrsph_ctx* ctx);

// A fused step whose neighbour lists overflowed under the consistency check
// was integrated in its pair pass; the prologue after it refused to commit
// it.  Put back the state the step started from: its sorted copy B (pos, va,
// oid) and the copy of A.ad the reorder kept, then the cell keys of that
// state for the next sort.
rrsph_status rollback_fused_step(rrsph_ctx* ctx) {
    const size_t n = ctx->P.n, pb = ctx->P.dim == 3 ? 32 : 16, vb = ctx->fp32 ? 16 : 32;
    if (!ctx->adB) return fail(ctx, RRSPH_ERR_HIP, "RRSPH_HIP: no copy of the step's input to restore");
    CopyPlan plan;
    plan.add(ctx->posB, ctx->posA, pb * n);
    plan.add(ctx->vaB, ctx->vaA, vb * n);
    plan.add(ctx->adB, ctx->adA, vb * n);
    plan.add(ctx->oidB, ctx->oidA, 4 * n);
    if (const rrsph_status st = enqueue_copies(ctx, plan); st != RRSPH_OK) return st;
    enqueue_state_keys(ctx);
    ctx->need_full = true;
    ctx->host_ctrl->rollback = 0;
    HIP_TRY(hipMemcpyAsync(ctx->ctrl, ctx->host_ctrl, sizeof(Ctrl), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RRSPH_OK;
}

// Read the control block after the stream drained and turn it into a status.
rrsph_status sync_status(rrsph_ctx* ctx) {
    HIP_TRY(hipMemcpyAsync(ctx->host_ctrl, ctx->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const Ctrl& c = *ctx->host_ctrl;
    const bool rerun = c.status == kStatusResort || c.status == kStatusRebuild || c.status == kStatusListOverflow;
    if (c.halted && c.status != 0 && !rerun) {
        if (c.rollback) {
            if (const rrsph_status st = rollback_fused_step(ctx); st != RRSPH_OK) return st;
        }
        return fail(ctx, (rrsph_status)c.status, status_message(c.status));
    }
    if (c.pending) {
        if (ctx->P.consistency_check && c.nbr_overflow) {
            // the last launched step was a fused one that overflowed: no
            // prologue ran after it, so do what it would have done (leave the
            // step uncommitted, put its input back, stay halted)
            Ctrl& h = *ctx->host_ctrl;
            h.pending = 0;
            h.halted = 1;
            h.status = RRSPH_ERR_TOO_MANY_NEIGHBOURS;
            if (h.wall_disappear) h.nwm_done = 0;
            if (const rrsph_status st = rollback_fused_step(ctx); st != RRSPH_OK) return st;
            return fail(ctx, RRSPH_ERR_TOO_MANY_NEIGHBOURS, status_message(RRSPH_ERR_TOO_MANY_NEIGHBOURS));
        }
        if (ctx->P.consistency_check && c.outside > 0 && ctx->P.consistency_treatment == 1)
            return fail(ctx, RRSPH_ERR_OUTSIDE_BOUNDARIES, status_message(RRSPH_ERR_OUTSIDE_BOUNDARIES));
    }
    return RRSPH_OK;
}

rrsph_status setup_params(rrsph_ctx* ctx, const rrsph_params_c& p) {
    StepParams& P = ctx->P;
    P = StepParams{};
    if (p.dim != 2 && p.dim != 3) return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: dim must be 2 or 3");
    const bool d3 = p.dim == 3;
    if (p.ntotal == 0 || p.max_neighbours < 2 || !(p.hsml > 0) || !(p.cell_scale_k > 0))
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: invalid sizes");
    if (p.nfluid > p.ntotal) return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: nfluid > ntotal");
    P.n = p.ntotal;
    P.nfluid = p.nfluid;
    P.max_nbr = p.max_neighbours;
    P.dim = p.dim;
    P.cell = p.cell_scale_k * p.hsml;
    // Input.cpp:45-70: at most 2^16 cells per axis in 2D, 2^10 in 3D
    const double tx = (p.x_maxgeom - p.x_mingeom) / P.cell, ty = (p.y_maxgeom - p.y_mingeom) / P.cell;
    const double tz = d3 ? (p.z_maxgeom - p.z_mingeom) / P.cell : 0.0;
    const double max_id = d3 ? 1024.0 : 65536.0;
    if (!(tx >= 0) || !(ty >= 0) || !(tz >= 0) || tx > max_id || ty > max_id || (d3 && tz >= max_id))
        return fail(ctx, RRSPH_ERR_TOO_MANY_CELLS, "can't make grid with so many cells");
    P.ncx = (int32_t)(uint64_t)tx + 1;
    P.ncy = (int32_t)(uint64_t)ty + 1;
    P.ncz = d3 ? (int32_t)(uint64_t)tz + 1 : 1;
    const uint64_t ncells = (uint64_t)P.ncx * (uint64_t)P.ncy * (uint64_t)P.ncz;
    if (ncells + 2 >= (1ull << 31)) return fail(ctx, RRSPH_ERR_TOO_MANY_CELLS, "can't make grid with so many cells");
    P.ncells = (uint32_t)ncells;
    P.xmin = p.x_mingeom;
    P.ymin = p.y_mingeom;
    P.zmin = p.z_mingeom;
    P.xmax = p.x_maxgeom;
    P.ymax = p.y_maxgeom;
    P.zmax = p.z_maxgeom;
    P.cut2 = P.cell * P.cell;
    P.h = p.hsml;
    P.mass = p.mass;
    P.rho0 = p.rho0;
    P.g = p.g;
    const double h = p.hsml, pi = p.pi;
    // SmoothingKernel.h:22-23, 93-94, 136-137, 194-195
    P.f_cubic = d3 ? double(3) / (double(2) * pi * (h * h * h)) : double(15) / (double(7) * pi * (h * h));
    P.f_gauss = d3 ? 1 / (pi * std::sqrt(pi) * (h * h * h)) : 1 / (pi * (h * h));
    P.f_wend = d3 ? 21 / (16 * pi * (h * h * h)) : 7 / (4 * pi * (h * h));
    P.f_desb = d3 ? 15 / (pi * (4 * h) * (4 * h) * (4 * h)) : 5 / (16 * pi * (h * h));
    P.skf_density = p.density_skf;
    P.skf_intf = p.intf_skf;
    P.skf_av = p.artificial_viscosity_skf;
    P.skf_xsph = p.average_velocity_skf;
    P.skf_ap = p.artificial_pressure_skf;
    P.eos_B = p.eos_c_sqr * p.rho0 / 7;
    P.density_treatment = p.density_treatment;
    if (p.density_normalization && p.density_treatment != 0)
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: density_normalization needs density_treatment 0");
    P.renorm = p.density_normalization ? 1 : 0;
    P.renorm_every = p.density_normalization_every ? p.density_normalization_every : 1u;
    P.dhc = p.density_delta_sph_coef * p.hsml * p.eos_c;
    P.intf_approx = p.intf_sph_approximation;
    P.art_pressure = p.artificial_pressure;
    P.ap_coef = p.artificial_pressure_coef;
    P.ap_index = p.artificial_pressure_index;
    P.ap_wdelta = host_w(p, p.delta, p.artificial_pressure_skf);
    P.av_on = p.artificial_viscosity;
    P.av_alpha = p.artificial_shear_visc;
    P.av_beta = p.artificial_bulk_visc;
    P.av_c = p.eos_sound_vel;
    const double etq = (double)0.1f;  // ArtificialViscosity.h:22 (a float literal)
    P.av_eta2 = (p.hsml * etq) * (p.hsml * etq);
    P.xsph_on = p.average_velocity;
    P.xsph_eps = p.average_velocity_coef;
    P.lj_on = p.boundary_treatment == 1;
    P.lj_r0 = 2 * p.hsml;
    P.lj_D = 5 * p.g * p.depth;
    P.dt_method = p.dt_correction_method;
    P.dt_dynamic_mode = p.dt_dynamic_mode;
    P.dt0 = p.dt;
    P.cfl = p.CFL_coef;
    P.c0 = p.eos_c;
    P.never_ending = (p.hsml / p.eos_c) * 1.E-6;
    P.consistency_check = p.consistency_check;
    P.consistency_treatment = p.consistency_treatment;
    P.boundary_motion = p.boundary_motion;
    P.nwm = p.nwm;
    P.nwm_axis = -1;
    if (p.nwm == 2 || p.nwm == 3 || p.nwm == 6) {
        if (p.nwm_direction == 1) P.nwm_axis = 0;
        else if (p.nwm_direction == 2) P.nwm_axis = 1;
        else if (p.nwm_direction == 3 && d3) P.nwm_axis = 2;
        else if (p.nwm_direction == 4 && d3 && p.boundary_motion)
            return fail(ctx, RRSPH_ERR_NOT_IMPLEMENTED, "RRSPH_HIP: radial wave makers are not implemented");
    } else if (p.nwm == 5) {
        P.nwm_axis = 0;
    }
    bool range_ok = !(p.nwm_particles_start < p.nfluid || p.nwm_particles_end < p.nfluid ||
                      p.nwm_particles_end < p.nwm_particles_start || p.nwm_particles_end > p.ntotal);
    if (!range_ok) P.nwm = 0;  // CLTimeIntegration.cpp:375-385: wrong range -> no wave maker
    P.nwm_start = p.nwm_particles_start;
    P.nwm_end = p.nwm_particles_end;
    P.nwm_time_start = p.nwm_time_start;
    P.nwm_freq = p.nwm_freq;
    P.nwm_delta = p.nwm_phase - p.nwm_freq * p.nwm_time_start;
    {
        const double kd = p.nwm_wave_number * p.depth, H = p.nwm_wave_magnitude;
        const double sh = std::sinh(kd), ch = std::cosh(kd);
        const double m1 = 2 * sh * sh / (sh * ch + kd);
        P.nwm_S0 = m1 != 0 ? H / m1 : 0;
        P.nwm_e2 = (m1 != 0 && sh != 0) ? (H * H / (32 * p.depth)) * (3 * ch / (sh * sh * sh) - 2 / m1) : 0;
        const double Hs = p.nwm_wave_magnitude, hd = p.depth;
        P.sol_H = Hs;
        P.sol_h = hd;
        P.sol_k = std::sqrt(3 * Hs / (4 * hd * hd * (Hs + hd)));
        P.sol_c = std::sqrt(p.g * (Hs + hd));
        P.sol_Tf = 2 / (P.sol_k * P.sol_c) * (3.8 + Hs / hd);
    }
    const int32_t segs = p.piston_segments > 1 ? p.piston_segments : 1;
    if (p.piston_nharm < 0 || segs > kMaxSegments || (int64_t)p.piston_nharm * segs > 32)
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT,
                    "RRSPH_HIP: piston_nharm * piston_segments must be at most 32, piston_segments at most 8");
    if (segs > 1 && (!d3 || p.nwm != 6 || !(p.piston_seg_zmax > p.piston_seg_zmin)))
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT,
                    "RRSPH_HIP: a segmented piston needs dim 3, nwm 6 and piston_seg_zmax > piston_seg_zmin");
    P.piston_nharm = p.piston_nharm;
    P.piston_segments = segs;
    P.piston_seg_z0 = segs > 1 ? p.piston_seg_zmin : 0.0;
    P.piston_seg_inv_w = segs > 1 ? segs / (p.piston_seg_zmax - p.piston_seg_zmin) : 0.0;
    P.piston_omega = p.piston_omega;
    for (int k = 0; k < 32; ++k) {
        P.piston_a[k] = p.piston_a[k];
        P.piston_b[k] = p.piston_b[k];
    }
    P.nblocks = grid_blocks(P.n, kBlock);
    P.blocks_per_xcd = (P.nblocks + kNumXcd - 1) / kNumXcd;
    P.ntiles = grid_blocks(P.n, kTile);
    P.tile_blocks_per_xcd = (grid_blocks(P.ntiles, kTilesPerBlock) + kNumXcd - 1) / kNumXcd;
    // radix passes over the key bits (dead key = ncells must fit)
    uint32_t bits = 0;
    while ((1ull << bits) <= (uint64_t)P.ncells) ++bits;
    ctx->passes = (bits + kRadixBits - 1) / kRadixBits;
    if (ctx->passes == 0) ctx->passes = 1;
    ctx->ntiles = grid_blocks(P.n, kSortTile);
    ctx->fp32 = p.precision == RRSPH_FP32;
    if (const char* e = std::getenv("RRSPH_HIP_SORT")) ctx->incremental = std::strcmp(e, "full") != 0;
    if (const char* e = std::getenv("RRSPH_HIP_RESORT_CAP")) {
        const long v = std::strtol(e, nullptr, 10);
        if (v >= 0 && v < (long)kResortCap) ctx->resort_cap = (uint32_t)v;
    }
    if (const char* e = std::getenv("RRSPH_HIP_RESORT_STAGE")) {  // test hook: force the binary-search placement
        const long v = std::strtol(e, nullptr, 10);
        if (v >= 0 && v <= (long)kBlock) ctx->resort_stage = (uint32_t)v;
    }
    ctx->fused = !(p.dt_correction_method == 2 && p.dt_dynamic_mode == 1);
    if (const char* e = std::getenv("RRSPH_HIP_PAIR_KERNEL")) {
        if (!std::strcmp(e, "general")) ctx->pair_impl = 0;
        else if (!std::strcmp(e, "tile")) ctx->pair_impl = 2;
        else if (!std::strcmp(e, "walk")) ctx->pair_impl = 3;
        else ctx->pair_impl = 1;
    }
    ctx->fast_pair = p.density_skf == 1 && p.intf_skf == 1 && p.artificial_viscosity_skf == 1 &&
                     p.average_velocity_skf == 1 && p.intf_sph_approximation == 2 && !p.artificial_pressure &&
                     (p.density_treatment == 0 || p.density_treatment == 1) && std::fabs(p.cell_scale_k - 2.0) == 0.0;
    // neighbour-list steps: the cubic configuration, 2D and 3D; skin as a
    // fraction of the cut-off (an entry holds a slot in kSlotBits bits in
    // 2D, kSlotBits3 in 3D)
    double skin_frac = 0.125;
    if (const char* e = std::getenv("RRSPH_HIP_LIST_SKIN")) skin_frac = std::strtod(e, nullptr);
    ctx->list_mode = ctx->fast_pair && skin_frac > 0.0 && skin_frac < 0.9 && P.n <= (d3 ? kSlotMask3 : kSlotMask);
    if (const char* e = std::getenv("RRSPH_HIP_LIST")) ctx->list_mode = ctx->list_mode && std::strcmp(e, "0") != 0;
    P.list_reach = P.cell * (1.0 + skin_frac);
    P.list_r2 = P.list_reach * P.list_reach * (1.0 + 1e-12);
    P.skin_half = 0.5 * P.cell * skin_frac * (1.0 - 1e-9);
    const double grow = std::pow(1.0 + skin_frac, d3 ? 3 : 2);
    P.list_cap = (uint32_t)std::ceil(p.max_neighbours * grow) + 8;
    // the list pass addresses the lists (list_cap + kListPad rows) with
    // 32-bit byte offsets: fewer rows where n is large (a build that finds
    // more entries for a particle runs its step without lists)
    const uint64_t rows_max = 0xffffffffull / (4ull * P.n);
    if (P.list_cap + kListPad > rows_max) {
        if (rows_max >= (uint64_t)p.max_neighbours + kListPad) P.list_cap = (uint32_t)(rows_max - kListPad);
        else ctx->list_mode = false;
    }
    if (d3 && !ctx->fast_pair)
        return fail(ctx, RRSPH_ERR_NOT_IMPLEMENTED,
                    "RRSPH_HIP: 3D supports cubic kernels for every term, approximation 2, no artificial pressure, "
                    "summation or continuity density");
    return RRSPH_OK;
}

void free_tape(AdjTape& T) {
    if (T.texec) (void)hipGraphExecDestroy(T.texec);
    if (T.tgraph) (void)hipGraphDestroy(T.tgraph);
    for (void* p : T.ck)
        if (p) (void)hipFree(p);
    T.ck.clear();
    for (auto& L : T.seg_lists) {
        void* lp[] = {L.nbl, L.nbc, L.keys};
        for (void* p : lp)
            if (p) (void)hipFree(p);
    }
    T.seg_lists.clear();
    if (T.iota) (void)hipFree(T.iota);
    if (T.nearf) (void)hipFree(T.nearf);
    void* ptrs[] = {T.seg, T.eta, T.zs, T.ws, T.eta_bar, T.partials, T.piston_sum, T.lam[0], T.lam[1], T.direct,
                    T.drho_adj, T.out_adj, T.flag, T.rec, T.uctrl, T.seg_keys, T.seg_perm};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    T = AdjTape{};
}

void free_all(rrsph_ctx* c) {
    drop_graph(c);
    free_tape(c->adj);
    void* ptrs[] = {c->posA, c->posB, c->vaA, c->vaB, c->adA, c->pred, c->der, c->xs, c->pred_scratch, c->pres,
                    c->pres_scratch, c->oidA,
                    c->oidB, c->key_next, c->kbuf[0], c->kbuf[1], c->vbuf[0], c->vbuf[1], c->hist, c->digit_total,
                    c->maxslots, c->LB, c->ctrl,
                    c->util_ctrl, c->red, c->st_r, c->st_v, c->st_rho, c->st_p, c->st_a, c->st_drho, c->st_av,
                    c->st_mu, c->st_itype, c->st_u32, c->skeyA, c->chg, c->chg_mask, c->wpre, c->pw,
                    c->pw_scratch, c->adB, c->nbl, c->nbc, c->rec, c->dens, c->home, c->dslots, c->cand};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->host_ctrl) (void)hipHostFree(c->host_ctrl);
    if (c->host_eta) (void)hipHostFree(c->host_eta);
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
}

// Cell keys of state A for the next step's sort.
void enqueue_state_keys(rrsph_ctx* ctx) {
    const uint32_t nb = grid_blocks(ctx->P.n, kBlock);
    if (ctx->P.dim == 3)
        state_keys3_kernel<<<nb, kBlock, 0, ctx->stream>>>(ctx->P, (const double4*)ctx->posA, ctx->key_next);
    else if (ctx->fp32)
        state_keys_kernel<float><<<nb, kBlock, 0, ctx->stream>>>(ctx->P, ctx->posA, (const float4*)ctx->vaA,
                                                                 ctx->key_next);
    else
        state_keys_kernel<double><<<nb, kBlock, 0, ctx->stream>>>(ctx->P, ctx->posA, (const double4*)ctx->vaA,
                                                                  ctx->key_next);
}

}  // namespace

// The reverse of one step on its re-sorted state (state precision R, adjoint
// arithmetic FP64): whole_step reverse, then the pair gather.
template <class R>
static void enqueue_reverse_step(rrsph_ctx* ctx, const AdjStep& S, Ctrl* uc, const double* lam_in, double* lam_out,
                                 double* psum, const uint32_t* keys, const uint32_t* perm,
                                 const AdjTape::Lists* lists, const StateView* src) {
    AdjTape& T = ctx->adj;
    const StepParams& P = ctx->P;
    const size_t n = P.n;
    const uint32_t nb = grid_blocks(P.n, kBlock);
    const uint32_t nb_xcd = P.blocks_per_xcd * kNumXcd;  // the pair gathers map blocks to XCDs
    if (keys) {  // the sort step k ran with, saved by the replay: reorder only
        ctx->sorted_keys = keys;
        ctx->sorted_perm = perm;
        enqueue_util_reorder<R>(ctx, src, uc);
    } else {
        enqueue_util_sort<R>(ctx, uc);
    }
    if (P.dim == 3) {
        StateB3<R> B = stateB3<R>(ctx, ctx->pred_scratch, ctx->pw_scratch, ctx->pres_scratch);
        AdjRec3<R>* rec = (AdjRec3<R>*)T.out_adj;
        // FP32 step over lists: near records relative to the home cells of
        // the build (its sorted keys) for the pair adjoint's walk
        float4* nearf = lists && sizeof(R) == 4 ? T.nearf : nullptr;
        adj_prep3<R><<<nb, kBlock, 0, ctx->stream>>>(P, S, B, lam_in, rec, T.direct, psum,
                                                     nearf ? lists->keys : nullptr, nearf);
        if (lists)
            adj_pair3<R, true><<<nb_xcd, kBlock, 0, ctx->stream>>>(P, S, B, ctx->sorted_keys, ctx->LB, lists->nbl,
                                                                   lists->nbc, rec, T.direct, lam_out, T.flag,
                                                                   nearf);
        else
            adj_pair3<R, false><<<nb_xcd, kBlock, 0, ctx->stream>>>(P, S, B, ctx->sorted_keys, ctx->LB, nullptr,
                                                                    nullptr, rec, T.direct, lam_out, T.flag,
                                                                    nullptr);
    } else {
        StateB<R> B = stateB<R>(ctx, ctx->pred_scratch, ctx->pres_scratch);
        adj_prep<R><<<nb, kBlock, 0, ctx->stream>>>(P, S, B, lam_in, T.out_adj, T.drho_adj, T.direct, psum);
        adj_pair<R><<<nb_xcd, kBlock, 0, ctx->stream>>>(P, S, B, ctx->sorted_keys, ctx->LB, T.out_adj, T.drho_adj,
                                                    T.direct, lam_out, T.flag);
    }
}

extern "C" {

rrsph_status rrsph_hip_create(const rrsph_params_c* params, rrsph_ctx** out) {
    if (!params || !out) return RRSPH_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    rrsph_ctx* ctx = new rrsph_ctx();
    if (params->struct_size != sizeof(rrsph_params_c)) {
        ctx->err = "RRSPH_HIP: rrsph_params_c size mismatch";
        *out = ctx;
        return RRSPH_ERR_INVALID_ARGUMENT;
    }
    ctx->prm = *params;
    *out = ctx;
    // parameters are validated on the host before any device is touched
    rrsph_status st = setup_params(ctx, *params);
    if (st != RRSPH_OK) return st;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(ctx, RRSPH_ERR_NO_DEVICE, "RRSPH_HIP: no HIP device");
    HIP_TRY(hipGetDevice(&ctx->device));
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->own_stream = true;
    const size_t n = ctx->P.n;
    const size_t r4 = ctx->fp32 ? sizeof(float4) : sizeof(double4);
    hipError_t e = hipSuccess;
    const size_t dim = ctx->P.dim;
    ctx->posA = dalloc<double2>(n * (dim == 3 ? 2 : 1), e);  // 3D: one double4 per particle
    ctx->posB = dalloc<double2>(n * (dim == 3 ? 2 : 1), e);
    ctx->vaA = dalloc<char>(n * r4, e);
    ctx->vaB = dalloc<char>(n * r4, e);
    ctx->adA = dalloc<char>(n * r4, e);
    ctx->pred = dalloc<char>(n * r4, e);
    ctx->der = dalloc<char>(n * r4, e);
    ctx->xs = dalloc<char>(n * r4, e);
    ctx->pred_scratch = dalloc<char>(n * r4, e);
    ctx->pres = dalloc<char>(n * (r4 / 4), e);
    ctx->pres_scratch = dalloc<char>(n * (r4 / 4), e);
    if (dim == 3) {
        ctx->pw = dalloc<char>(n * (r4 / 4), e);
        ctx->pw_scratch = dalloc<char>(n * (r4 / 4), e);
    }
    // the fused pair pass integrates in the same launch; with the consistency
    // check a neighbour overflow undoes the step from B and this copy of A.ad
    if (ctx->fused && ctx->P.consistency_check) ctx->adB = dalloc<char>(n * r4, e);
    ctx->oidA = dalloc<uint32_t>(n, e);
    ctx->oidB = dalloc<uint32_t>(n, e);
    ctx->key_next = dalloc<uint32_t>(n, e);
    ctx->kbuf[0] = dalloc<uint32_t>(n, e);
    ctx->kbuf[1] = dalloc<uint32_t>(n, e);
    ctx->vbuf[0] = dalloc<uint32_t>(n, e);
    ctx->vbuf[1] = dalloc<uint32_t>(n, e);
    ctx->hist = dalloc<uint32_t>((size_t)kRadix * ctx->ntiles, e);
    ctx->digit_total = dalloc<uint32_t>(kRadix, e);
    ctx->maxslots = dalloc<unsigned long long>(4 * kMaxSlots, e);  // second half: utility sorts
    ctx->LB = dalloc<uint32_t>((size_t)ctx->P.ncells + 2, e);
    ctx->ctrl = dalloc<Ctrl>(1, e);
    ctx->util_ctrl = dalloc<Ctrl>(1, e);
    ctx->red = dalloc<double>(kDiagSlots * (size_t)kDiagMaxBlocks + 64, e);
    ctx->skeyA = dalloc<uint32_t>(n, e);
    ctx->chg = dalloc<unsigned long long>(kResortCap, e);
    ctx->chg_mask = dalloc<unsigned long long>(n / kWave + 1, e);
    ctx->wpre = dalloc<uint32_t>(n / kWave + 1, e);
    ctx->P.chg_mask = ctx->chg_mask;
    if (ctx->list_mode) {
        ctx->nbl = dalloc<uint32_t>((size_t)(ctx->P.list_cap + kListPad) * n, e);
        ctx->nbc = dalloc<uint32_t>(n, e);
        ctx->rec = dalloc<ListRec<double>>(n, e);
        if (dim == 3) ctx->dens = dalloc<double4>(n, e);
        if (dim == 3) ctx->cand = dalloc<float4>(n, e);
        ctx->home = dalloc<uint32_t>(n, e);
        ctx->dslots = dalloc<unsigned long long>(2 * kMaxSlots, e);
    }
    if (e != hipSuccess) return fail(ctx, RRSPH_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    if (ctx->dslots) HIP_TRY(hipMemset(ctx->dslots, 0, 2 * kMaxSlots * sizeof(unsigned long long)));
    if (ctx->nbl) HIP_TRY(hipMemset(ctx->nbl, 0, (size_t)(ctx->P.list_cap + kListPad) * n * sizeof(uint32_t)));
    HIP_TRY(hipHostMalloc((void**)&ctx->host_ctrl, sizeof(Ctrl), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void**)&ctx->host_eta, kMaxGauges * sizeof(double), hipHostMallocDefault));
    HIP_TRY(scratch_address(ctx->device, &ctx->scratch));
    std::memset(ctx->host_ctrl, 0, sizeof(Ctrl));
    HIP_TRY(hipMemcpy(ctx->util_ctrl, ctx->host_ctrl, sizeof(Ctrl), hipMemcpyHostToDevice));
    return RRSPH_OK;
}

void rrsph_hip_destroy(rrsph_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    delete ctx;
}

const char* rrsph_hip_last_error(const rrsph_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

static rrsph_status ensure_staging(rrsph_ctx* ctx) {
    if (ctx->st_r) return RRSPH_OK;
    const size_t n = ctx->P.n;
    hipError_t e = hipSuccess;
    const size_t dim = ctx->P.dim;
    ctx->st_r = dalloc<double>(dim * n, e);
    ctx->st_v = dalloc<double>(dim * n, e);
    ctx->st_rho = dalloc<double>(n, e);
    ctx->st_p = dalloc<double>(n, e);
    ctx->st_a = dalloc<double>(dim * n, e);
    ctx->st_drho = dalloc<double>(n, e);
    ctx->st_av = dalloc<double>(dim * n, e);
    ctx->st_mu = dalloc<double>(n, e);
    ctx->st_itype = dalloc<int32_t>(n, e);
    ctx->st_u32 = dalloc<uint32_t>(n, e);
    if (e != hipSuccess) return fail(ctx, RRSPH_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    return RRSPH_OK;
}

rrsph_status rrsph_hip_upload(rrsph_ctx* ctx, const double* r, const double* v, const double* rho,
                              const int32_t* itype, uint32_t ntotal) {
    if (!ctx || !ctx->stream) return RRSPH_ERR_INVALID_ARGUMENT;
    if (!r || !v || !rho || !itype || ntotal != ctx->P.n)
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "dump corrupted: dump ntotal doesn't match params.ntotal!");
    rrsph_status st = ensure_staging(ctx);
    if (st != RRSPH_OK) return st;
    const size_t n = ctx->P.n;
    const size_t dim = ctx->P.dim;
    HIP_TRY(hipMemcpyAsync(ctx->st_r, r, dim * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->st_v, v, dim * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->st_rho, rho, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->st_itype, itype, n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    Ctrl c{};
    c.time = ctx->prm.start_time;
    c.dt = ctx->prm.dt;
    c.t_end = INFINITY;
    *ctx->host_ctrl = c;
    HIP_TRY(hipMemcpyAsync(ctx->ctrl, ctx->host_ctrl, sizeof(Ctrl), hipMemcpyHostToDevice, ctx->stream));
    const uint32_t nb = grid_blocks(ctx->P.n, kBlock);
    if (dim == 3) {
        if (ctx->fp32)
            init3_kernel<float><<<nb, kBlock, 0, ctx->stream>>>(ctx->P, ctx->st_r, ctx->st_v, ctx->st_rho,
                                                                ctx->st_itype, stateA3<float>(ctx), ctx->key_next,
                                                                ctx->ctrl);
        else
            init3_kernel<double><<<nb, kBlock, 0, ctx->stream>>>(ctx->P, ctx->st_r, ctx->st_v, ctx->st_rho,
                                                                 ctx->st_itype, stateA3<double>(ctx), ctx->key_next,
                                                                 ctx->ctrl);
    } else if (ctx->fp32) {
        init_kernel<float><<<nb, kBlock, 0, ctx->stream>>>(ctx->P, ctx->st_r, ctx->st_v, ctx->st_rho, ctx->st_itype,
                                                           stateA<float>(ctx), ctx->key_next, ctx->ctrl);
    } else {
        init_kernel<double><<<nb, kBlock, 0, ctx->stream>>>(ctx->P, ctx->st_r, ctx->st_v, ctx->st_rho,
                                                            ctx->st_itype, stateA<double>(ctx), ctx->key_next,
                                                            ctx->ctrl);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->uploaded = true;
    ctx->stepped = false;
    ctx->fields_valid = false;
    ctx->need_full = true;
    // step indices restart: forget the overflow history
    ctx->full_until = ctx->launched;
    ctx->last_fallback = ~0ull;
    ctx->full_run = ctx->full_run_end = 0;
    return RRSPH_OK;
}

// Enqueue nsteps steps: the first after an upload (or a re-sort overflow)
// with the full radix sort, the others with the incremental re-sort.
static uint64_t committed_steps(const Ctrl& c);

// Whether the next list step is a build: after an upload, a restore or a
// run without lists, and every list_interval steps.
static bool next_is_build(const rrsph_ctx* ctx, uint64_t step) {
    return ctx->need_full || step - ctx->build_at >= ctx->list_interval;
}

static rrsph_status launch_steps(rrsph_ctx* ctx, uint32_t nsteps) {
    if (ctx->use_graph && !ctx->gexec) {
        rrsph_status st = build_graph(ctx);
        if (st != RRSPH_OK) return st;
    }
    if (list_active(ctx)) {
        // step indices from the committed count of the last synchronisation
        const uint64_t base = committed_steps(*ctx->host_ctrl);
        for (uint32_t k = 0; k < nsteps; ++k) {
            const bool build = next_is_build(ctx, base + k);
            if (ctx->use_graph) HIP_TRY(hipGraphLaunch(build ? ctx->gexec_build : ctx->gexec_list, ctx->stream));
            else enqueue_list_step(ctx, nullptr, build);
            if (build) {
                ctx->build_at = base + k;
                ++ctx->n_build;
                ++ctx->n_full;
            } else {
                ++ctx->n_list;
            }
            ++ctx->launched;
            ctx->need_full = false;
        }
        HIP_TRY(hipGetLastError());
        if (nsteps) {
            ctx->stepped = true;
            ctx->fields_valid = true;
        }
        return RRSPH_OK;
    }
    for (uint32_t k = 0; k < nsteps; ++k) {
        const bool full = ctx->need_full || !ctx->incremental || ctx->launched < ctx->full_until;
        if (ctx->use_graph) HIP_TRY(hipGraphLaunch(full ? ctx->gexec_full : ctx->gexec, ctx->stream));
        else enqueue_step(ctx, nullptr, full);
        // where this step's sort lands: the last radix pass's buffers, or the re-sort's
        const uint32_t fb = (ctx->passes - 1) & 1u;
        ctx->last_keys = full ? ctx->kbuf[fb] : ctx->kbuf[0];
        ctx->last_perm = full ? ctx->vbuf[fb] : ctx->vbuf[0];
        ++(full ? ctx->n_full : ctx->n_incremental);
        ++ctx->launched;
        ctx->need_full = false;
    }
    HIP_TRY(hipGetLastError());
    if (nsteps) {
        ctx->stepped = true;
        ctx->fields_valid = true;
    }
    return RRSPH_OK;
}

static uint64_t committed_steps(const Ctrl& c) { return c.step + (c.pending ? 1 : 0); }

// One list step with the build decided by the caller (the adjoint's tapes
// follow their own schedule); `abs_step` is its step index since the upload.
// A build's sort lands in kbuf / vbuf of the last radix pass.
static rrsph_status launch_list_step(rrsph_ctx* ctx, bool build, uint64_t abs_step) {
    if (ctx->use_graph && !ctx->gexec) {
        rrsph_status st = build_graph(ctx);
        if (st != RRSPH_OK) return st;
    }
    if (ctx->use_graph) HIP_TRY(hipGraphLaunch(build ? ctx->gexec_build : ctx->gexec_list, ctx->stream));
    else enqueue_list_step(ctx, nullptr, build);
    HIP_TRY(hipGetLastError());
    if (build) {
        const uint32_t fb = (ctx->passes - 1) & 1u;
        ctx->last_keys = ctx->kbuf[fb];
        ctx->last_perm = ctx->vbuf[fb];
        ctx->build_at = abs_step;
        ++ctx->n_build;
        ++ctx->n_full;
    } else {
        ++ctx->n_list;
    }
    ++ctx->launched;
    ctx->need_full = false;
    ctx->stepped = true;
    ctx->fields_valid = true;
    return RRSPH_OK;
}

// A step overflowed the incremental re-sort at step index `at`.  Overflows
// in close succession mean many particles change cell every step: keep to
// the full sort for a while instead of paying a host round trip (and the
// no-op tail of the launched batch) per step, then try again.  A lone
// overflow, e.g. the first step after an upload, costs one round trip.
// Distances are in committed steps, not launches: a halted batch still
// counts its no-op launches.
static void note_resort_overflow(rrsph_ctx* ctx, uint64_t at) {
    ++ctx->n_fallback;
    if (ctx->last_fallback != ~0ull && at - ctx->last_fallback <= kFallbackWindow) {
        const bool again = ctx->full_run && at - ctx->full_run_end <= kFallbackWindow;
        ctx->full_run = again ? std::min(2 * ctx->full_run, kFullSortRunMax) : kFullSortRun;
        ctx->full_until = ctx->launched + ctx->full_run;
        ctx->full_run_end = at + ctx->full_run;
    }
    ctx->last_fallback = at;
    ctx->need_full = true;  // the overflowing step itself runs with the full sort
}

// A list step found its lists possibly stale at step index `at` (the
// displacements since the build reached half the skin): build every
// 0.8 x that many steps from now on, until the measured rate says more.
static void note_list_stale(rrsph_ctx* ctx, uint64_t at) {
    ++ctx->n_list_halts;
    const uint64_t age = at > ctx->build_at ? at - ctx->build_at : 1;
    ctx->list_interval = std::max<uint64_t>(1, age * 4 / 5);
    ctx->need_full = true;  // the halted step runs as a build
}

// After a clean batch: the interval between builds from the displacement
// per step (80% of the steps that use up half the skin), the larger of the
// mean since the last build and the last step's.
static void adapt_list_interval(rrsph_ctx* ctx) {
    const Ctrl& c = *ctx->host_ctrl;
    const uint64_t now = committed_steps(c);
    // disp_acc covers the steps after the build up to the one before the last
    const uint64_t counted = now > ctx->build_at + 1 ? now - ctx->build_at - 1 : 0;
    double rate = c.disp_last;
    if (counted > 0) rate = std::max(rate, c.disp_acc / (double)counted);
    const double k = rate > 0.0 ? 0.8 * ctx->P.skin_half / rate : 1e9;
    ctx->list_interval = (uint64_t)std::min(2048.0, std::max(1.0, std::floor(k)));
}

// After a synchronised launch: a step that halted for a reason the host
// handles (more changed cells than the incremental re-sort takes; lists
// possibly stale; a list out of room) did so before touching the state;
// clear the halt and set up the re-run of that step.
static rrsph_status resume_halt(rrsph_ctx* ctx, bool* resumed) {
    Ctrl& c = *ctx->host_ctrl;
    *resumed = false;
    if (!c.halted) return RRSPH_OK;
    const uint64_t at = committed_steps(c);  // the step that halted
    if (c.status == kStatusResort) {
        note_resort_overflow(ctx, at);
    } else if (c.status == kStatusRebuild) {
        note_list_stale(ctx, at);
    } else if (c.status == kStatusListOverflow) {
        ctx->list_mode = false;  // more candidates than a list holds: sort every step from here
        ctx->need_full = true;
    } else {
        return RRSPH_OK;
    }
    c.halted = 0;
    c.status = 0;
    c.list_overflow = 0;
    HIP_TRY(hipMemcpyAsync(ctx->ctrl, ctx->host_ctrl, sizeof(Ctrl), hipMemcpyHostToDevice, ctx->stream));
    *resumed = true;
    return RRSPH_OK;
}

// Run nsteps steps to completion (or to a halt) and report the status.
// List steps go in batches of about twice the build interval, so that a
// halt wastes few launches and the interval follows the flow.
static rrsph_status run_steps(rrsph_ctx* ctx, uint32_t nsteps) {
    const uint64_t target = committed_steps(*ctx->host_ctrl) + nsteps;
    while (true) {
        const uint64_t now = committed_steps(*ctx->host_ctrl);
        if (now >= target) return RRSPH_OK;
        uint64_t chunk = target - now;
        if (list_active(ctx)) chunk = std::min<uint64_t>(chunk, std::min<uint64_t>(256, std::max<uint64_t>(16, 2 * ctx->list_interval)));
        rrsph_status st = launch_steps(ctx, (uint32_t)chunk);
        if (st != RRSPH_OK) return st;
        st = sync_status(ctx);
        if (st != RRSPH_OK) return st;
        bool resumed = false;
        st = resume_halt(ctx, &resumed);
        if (st != RRSPH_OK) return st;
        if (!resumed && ctx->host_ctrl->halted) return RRSPH_OK;  // stopped at t_end
        if (!resumed && list_active(ctx)) adapt_list_interval(ctx);
    }
}

rrsph_status rrsph_hip_step(rrsph_ctx* ctx, uint32_t nsteps) {
    if (!ctx || !ctx->uploaded) return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: no state uploaded");
    return run_steps(ctx, nsteps);
}

rrsph_status rrsph_hip_run_until(rrsph_ctx* ctx, double t_end, uint64_t max_steps, uint64_t* steps_taken) {
    if (!ctx || !ctx->uploaded) return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: no state uploaded");
    // set t_end on the device (the prologue halts with status 0 past it)
    HIP_TRY(hipMemcpyAsync(ctx->host_ctrl, ctx->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const uint64_t step0 = ctx->host_ctrl->step + (ctx->host_ctrl->pending ? 1 : 0);
    if (ctx->host_ctrl->halted && ctx->host_ctrl->status == 0) ctx->host_ctrl->halted = 0;
    ctx->host_ctrl->t_end = t_end;
    HIP_TRY(hipMemcpyAsync(ctx->ctrl, ctx->host_ctrl, sizeof(Ctrl), hipMemcpyHostToDevice, ctx->stream));
    uint64_t done = 0;
    rrsph_status st = RRSPH_OK;
    while (done < max_steps) {
        const uint32_t chunk = (uint32_t)std::min<uint64_t>(max_steps - done, 256);
        st = run_steps(ctx, chunk);
        done += chunk;
        if (st != RRSPH_OK || ctx->host_ctrl->halted) break;
    }
    const uint64_t step1 = ctx->host_ctrl->step + (ctx->host_ctrl->pending ? 1 : 0);
    if (steps_taken) *steps_taken = step1 - step0;
    if (list_active(ctx)) ctx->need_full = true;  // a list build launched after the stop did not run
    // clear the stop so later calls run
    ctx->host_ctrl->t_end = INFINITY;
    if (ctx->host_ctrl->halted && ctx->host_ctrl->status == 0) ctx->host_ctrl->halted = 0;
    HIP_TRY(hipMemcpyAsync(ctx->ctrl, ctx->host_ctrl, sizeof(Ctrl), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return st;
}

rrsph_status rrsph_hip_get_time(rrsph_ctx* ctx, double* time, double* dt, uint64_t* step) {
    if (!ctx) return RRSPH_ERR_INVALID_ARGUMENT;
    HIP_TRY(hipMemcpyAsync(ctx->host_ctrl, ctx->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const Ctrl& c = *ctx->host_ctrl;
    if (time) *time = c.pending ? c.time + c.dt_new : c.time;
    if (dt) *dt = c.pending ? c.dt_new : c.dt;
    if (step) *step = c.step + (c.pending ? 1 : 0);
    return RRSPH_OK;
}

// The current state in original particle order into the device staging
// arrays st_r, st_v, st_rho, st_p, st_itype (what rrsph_hip_download copies).
static rrsph_status enqueue_unpack(rrsph_ctx* ctx) {
    const size_t dim = ctx->P.dim;
    const uint32_t nb = grid_blocks(ctx->P.n, kBlock);
    if (dim == 3) {
        if (ctx->fp32)
            download3_kernel<float><<<nb, kBlock, 0, ctx->stream>>>(
                ctx->P, stateA3<float>(ctx), ctx->stepped ? (const float*)ctx->pres : nullptr, ctx->st_r, ctx->st_v,
                ctx->st_rho, ctx->st_p, ctx->st_itype);
        else
            download3_kernel<double><<<nb, kBlock, 0, ctx->stream>>>(
                ctx->P, stateA3<double>(ctx), ctx->stepped ? (const double*)ctx->pres : nullptr, ctx->st_r,
                ctx->st_v, ctx->st_rho, ctx->st_p, ctx->st_itype);
    } else if (ctx->fp32) {
        download_kernel<float><<<nb, kBlock, 0, ctx->stream>>>(
            ctx->P, stateA<float>(ctx), ctx->stepped ? (const float*)ctx->pres : nullptr, ctx->st_r, ctx->st_v,
            ctx->st_rho, ctx->st_p, ctx->st_itype);
    } else {
        download_kernel<double><<<nb, kBlock, 0, ctx->stream>>>(
            ctx->P, stateA<double>(ctx), ctx->stepped ? (const double*)ctx->pres : nullptr, ctx->st_r, ctx->st_v,
            ctx->st_rho, ctx->st_p, ctx->st_itype);
    }
    HIP_TRY(hipGetLastError());
    return RRSPH_OK;
}

rrsph_status rrsph_hip_download(rrsph_ctx* ctx, uint32_t fields, double* r, double* v, double* rho, double* p,
                                int32_t* itype) {
    if (!ctx || !ctx->uploaded) return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: no state uploaded");
    const size_t n = ctx->P.n, dim = ctx->P.dim;
    if (const rrsph_status st = enqueue_unpack(ctx); st != RRSPH_OK) return st;
    if ((fields & RRSPH_FIELD_R) && r)
        HIP_TRY(hipMemcpyAsync(r, ctx->st_r, dim * n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if ((fields & RRSPH_FIELD_V) && v)
        HIP_TRY(hipMemcpyAsync(v, ctx->st_v, dim * n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if ((fields & RRSPH_FIELD_RHO) && rho)
        HIP_TRY(hipMemcpyAsync(rho, ctx->st_rho, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if ((fields & RRSPH_FIELD_P) && p)
        HIP_TRY(hipMemcpyAsync(p, ctx->st_p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if ((fields & RRSPH_FIELD_ITYPE) && itype)
        HIP_TRY(hipMemcpyAsync(itype, ctx->st_itype, n * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RRSPH_OK;
}

rrsph_status rrsph_hip_set_export_fields(rrsph_ctx* ctx, int32_t enable) {
    if (!ctx) return RRSPH_ERR_INVALID_ARGUMENT;
    if ((bool)enable != ctx->export_fields) {
        ctx->export_fields = enable != 0;
        drop_graph(ctx);
    }
    return RRSPH_OK;
}

rrsph_status rrsph_hip_fields(rrsph_ctx* ctx, double* a, double* drho, double* av, double* arvmu) {
    if (!ctx || !ctx->stepped) return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: no step taken yet");
    if (!ctx->fields_valid)
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: no step taken since the state was restored");
    if (ctx->fused && !ctx->export_fields)
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: enable rrsph_hip_set_export_fields before stepping");
    const size_t n = ctx->P.n, dim = ctx->P.dim;
    const uint32_t nb = grid_blocks(ctx->P.n, kBlock);
    if (dim == 3 && ctx->fp32)
        fields3_kernel<float><<<nb, kBlock, 0, ctx->stream>>>(ctx->P.n, ctx->oidA, (const float4*)ctx->der,
                                                              (const float4*)ctx->xs, ctx->st_a, ctx->st_drho,
                                                              ctx->st_av, ctx->st_mu);
    else if (dim == 3)
        fields3_kernel<double><<<nb, kBlock, 0, ctx->stream>>>(ctx->P.n, ctx->oidA, (const double4*)ctx->der,
                                                               (const double4*)ctx->xs, ctx->st_a, ctx->st_drho,
                                                               ctx->st_av, ctx->st_mu);
    else if (ctx->fp32)
        fields_kernel<float><<<nb, kBlock, 0, ctx->stream>>>(ctx->P.n, ctx->oidA, (const float4*)ctx->der,
                                                             (const float4*)ctx->xs, ctx->st_a, ctx->st_drho,
                                                             ctx->st_av, ctx->st_mu);
    else
        fields_kernel<double><<<nb, kBlock, 0, ctx->stream>>>(ctx->P.n, ctx->oidA, (const double4*)ctx->der,
                                                              (const double4*)ctx->xs, ctx->st_a, ctx->st_drho,
                                                              ctx->st_av, ctx->st_mu);
    HIP_TRY(hipGetLastError());
    if (a) HIP_TRY(hipMemcpyAsync(a, ctx->st_a, dim * n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (drho) HIP_TRY(hipMemcpyAsync(drho, ctx->st_drho, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (av) HIP_TRY(hipMemcpyAsync(av, ctx->st_av, dim * n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (arvmu) HIP_TRY(hipMemcpyAsync(arvmu, ctx->st_mu, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RRSPH_OK;
}

rrsph_status rrsph_hip_cell_keys(rrsph_ctx* ctx, uint32_t* keys) {
    if (!ctx || !ctx->uploaded || !keys) return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: bad arguments");
    const uint32_t nb = grid_blocks(ctx->P.n, kBlock);
    if (ctx->P.dim == 3)
        morton3_kernel<<<nb, kBlock, 0, ctx->stream>>>(ctx->P, (const double4*)ctx->posA, ctx->oidA, ctx->st_u32);
    else
        morton_kernel<<<nb, kBlock, 0, ctx->stream>>>(ctx->P, ctx->posA, ctx->oidA, ctx->st_u32);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(keys, ctx->st_u32, ctx->P.n * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RRSPH_OK;
}

rrsph_status rrsph_hip_neighbours(rrsph_ctx* ctx, uint32_t* out, uint32_t max_nbr) {
    if (!ctx || !ctx->uploaded || !out || max_nbr < 2)
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: bad arguments");
    const size_t n = ctx->P.n;
    uint32_t* dout = nullptr;
    HIP_TRY(hipMalloc((void**)&dout, n * (size_t)max_nbr * sizeof(uint32_t)));
    Ctrl zero{};
    HIP_TRY(hipMemcpyAsync(ctx->util_ctrl, &zero, sizeof(Ctrl), hipMemcpyHostToDevice, ctx->stream));
    const uint32_t nb = grid_blocks(ctx->P.n, kBlock);
    if (ctx->P.dim == 3 && ctx->fp32) {
        enqueue_util_sort<float>(ctx);
        neighbours3_kernel<float><<<nb, kBlock, 0, ctx->stream>>>(
            ctx->P, stateB3<float>(ctx, ctx->pred_scratch, ctx->pw_scratch, ctx->pres_scratch), ctx->sorted_keys,
            ctx->LB, dout, max_nbr, ctx->util_ctrl);
    } else if (ctx->P.dim == 3) {
        enqueue_util_sort<double>(ctx);
        neighbours3_kernel<double><<<nb, kBlock, 0, ctx->stream>>>(
            ctx->P, stateB3<double>(ctx, ctx->pred_scratch, ctx->pw_scratch, ctx->pres_scratch), ctx->sorted_keys,
            ctx->LB, dout, max_nbr, ctx->util_ctrl);
    } else if (ctx->fp32) {
        enqueue_util_sort<float>(ctx);
        neighbours_kernel<float><<<nb, kBlock, 0, ctx->stream>>>(ctx->P, stateB<float>(ctx, ctx->pred_scratch, ctx->pres_scratch),
                                                                 ctx->sorted_keys, ctx->LB, dout, max_nbr,
                                                                 ctx->util_ctrl);
    } else {
        enqueue_util_sort<double>(ctx);
        neighbours_kernel<double><<<nb, kBlock, 0, ctx->stream>>>(ctx->P, stateB<double>(ctx, ctx->pred_scratch, ctx->pres_scratch),
                                                                  ctx->sorted_keys, ctx->LB, dout, max_nbr,
                                                                  ctx->util_ctrl);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, dout, n * (size_t)max_nbr * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->host_ctrl, ctx->util_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipFree(dout));
    const uint32_t overflow = ctx->host_ctrl->nbr_overflow;
    // restore the main control block mirror
    HIP_TRY(hipMemcpy(ctx->host_ctrl, ctx->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    if (overflow && ctx->P.consistency_check)
        return fail(ctx, RRSPH_ERR_TOO_MANY_NEIGHBOURS, status_message(RRSPH_ERR_TOO_MANY_NEIGHBOURS));
    return RRSPH_OK;
}

rrsph_status rrsph_hip_diagnostics(rrsph_ctx* ctx, rrsph_diagnostics* out) {
    if (!ctx || !ctx->uploaded || !out) return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: bad arguments");
    const uint32_t nblk = std::min<uint32_t>(kDiagMaxBlocks, grid_blocks(ctx->P.n, kBlock));
    if (ctx->P.dim == 3 && ctx->fp32)
        diag_partials3<float><<<nblk, kBlock, 0, ctx->stream>>>(ctx->P, stateA3<float>(ctx), ctx->red);
    else if (ctx->P.dim == 3)
        diag_partials3<double><<<nblk, kBlock, 0, ctx->stream>>>(ctx->P, stateA3<double>(ctx), ctx->red);
    else if (ctx->fp32)
        diag_partials<float><<<nblk, kBlock, 0, ctx->stream>>>(ctx->P, stateA<float>(ctx), ctx->red);
    else
        diag_partials<double><<<nblk, kBlock, 0, ctx->stream>>>(ctx->P, stateA<double>(ctx), ctx->red);
    diag_final<<<1, kBlock, 0, ctx->stream>>>(ctx->red, nblk);
    HIP_TRY(hipGetLastError());
    double h[kDiagSlots];
    HIP_TRY(hipMemcpyAsync(h, ctx->red + (size_t)kDiagSlots * kDiagMaxBlocks, sizeof(h), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    double t, dt;
    uint64_t step;
    rrsph_status st = rrsph_hip_get_time(ctx, &t, &dt, &step);
    if (st != RRSPH_OK) return st;
    out->time = t;
    out->dt = dt;
    out->step = step;
    out->n_fluid_alive = (uint32_t)h[0];
    out->n_outside = (uint32_t)h[1];
    out->mass_fluid = h[0] * ctx->P.mass;
    out->kinetic_energy = h[2];
    out->potential_energy = h[3];
    out->internal_energy = h[4];
    out->max_speed = h[5];
    out->rho_min = -h[6];
    out->rho_max = h[7];
    return RRSPH_OK;
}

rrsph_status rrsph_hip_gauges(rrsph_ctx* ctx, const double* x_gauges, uint32_t ngauges, double search_n,
                              int32_t ptype, double* out) {
    if (!ctx || !ctx->uploaded || (!x_gauges && ngauges) || (!out && ngauges))
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: bad arguments");
    if (ngauges == 0) return RRSPH_OK;
    if (ngauges > 4096) return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: at most 4096 gauges");
    if (ctx->P.dim == 3) return fail(ctx, RRSPH_ERR_NOT_IMPLEMENTED, "RRSPH_HIP: gauges are 2D (x-y layers)");
    double* dx = ctx->red;  // ngauges inputs followed by ngauges results
    HIP_TRY(hipMemcpyAsync(dx, x_gauges, ngauges * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemsetAsync(dx + ngauges, 0, ngauges * sizeof(double), ctx->stream));
    const double radius = search_n * ctx->P.h;
    const uint32_t nblk = std::min<uint32_t>(1024, grid_blocks(ctx->P.n, kBlock));
    dim3 grid(nblk, ngauges);
    if (ctx->fp32)
        gauge_kernel<float><<<grid, kBlock, 0, ctx->stream>>>(ctx->P.n, stateA<float>(ctx), dx, radius, ptype,
                                                              (unsigned long long*)(dx + ngauges));
    else
        gauge_kernel<double><<<grid, kBlock, 0, ctx->stream>>>(ctx->P.n, stateA<double>(ctx), dx, radius, ptype,
                                                               (unsigned long long*)(dx + ngauges));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, dx + ngauges, ngauges * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RRSPH_OK;
}

rrsph_status rrsph_hip_probe(rrsph_ctx* ctx, const double* points, uint32_t npoints, uint32_t field,
                             double* out) {
    if (!ctx || !ctx->uploaded || (npoints && (!points || !out)))
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: bad arguments");
    const StepParams& P = ctx->P;
    const uint32_t dim = (uint32_t)P.dim;
    // fatp.cpp:92-119 (check_input): a field the layer holds, a point in the geometry box
    if (field > RRSPH_PROBE_RHO || (field == RRSPH_PROBE_VZ && dim != 3))
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "wrong target value");
    if (npoints == 0) return RRSPH_OK;
    if (npoints > RRSPH_MAX_PROBES) return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: too many probe points");
    const double lo[3] = {P.xmin, P.ymin, P.zmin}, hi[3] = {P.xmax, P.ymax, P.zmax};
    for (uint32_t g = 0; g < npoints; ++g)
        for (uint32_t d = 0; d < dim; ++d) {
            const double x = points[(size_t)g * dim + d];
            if (!(lo[d] <= x && lo[d] + (hi[d] - lo[d]) >= x))
                return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "point is outside bounds");
        }
    // the field as a dump layer holds it
    if (const rrsph_status st = enqueue_unpack(ctx); st != RRSPH_OK) return st;
    const double* fld = field == RRSPH_PROBE_P ? ctx->st_p : field == RRSPH_PROBE_RHO ? ctx->st_rho : ctx->st_v;
    const uint32_t stride = field <= RRSPH_PROBE_VZ ? dim : 1u, comp = field <= RRSPH_PROBE_VZ ? field : 0u;
    // summation density of the current positions on the sorted copy (never halted)
    Ctrl zero{};
    HIP_TRY(hipMemcpyAsync(ctx->util_ctrl, &zero, sizeof(Ctrl), hipMemcpyHostToDevice, ctx->stream));
    double* dpts = ctx->red;  // dim * npoints points, then npoints results
    double* dout = dpts + (size_t)dim * npoints;
    HIP_TRY(hipMemcpyAsync(dpts, points, (size_t)dim * npoints * sizeof(double), hipMemcpyHostToDevice,
                           ctx->stream));
    const uint32_t nb_xcd = P.blocks_per_xcd * kNumXcd, pb = grid_blocks(npoints, kWave);
    auto run = [&](auto tag) {
        using R = decltype(tag);
        enqueue_util_sort<R>(ctx);
        if (dim == 3) {
            StateB3<R> B = stateB3<R>(ctx, ctx->pred_scratch, ctx->pw_scratch, ctx->pres_scratch);
            density_sum3_kernel<R><<<nb_xcd, kBlock, 0, ctx->stream>>>(P, ctx->util_ctrl, B, ctx->sorted_keys,
                                                                       ctx->LB);
            probe3_kernel<R><<<pb, kWave, 0, ctx->stream>>>(P, B, ctx->LB, dpts, npoints, fld, stride, comp, dout);
        } else {
            StateB<R> B = stateB<R>(ctx, ctx->pred_scratch, ctx->pres_scratch);
            density_sum_kernel<R><<<nb_xcd, kBlock, 0, ctx->stream>>>(P, ctx->util_ctrl, B, ctx->sorted_keys,
                                                                      ctx->LB);
            probe_kernel<R><<<pb, kWave, 0, ctx->stream>>>(P, B, ctx->LB, dpts, npoints, fld, stride, comp, dout);
        }
    };
    if (ctx->fp32) run(float{});
    else run(double{});
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, dout, npoints * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    Ctrl u{};
    HIP_TRY(hipMemcpyAsync(&u, ctx->util_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    // the layer's find_neighbours (fatp.cpp:136-137) throws on overflow with the check on
    if (u.nbr_overflow && P.consistency_check)
        return fail(ctx, RRSPH_ERR_TOO_MANY_NEIGHBOURS, status_message(RRSPH_ERR_TOO_MANY_NEIGHBOURS));
    return RRSPH_OK;
}

void* rrsph_hip_stream(rrsph_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

rrsph_status rrsph_hip_set_graph(rrsph_ctx* ctx, int32_t enable) {
    if (!ctx) return RRSPH_ERR_INVALID_ARGUMENT;
    ctx->use_graph = enable != 0;
    if (!ctx->use_graph) drop_graph(ctx);
    return RRSPH_OK;
}

uint32_t rrsph_hip_kernel_class_count(void) { return K_NUM_CLASSES; }
const char* rrsph_hip_kernel_class_name(uint32_t k) { return k < K_NUM_CLASSES ? kClassNames[k] : ""; }

rrsph_status rrsph_hip_profile_steps(rrsph_ctx* ctx, uint32_t nsteps, double* avg_ms, uint32_t n_classes,
                                     double* total_ms) {
    if (!ctx || !ctx->uploaded) return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: no state uploaded");
    Profiler prof;
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    HIP_TRY(hipEventRecord(t0, ctx->stream));
    const uint64_t first = committed_steps(*ctx->host_ctrl);
    for (uint32_t k = 0; k < nsteps; ++k) {
        if (list_active(ctx)) {
            const bool build = next_is_build(ctx, first + k);
            enqueue_list_step(ctx, &prof, build);
            if (build) {
                ctx->build_at = first + k;
                ++ctx->n_build;
                ++ctx->n_full;
            } else {
                ++ctx->n_list;
            }
        } else {
            const bool full = ctx->need_full || !ctx->incremental || ctx->launched < ctx->full_until;
            enqueue_step(ctx, &prof, full);
            ++(full ? ctx->n_full : ctx->n_incremental);
        }
        ++ctx->launched;
        ctx->need_full = false;
    }
    HIP_TRY(hipEventRecord(t1, ctx->stream));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (nsteps) {
        ctx->stepped = true;
        ctx->fields_valid = true;
    }
    std::vector<double> sum(K_NUM_CLASSES, 0.0);
    std::vector<uint32_t> cnt(K_NUM_CLASSES, 0);
    for (auto& r : prof.rec) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, r.second.first, r.second.second);
        sum[r.first] += ms;
        cnt[r.first] += 1;
        (void)hipEventDestroy(r.second.first);
        (void)hipEventDestroy(r.second.second);
    }
    for (uint32_t k = 0; k < n_classes && k < K_NUM_CLASSES; ++k) avg_ms[k] = cnt[k] ? sum[k] / cnt[k] : 0.0;
    float tot = 0;
    (void)hipEventElapsedTime(&tot, t0, t1);
    if (total_ms) *total_ms = tot;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    rrsph_status st = sync_status(ctx);
    if (st != RRSPH_OK) return st;
    bool resumed = false;
    st = resume_halt(ctx, &resumed);
    if (st != RRSPH_OK || !resumed) return st;
    // finish the steps after a halted one, unprofiled
    return run_steps(ctx, (uint32_t)(first + nsteps - committed_steps(*ctx->host_ctrl)));
}

rrsph_status rrsph_hip_time_steps(rrsph_ctx* ctx, uint32_t nsteps, double* total_ms) {
    if (!ctx || !ctx->uploaded) return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: no state uploaded");
    if (ctx->use_graph && !ctx->gexec) {
        rrsph_status st = build_graph(ctx);
        if (st != RRSPH_OK) return st;
    }
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    HIP_TRY(hipEventRecord(t0, ctx->stream));
    rrsph_status st = run_steps(ctx, nsteps);
    HIP_TRY(hipEventRecord(t1, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, t0, t1);
    if (total_ms) *total_ms = ms;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return st;
}

rrsph_status rrsph_hip_sort_stats(rrsph_ctx* ctx, uint64_t* incremental_steps, uint64_t* full_steps,
                                  uint64_t* fallbacks, uint32_t* last_changed) {
    if (!ctx) return RRSPH_ERR_INVALID_ARGUMENT;
    if (incremental_steps) *incremental_steps = ctx->n_incremental;
    if (full_steps) *full_steps = ctx->n_full;
    if (fallbacks) *fallbacks = ctx->n_fallback;
    if (last_changed) *last_changed = ctx->host_ctrl ? ctx->host_ctrl->resort_count : 0;
    return RRSPH_OK;
}

rrsph_status rrsph_hip_list_stats(rrsph_ctx* ctx, uint64_t* list_steps, uint64_t* builds, uint64_t* stale_halts,
                                  uint32_t* interval) {
    if (!ctx) return RRSPH_ERR_INVALID_ARGUMENT;
    if (list_steps) *list_steps = ctx->n_list;
    if (builds) *builds = ctx->n_build;
    if (stale_halts) *stale_halts = ctx->n_list_halts;
    if (interval) *interval = ctx->list_mode ? (uint32_t)ctx->list_interval : 0u;
    return RRSPH_OK;
}

// ---------------------------------------------------------------- differentiation
static rrsph_status to_gauge_spec(rrsph_ctx* ctx, const rrsph_gauge_spec* in, GaugeSpec3& G) {
    if (!in || in->n > kMaxGauges || !(in->radius > 0) || !(in->tau > 0))
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: bad gauge spec");
    G = GaugeSpec3{};
    G.g.n = in->n;
    G.g.ptype = in->ptype;
    G.g.radius = in->radius;
    G.g.tau = in->tau;
    G.g.y_ref = in->y_ref;
    for (uint32_t g = 0; g < in->n; ++g) {
        G.g.x[g] = in->x[g];
        G.z[g] = in->z[g];
    }
    return RRSPH_OK;
}

// smooth gauges of state A into eta/zs/ws (device, kMaxGauges entries each)
// (with ctrl: into row ctrl->step - base of tapes of nsteps rows)
static rrsph_status enqueue_smooth_gauges(rrsph_ctx* ctx, const GaugeSpec3& G, double* partials, double* eta,
                                          double* zs, double* ws, const Ctrl* ctrl = nullptr, uint64_t base = 0,
                                          uint32_t nsteps = 0) {
    const uint32_t nb = gauge_blocks(ctx->P.n);  // partials holds 2 kMaxGauges nb doubles
    if (ctx->P.dim == 3)
        smooth_gauge_partials3<<<nb, kBlock, 0, ctx->stream>>>(ctx->P.n, (const double4*)ctx->posA, G, partials);
    else if (ctx->fp32)
        smooth_gauge_partials<float><<<nb, kBlock, 0, ctx->stream>>>(ctx->P.n, ctx->posA, (const float4*)ctx->vaA,
                                                                      G.g, partials);
    else
        smooth_gauge_partials<double><<<nb, kBlock, 0, ctx->stream>>>(ctx->P.n, ctx->posA, (const double4*)ctx->vaA,
                                                                       G.g, partials);
    smooth_gauge_final<<<std::max(G.g.n, 1u), kBlock, 0, ctx->stream>>>(nb, G.g, partials, eta, zs, ws, ctrl, base,
                                                                         nsteps);
    HIP_TRY(hipGetLastError());
    return RRSPH_OK;
}

static rrsph_status adjoint_supported(rrsph_ctx* ctx) {
    const StepParams& P = ctx->P;
    if (!ctx->fast_pair || P.density_treatment != 1)
        return fail(ctx, RRSPH_ERR_NOT_IMPLEMENTED,
                    "RRSPH_HIP: the adjoint needs continuity density, cubic kernels, approximation 2 and no "
                    "artificial pressure");
    if (P.dt_method == 2 && P.dt_dynamic_mode == 1)
        return fail(ctx, RRSPH_ERR_NOT_IMPLEMENTED, "RRSPH_HIP: the adjoint needs dt independent of the state");
    if (P.nwm == 4 || P.nwm == 5)
        return fail(ctx, RRSPH_ERR_NOT_IMPLEMENTED, "RRSPH_HIP: the adjoint supports the piston wave makers only");
    return RRSPH_OK;
}

// one saved state: posA | vaA | adA | oidA | pres | ctrl
// bytes per particle of the FP64 position array (2D double2, 3D double4)
// and of one real4 state array (va, ad: FP32 or FP64)
static size_t pos_bytes(const rrsph_ctx* ctx) { return ctx->P.dim == 3 ? 32 : 16; }
static size_t real4_bytes(const rrsph_ctx* ctx) { return ctx->fp32 ? 16 : 32; }
// bytes of one saved state: pos | va | ad | oid | pressure | control block
// (the pressure of the step that produced the state, in its slot order: the
// reference's p array, which download() returns)
static size_t state_bytes(const rrsph_ctx* ctx) {
    return (pos_bytes(ctx) + 2 * real4_bytes(ctx) + 4 + real4_bytes(ctx) / 4) * ctx->P.n + sizeof(Ctrl);
}

// the blob of one saved state: posA | vaA | adA | oidA | pres | ctrl
static CopyPlan state_plan(rrsph_ctx* ctx, void* blob, bool save) {
    const size_t n = ctx->P.n, pb = pos_bytes(ctx), vb = real4_bytes(ctx);
    char* d = (char*)blob;
    void* arrays[6] = {ctx->posA, ctx->vaA, ctx->adA, ctx->oidA, ctx->pres, ctx->ctrl};
    const size_t bytes[6] = {pb * n, vb * n, vb * n, 4 * n, vb / 4 * n, sizeof(Ctrl)};
    CopyPlan plan;
    size_t off = 0;
    for (int k = 0; k < 6; ++k) {
        if (save) plan.add(arrays[k], d + off, bytes[k]);
        else plan.add(d + off, arrays[k], bytes[k]);
        off += bytes[k];
    }
    return plan;
}

static rrsph_status save_state(rrsph_ctx* ctx, void* dst) { return enqueue_copies(ctx, state_plan(ctx, dst, true)); }

// The arrays of the state saved in blob, in place (layout of state_plan).
static StateView state_view(const rrsph_ctx* ctx, void* blob) {
    const size_t n = ctx->P.n, pb = pos_bytes(ctx), vb = real4_bytes(ctx);
    char* d = (char*)blob;
    return StateView{d, d + pb * n, d + (pb + vb) * n, (uint32_t*)(d + (pb + 2 * vb) * n)};
}

static rrsph_status restore_state(rrsph_ctx* ctx, const void* src, bool keys = true) {
    if (const rrsph_status st = enqueue_copies(ctx, state_plan(ctx, const_cast<void*>(src), false)); st != RRSPH_OK)
        return st;
    ctx->fields_valid = false;  // der / xs belong to whatever step ran last
    // the next step's cell keys from the restored positions, and a full sort
    ctx->need_full = true;
    if (!keys) return RRSPH_OK;  // the caller reorders with a saved sort and launches no step
    enqueue_state_keys(ctx);
    HIP_TRY(hipGetLastError());
    return RRSPH_OK;
}

// restore_state, then bring the host mirror of the control block up to date
static rrsph_status restore_state_sync(rrsph_ctx* ctx, const void* src) {
    rrsph_status st = restore_state(ctx, src);
    if (st != RRSPH_OK) return st;
    HIP_TRY(hipMemcpyAsync(ctx->host_ctrl, ctx->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RRSPH_OK;
}

// One forward tape step as the graph of a sorted step with the incremental
// re-sort, then its record and gauges (rows of the step index - base),
// when that is what launch_steps would launch; returns false (nothing
// launched) otherwise.  A launch after a kernel launched on its own (the
// record, the gauges) waits several microseconds longer for the stream
// than a kernel inside a graph, so the step costs about 20 us less.
static rrsph_status launch_taped_step(rrsph_ctx* ctx, uint64_t base, bool& launched) {
    AdjTape& T = ctx->adj;
    launched = false;
    if (!ctx->use_graph || ctx->need_full || !ctx->incremental || ctx->launched < ctx->full_until) return RRSPH_OK;
    if (!T.texec || T.tbase != base) {
        if (T.texec) (void)hipGraphExecDestroy(T.texec);
        if (T.tgraph) (void)hipGraphDestroy(T.tgraph);
        T.texec = nullptr;
        T.tgraph = nullptr;
        HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        enqueue_step(ctx, nullptr, false);
        adj_record<<<1, 1, 0, ctx->stream>>>(ctx->ctrl, T.rec, base, T.nsteps);
        const rrsph_status st = enqueue_smooth_gauges(ctx, T.spec, T.partials, T.eta, T.zs, T.ws, ctx->ctrl, base,
                                                      T.nsteps);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (st != RRSPH_OK || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return st != RRSPH_OK ? st : fail(ctx, RRSPH_ERR_HIP, std::string("hipStreamEndCapture: ") +
                                                                      hipGetErrorString(e));
        }
        T.tgraph = g;
        HIP_TRY(hipGraphInstantiate(&T.texec, T.tgraph, nullptr, nullptr, 0));
        T.tbase = base;
    }
    HIP_TRY(hipGraphLaunch(T.texec, ctx->stream));
    // launch_steps' bookkeeping of a step with the incremental re-sort
    ctx->last_keys = ctx->kbuf[0];
    ctx->last_perm = ctx->vbuf[0];
    ++ctx->n_incremental;
    ++ctx->launched;
    ctx->need_full = false;
    ctx->stepped = true;
    ctx->fields_valid = true;
    launched = true;
    return RRSPH_OK;
}

constexpr uint32_t kHaltCheck = 64;  // tape steps between two looks at the control block

// Launch steps [k0, k1) of the tape, reading the control block back every
// kHaltCheck steps only: every step's record, the checkpoints (forward) or
// segment states (replay; slot k - kslot of seg_base), and the gauges
// (forward).  Returns the status read after the last step (or after the
// look that found a halt).
static rrsph_status adjoint_run(rrsph_ctx* ctx, uint32_t k0, uint32_t k1, bool forward, void* seg_base,
                                uint32_t kslot, uint64_t abs0) {
    AdjTape& T = ctx->adj;
    const size_t n = ctx->P.n;
    // the forward keeps the states of the last segment (of every step with
    // keep_all): the reverse pass starts there and needs no replay of it
    const uint32_t last0 = T.keep_all ? 0 : (T.nsteps - 1) / T.every * T.every;
    const uint64_t base = abs0 - k0;  // step index of tape step 0
    // the copies after a step (its sort, its build's lists) wait for the
    // state copy before the next step and go in the same launch when the
    // ranges fit one; the next step's sort overwrites them only after that
    CopyPlan pending;
    auto flush = [&]() -> rrsph_status {
        if (pending.n == 0) return RRSPH_OK;
        const CopyPlan p = pending;
        pending.n = 0;
        return enqueue_copies(ctx, p);
    };
    auto save = [&](void* dst) -> rrsph_status {
        const CopyPlan p = state_plan(ctx, dst, true);
        if (pending.n + p.n > kCopyRanges) {
            if (const rrsph_status s2 = flush(); s2 != RRSPH_OK) return s2;
        }
        for (int i = 0; i < p.n; ++i) pending.r[pending.n++] = p.r[i];
        return flush();
    };
    for (uint32_t k = k0; k < k1; ++k) {
        rrsph_status st;
        // after a halt the rest of the range is launches that do nothing:
        // look every kHaltCheck steps, so that a halt early in a long range
        // costs at most that many (the caller resumes from the halted step)
        if (k > k0 && (k - k0) % kHaltCheck == 0) {
            st = flush();
            if (st != RRSPH_OK) return st;
            HIP_TRY(hipMemcpyAsync(ctx->host_ctrl, ctx->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            if (ctx->host_ctrl->halted) return sync_status(ctx);
        }
        if (forward && k % T.every == 0) {
            if (T.lists && k > 0) {  // the build interval from the displacements so far
                HIP_TRY(hipMemcpyAsync(ctx->host_ctrl, ctx->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
                HIP_TRY(hipStreamSynchronize(ctx->stream));
                if (!ctx->host_ctrl->halted) adapt_list_interval(ctx);
            }
            st = save(T.ck[k / T.every]);
            if (st != RRSPH_OK) return st;
        }
        const bool keep = forward ? k >= last0 : true;  // seg holds this step's state
        const size_t slot = forward ? k - last0 : k - kslot;
        if (keep) {
            st = save((char*)seg_base + slot * T.blob);
            if (st != RRSPH_OK) return st;
            if (!forward && k + 1 == k1) {  // the state before the segment's last step is all the adjoint needs
                // that step is not launched: a list step reads the lists of the segment's last build; a
                // step that would build sorts its state in the reverse pass
                if (T.lists) T.seg_build[slot] = T.builds[k] ? -1 : (int32_t)T.nseg_lists - 1;
                break;
            }
        }
        bool build = false, taped = false;
        if (T.lists) {
            const uint64_t abs = abs0 + (k - k0);
            build = forward ? (k % T.every == 0 || k == T.force_build || abs - ctx->build_at >= ctx->list_interval)
                            : T.builds[k] != 0;
            if (forward) T.builds[k] = build ? 1 : 0;
            st = launch_list_step(ctx, build, abs);
        } else {
            st = forward ? launch_taped_step(ctx, base, taped) : RRSPH_OK;
            if (st == RRSPH_OK && !taped) st = launch_steps(ctx, 1);
        }
        if (st != RRSPH_OK) return st;
        if (keep) {
            if (slot == 0 || !T.lists) T.nseg_lists = 0;
            CopyPlan plan;
            if (!T.lists || build) {  // keep the sort of a step whose state seg holds
                plan.add(ctx->last_keys, T.seg_keys + slot * n, n * sizeof(uint32_t));
                plan.add(ctx->last_perm, T.seg_perm + slot * n, n * sizeof(uint32_t));
                T.seg_sorted[slot] = 1;
            } else {
                T.seg_sorted[slot] = 0;  // its state is in its build's slot order
            }
            if (T.lists && build) {  // keep the build's lists and sorted keys
                if (T.nseg_lists == T.seg_lists.size()) {
                    AdjTape::Lists L;
                    hipError_t e = hipSuccess;
                    L.nbl = dalloc<uint32_t>((size_t)(ctx->P.list_cap + kListPad) * n, e);
                    L.nbc = dalloc<uint32_t>(n, e);
                    L.keys = dalloc<uint32_t>(n, e);
                    T.seg_lists.push_back(L);
                    if (e != hipSuccess)
                        return fail(ctx, RRSPH_ERR_HIP, std::string("hipMalloc (adjoint lists): ") +
                                                            hipGetErrorString(e));
                }
                const AdjTape::Lists& L = T.seg_lists[T.nseg_lists];
                HIP_TRY(hipMemcpyAsync(L.nbl, ctx->nbl, (size_t)(ctx->P.list_cap + kListPad) * n * sizeof(uint32_t),
                                       hipMemcpyDeviceToDevice, ctx->stream));
                plan.add(ctx->nbc, L.nbc, n * sizeof(uint32_t));
                plan.add(ctx->last_keys, L.keys, n * sizeof(uint32_t));
                ++T.nseg_lists;
            }
            if (T.lists) T.seg_build[slot] = (int32_t)T.nseg_lists - 1;
            pending = plan;  // empty before: the state copy before the step took it
        }
        if (forward && !taped) {
            adj_record<<<1, 1, 0, ctx->stream>>>(ctx->ctrl, T.rec, base, T.nsteps);
            st = enqueue_smooth_gauges(ctx, T.spec, T.partials, T.eta, T.zs, T.ws, ctx->ctrl, base, T.nsteps);
            if (st != RRSPH_OK) return st;
        }
    }
    if (const rrsph_status st = flush(); st != RRSPH_OK) return st;
    return sync_status(ctx);
}

// adjoint_run with its halts handled: a step that changed more cells than
// the incremental re-sort takes, or (forward, list steps) found its lists
// possibly stale, halts before touching the state, and the launches after
// it do nothing; clear the halt and resume the range from that step, which
// runs with the full sort or as a build (its record, checkpoint, segment
// state and gauges are written again).
static rrsph_status adjoint_run_checked(rrsph_ctx* ctx, uint32_t k0, uint32_t k1, bool forward, void* seg_base) {
    HIP_TRY(hipMemcpyAsync(ctx->host_ctrl, ctx->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const uint64_t start = committed_steps(*ctx->host_ctrl);
    AdjTape& T = ctx->adj;
    uint32_t from = k0;
    while (true) {
        rrsph_status st = adjoint_run(ctx, from, k1, forward, seg_base, k0, start + (from - k0));
        if (st != RRSPH_OK) return st;
        Ctrl& c = *ctx->host_ctrl;
        if (!c.halted) return RRSPH_OK;
        const bool resort = c.status == kStatusResort;
        const bool stale = c.status == kStatusRebuild && forward && T.lists;
        if (c.status == kStatusListOverflow)
            return fail(ctx, RRSPH_ERR_NOT_IMPLEMENTED,
                        "RRSPH_HIP: a neighbour list ran out of room during the adjoint's tape; run it with "
                        "RRSPH_HIP_LIST=0");
        if (!resort && !stale) return RRSPH_OK;
        const uint32_t at = k0 + (uint32_t)(committed_steps(c) - start);  // the step that halted
        // the resumed step runs with the full sort or builds, so every resume advances
        if (at < from || at >= k1 || (at == from && from != k0))
            return fail(ctx, RRSPH_ERR_HIP, "RRSPH_HIP: halt outside the replayed range");
        if (resort) {
            note_resort_overflow(ctx, committed_steps(c));
        } else {
            note_list_stale(ctx, committed_steps(c));
            T.force_build = at;
        }
        c.halted = 0;
        c.status = 0;
        HIP_TRY(hipMemcpyAsync(ctx->ctrl, ctx->host_ctrl, sizeof(Ctrl), hipMemcpyHostToDevice, ctx->stream));
        from = at;
    }
}

rrsph_status rrsph_hip_smooth_gauges(rrsph_ctx* ctx, const rrsph_gauge_spec* spec, double* eta) {
    if (!ctx || !ctx->uploaded || !eta) return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: bad arguments");
    GaugeSpec3 G;
    rrsph_status st = to_gauge_spec(ctx, spec, G);
    if (st != RRSPH_OK) return st;
    const uint32_t nb = grid_blocks(ctx->P.n, kBlock);
    double* buf = nullptr;
    HIP_TRY(hipMalloc((void**)&buf, (2 * kMaxGauges * (size_t)nb + 3 * kMaxGauges) * sizeof(double)));
    double* out = buf + 2 * kMaxGauges * (size_t)nb;
    st = enqueue_smooth_gauges(ctx, G, buf, out, out + kMaxGauges, out + 2 * kMaxGauges);
    if (st == RRSPH_OK) {
        HIP_TRY(hipMemcpyAsync(eta, out, G.g.n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    HIP_TRY(hipFree(buf));
    return st;
}

rrsph_status rrsph_hip_smooth_gauges_noalloc(rrsph_ctx* ctx, const rrsph_gauge_spec* spec, double* eta) {
    if (!ctx || !ctx->uploaded || !eta || !spec)
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: bad arguments");
    if (spec->n > kMaxGauges)
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT,
                    "RRSPH_HIP: rrsph_hip_smooth_gauges_noalloc takes at most 64 gauges (its scratch holds 64)");
    GaugeSpec3 G;
    rrsph_status st = to_gauge_spec(ctx, spec, G);
    if (st != RRSPH_OK) return st;
    if (!ctx->scratch || !ctx->host_eta) return fail(ctx, RRSPH_ERR_HIP, "RRSPH_HIP: no scratch on this device");
    // every partial and result slot used below is written before it is read:
    // nothing in the scratch needs clearing
    double* partials = ctx->scratch;
    double* out = ctx->scratch + 2 * (size_t)kMaxGauges * kGaugeBlocks;
    st = enqueue_smooth_gauges(ctx, G, partials, out, out + kMaxGauges, out + 2 * kMaxGauges);
    if (st != RRSPH_OK) return st;
    HIP_TRY(hipMemcpyAsync(ctx->host_eta, out, G.g.n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::memcpy(eta, ctx->host_eta, G.g.n * sizeof(double));
    return RRSPH_OK;
}

rrsph_status rrsph_hip_set_piston(rrsph_ctx* ctx, double omega, uint32_t nharm, const double* a, const double* b) {
    if (!ctx) return RRSPH_ERR_INVALID_ARGUMENT;
    const uint32_t total = nharm * (uint32_t)ctx->P.piston_segments;  // paddle-major
    if (nharm > RRSPH_MAX_HARMONICS || total > RRSPH_MAX_HARMONICS || (nharm && (!a || !b)))
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: bad piston coefficients");
    ctx->prm.piston_omega = omega;
    ctx->prm.piston_nharm = (int32_t)nharm;
    ctx->P.piston_omega = omega;
    ctx->P.piston_nharm = (int32_t)nharm;
    for (uint32_t k = 0; k < RRSPH_MAX_HARMONICS; ++k) {
        const double ak = k < total ? a[k] : 0.0, bk = k < total ? b[k] : 0.0;
        ctx->prm.piston_a[k] = ak;
        ctx->prm.piston_b[k] = bk;
        ctx->P.piston_a[k] = ak;
        ctx->P.piston_b[k] = bk;
    }
    drop_graph(ctx);  // the captured steps hold the old coefficients
    ctx->adj.valid = false;
    return RRSPH_OK;
}

// The 2D adjoint's tapes and replays run the steps without neighbour lists
// (a replay must repeat its step exactly, whatever the list schedule was;
// the 3D tapes record their schedule instead).  Entering and leaving that
// state forces a full sort and a list build.
struct ListSuspend {
    rrsph_ctx* c;
    bool was;
    explicit ListSuspend(rrsph_ctx* ctx) : c(ctx), was(ctx->list_suspended) {
        if (c->list_mode && !was) c->need_full = true;
        c->list_suspended = true;
    }
    ~ListSuspend() {
        c->list_suspended = was;
        if (c->list_mode && !was) c->need_full = true;
    }
};

static rrsph_status adjoint_forward_impl(rrsph_ctx* ctx, uint32_t nsteps, uint32_t checkpoint_every,
                                         const rrsph_gauge_spec* spec, double* eta);
static rrsph_status adjoint_backward_impl(rrsph_ctx* ctx, const double* eta_bar, double* grad_a, double* grad_b,
                                          double* grad_r0, double* grad_v0, double* grad_rho0);

rrsph_status rrsph_hip_adjoint_forward(rrsph_ctx* ctx, uint32_t nsteps, uint32_t checkpoint_every,
                                       const rrsph_gauge_spec* spec, double* eta) {
    if (!ctx) return RRSPH_ERR_INVALID_ARGUMENT;
    if (ctx->P.dim == 3) {  // 3D tapes run list steps when the context does (their schedule recorded)
        rrsph_status st = adjoint_forward_impl(ctx, nsteps, checkpoint_every, spec, eta);
        ctx->need_full = true;
        return st;
    }
    ListSuspend guard(ctx);
    return adjoint_forward_impl(ctx, nsteps, checkpoint_every, spec, eta);
}

rrsph_status rrsph_hip_adjoint_backward(rrsph_ctx* ctx, const double* eta_bar, double* grad_a, double* grad_b,
                                        double* grad_r0, double* grad_v0, double* grad_rho0) {
    if (!ctx) return RRSPH_ERR_INVALID_ARGUMENT;
    if (ctx->P.dim == 3) {
        rrsph_status st = adjoint_backward_impl(ctx, eta_bar, grad_a, grad_b, grad_r0, grad_v0, grad_rho0);
        ctx->need_full = true;
        return st;
    }
    ListSuspend guard(ctx);
    return adjoint_backward_impl(ctx, eta_bar, grad_a, grad_b, grad_r0, grad_v0, grad_rho0);
}

static rrsph_status adjoint_forward_impl(rrsph_ctx* ctx, uint32_t nsteps, uint32_t checkpoint_every,
                                         const rrsph_gauge_spec* spec, double* eta) {
    if (!ctx || !ctx->uploaded || nsteps == 0 || checkpoint_every == 0 || !eta)
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: bad arguments");
    rrsph_status st = adjoint_supported(ctx);
    if (st != RRSPH_OK) return st;
    GaugeSpec3 G;
    st = to_gauge_spec(ctx, spec, G);
    if (st != RRSPH_OK) return st;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    free_tape(ctx->adj);
    AdjTape& T = ctx->adj;
    const size_t n = ctx->P.n;
    const bool d3 = ctx->P.dim == 3;
    T.nsteps = nsteps;
    T.every = std::min(checkpoint_every, nsteps);
    T.spec = G;
    T.blob = (state_bytes(ctx) + 255) / 256 * 256;
    const uint32_t nck = (nsteps + T.every - 1) / T.every;
    const uint32_t nb = grid_blocks(ctx->P.n, kBlock);
    hipError_t e = hipSuccess;
    T.ck.assign(nck + 1, nullptr);
    for (auto& p : T.ck) p = dalloc<char>(T.blob, e);
    T.lists = list_active(ctx) && d3;
    {
        // every step's state and sort when they fit in half the free memory
        // (RRSPH_HIP_ADJ_STORE=segments keeps one segment's, and replays)
        size_t free_b = 0, total_b = 0;
        const char* store = std::getenv("RRSPH_HIP_ADJ_STORE");
        const bool segments = store && std::strcmp(store, "segments") == 0;
        const double all_b = (double)nsteps * (double)(T.blob + 2 * n * sizeof(uint32_t));
        T.keep_all = !T.lists && !segments && hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
                     all_b <= 0.5 * (double)free_b;
    }
    uint32_t cap = T.keep_all ? nsteps : T.every;  // states seg holds
    const bool alloc_ok = e == hipSuccess;
    T.seg = dalloc<char>(T.blob * cap, e);
    T.seg_keys = dalloc<uint32_t>(n * cap, e);
    T.seg_perm = dalloc<uint32_t>(n * cap, e);
    if (alloc_ok && e != hipSuccess && T.keep_all) {  // every state did not fit after all: one segment's
        void* got[] = {T.seg, T.seg_keys, T.seg_perm};
        for (void* p : got)
            if (p) (void)hipFree(p);
        (void)hipGetLastError();
        e = hipSuccess;
        T.keep_all = false;
        cap = T.every;
        T.seg = dalloc<char>(T.blob * cap, e);
        T.seg_keys = dalloc<uint32_t>(n * cap, e);
        T.seg_perm = dalloc<uint32_t>(n * cap, e);
    }
    T.seg_sorted.assign(cap, 0);
    T.builds.assign(nsteps, 0);
    T.seg_build.assign(cap, -1);
    T.nseg_lists = 0;
    T.force_build = ~0ull;
    if (T.lists) T.iota = dalloc<uint32_t>(n, e);
    if (T.lists && ctx->fp32) T.nearf = dalloc<float4>(n, e);
    T.eta = dalloc<double>((size_t)nsteps * kMaxGauges, e);
    T.zs = dalloc<double>((size_t)nsteps * kMaxGauges, e);
    T.ws = dalloc<double>((size_t)nsteps * kMaxGauges, e);
    T.eta_bar = dalloc<double>((size_t)nsteps * kMaxGauges, e);
    T.partials = dalloc<double>(2 * kMaxGauges * (size_t)nb, e);
    T.piston_sum = dalloc<double>((size_t)nsteps * kMaxSegments, e);
    T.lam[0] = dalloc<double>((d3 ? kLam3 : 8) * n, e);
    T.lam[1] = dalloc<double>((d3 ? kLam3 : 8) * n, e);
    T.direct = dalloc<double>((d3 ? kDirect3 : 5) * n, e);
    T.drho_adj = dalloc<double>(n, e);
    T.out_adj = dalloc<double4>((d3 ? sizeof(AdjRec3<double>) / sizeof(double4) : 1) * n, e);
    T.flag = dalloc<uint32_t>(1, e);
    if (e != hipSuccess) {
        free_tape(ctx->adj);
        return fail(ctx, RRSPH_ERR_HIP, std::string("hipMalloc (adjoint tape): ") + hipGetErrorString(e));
    }
    T.rec = dalloc<AdjStep>(nsteps, e);
    T.uctrl = dalloc<Ctrl>(nsteps, e);
    if (e != hipSuccess) {
        free_tape(ctx->adj);
        return fail(ctx, RRSPH_ERR_HIP, std::string("hipMalloc (adjoint tape): ") + hipGetErrorString(e));
    }
    if (T.lists) {
        std::vector<uint32_t> id(n);
        for (size_t i = 0; i < n; ++i) id[i] = (uint32_t)i;
        HIP_TRY(hipMemcpy(T.iota, id.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    st = adjoint_run_checked(ctx, 0, nsteps, true, T.seg);
    if (st != RRSPH_OK) return st;
    if (ctx->host_ctrl->halted)
        return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: the run halted during the adjoint forward");
    st = save_state(ctx, T.ck[nck]);
    if (st != RRSPH_OK) return st;
    T.seg_last = true;
    std::vector<double> h((size_t)nsteps * kMaxGauges);
    HIP_TRY(hipMemcpyAsync(h.data(), T.eta, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    T.hrec.resize(nsteps);
    HIP_TRY(hipMemcpyAsync(T.hrec.data(), T.rec, nsteps * sizeof(AdjStep), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (uint32_t k = 0; k < nsteps; ++k)
        for (uint32_t g = 0; g < G.g.n; ++g) eta[(size_t)k * G.g.n + g] = h[(size_t)k * kMaxGauges + g];
    {
        std::vector<Ctrl> uc(nsteps, Ctrl{});
        for (uint32_t k = 0; k < nsteps; ++k) uc[k].cur_dt = T.hrec[k].dt_pred;
        HIP_TRY(hipMemcpyAsync(T.uctrl, uc.data(), nsteps * sizeof(Ctrl), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    T.valid = true;
    return RRSPH_OK;
}

// dL/dx of the loss term of the gauges of state A, or of the saved state
// src (lam += ...)
static void enqueue_gauge_grad(rrsph_ctx* ctx, uint32_t k, double* lam, const StateView* src = nullptr) {
    AdjTape& T = ctx->adj;
    const uint32_t nb = grid_blocks(ctx->P.n, kBlock);
    const size_t o = (size_t)k * kMaxGauges;
    const void* pos = src ? src->pos : (const void*)ctx->posA;
    const void* va = src ? src->va : ctx->vaA;
    const uint32_t* oid = src ? src->oid : ctx->oidA;
    if (ctx->P.dim == 3)
        smooth_gauge_grad3<<<nb, kBlock, 0, ctx->stream>>>(ctx->P.n, (const double4*)pos, oid, T.spec,
                                                           T.eta_bar + o, T.zs + o, T.ws + o, lam);
    else if (ctx->fp32)
        smooth_gauge_grad<float><<<nb, kBlock, 0, ctx->stream>>>(ctx->P.n, (const double2*)pos, (const float4*)va,
                                                                 oid, T.spec.g, T.eta_bar + o, T.zs + o,
                                                                 T.ws + o, lam);
    else
        smooth_gauge_grad<double><<<nb, kBlock, 0, ctx->stream>>>(ctx->P.n, (const double2*)pos, (const double4*)va,
                                                                  oid, T.spec.g, T.eta_bar + o, T.zs + o,
                                                                  T.ws + o, lam);
}

static rrsph_status adjoint_backward_impl(rrsph_ctx* ctx, const double* eta_bar, double* grad_a, double* grad_b,
                                          double* grad_r0, double* grad_v0, double* grad_rho0) {
    if (!ctx || !eta_bar) return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: bad arguments");
    AdjTape& T = ctx->adj;
    if (!T.valid) return fail(ctx, RRSPH_ERR_INVALID_ARGUMENT, "RRSPH_HIP: run rrsph_hip_adjoint_forward first");
    const StepParams& P = ctx->P;
    const bool d3 = P.dim == 3;
    const size_t n = P.n;
    const uint32_t lw = d3 ? kLam3 : 8;  // doubles of adjoint state per particle
    const uint32_t N = T.nsteps, G = T.spec.g.n;
    const uint32_t nb = grid_blocks(P.n, kBlock);
    {
        std::vector<double> h((size_t)N * kMaxGauges, 0.0);
        for (uint32_t k = 0; k < N; ++k)
            for (uint32_t g = 0; g < G; ++g) h[(size_t)k * kMaxGauges + g] = eta_bar[(size_t)k * G + g];
        HIP_TRY(hipMemcpyAsync(T.eta_bar, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    HIP_TRY(hipMemsetAsync(T.piston_sum, 0, (size_t)N * kMaxSegments * sizeof(double), ctx->stream));
    HIP_TRY(hipMemsetAsync(T.lam[0], 0, lw * n * sizeof(double), ctx->stream));
    HIP_TRY(hipMemsetAsync(T.flag, 0, sizeof(uint32_t), ctx->stream));
    const uint32_t nck = (N + T.every - 1) / T.every;
    // the loss term of the final state
    rrsph_status st = restore_state(ctx, T.ck[nck]);
    if (st != RRSPH_OK) return st;
    enqueue_gauge_grad(ctx, N - 1, T.lam[0]);
    HIP_TRY(hipGetLastError());
    int cur = 0;
    for (int seg = (int)nck - 1; seg >= 0; --seg) {
        const uint32_t n0 = (uint32_t)seg * T.every, n1 = std::min(n0 + T.every, N);
        // replay the segment from its checkpoint, keeping the state before each
        // step (the last segment's states are there already after a forward)
        if (!T.keep_all && !(T.seg_last && seg == (int)nck - 1)) {
            std::fill(T.seg_sorted.begin(), T.seg_sorted.end(), (uint8_t)0);
            std::fill(T.seg_build.begin(), T.seg_build.end(), -1);
            st = restore_state(ctx, T.ck[seg]);
            if (st != RRSPH_OK) return st;
            st = adjoint_run_checked(ctx, n0, n1, false, T.seg);
            if (st != RRSPH_OK) return st;
        }
        T.seg_last = false;
        for (int k = (int)n1 - 1; k >= (int)n0; --k) {
            const size_t slot = (size_t)(T.keep_all ? k : k - (int)n0);
            const bool sorted = T.seg_sorted[slot] != 0;  // the replay launched step k with a sort
            // list steps: the build whose lists step k read, its state in that build's slot order
            const int32_t b = T.lists ? T.seg_build[slot] : -1;
            const AdjTape::Lists* lists = b >= 0 ? &T.seg_lists[(size_t)b] : nullptr;
            // a step reordered with a saved sort reads its state where it was
            // saved; one that sorts again needs it in A (and its cell keys)
            char* blob = (char*)T.seg + slot * T.blob;
            const bool in_place = sorted || lists;
            const StateView view = state_view(ctx, blob);
            const StateView* src = in_place ? &view : nullptr;
            if (!in_place) {
                st = restore_state(ctx, blob, true);
                if (st != RRSPH_OK) return st;
            }
            // sort and reorder state k exactly as its step did (predict with its dt)
            const uint32_t* sk = sorted ? T.seg_keys + slot * n : (lists ? lists->keys : nullptr);
            const uint32_t* sp = sorted ? T.seg_perm + slot * n : (lists ? T.iota : nullptr);
            if (ctx->fp32)
                enqueue_reverse_step<float>(ctx, T.hrec[k], T.uctrl + k, T.lam[cur], T.lam[1 - cur],
                                            T.piston_sum + (size_t)k * kMaxSegments, sk, sp, lists, src);
            else
                enqueue_reverse_step<double>(ctx, T.hrec[k], T.uctrl + k, T.lam[cur], T.lam[1 - cur],
                                             T.piston_sum + (size_t)k * kMaxSegments, sk, sp, lists, src);
            if (k >= 1) enqueue_gauge_grad(ctx, k - 1, T.lam[1 - cur], src);  // loss term of state k
            HIP_TRY(hipGetLastError());
            cur = 1 - cur;
        }
    }
    // piston coefficients of paddle g: dV_g/da_{g,k} = k w cos(k w t), dV_g/db_{g,k} = -k w sin(k w t)
    std::vector<double> ps((size_t)N * kMaxSegments);
    HIP_TRY(hipMemcpyAsync(ps.data(), T.piston_sum, ps.size() * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    std::vector<double> lam(lw * n);
    HIP_TRY(hipMemcpyAsync(lam.data(), T.lam[cur], lam.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    uint32_t flag = 0;
    HIP_TRY(hipMemcpyAsync(&flag, T.flag, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const int nh = P.nwm == 6 ? std::max(P.piston_nharm, 0) : 0;
    for (int g = 0; g < P.piston_segments; ++g) {
        for (int k = 0; k < nh; ++k) {
            double ga = 0, gb = 0;
            const double kw = (k + 1) * P.piston_omega;
            for (uint32_t s = 0; s < N; ++s) {
                const double v = ps[(size_t)s * kMaxSegments + g];
                ga += v * kw * std::cos(kw * T.hrec[s].t);
                gb -= v * kw * std::sin(kw * T.hrec[s].t);
            }
            if (grad_a) grad_a[g * nh + k] = ga;
            if (grad_b) grad_b[g * nh + k] = gb;
        }
    }
    const uint32_t D = d3 ? 3 : 2;
    const uint32_t vo = d3 ? 3 : 2, ro = d3 ? 6 : 4;  // offsets of v and rho in lam
    for (size_t i = 0; i < n; ++i) {
        for (uint32_t c = 0; c < D; ++c) {
            if (grad_r0) grad_r0[D * i + c] = lam[lw * i + c];
            if (grad_v0) grad_v0[D * i + c] = lam[lw * i + vo + c];
        }
        if (grad_rho0) grad_rho0[i] = lam[lw * i + ro];
    }
    // leave the context as the forward left it (restore_state makes the next
    // step sort, and build when it runs list steps)
    st = restore_state_sync(ctx, T.ck[nck]);
    if (st != RRSPH_OK) return st;
    if (flag) return fail(ctx, RRSPH_ERR_TOO_MANY_NEIGHBOURS, status_message(RRSPH_ERR_TOO_MANY_NEIGHBOURS));
    return RRSPH_OK;
}

uint32_t RRSPH_HIP_GetSpecificVersionMajor(void) { return RRSPH_HIP_VERSION_MAJOR; }
uint32_t RRSPH_HIP_GetSpecificVersionMinor(void) { return RRSPH_HIP_VERSION_MINOR; }
uint32_t RRSPH_HIP_GetSpecificVersionPatch(void) { return RRSPH_HIP_VERSION_PATCH; }
const char* RRSPH_HIP_GetSpecificVersionName(void) { return "RRSPH_HIP"; }

}  // extern "C"
