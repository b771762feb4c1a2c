// learn.cpp -- host scheduler behind the C ABI (include/ltl_amd.h).
//
// Mirrors the reference's LTLI() (code/ltli6463.cu:1043-1497): atoms first,
// then cost levels c1+1..maxCost, within a level the operators in the order
// ~, &, |, X, F, G, U, binary operators over the splits i = c1..(c-cop)/2
// (left cost i, right cost c-cop-i), overflow into check-only mode and the
// "last round" rule.  Differences by design:
//   - device memory is one pool per context, reused across calls;
//   - kernels of a level are enqueued back to back; the host synchronises
//     once per level (the reference copies the result index back after
//     every launch, code/ltli6463.cu:1232 etc.);
//   - the result is the accepted candidate with the smallest canonical id
//     (deterministic; the reference keeps the first atomicCAS winner,
//     code/ltli6463.cu:243).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/ltl_amd.h"
#include "formula.hpp"
#include "kernels.hpp"
#include "batch.hpp"
#include "scorer.hpp"

using namespace ltl;

namespace {

// (the dedup_path_* entries count the chunks each dedup path took, with or
// without profiling, and hold no time: that is in decide and scan_finish)
// (gen_binary: the &, | and U blocks of a split generated in one pass)
enum { K_GEN1, K_GEN2, K_GENU, K_DECIDE, K_SCAN, K_EMIT1, K_EMIT2, K_EMITU, K_CHK1, K_CHK2, K_CHKU,
       K_TSTAT, K_SCORE, K_SORT, K_CHKR, K_LOG, K_EMITM, K_PSMALL, K_PLB, K_PFOUR, K_GENB, K_NKIND };
const char *kKindName[K_NKIND] = {"gen_unary", "gen_and_or", "gen_until", "decide",
                                  "scan_finish", "emit_unary", "emit_and_or", "emit_until",
                                  "check_unary", "check_and_or", "check_until",
                                  "tile_stats", "score_tiles", "sort_tiles", "check_ranked", "log_rows",
                                  "emit_meta", "dedup_path_small", "dedup_path_lookback", "dedup_path_four",
                                  "gen_binary"};

struct Block {
    uint64_t id_base, size;
    int op;
    int64_t ls, nL, rs, nR;
    int level;
};

struct PendingEvent {
    int kind;
    hipEvent_t a, b;
    uint64_t alg, min, items;
    const long long *count;     // emit: survivors of the chunk (pinned), bytes per survivor below
    uint64_t alg_per, min_per;
};

uint64_t splitmix(uint64_t x) {
    uint64_t z = x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Fingerprint keys (two 32-bit halves) with odd halves: two sets that differ
// in one word half only never share a fingerprint (an odd key times a
// nonzero difference below 2^32 is nonzero mod 2^64); any pair of different
// sets collides with probability <= 2^-31 per fingerprint over the keys.
uint64_t fp_key(uint64_t x) { return splitmix(x) | 0x0000000100000001ull; }

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

constexpr int kCountSlots = 1 << 17;

// Nodes of a formula walk read back from the device (k_tree).
constexpr int32_t kTreeCap = 1024;

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct ltl_ctx {
    ltl_options opt{};
    hipStream_t stream = nullptr;
    char *pool = nullptr;
    size_t pool_bytes = 0;
    // Workgroup regions of the batched launch whose tables are known empty
    // (each instance clears the slots it used): the first clean_n regions of
    // clean_region bytes at clean_base, with clean_tslots table slots.  Any
    // other use of the pool (ltl_learn stores sets there) sets clean_n to 0.
    char *clean_base = nullptr;
    uint64_t clean_region = 0;
    uint32_t clean_tslots = 0;
    int clean_n = 0;
    // A fixed-size dedup table left by the last call (tab_ok: the call ended
    // without error and nothing has written the pool since): its address,
    // slots - 1 and epoch.  The next call with the same table takes the next
    // epoch instead of clearing it (kEpochs calls per clear).
    bool tab_ok = false;
    uint64_t *tab_ptr = nullptr;
    uint32_t tab_mask = 0, tab_epoch = 0;
    // Device memory of the current call: the pool, or the library's static
    // scratch region of this device (ltl_learn_noalloc; its address from
    // set-up, noalloc_scratch()).
    char *arena = nullptr;
    size_t arena_bytes = 0;
    char *scratch = nullptr;
    DevState *h_state = nullptr;  // pinned

    // layout of the last call
    int W = 0, P = 0, A = 0, nw = 0, B = 1, K = 1;  // B = 1: packed, B = 0: K words per trace
    std::vector<int> pk_word, pk_off;  // packed layout: word and bit offset of each trace
    std::vector<int> sum_order;        // summary bit -> trace
    std::vector<uint8_t> lens;         // trace lengths of the last call
    int64_t capacity = 0;
    int64_t ws = 0;  // word stride of the word-major store (capacity rounded up)
    int64_t CH = 0;  // candidate slots per chunk
    uint32_t tmask = 0;           // slots in use - 1 (the table grows between levels)
    size_t tab_max_slots = 0;     // slots carved for the table
    WordAux *d_aux = nullptr;
    DevState *d_state = nullptr;
    uint64_t *d_store = nullptr, *d_tab = nullptr, *d_fp2 = nullptr, *d_sum = nullptr, *d_key = nullptr;
    SumAux *d_saux = nullptr;
    int nsw = 0;
    uint8_t *d_op = nullptr;
    int32_t *d_left = nullptr, *d_right = nullptr;
    uint32_t *t_slot = nullptr, *t_cnt = nullptr, *t_off = nullptr, *t_list = nullptr;
    uint64_t *t_fp2 = nullptr, *t_mask = nullptr;
    uint8_t *t_succ = nullptr;    // per candidate: accepted successors (deferred level)
    uint64_t *d_stage = nullptr;  // atom words and summaries, contiguous, before the scatter
    uint32_t *d_kinfo = nullptr;  // per trace: relaxed-key bit range (dedup modes 1-3)
    uint64_t *d_lb = nullptr;     // look-back status words of a chunk's workgroups + ticket
    // next level's ~/X/F/G fingerprints from a level's emits (8 values of
    // nx_stride entries; ChunkArgs::nx_fp)
    uint64_t *d_nxfp = nullptr;
    int64_t nx_stride = 0;
    int64_t lb_words = 0;
    uint32_t lb_epoch = 0;        // chunks deduplicated by look-back in this call
    uint64_t *d_ptrace = nullptr; // packed layout: per packed trace, trace | offset << 32 | length << 40
    int32_t *d_gword = nullptr;   // packed layout: per summary group, its first word; then its last
    void *d_scan = nullptr;
    size_t scan_bytes = 0;
    std::vector<int64_t> lvl_s, lvl_e;
    // Stored sets [lazy_s, lazy_e) of the last call were emitted with
    // metadata only (a deferred level); their words and summaries are
    // recomputed from the parents by the first ltl_store_read or
    // ltl_store_summaries.
    int64_t lazy_s = 0, lazy_e = 0;
    ChunkArgs lazy_args{};
    // Stored sets [lzw_s, lzw_e) (the level below the last generated one)
    // were emitted with summaries and metadata but without words; the check
    // kernels recompute their words from the parents, and launch_materialize
    // writes them before anything reads them in bulk.
    int64_t lzw_s = 0, lzw_e = 0;
    ChunkArgs lzw_args{};

    // profiling
    ltl_kernel_stat stats[K_NKIND]{};
    std::vector<PendingEvent> pending;
    std::vector<hipEvent_t> free_events;
    long long *h_counts = nullptr;  // pinned, one slot per profiled emit of a level
    // pinned staging of small per-level uploads: two halves, each refilled
    // only after the copy that last read it has completed (its event)
    char *h_stage = nullptr;
    size_t h_stage_cap = 0;
    char *h_init = nullptr;       // pinned staging of a call's initial uploads
    size_t h_init_cap = 0;
    int h_stage_i = 0;
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    int n_counts = 0;
    ltl_batch_params batch{1024, 4, 0, 0, (int64_t)1 << 18};
    ltl_batch_stats batch_stats{};
    std::vector<ltl_ctx *> helpers;  // contexts for the whole-device calls of a batch
    int n_cu = 0;
    int64_t chunk_counter = 0;

    // learned order and enumeration log
    ltl_scorer *scorer = nullptr;
    float *d_tstat = nullptr;     // tile statistics of the stored levels
    int64_t tstat_cap = 0;        // tiles
    int64_t *d_toff = nullptr;    // first tile of each level
    int toff_cap = 0;
    LBlock *d_blk = nullptr;      // binary blocks of the current level
    int blk_cap = 0;
    float *d_keys = nullptr;      // scores: 2 x ent_cap
    uint32_t *d_vals = nullptr;   // tile-pair entries: 2 x ent_cap
    int64_t ent_cap = 0;
    void *d_sort = nullptr;
    size_t sort_cap = 0;
    float *d_log = nullptr;       // log rows
    int32_t *d_tree = nullptr;    // the found formula's walk (1 + 3 * kTreeCap), in the call's memory
    int64_t log_cap = 0, log_n = 0;
    int32_t log_per_block = 0;
    // per-launch log (ltl_set_launch_log)
    ltl_launch_fn launch_fn = nullptr;
    void *launch_user = nullptr;
};

static int32_t hip_fail(hipError_t e) { return e == hipErrorOutOfMemory ? LTL_E_OOM : LTL_E_HIP; }

// Stream-ordered upload of a small host buffer through pinned staging (an
// asynchronous copy from pageable memory would hold the host until the
// stream reaches it).
static hipError_t upload_pinned(ltl_ctx *c, void *dst, const void *src, size_t bytes, hipStream_t s);
// With LTL_DEBUG set in the environment, a failing HIP call is reported on
// stderr with its source line.
static void report_hip(hipError_t e, const char *what, int line) {
    static const bool on = getenv("LTL_DEBUG") != nullptr;
    if (on) fprintf(stderr, "ltl: %s failed at learn.cpp:%d: %s\n", what, line, hipGetErrorString(e));
}
#define HIP_TRY(x)                                   \
    do {                                             \
        hipError_t e_ = (x);                         \
        if (e_ != hipSuccess) {                      \
            report_hip(e_, #x, __LINE__);            \
            return hip_fail(e_);                     \
        }                                            \
    } while (0)

extern "C" void ltl_default_options(ltl_options *o) {
    memset(o, 0, sizeof(*o));
    for (auto &c : o->cost) c = 1;
    o->max_cost = 10;
    o->neg_type = 1;
    o->dedup_mode = LTL_DEDUP_EXACT;
    o->order = LTL_ORDER_PLAIN;
    o->precheck = 1;
}

static int32_t check_options(const ltl_options *o) {
    for (int i = 0; i < 8; ++i)
        if (o->cost[i] < 1 || o->cost[i] > SHRT_MAX) return LTL_E_ARG;
    if (o->max_cost < 1 || o->max_cost > SHRT_MAX) return LTL_E_ARG;
    if (o->neg_type < 1 || o->neg_type > 2) return LTL_E_ARG;
    if (o->dedup_mode > LTL_DEDUP_REF3) return LTL_E_ARG;
    if (o->hamming_k < 0) return LTL_E_ARG;
    if (o->hash_slots_log2 < 0 || o->hash_slots_log2 > 31) return LTL_E_ARG;
    if (o->order != LTL_ORDER_PLAIN && o->order != LTL_ORDER_LEARNED) return LTL_E_ARG;
    if (o->profile < 0 || o->profile > 3) return LTL_E_ARG;
    return 0;
}

extern "C" ltl_ctx *ltl_create(const ltl_options *opt, int32_t *status) {
    auto fail = [&](int32_t s) -> ltl_ctx * { if (status) *status = s; return nullptr; };
    ltl_options o;
    if (opt) o = *opt; else ltl_default_options(&o);
    if (check_options(&o)) return fail(LTL_E_ARG);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(LTL_E_NODEVICE);
    if (o.device < 0 || o.device >= n) return fail(LTL_E_ARG);
    if (hipSetDevice(o.device) != hipSuccess) return fail(LTL_E_HIP);
    auto *c = new ltl_ctx;
    c->opt = o;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void **)&c->h_state, sizeof(DevState), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->h_counts, kCountSlots * sizeof(long long), hipHostMallocDefault) != hipSuccess) {
        ltl_destroy(c);
        return fail(LTL_E_HIP);
    }
    for (int k = 0; k < K_NKIND; ++k) snprintf(c->stats[k].name, sizeof(c->stats[k].name), "%s", kKindName[k]);
    // The device address of this GPU's scratch region (ltl_learn_noalloc),
    // taken once per device with the device current; null if unavailable.
    if (noalloc_scratch(o.device, &c->scratch) != hipSuccess) c->scratch = nullptr;
    if (hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, o.device) != hipSuccess) c->n_cu = 256;
    if (status) *status = 0;
    return c;
}

extern "C" int32_t ltl_set_options(ltl_ctx *c, const ltl_options *o) {
    if (!c || !o || check_options(o)) return LTL_E_ARG;
    if (o->device != c->opt.device) return LTL_E_ARG;
    c->opt = *o;
    return 0;
}

extern "C" int32_t ltl_get_options(const ltl_ctx *c, ltl_options *o) {
    if (!c || !o) return LTL_E_ARG;
    *o = c->opt;
    return 0;
}

extern "C" void ltl_destroy(ltl_ctx *c) {
    if (!c) return;
    // (errors here are reported with LTL_DEBUG and then cleared: nothing is
    // left for a later call's launch checks to pick up)
    auto chk = [](hipError_t e, const char *what) {
        if (e != hipSuccess) report_hip(e, what, __LINE__);
    };
    for (ltl_ctx *h : c->helpers) ltl_destroy(h);
    chk(hipSetDevice(c->opt.device), "hipSetDevice");
    if (c->stream) chk(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    for (auto &p : c->pending) {
        chk(hipEventDestroy(p.a), "hipEventDestroy(pending)");
        chk(hipEventDestroy(p.b), "hipEventDestroy(pending)");
    }
    for (auto e : c->free_events) chk(hipEventDestroy(e), "hipEventDestroy(free)");
    if (c->pool) chk(hipFree(c->pool), "hipFree(pool)");
    for (void *p : {(void *)c->d_tstat, (void *)c->d_toff, (void *)c->d_blk, (void *)c->d_keys, (void *)c->d_vals,
                    c->d_sort, (void *)c->d_log})
        if (p) chk(hipFree(p), "hipFree");
    if (c->h_state) chk(hipHostFree(c->h_state), "hipHostFree(state)");
    if (c->h_counts) chk(hipHostFree(c->h_counts), "hipHostFree(counts)");
    if (c->h_stage) chk(hipHostFree(c->h_stage), "hipHostFree(stage)");
    if (c->h_init) chk(hipHostFree(c->h_init), "hipHostFree(init)");
    for (hipEvent_t e : c->stage_ev)
        if (e) chk(hipEventDestroy(e), "hipEventDestroy(stage)");
    if (c->stream) chk(hipStreamDestroy(c->stream), "hipStreamDestroy");
    (void)hipGetLastError();
    delete c;
}

extern "C" void ltl_result_free(ltl_result *r) {
    if (!r) return;
    free(r->formula);
    free(r->nodes);
    r->formula = nullptr;
    r->nodes = nullptr;
}

// ---------------------------------------------------------------------------
// Pool: reserved once per context at the budget, carved per call.
// ---------------------------------------------------------------------------
static int32_t ensure_pool(ltl_ctx *c) {
    size_t want = c->opt.hbm_budget_bytes;
    // (budget 0: the pool sized from the free memory at the first call is
    // kept, without a memory query per call)
    if (!want && c->pool) return 0;
    if (!want) {
        size_t fr = 0, tot = 0;
        HIP_TRY(hipMemGetInfo(&fr, &tot));
        want = fr / 4 * 3;
    }
    want = align_up(want, 1 << 21);
    if (c->pool && c->pool_bytes >= want) return 0;
    if (c->pool) { hipFree(c->pool); c->pool = nullptr; c->pool_bytes = 0; }
    c->clean_n = 0;
    c->tab_ok = false;
    HIP_TRY(hipMalloc((void **)&c->pool, want));
    c->pool_bytes = want;
    return 0;
}

static int32_t carve(ltl_ctx *c) {
    const size_t budget = c->arena_bytes;
    const size_t per_cs = (size_t)c->nw * 8 + (size_t)c->nsw * 32 + 25;
    // chunk: per-candidate temporaries (19 B each) within 1/32 of the pool
    int64_t ch = 1 << 26;
    while (ch > 4096 && (size_t)ch * 19 > budget / 32) ch >>= 1;
    // A fixed table keeps room for the stored sets: a chunk takes at most a
    // quarter of its slots.
    if (c->opt.hash_slots_log2)
        while (ch > 4096 && ch > ((int64_t)1 << c->opt.hash_slots_log2) / 4) ch >>= 1;
    c->CH = ch;
    const size_t tiles = (size_t)ch / 64;
    c->scan_bytes = scan_temp_bytes((int)tiles);
    size_t fixed = align_up((size_t)c->nw * sizeof(WordAux), 256) + align_up(sizeof(DevState), 256) +
                   align_up((size_t)c->nsw * sizeof(SumAux), 256) +
                   2 * align_up((size_t)ch * 4, 256) + align_up((size_t)ch * 8, 256) + align_up((size_t)ch, 256) +
                   align_up(tiles * 16, 256) +
                   2 * align_up(tiles * 4, 256) + align_up(c->scan_bytes, 256) +
                   align_up((size_t)(c->nw + 4 * c->nsw) * c->A * 8, 256) + align_up((size_t)c->W * 4, 256) +
                   align_up((size_t)c->W * 8, 256) + align_up((tiles / 4 + 2) * 8, 256) +
                   align_up((1 + 3 * (size_t)kTreeCap) * 4, 256) + align_up((size_t)c->nsw * 8, 256) +
                   align_up(8 * 8 * ((size_t)ch / 8 + 128), 256) + 15 * 256;
    if (fixed >= budget) return LTL_E_OOM;
    const size_t R = budget - fixed;
    // Table: load stays <= 1/2 with a full chunk pending (cap + ch <= slots/2).
    // Without a fixed size, start from the largest table within half the
    // remaining pool and halve while the smaller table still holds every set
    // the pool can store.
    size_t slots;
    if (c->opt.hash_slots_log2) {
        slots = (size_t)1 << c->opt.hash_slots_log2;
        if (slots * 16 >= R) return LTL_E_OOM;
    } else {
        slots = (size_t)1 << 31;
        while (slots > 4096 && slots * 16 > R / 2) slots >>= 1;
        while (slots > 4096) {
            const size_t h = slots / 2;
            const int64_t cap_h = (int64_t)((R - h * 16) / per_cs);
            if ((int64_t)(h / 2) - ch < cap_h) break;
            slots = h;
        }
    }
    int64_t cap = std::min<int64_t>((int64_t)((R - slots * 16) / per_cs), (int64_t)(slots / 2) - ch);
    cap = std::min<int64_t>(cap, (int64_t)INT32_MAX - 64) / 64 * 64;
    if (cap < 64) return LTL_E_OOM;
    if (c->opt.capacity > 0) cap = std::min<int64_t>(cap, c->opt.capacity);
    c->capacity = cap;
    c->tmask = (uint32_t)(slots - 1);
    c->tab_max_slots = slots;

    char *p = c->arena;
    auto take = [&](size_t n) { char *q = p; p += align_up(n, 256); return q; };
    c->d_aux = (WordAux *)take((size_t)c->nw * sizeof(WordAux));
    c->d_tree = (int32_t *)take((1 + 3 * (size_t)kTreeCap) * 4);
    c->d_state = (DevState *)take(sizeof(DevState));
    c->d_saux = (SumAux *)take((size_t)c->nsw * sizeof(SumAux));
    c->t_slot = (uint32_t *)take((size_t)ch * 4);
    c->t_list = (uint32_t *)take((size_t)ch * 4);
    c->t_fp2 = (uint64_t *)take((size_t)ch * 8);
    c->t_succ = (uint8_t *)take((size_t)ch);
    c->t_mask = (uint64_t *)take(tiles * 16);
    c->t_cnt = (uint32_t *)take(tiles * 4);
    c->t_off = (uint32_t *)take(tiles * 4);
    c->d_scan = take(c->scan_bytes);
    c->d_stage = (uint64_t *)take((size_t)(c->nw + 4 * c->nsw) * c->A * 8);
    c->d_kinfo = (uint32_t *)take((size_t)c->W * 4);
    c->d_ptrace = (uint64_t *)take((size_t)c->W * 8);
    c->d_gword = (int32_t *)take((size_t)c->nsw * 8);
    c->lb_words = (int64_t)(tiles / 4 + 2);  // a status word per workgroup, the ticket in the last
    c->d_lb = (uint64_t *)take((size_t)c->lb_words * 8);
    // (a level of at most CH / 8 sets: what the fused ~/X/F/G pass over it takes)
    c->nx_stride = ch / 8 + 128;
    c->d_nxfp = (uint64_t *)take(8 * 8 * (size_t)c->nx_stride);
    c->d_tab = (uint64_t *)take(slots * 16);
    const int64_t capr = align_up((size_t)cap, 64);
    c->ws = 64;  // word stride inside a 64-set tile (tile-major store)
    c->d_op = (uint8_t *)take((size_t)capr);
    c->d_left = (int32_t *)take((size_t)capr * 4);
    c->d_right = (int32_t *)take((size_t)capr * 4);
    c->d_fp2 = (uint64_t *)take((size_t)capr * 8);
    c->d_key = (uint64_t *)take((size_t)capr * 8);
    c->d_store = (uint64_t *)take((size_t)capr * c->nw * 8);
    c->d_sum = (uint64_t *)take((size_t)capr * c->nsw * 32);
    if ((size_t)(p - c->arena) > budget) return LTL_E_OOM;
    return 0;
}

static hipError_t upload_pinned(ltl_ctx *c, void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    hipError_t e;
    if (bytes > c->h_stage_cap) {
        for (hipEvent_t ev : c->stage_ev)
            if (ev && (e = hipEventSynchronize(ev)) != hipSuccess) return e;
        if (c->h_stage) hipHostFree(c->h_stage);
        c->h_stage = nullptr;
        c->h_stage_cap = 0;
        const size_t cap = std::max<size_t>(bytes, (size_t)1 << 16);
        if ((e = hipHostMalloc((void **)&c->h_stage, 2 * cap, hipHostMallocDefault)) != hipSuccess) return e;
        c->h_stage_cap = cap;
    }
    const int i = c->h_stage_i;
    c->h_stage_i ^= 1;
    if (!c->stage_ev[i]) {
        if ((e = hipEventCreateWithFlags(&c->stage_ev[i], hipEventDisableTiming)) != hipSuccess) return e;
    } else if ((e = hipEventSynchronize(c->stage_ev[i])) != hipSuccess) {
        return e;
    }
    char *buf = c->h_stage + (size_t)i * c->h_stage_cap;
    memcpy(buf, src, bytes);
    if ((e = hipMemcpyAsync(dst, buf, bytes, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    return hipEventRecord(c->stage_ev[i], s);
}

// ---------------------------------------------------------------------------
// Profiling
// ---------------------------------------------------------------------------
static hipEvent_t get_event(ltl_ctx *c) {
    if (!c->free_events.empty()) {
        hipEvent_t e = c->free_events.back();
        c->free_events.pop_back();
        return e;
    }
    hipEvent_t e;
    hipEventCreate(&e);
    return e;
}

static void drain_events(ltl_ctx *c) {
    for (auto &p : c->pending) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, p.a, p.b);
        if (p.count) {
            p.items = (uint64_t)*p.count;
            p.alg = p.items * p.alg_per;
            p.min = p.items * p.min_per;
        }
        auto &s = c->stats[p.kind];
        s.launches++;
        s.seconds += ms * 1e-3;
        s.alg_bytes += p.alg;
        s.min_bytes += p.min;
        s.items += p.items;
        c->free_events.push_back(p.a);
        c->free_events.push_back(p.b);
    }
    c->pending.clear();
    c->n_counts = 0;
}

// Kernel kinds that stream the store (generation and emit): the ones
// profile = 2 times.
static bool streaming_kind(int kind) {
    return kind == K_GEN1 || kind == K_GEN2 || kind == K_GENU || kind == K_GENB || kind == K_EMIT1 || kind == K_EMIT2 ||
           kind == K_EMITU || kind == K_EMITM;
}

// Kernel kinds timed with events: profile = 1 every kind, 2 the streaming
// kinds, 3 the kinds of profile_kinds.
static bool profiled(const ltl_ctx *c, int kind) {
    switch (c->opt.profile) {
    case 1: return true;
    case 2: return streaming_kind(kind);
    case 3: return kind < 32 && ((c->opt.profile_kinds >> kind) & 1u);
    default: return false;
    }
}

// Launch, bracketed by HIP events on the library's stream when its kind is
// profiled (each timed launch costs a dispatch barrier: about 5 us, on
// 500+500 traces 1.4 ms per call with every streaming launch timed).
template <class F>
static hipError_t timed(ltl_ctx *c, int kind, uint64_t alg, uint64_t mn, uint64_t items, F &&launch) {
    if (!profiled(c, kind)) return launch();
    PendingEvent pe{kind, get_event(c), get_event(c), alg, mn, items, nullptr, 0, 0};
    hipEventRecord(pe.a, c->stream);
    hipError_t e = launch();
    hipEventRecord(pe.b, c->stream);
    c->pending.push_back(pe);
    return e;
}

// ---------------------------------------------------------------------------
// Per-instance host preparation (shared by ltl_learn and ltl_learn_batch)
// ---------------------------------------------------------------------------
namespace {

struct Prepared {
    int B, nw, nsw;
    std::vector<WordAux> aux;          // [nw]
    std::vector<SumAux> saux;          // [nsw]
    std::vector<uint64_t> atom_words;  // [nw][A]
    std::vector<uint64_t> atom_sums;   // [4*nsw][A]
    // packed layout (B = 1)
    std::vector<uint64_t> ptrace;      // per packed trace: trace | offset << 32 | length << 40
    std::vector<int> pk_word, pk_off;  // per trace: word and bit offset
    std::vector<int> order;            // summary bit -> trace (packing order)
    std::vector<int32_t> gword;        // per summary group: first word, then (at nsw + g) last word
};

int words_per_trace(const ltl_problem *pb) { return pb->words_per_trace > 1 ? pb->words_per_trace : 1; }

// Layout code: 1 = traces packed whole into words, 0 = K words per trace.
int slot_width(const ltl_problem *pb) { return words_per_trace(pb) > 1 ? 0 : 1; }

// Packed layout (B = 1): the traces go into words by best fit in decreasing
// length order; the traces of a word take its bits from the top down (the
// first at [64 - L1, 64)), each with its positions in descending bit order.
// The packing order (word by word, top down) numbers the summary bits: group
// g = packed traces 64g .. 64g + 63, which may start or end inside a word.
// Fingerprint keys per word from `seed` (kDefaultSeed unless a call is
// repeated after a fingerprint collision).
constexpr uint64_t kDefaultSeed = 0x5eed1e55c0ffee11ull;

void prepare_packed(const ltl_problem *pb, Prepared &q, uint64_t seed = kDefaultSeed) {
    const int P = pb->num_pos, W = pb->num_pos + pb->num_neg, A = pb->num_atoms;
    q.B = 1;
    q.nsw = (W + 63) / 64;
    q.aux.clear();
    q.pk_word.assign(W, 0);
    q.pk_off.assign(W, 0);
    q.ptrace.assign(W, 0);
    std::vector<int> &order = q.order;  // packed index -> trace
    order.clear();
    order.reserve(W);
    // Best fit in decreasing length order: each trace goes to the fullest
    // word it still fits in (word lists by free bits, 0..64).
    std::vector<int> idx(W);
    for (int i = 0; i < W; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return pb->lens[a] > pb->lens[b]; });
    std::vector<std::vector<int>> bins;
    std::vector<int> used;
    std::vector<std::vector<int>> by_free(65);
    for (int i : idx) {
        const int L = pb->lens[i];
        int r = L;
        while (r <= 64 && by_free[r].empty()) ++r;
        int b;
        if (r > 64) {
            b = (int)bins.size();
            bins.emplace_back();
            used.push_back(0);
        } else {
            b = by_free[r].back();
            by_free[r].pop_back();
        }
        bins[b].push_back(i);
        used[b] += L;
        if (used[b] < 64) by_free[64 - used[b]].push_back(b);
    }
    std::vector<int32_t> tword;  // packed trace -> word
    tword.reserve(W);
    for (size_t b = 0; b < bins.size(); ++b) {
        const int w = (int)q.aux.size();
        const int t0 = (int)order.size();
        for (size_t k = 0; k < bins[b].size(); ++k) tword.push_back(w);
        WordAux x;
        memset(&x, 0, sizeof(x));
        x.k1 = fp_key(seed ^ (2ull * w));
        x.k2 = fp_key(~seed ^ (2ull * w + 1));
        x.sinfo = make_sinfo((uint32_t)(t0 / 64), t0 % 64, (int)bins[b].size());
        int top = 64;
        for (int i : bins[b]) {
            const int L = pb->lens[i], o = top - L;
            top = o;
            const uint64_t m = len_mask(L) << o;
            x.lm |= m;
            x.inner |= m & ~(1ull << o);
            x.chk |= 1ull << (o + L - 1);
            if (i < P) x.tgt |= 1ull << (o + L - 1);
            q.pk_word[i] = w;
            q.pk_off[i] = o;
            q.ptrace[order.size()] = (uint64_t)(uint32_t)i | ((uint64_t)o << 32) | ((uint64_t)L << 40);
            order.push_back(i);
        }
        q.aux.push_back(x);
    }
    q.nw = (int)q.aux.size();
    q.gword.assign(2 * (size_t)q.nsw, 0);
    for (int g = 0; g < q.nsw; ++g) {
        q.gword[g] = tword[64 * g];
        q.gword[q.nsw + g] = tword[std::min(W, 64 * g + 64) - 1];
    }
    // The emit kernels shift each trace's summary bits in, in packing order:
    // the t-th of the n traces of group g sits at bit n - 1 - t of summary
    // word g.  order[] becomes summary bit -> trace.
    std::vector<int> packed(order);
    for (int g = 0; g < q.nsw; ++g) {
        const int t0 = 64 * g, n = std::min(W, t0 + 64) - t0;
        for (int p = 0; p < n; ++p) order[t0 + n - 1 - p] = packed[t0 + p];
    }
    q.saux.assign(q.nsw, SumAux{0, 0, 0});
    for (int t = 0; t < W; ++t) {
        q.saux[t / 64].val |= 1ull << (t % 64);
        if (order[t] < P) q.saux[t / 64].tgt |= 1ull << (t % 64);
        if (pb->lens[order[t]] >= 2) q.saux[t / 64].val2 |= 1ull << (t % 64);
    }
    q.atom_words.assign((size_t)q.nw * A, 0);
    const int nsw = q.nsw;
    q.atom_sums.assign((size_t)4 * nsw * A, 0);
    for (int a = 0; a < A; ++a)
        for (int t = 0; t < W; ++t) {
            const int i = order[t], L = pb->lens[i];
            const uint64_t m = len_mask(L), x = pb->atoms[(size_t)a * W + i] & m;
            q.atom_words[(size_t)q.pk_word[i] * A + a] |= (__builtin_bitreverse64(x) >> (64 - L)) << q.pk_off[i];
            const uint64_t bit = 1ull << (t % 64);
            const int j = t / 64;
            if (x & 1) q.atom_sums[(size_t)(0 * nsw + j) * A + a] |= bit;
            if (x & 2) q.atom_sums[(size_t)(1 * nsw + j) * A + a] |= bit;
            if (x) q.atom_sums[(size_t)(2 * nsw + j) * A + a] |= bit;
            if (x == m) q.atom_sums[(size_t)(3 * nsw + j) * A + a] |= bit;
        }
}

// K words per trace (B = 0): word k of trace i is word i * K + k; only word
// 0 of a trace takes part in the target check.
void prepare_multiword(const ltl_problem *pb, Prepared &q, uint64_t seed = kDefaultSeed) {
    const int P = pb->num_pos, W = pb->num_pos + pb->num_neg, A = pb->num_atoms, K = words_per_trace(pb);
    q.B = 0;
    q.nw = W * K;
    q.nsw = (W + 63) / 64;
    q.aux.assign(q.nw, WordAux{});
    for (int w = 0; w < q.nw; ++w) {
        WordAux &x = q.aux[w];
        memset(&x, 0, sizeof(x));
        const int i = w / K, k = w % K;
        x.k1 = fp_key(seed ^ (2ull * w));
        x.k2 = fp_key(~seed ^ (2ull * w + 1));
        x.lm = word_mask(pb->lens[i], k);
        x.chk = k == 0 ? 1ull : 0ull;
        x.tgt = k == 0 && i < P ? 1ull : 0ull;
    }
    q.atom_words.assign((size_t)q.nw * A, 0);
    const int nsw = q.nsw;
    q.atom_sums.assign((size_t)4 * nsw * A, 0);
    q.saux.assign(nsw, SumAux{0, 0, 0});
    for (int i = 0; i < W; ++i) {
        q.saux[i / 64].val |= 1ull << (i % 64);
        if (i < P) q.saux[i / 64].tgt |= 1ull << (i % 64);
        if (pb->lens[i] >= 2) q.saux[i / 64].val2 |= 1ull << (i % 64);
    }
    for (int a = 0; a < A; ++a)
        for (int i = 0; i < W; ++i) {
            bool nz = false, fl = true;
            uint64_t x0 = 0;
            for (int k = 0; k < K; ++k) {
                const uint64_t m = word_mask(pb->lens[i], k);
                const uint64_t x = pb->atoms[((size_t)a * W + i) * K + k] & m;
                q.atom_words[(size_t)(i * K + k) * A + a] = x;
                if (k == 0) x0 = x;
                nz |= x != 0;
                fl &= x == m;
            }
            const uint64_t bit = 1ull << (i % 64);
            const int j = i / 64;
            if (x0 & 1) q.atom_sums[(size_t)(0 * nsw + j) * A + a] |= bit;
            if (x0 & 2) q.atom_sums[(size_t)(1 * nsw + j) * A + a] |= bit;
            if (nz) q.atom_sums[(size_t)(2 * nsw + j) * A + a] |= bit;
            if (fl) q.atom_sums[(size_t)(3 * nsw + j) * A + a] |= bit;
        }
}

// Layout code B: 1 = packed words, 0 = K words per trace.
void prepare_instance(const ltl_problem *pb, int B, Prepared &q, uint64_t seed = kDefaultSeed) {
    if (B == 0) prepare_multiword(pb, q, seed);
    else prepare_packed(pb, q, seed);
}

// Validate an instance; on an accepted atom fill *res (the reference checks
// atoms on the host before any device work, code/ltli6463.cu:1100-1136).
// Returns 1 if res is final, 0 to continue, or a negative status.
int32_t precheck_host(const ltl_options &o, const ltl_problem *pb, ltl_result *res, AtomNames &names) {
    const int P = pb->num_pos, N = pb->num_neg, W = P + N, A = pb->num_atoms, K = words_per_trace(pb);
    if (P < 0 || N < 0 || W < 1 || A < 1 || A > 64 || !pb->lens || !pb->atoms || K > 4) return LTL_E_ARG;
    for (int i = 0; i < W; ++i)
        if (pb->lens[i] < 1 || pb->lens[i] > 64 * K) return LTL_E_LIMIT;
    // the reference's relaxed keys are defined on one word per trace
    if (K > 1 && o.dedup_mode != LTL_DEDUP_EXACT) return LTL_E_ARG;
    // The learned order tests every candidate of a level in its check
    // passes, which is the reference's outcome only with exact dedup
    // (relaxed keys test key-new candidates only, code/ltli6463.cu:231-243).
    if (o.order == LTL_ORDER_LEARNED && o.dedup_mode != LTL_DEDUP_EXACT) {
        int64_t len_sum = 0;
        for (int i = 0; i < W; ++i) len_sum += pb->lens[i];
        if (len_sum > 126) return LTL_E_ARG;
    }
    if (!atom_names(pb->names, A, names)) return LTL_E_ARG;
    res->cost = o.cost[0];
    std::vector<uint64_t> cs(W);
    for (int a = 0; a < A; ++a) {
        for (int i = 0; i < W; ++i) cs[i] = pb->atoms[((size_t)a * W + i) * K] & word_mask(pb->lens[i], 0);
        if (ltl_mismatches(cs.data(), P, N) <= o.hamming_k) {
            res->status = LTL_FOUND;
            res->all_ltls = (uint64_t)a + 1;
            res->found_id = a;
            res->n_nodes = 1;
            res->nodes = (ltl_node *)malloc(sizeof(ltl_node));
            res->nodes[0] = {LTL_ATOM, a, -1, -1};
            res->formula = strdup(names[a].c_str());
            return 1;
        }
    }
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Enumeration
// ---------------------------------------------------------------------------
namespace {

struct Run {
    ltl_ctx *c;
    const ltl_problem *pb;
    ChunkArgs base{};
    uint64_t all_ltls = 0;
    uint64_t candidates = 0, alg_bytes = 0;
    bool otf = false;         // overflow seen at a previous level end
    bool check_only = false;  // blocks are tested only (precheck pass)
    bool count_only = false;  // blocks are only counted (candidates of a level)
    uint64_t counted = 0;
    bool ranked = false;      // check pass: binary blocks go to the ranked rounds
    bool logging = false;     // check pass: binary blocks are also logged
    bool skip_binary = false; // check pass: binary blocks were tested ahead of the level
    bool only_binary = false; // check pass: test the binary blocks only (lookahead)
    // Plain order: the binary blocks of this check pass were tested by the
    // lookahead ahead of the level; they are registered without a launch,
    // and the lookahead's first hit (index among the pass's binary blocks,
    // offset in the block; -1: none) becomes a candidate id of this pass.
    // Likewise for the lookahead's smallest candidate whose complement is
    // accepted (la_neg_block, la_neg_off -> la_neg).
    bool reuse_binary = false;
    int64_t la_block = -1, la_neg_block = -1;
    uint64_t la_off = 0, la_neg_off = 0;
    int64_t bin_seen = 0;             // binary blocks registered in the pass
    uint64_t la_found = ~0ull;        // the hit's id in this pass
    uint64_t la_neg = ~0ull;
    uint64_t checked = 0;
    std::vector<Block> blocks;
    std::vector<LBlock> lblk;  // binary blocks of the current check pass
    int64_t n_ent = 0;         // their tile pairs
    int cur_level = 0;         // level of the blocks being registered
    // Deferred level [defer_s, defer_e) (stored without words and
    // summaries): check passes do not launch the unary blocks over it; the
    // successors were tested in its generation (succ_mask), and the blocks
    // are listed here as (first candidate id, successor index).
    int64_t defer_s = -1, defer_e = -1;
    int succ_mask = 0;
    std::vector<std::pair<uint64_t, int>> succ_blocks;
    int64_t level_stored = 0;  // sets stored when the current level started (launch log)

    // A block of candidates, in enumeration order; `launched`: its kernels
    // are queued in this pass.
    void reg(const Block &b, bool launched = true) {
        blocks.push_back(b);
        if (c->launch_fn) {
            const ltl_launch_record rec{b.level,     b.op,         check_only ? 0 : 1, launched ? 1 : 0,
                                        b.id_base, level_stored, b.size};
            c->launch_fn(&rec, c->launch_user);
        }
    }

    // Check-only pass over one block: no temporaries, so no chunking beyond
    // keeping the grid within limits.
    int32_t run_check(int op, ChunkArgs a) {
        ltl_ctx *c = this->c;
        const bool bin = op == LTL_AND || op == LTL_OR || op == LTL_UNTIL;
        const uint64_t sv = (uint64_t)c->nsw * 8, cs = (uint64_t)c->nw * 8;
        uint64_t cand, alg, mn;
        int kind;
        if (bin) {
            a.ush = (int32_t)((a.rs + a.b0) & 63);
            a.nbs = (int32_t)((a.b1 - a.b0 + a.ush + 63) / 64);
            a.n_tiles = (int32_t)((a.a1 - a.a0) * a.nbs);
            const uint64_t pairs = (uint64_t)(a.a1 - a.a0) * (uint64_t)(a.b1 - a.b0);
            const uint64_t per = op == LTL_UNTIL ? cs : sv;
            cand = op == LTL_UNTIL ? 2 * pairs : pairs;
            alg = pairs * 2 * per;
            mn = (uint64_t)(a.a1 - a.a0 + a.b1 - a.b0) * per;
            kind = op == LTL_UNTIL ? K_CHKU : K_CHK2;
        } else {
            a.nbs = 1;
            a.ush = (int32_t)((a.ls + a.a0) & 63);
            a.n_tiles = (int32_t)((a.a1 - a.a0 + a.ush + 63) / 64);
            cand = (uint64_t)(a.a1 - a.a0);
            alg = mn = cand * sv;
            kind = K_CHK1;
        }
        checked += cand;
        HIP_TRY(timed(c, kind, alg, mn, cand, [&] { return launch_check(op, c->B, a, c->stream); }));
        return 0;
    }

    int32_t run_chunk(int op, ChunkArgs a) {
        ltl_ctx *c = this->c;
        a.chunk_idx = ++c->chunk_counter;
        const bool bin = op == LTL_AND || op == LTL_OR || op == LTL_UNTIL;
        const int nd = op == LTL_UNTIL ? 2 : 1;
        uint64_t cand, alg, mn;
        const uint64_t cs = (uint64_t)c->nw * 8;
        if (bin) {
            a.ush = (int32_t)((a.rs + a.b0) & 63);
            a.nbs = (int32_t)((a.b1 - a.b0 + a.ush + 63) / 64);
            a.n_tiles = (int32_t)((a.a1 - a.a0) * a.nbs);
            const uint64_t pairs = (uint64_t)(a.a1 - a.a0) * (uint64_t)(a.b1 - a.b0);
            cand = pairs * nd;
            alg = pairs * 2 * cs + cand * 20;
            mn = (uint64_t)(a.a1 - a.a0 + a.b1 - a.b0) * cs + cand * 20;
        } else {
            a.nbs = 1;
            a.ush = (int32_t)((a.ls + a.a0) & 63);
            a.n_tiles = (int32_t)((a.a1 - a.a0 + a.ush + 63) / 64);
            cand = (uint64_t)(a.a1 - a.a0);
            alg = cand * (cs + 20);
            mn = alg;
        }
        candidates += cand;
        alg_bytes += alg;
        const int gk = op == LTL_UNTIL ? K_GENU : bin ? K_GEN2 : K_GEN1;
        HIP_TRY(timed(c, gk, alg, mn, cand, [&] { return launch_gen(op, c->B, a, c->stream); }));
        if (otf) return 0;
        return store_chunk(op, a, cand);
    }

    // decide, scan, finish, compact and emit of a generated chunk (its
    // temporaries at a.t_slot, a.t_fp2, ...).
    int32_t store_chunk(int op, const ChunkArgs &a, uint64_t cand) {
        ltl_ctx *c = this->c;
        const bool bin = op == LTL_AND || op == LTL_OR || op == LTL_UNTIL;
        const uint64_t cs = (uint64_t)c->nw * 8;
        const int ek = a.deferred ? K_EMITM : op == LTL_UNTIL ? K_EMITU : bin ? K_EMIT2 : K_EMIT1;
        // Small chunks: one workgroup; mid-size chunks: one look-back launch;
        // large ones: decide, device scan, finish and compact (LTL_DEDUP_SCAN
        // in the environment: the four-kernel path for every chunk).
        static const bool scan_path = getenv("LTL_DEDUP_SCAN") != nullptr;
        if (!scan_path && a.n_tiles <= kSmallTiles) {
            // (one workgroup: the launches, not the work, would set the time)
            HIP_TRY(timed(c, K_DECIDE, cand * 28, cand * 28, cand, [&] { return launch_dedup_small(op, a, c->stream); }));
            c->stats[K_PSMALL].launches++;
            c->stats[K_PSMALL].items += cand;
        } else if (!scan_path && a.n_tiles <= kLbMaxTiles && c->lb_epoch + 1 < (1u << 30)) {
            ChunkArgs b = a;
            b.lb_epoch = ++c->lb_epoch;
            HIP_TRY(timed(c, K_DECIDE, cand * 28, cand * 28, cand, [&] { return launch_dedup_lb(op, b, c->stream); }));
            c->stats[K_PLB].launches++;
            c->stats[K_PLB].items += cand;
        } else {
            c->stats[K_PFOUR].launches++;
            c->stats[K_PFOUR].items += cand;
            HIP_TRY(timed(c, K_DECIDE, cand * 28, cand * 28, cand, [&] { return launch_decide(op, a, c->stream); }));
            HIP_TRY(timed(c, K_SCAN, (uint64_t)a.n_tiles * 8, (uint64_t)a.n_tiles * 8, (uint64_t)a.n_tiles, [&] {
                hipError_t e = launch_scan(a.t_cnt, a.t_off, a.n_tiles, c->d_scan, c->scan_bytes, c->stream);
                if (e != hipSuccess) return e;
                return launch_finish(a, c->stream);
            }));
            HIP_TRY(timed(c, K_SCAN, 0, 0, 0, [&] { return launch_compact(op, a, c->stream); }));
        }
        // relaxed keys: the accepted survivors (and, in the chunk that
        // overflowed the store, every accepted candidate after the first
        // survivor that did not fit) lower the result
        if (a.kmode) HIP_TRY(timed(c, K_DECIDE, 0, 0, 0, [&] { return launch_rx_tail(op, a, c->stream); }));
        HIP_TRY(timed(c, ek, 0, 0, 0, [&] { return launch_emit(op, c->B, a, c->stream); }));
        if (profiled(c, ek) && c->n_counts < kCountSlots) {
            // survivors of this chunk -> bytes of its emit (read after the level sync)
            long long *slot = c->h_counts + c->n_counts++;
            HIP_TRY(hipMemcpyAsync(slot, &c->d_state->chunk_total, sizeof(long long), hipMemcpyDeviceToHost,
                                   c->stream));
            PendingEvent &pe = c->pending.back();
            pe.count = slot;
            // operand words read, words and summaries written, metadata (a
            // deferred level: metadata and the successor flag only)
            const uint64_t sv = (uint64_t)c->nsw * 32;
            if (a.deferred) {
                pe.alg_per = pe.min_per = 17 + 8 + 4 + 12 + (a.t_succ ? 1 : 0);
            } else {
                pe.alg_per = (bin ? 2 : 1) * cs + cs + sv + 17 + 8 + 4 + 12;
                pe.min_per = cs + sv + 17 + 8 + 4 + 12;
            }
        }
        return 0;
    }

    // X, F and G of one operand level (equal costs) in one generation pass,
    // then the three blocks' decide ... emit one after the other, in
    // canonical order; each block's kernels stop only for an accepted
    // candidate up to its own end.  Falls back to three passes when the
    // three blocks do not fit one chunk's temporaries, in check-only or
    // counting passes, and for relaxed keys.
    int32_t unary_xfg(int64_t s, int64_t e) {
        const int64_t n = e - s;
        const int64_t tiles = (n + 63 + 63) / 64, slots = tiles * 64;
        if (n <= 0 || check_only || count_only || base.kmode || 3 * slots > c->CH) {
            int32_t r = unary(LTL_NEXT, s, e);
            if (!r) r = unary(LTL_FIN, s, e);
            if (!r) r = unary(LTL_GLOB, s, e);
            return r;
        }
        const int ops[3] = {LTL_NEXT, LTL_FIN, LTL_GLOB};
        for (int j = 0; j < 3; ++j) reg({all_ltls + (uint64_t)j * n, (uint64_t)n, ops[j], s, n, 0, 1, cur_level});
        ChunkArgs a = base;
        a.id_base = all_ltls;
        a.ls = s; a.rs = 0; a.nR = 1;
        a.a0 = 0; a.a1 = n;
        a.b0 = a.b1 = 0;
        a.nbs = 1;
        a.ush = (int32_t)((a.ls + a.a0) & 63);
        a.n_tiles = (int32_t)((n + a.ush + 63) / 64);
        a.chunk_idx = ++c->chunk_counter;
        const uint64_t cs = (uint64_t)c->nw * 8;
        const uint64_t alg = (uint64_t)n * (cs + 60);
        candidates += 3 * (uint64_t)n;
        alg_bytes += alg;
        HIP_TRY(timed(c, K_GEN1, alg, alg, 3 * (uint64_t)n,
                      [&] { return launch_gen_xfg(a, n, slots, false, 0, c->stream); }));
        if (!otf)
            for (int j = 0; j < 3; ++j) {
                ChunkArgs b = a;
                b.id_base = all_ltls + (uint64_t)j * n;
                b.id_end = b.id_base + (uint64_t)n;
                if (j) b.chunk_idx = ++c->chunk_counter;
                b.t_slot = c->t_slot + j * slots;
                b.t_fp2 = c->t_fp2 + j * slots;
                b.t_list = c->t_list + j * slots;
                b.t_mask = c->t_mask + j * tiles;
                b.t_cnt = c->t_cnt + j * tiles;
                b.t_off = c->t_off + j * tiles;
                if (a.t_succ) b.t_succ = c->t_succ + j * slots;
                if (int32_t r = store_chunk(ops[j], b, (uint64_t)n)) return r;
            }
        all_ltls += 3 * (uint64_t)n;
        return 0;
    }

    // Chunk temporaries from offset `off` on (candidates; tiles = off / 64):
    // the regions below it hold the X, F and G candidates of a fused pass
    // while the & and | blocks run.
    int64_t tmp_off = 0;
    void set_tmp_off(int64_t off) {
        tmp_off = off;
        base.t_slot = c->t_slot + off;
        base.t_fp2 = c->t_fp2 + off;
        base.t_list = c->t_list + off;
        base.t_mask = c->t_mask + off / 64;
        base.t_cnt = c->t_cnt + off / 64;
        base.t_off = c->t_off + off / 64;
        if (base.t_succ) base.t_succ = c->t_succ + off;
        if (base.t_acc) base.t_acc = c->t_succ + off;
    }
    int64_t ch_cap() const { return c->CH - tmp_off; }

    // ~, X, F and G of one operand level in one generation pass (a tested
    // level, equal costs): the negation block, first in the level, is
    // decided and emitted at once; the X, F and G candidates wait in their
    // temporaries (regions 1-3) until finish_xfg, after the & and | blocks
    // that come between them in canonical order.  Their ids (from id_xfg)
    // are larger than every & and | id, so the table's smallest-id rule
    // gives the outcome of the canonical sequence.
    // First store index of the level whose emits computed the next level's
    // ~/X/F/G fingerprints (ChunkArgs::nx_fp), -1: none; unary_nxfg over that
    // level takes them instead of generating the four blocks.
    int64_t nx_from = -1;
    struct PendXfg {
        bool on = false;
        int64_t n = 0, tiles = 0, slots = 0;
        uint64_t id_xfg = 0;
        ChunkArgs a{};
    } pend;
    int32_t unary_nxfg(int64_t s, int64_t e, uint64_t id_xfg) {
        const int64_t n = e - s;
        const int64_t tiles = (n + 63 + 63) / 64, slots = tiles * 64;
        reg({all_ltls, (uint64_t)n, LTL_NOT, s, n, 0, 1, cur_level});
        ChunkArgs a = base;
        a.id_base = all_ltls;
        a.ls = s; a.rs = 0; a.nR = 1;
        a.a0 = 0; a.a1 = n;
        a.b0 = a.b1 = 0;
        a.nbs = 1;
        a.ush = (int32_t)((a.ls + a.a0) & 63);
        a.n_tiles = (int32_t)((n + a.ush + 63) / 64);
        a.chunk_idx = ++c->chunk_counter;
        const uint64_t cs = (uint64_t)c->nw * 8;
        const uint64_t alg = (uint64_t)n * (cs + 80);
        candidates += 4 * (uint64_t)n;
        alg_bytes += alg;
        // (the operand level's emits left the four blocks' fingerprints: no
        // second read of the level, only the inserts)
        const bool from_fp = nx_from == s && !a.t_succ;
        nx_from = -1;
        HIP_TRY(timed(c, K_GEN1, alg, alg, 4 * (uint64_t)n, [&] {
            return from_fp ? launch_insert_xfg(a, n, slots, id_xfg, c->d_nxfp, c->nx_stride, c->stream)
                           : launch_gen_xfg(a, n, slots, true, id_xfg, c->stream);
        }));
        ChunkArgs b = a;
        b.id_end = b.id_base + (uint64_t)n;
        if (int32_t r = store_chunk(LTL_NOT, b, (uint64_t)n)) return r;
        all_ltls += (uint64_t)n;
        pend = PendXfg{true, n, tiles, slots, id_xfg, a};
        set_tmp_off(4 * slots);
        return 0;
    }
    int32_t finish_xfg() {
        set_tmp_off(0);
        if (!pend.on) return 0;
        pend.on = false;
        const int64_t n = pend.n, s = pend.a.ls;
        if (all_ltls != pend.id_xfg) {  // the & and | blocks did not add up
            report_hip(hipErrorUnknown, "finish_xfg", __LINE__);
            return LTL_E_HIP;
        }
        const int ops[3] = {LTL_NEXT, LTL_FIN, LTL_GLOB};
        for (int j = 0; j < 3; ++j) reg({all_ltls + (uint64_t)j * n, (uint64_t)n, ops[j], s, n, 0, 1, cur_level});
        for (int j = 0; j < 3; ++j) {
            ChunkArgs b = pend.a;
            b.id_base = all_ltls + (uint64_t)j * n;
            b.id_end = b.id_base + (uint64_t)n;
            b.chunk_idx = ++c->chunk_counter;
            b.t_slot = c->t_slot + (j + 1) * pend.slots;
            b.t_fp2 = c->t_fp2 + (j + 1) * pend.slots;
            b.t_list = c->t_list + (j + 1) * pend.slots;
            b.t_mask = c->t_mask + (j + 1) * pend.tiles;
            b.t_cnt = c->t_cnt + (j + 1) * pend.tiles;
            b.t_off = c->t_off + (j + 1) * pend.tiles;
            if (pend.a.t_succ) b.t_succ = c->t_succ + (j + 1) * pend.slots;
            if (int32_t r = store_chunk(ops[j], b, (uint64_t)n)) return r;
        }
        all_ltls += 3 * (uint64_t)n;
        return 0;
    }

    // Temporaries of a chunk at candidate slot `off` (the region scheme of
    // set_tmp_off, for one chunk's arguments).
    void region(ChunkArgs &a, int64_t off) const {
        a.t_slot = c->t_slot + off;
        a.t_fp2 = c->t_fp2 + off;
        a.t_list = c->t_list + off;
        a.t_mask = c->t_mask + off / 64;
        a.t_cnt = c->t_cnt + off / 64;
        a.t_off = c->t_off + off / 64;
        a.t_succ = base.t_succ ? c->t_succ + off : nullptr;
    }

    // Binary blocks of a tested level with equal costs of &, | and U
    // (level_blocks, fused path): per split (li, ri) one generation pass for
    // its & and | blocks and, with_u, its U block; the & block is decided
    // and emitted at once (every earlier block is committed), the | and U
    // blocks wait in their regions for their canonical positions
    // (finish_or after the last & block, finish_u after X, F and G).  Every
    // block is one chunk, so the whole level's binary candidates must fit the
    // temporaries (fused_fits).  id_u: the first U id of the level.
    struct PendBin {
        int op;
        ChunkArgs a;
        uint64_t size;
    };
    std::vector<PendBin> pend_or, pend_u;
    struct Split2 {
        int li, ri;
        int64_t ls, nL, rs, nR, slots;
    };
    std::vector<Split2> splits_of(int cost, int c1, int cb) const {
        std::vector<Split2> v;
        for (int i = c1; 2 * i <= cost - cb; ++i) {
            const int j = cost - i - cb;
            const int64_t ls = c->lvl_s[i], nL = c->lvl_e[i] - ls, rs = c->lvl_s[j], nR = c->lvl_e[j] - rs;
            if (nL <= 0 || nR <= 0) continue;
            const int64_t nbs = (nR + (rs & 63) + 63) / 64;
            v.push_back({i, j, ls, nL, rs, nR, nL * nbs * 64});
        }
        return v;
    }
    // Candidate slots the fused blocks of the level need (with_u: U too).
    int64_t fused_slots(const std::vector<Split2> &v, bool with_u) const {
        int64_t t = 0;
        for (const auto &sp : v) t += (with_u ? 4 : 2) * sp.slots;
        return t;
    }
    int32_t binary_fused(const std::vector<Split2> &v, bool with_u, uint64_t id_u) {
        uint64_t and_total = 0;
        for (const auto &sp : v) and_total += (uint64_t)sp.nL * (uint64_t)sp.nR;
        uint64_t id_or = all_ltls + and_total;
        int64_t off = tmp_off;
        const uint64_t cs = (uint64_t)c->nw * 8;
        for (const auto &sp : v) {
            const uint64_t size = (uint64_t)sp.nL * (uint64_t)sp.nR;
            ChunkArgs a = base;
            a.id_base = all_ltls;
            a.ls = sp.ls; a.rs = sp.rs; a.nR = sp.nR;
            a.a0 = 0; a.a1 = sp.nL;
            a.b0 = 0; a.b1 = sp.nR;
            a.ush = (int32_t)(sp.rs & 63);
            a.nbs = (int32_t)((sp.nR + a.ush + 63) / 64);
            a.n_tiles = (int32_t)(sp.nL * a.nbs);
            a.chunk_idx = ++c->chunk_counter;
            region(a, off);
            const FusedBin f{id_or, id_u, sp.slots, 2 * sp.slots};
            const uint64_t cand = (with_u ? 4 : 2) * size;
            const uint64_t alg = size * 2 * cs + cand * 20;
            const uint64_t mn = (uint64_t)(sp.nL + sp.nR) * cs + cand * 20;
            candidates += cand;
            alg_bytes += alg;
            HIP_TRY(timed(c, K_GENB, alg, mn, cand,
                          [&] { return launch_gen_fused(a, f, with_u, c->stream); }));
            reg({all_ltls, size, LTL_AND, sp.ls, sp.nL, sp.rs, sp.nR, cur_level});
            if (!otf)
                if (int32_t r = store_chunk(LTL_AND, a, size)) return r;
            all_ltls += size;
            ChunkArgs b = a;
            b.id_base = id_or;
            region(b, off + sp.slots);
            pend_or.push_back({LTL_OR, b, size});
            if (with_u) {
                ChunkArgs u = a;
                u.id_base = id_u;
                region(u, off + 2 * sp.slots);
                pend_u.push_back({LTL_UNTIL, u, 2 * size});
            }
            off += (with_u ? 4 : 2) * sp.slots;
            id_or += size;
            id_u += 2 * size;
        }
        return 0;
    }
    // Decide ... emit of the pending blocks of one kind, in order.
    int32_t finish_pending(std::vector<PendBin> &v) {
        for (auto &p : v) {
            reg({all_ltls, p.size, p.op, p.a.ls, p.a.a1 - p.a.a0, p.a.rs, p.a.nR, cur_level});
            if (all_ltls != p.a.id_base) {  // (ids fixed at generation must be the canonical ones)
                report_hip(hipErrorUnknown, "finish_pending", __LINE__);
                return LTL_E_HIP;
            }
            p.a.chunk_idx = ++c->chunk_counter;
            if (!otf)
                if (int32_t r = store_chunk(p.op, p.a, p.size)) return r;
            all_ltls += p.size;
        }
        v.clear();
        return 0;
    }

    int32_t unary(int op, int64_t s, int64_t e) {
        const int64_t n = e - s;
        if (n <= 0 || (check_only && only_binary)) return 0;
        if (count_only) {
            counted += (uint64_t)n;
            return 0;
        }
        const bool deferred_succ = check_only && s == defer_s && e == defer_e;
        reg({all_ltls, (uint64_t)n, op, s, n, 0, 1, cur_level}, !deferred_succ);
        if (deferred_succ) {
            const int j = op == LTL_NOT ? 0 : op == LTL_NEXT ? 1 : op == LTL_FIN ? 2 : 3;
            succ_blocks.push_back({all_ltls, j});
            checked += (uint64_t)n;
            all_ltls += (uint64_t)n;
            return 0;
        }
        // (a chunk's temporaries hold CH candidates, the tile shift up to 63)
        const int64_t step = check_only ? ((int64_t)1 << 36) : ch_cap() - 64;
        for (int64_t a0 = 0; a0 < n; a0 += step) {
            ChunkArgs a = base;
            a.id_base = all_ltls;
            a.ls = s; a.rs = 0; a.nR = 1;
            a.a0 = a0; a.a1 = std::min<int64_t>(n, a0 + step);
            a.b0 = a.b1 = 0;
            if (int32_t r = check_only ? run_check(op, a) : run_chunk(op, a)) return r;
        }
        all_ltls += (uint64_t)n;
        return 0;
    }

    // Binary block of left level li and right level ri.
    int32_t binary(int op, int li, int ri) {
        const int64_t ls = c->lvl_s[li], le = c->lvl_e[li], rs = c->lvl_s[ri], re = c->lvl_e[ri];
        const int64_t nL = le - ls, nR = re - rs;
        if (nL <= 0 || nR <= 0 || (check_only && skip_binary)) return 0;
        const int nd = op == LTL_UNTIL ? 2 : 1;
        const uint64_t size = (uint64_t)nL * (uint64_t)nR * nd;
        if (count_only) {
            counted += size;
            return 0;
        }
        reg({all_ltls, size, op, ls, nL, rs, nR, cur_level}, !(check_only && reuse_binary));
        if (check_only && reuse_binary) {
            if (bin_seen == la_block) la_found = all_ltls + la_off;
            if (bin_seen == la_neg_block) la_neg = all_ltls + la_neg_off;
            ++bin_seen;
            all_ltls += size;
            return 0;
        }
        if (check_only && (ranked || logging)) {
            const int64_t nTR = (nR + 63) / 64;
            lblk.push_back({op, li, ri, 0, ls, nL, rs, nR, all_ltls, nTR, n_ent});
            n_ent += (nL + 63) / 64 * nTR;
            if (ranked) {
                all_ltls += size;
                return 0;
            }
        }
        if (check_only) {
            const int64_t nbs = (nR + 63 + 63) / 64;  // with the lane shift
            const int64_t rows = std::max<int64_t>(4, ((int64_t)1 << 30) / nbs * 4);
            for (int64_t a0 = 0; a0 < nL; a0 += rows) {
                ChunkArgs a = base;
                a.id_base = all_ltls;
                a.ls = ls; a.rs = rs; a.nR = nR;
                a.a0 = a0; a.a1 = std::min<int64_t>(nL, a0 + rows);
                a.b0 = 0; a.b1 = nR;
                if (int32_t r = run_check(op, a)) return r;
            }
            all_ltls += size;
            return 0;
        }
        const int64_t row_slots = (nR + 63 + 63) / 64 * 64 * nd;  // stripes of a row, with the lane shift
        if (row_slots <= ch_cap()) {
            const int64_t rows = std::max<int64_t>(1, ch_cap() / row_slots);
            for (int64_t a0 = 0; a0 < nL; a0 += rows) {
                ChunkArgs a = base;
                a.id_base = all_ltls;
                a.ls = ls; a.rs = rs; a.nR = nR;
                a.a0 = a0; a.a1 = std::min<int64_t>(nL, a0 + rows);
                a.b0 = 0; a.b1 = nR;
                if (int32_t r = run_chunk(op, a)) return r;
            }
        } else {
            const int64_t cols = std::max<int64_t>(64, (ch_cap() / nd - 64) / 64 * 64);
            for (int64_t a0 = 0; a0 < nL; ++a0)
                for (int64_t b0 = 0; b0 < nR; b0 += cols) {
                    ChunkArgs a = base;
                    a.id_base = all_ltls;
                    a.ls = ls; a.rs = rs; a.nR = nR;
                    a.a0 = a0; a.a1 = a0 + 1;
                    a.b0 = b0; a.b1 = std::min<int64_t>(nR, b0 + cols);
                    if (int32_t r = run_chunk(op, a)) return r;
                }
        }
        all_ltls += size;
        return 0;
    }
};

struct Meta {
    int op;
    int32_t l, r;
};

int32_t read_meta(ltl_ctx *c, int64_t k, Meta &m) {
    uint8_t op;
    HIP_TRY(hipMemcpy(&op, c->d_op + k, 1, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&m.l, c->d_left + k, 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&m.r, c->d_right + k, 4, hipMemcpyDeviceToHost));
    m.op = op;
    return 0;
}


// The nodes below the root as the walk of build_tree visits them, read on the
// device (one launch and one copy); walk[0] < 0 when the walk did not finish.
int32_t tree_walk(ltl_ctx *c, int32_t l, int32_t r, int64_t n_stored, std::vector<int32_t> &walk) {
    HIP_TRY(launch_tree(c->d_op, c->d_left, c->d_right, n_stored, l, r, c->d_tree, kTreeCap, c->stream));
    walk.assign(1 + 3 * (size_t)kTreeCap, 0);
    HIP_TRY(hipMemcpyAsync(walk.data(), c->d_tree, walk.size() * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// Build the node array (breadth first) from the root triple and the store's
// parent indices; store indices below the atom count are atoms.
int32_t build_tree(ltl_ctx *c, int op, int32_t l, int32_t r, int atom, std::vector<ltl_node> &nodes,
                   std::vector<int64_t> &src, int64_t n_stored) {
    nodes.clear();
    src.clear();
    // The nodes' metadata from a walk on the device; should it not finish,
    // small stores: the parent arrays in one copy each, else three
    // synchronous copies per node.
    std::vector<int32_t> walk;
    if (op != LTL_ATOM && n_stored > 0)
        if (int32_t e = tree_walk(c, l, r, n_stored, walk)) return e;
    const bool walked = !walk.empty() && walk[0] >= 0;
    size_t wi = 0;
    std::vector<uint8_t> ops;
    std::vector<int32_t> lefts, rights;
    if (!walked && n_stored > 0 && n_stored <= ((int64_t)1 << 20)) {
        ops.resize(n_stored);
        lefts.resize(n_stored);
        rights.resize(n_stored);
        HIP_TRY(hipMemcpy(ops.data(), c->d_op, (size_t)n_stored, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(lefts.data(), c->d_left, (size_t)n_stored * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(rights.data(), c->d_right, (size_t)n_stored * 4, hipMemcpyDeviceToHost));
    }
    nodes.push_back({op, op == LTL_ATOM ? atom : -1, -1, -1});
    src.push_back(-1);
    struct Q { int node; int32_t l, r; };
    std::vector<Q> q{{0, l, r}};
    for (size_t h = 0; h < q.size(); ++h) {
        Q cur = q[h];
        int32_t kids[2] = {cur.l, cur.r};
        for (int s = 0; s < 2; ++s) {
            if (kids[s] < 0) continue;
            Meta m;
            if (walked) {
                if ((int64_t)wi >= walk[0]) return LTL_E_HIP;  // (the walks differ: not expected)
                m.op = walk[1 + 3 * wi];
                m.l = walk[2 + 3 * wi];
                m.r = walk[3 + 3 * wi];
                ++wi;
            } else if (kids[s] < (int64_t)ops.size()) {
                m.op = ops[kids[s]];
                m.l = lefts[kids[s]];
                m.r = rights[kids[s]];
            } else if (int32_t e = read_meta(c, kids[s], m)) {
                return e;
            }
            const int idx = (int)nodes.size();
            nodes.push_back({m.op, m.op == LTL_ATOM ? kids[s] : -1, -1, -1});
            src.push_back(kids[s]);
            if (s == 0) nodes[cur.node].left = idx; else nodes[cur.node].right = idx;
            if (m.op != LTL_ATOM) q.push_back({idx, m.l, m.r});
            if (nodes.size() > 100000) return LTL_E_LIMIT;
        }
    }
    return 0;
}

std::string node_string(const std::vector<ltl_node> &n, int i, const AtomNames &names) {
    const ltl_node &x = n[i];
    switch (x.op) {
    case LTL_ATOM: return names[x.atom];
    case LTL_NOT: return "~(" + node_string(n, x.left, names) + ")";
    case LTL_NEXT: return "X(" + node_string(n, x.left, names) + ")";
    case LTL_FIN: return "F(" + node_string(n, x.left, names) + ")";
    case LTL_GLOB: return "G(" + node_string(n, x.left, names) + ")";
    default: {
        const char *mid = x.op == LTL_AND ? ")&(" : x.op == LTL_OR ? ")|(" : ")U(";
        return "(" + node_string(n, x.left, names) + mid + node_string(n, x.right, names) + ")";
    }
    }
}

}  // namespace

// One learning call.  noalloc: the device memory of the call is this
// device's static scratch region (no allocation of any kind on the device,
// no pinned staging, no events); plain order only.  seed: the fingerprint
// keys; weak_fp1: (tests only) fingerprint 1 keyed with zeros, which makes
// its collisions between different sets common.
static int32_t learn_once(ltl_ctx *c, const ltl_problem *pb, ltl_result *res, bool noalloc, uint64_t seed,
                          bool weak_fp1) {
    if (!res) return LTL_E_ARG;
    memset(res, 0, sizeof(*res));
    res->found_id = -1;
    if (!c || !pb) return res->status = LTL_E_ARG;
    const double t0 = now_s();
    const ltl_options &o = c->opt;
    if (o.order == LTL_ORDER_LEARNED && !c->scorer) return res->status = LTL_E_ARG;  // no scorer attached
    // (the learned order, the log and profiling allocate buffers or events)
    if (noalloc && (o.order != LTL_ORDER_PLAIN || c->log_cap > 0 || o.profile)) return res->status = LTL_E_ARG;
    if (noalloc && !c->scratch) return res->status = LTL_E_HIP;
    // An error left by an earlier HIP call of this thread (kernel launches
    // are checked with hipGetLastError) is cleared here, reported with
    // LTL_DEBUG.
    if (hipError_t stale = hipGetLastError(); stale != hipSuccess) report_hip(stale, "an earlier call", __LINE__);
    AtomNames names;
    for (auto &s : c->stats) { s.launches = 0; s.seconds = 0; s.alg_bytes = s.min_bytes = s.items = 0; }
    {
        const int32_t pre = precheck_host(o, pb, res, names);
        if (pre < 0) return res->status = pre;
        if (pre == 1) {
            res->seconds = now_s() - t0;
            c->W = 0;
            return res->status;
        }
    }
    const int P = pb->num_pos, N = pb->num_neg, W = P + N, A = pb->num_atoms;
    (void)N;
    int maxlen = 0;
    for (int i = 0; i < W; ++i) maxlen = std::max<int>(maxlen, pb->lens[i]);
    const int c1 = o.cost[0], c2 = o.cost[1], c3 = o.cost[2], c4 = o.cost[3];
    const int c5 = o.cost[4], c6 = o.cost[5], c7 = o.cost[6], c8 = o.cost[7];
    const int maxc = o.max_cost;
    const uint64_t all_ltls = (uint64_t)A;  // atoms (code/ltli6463.cu:1134)

    // Layout: traces packed whole into words (B = 1), or K words per trace
    // (B = 0) for traces longer than 64 positions.  Per-word keys and masks,
    // summary masks, packed atoms and their summaries.
    if (hipSetDevice(o.device) != hipSuccess) return res->status = LTL_E_HIP;
    (void)maxlen;
    const int K = words_per_trace(pb);
    const int B = K > 1 ? 0 : 1;
    Prepared prep;
    prepare_instance(pb, B, prep, seed);
    if (weak_fp1)
        for (auto &x : prep.aux) x.k1 = 0;
    c->B = B; c->K = K; c->W = W; c->P = P; c->A = A;
    c->nw = prep.nw;
    c->nsw = prep.nsw;
    c->pk_word.swap(prep.pk_word);
    if (B == 1) c->sum_order.swap(prep.order);
    else {
        c->sum_order.resize(W);
        for (int i = 0; i < W; ++i) c->sum_order[i] = i;
    }
    c->pk_off.swap(prep.pk_off);
    c->lens.assign(pb->lens, pb->lens + W);
    c->lazy_s = c->lazy_e = 0;
    c->lzw_s = c->lzw_e = 0;
    if (noalloc) {
        c->arena = c->scratch;
        c->arena_bytes = LTL_NOALLOC_SCRATCH_BYTES;
        // an instance whose buffers do not fit the scratch region is beyond
        // this entry point's limits
        if (carve(c) || c->capacity < A) {
            report_hip(hipErrorOutOfMemory, "ltl_learn_noalloc: instance beyond the scratch region", __LINE__);
            return res->status = LTL_E_LIMIT;
        }
    } else {
        if (int32_t e = ensure_pool(c)) return res->status = e;
        c->arena = c->pool;
        c->arena_bytes = c->pool_bytes;
        c->clean_n = 0;  // (the batched launch's regions are overwritten)
        if (int32_t e = carve(c)) return res->status = e;
        if (c->capacity < A) return res->status = LTL_E_OOM;
    }
    const int nsw = c->nsw;
    std::vector<WordAux> &aux = prep.aux;
    std::vector<SumAux> &saux = prep.saux;
    std::vector<uint64_t> &atom_words = prep.atom_words, &sum0 = prep.atom_sums;
    DevState st0{};
    st0.found = kEmpty;
    st0.rx_found = kEmpty;
    st0.neg_found[0] = st0.neg_found[1] = kEmpty;
    for (auto &m : st0.succ_min) m = kEmpty;
    st0.n_stored = A;
    st0.overflow_chunk = LLONG_MAX;
    hipStream_t s = c->stream;
    // The call's uploads go through one pinned buffer (asynchronous copies
    // from pageable memory would each hold the host); the stream is idle
    // here, so the buffer is free.
    const size_t init_bytes = align_up(aux.size() * sizeof(WordAux), 256) + align_up(saux.size() * sizeof(SumAux), 256) +
                              align_up((atom_words.size() + sum0.size()) * 8, 256) + align_up(sizeof(st0), 256) +
                              align_up((size_t)W * 8, 256) + align_up(prep.gword.size() * 4, 256);
    HIP_TRY(hipStreamSynchronize(s));
    if (!noalloc && init_bytes > c->h_init_cap) {
        if (c->h_init) hipHostFree(c->h_init);
        c->h_init = nullptr;
        c->h_init_cap = 0;
        HIP_TRY(hipHostMalloc((void **)&c->h_init, init_bytes, hipHostMallocDefault));
        c->h_init_cap = init_bytes;
    }
    size_t init_off = 0;
    auto upload_init = [&](void *dst, const void *src, size_t n) -> hipError_t {
        if (n == 0) return hipSuccess;
        // (no staging buffer: a copy from pageable memory, waited for below)
        if (noalloc) return hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s);
        memcpy(c->h_init + init_off, src, n);
        const hipError_t e = hipMemcpyAsync(dst, c->h_init + init_off, n, hipMemcpyHostToDevice, s);
        init_off += align_up(n, 256);
        return e;
    };
    HIP_TRY(upload_init(c->d_aux, aux.data(), aux.size() * sizeof(WordAux)));
    HIP_TRY(upload_init(c->d_saux, saux.data(), saux.size() * sizeof(SumAux)));
    // Atom words and summaries: one contiguous copy, then scattered into the
    // word-major store.  (`kinfo` outlives its copy: the first level ends
    // with a stream sync.)
    std::vector<uint32_t> kinfo;
    std::vector<uint64_t> stage(atom_words);
    stage.insert(stage.end(), sum0.begin(), sum0.end());
    HIP_TRY(upload_init(c->d_stage, stage.data(), stage.size() * 8));
    HIP_TRY(launch_scatter(c->d_store, c->ws, c->d_stage, A, c->nw, s));
    HIP_TRY(launch_scatter(c->d_sum, c->ws, c->d_stage + (size_t)c->nw * A, A, (int64_t)4 * nsw, s));
    HIP_TRY(upload_init(c->d_state, &st0, sizeof(st0)));
    // Without a fixed size the table starts small and grows before a level
    // is generated (clearing 2^31 slots would cost milliseconds per call).
    if (!o.hash_slots_log2) {
        size_t slots = (size_t)1 << 16;
        while (slots < c->tab_max_slots && slots < 4 * (size_t)A) slots <<= 1;
        c->tmask = (uint32_t)(std::min(slots, c->tab_max_slots) - 1);
    }
    // The table: cleared, or (a fixed-size table in the context's pool, as
    // the last call left it) the next epoch's.
    const bool tab_reuse = !noalloc && o.hash_slots_log2 && c->tab_ok && c->tab_ptr == c->d_tab &&
                           c->tab_mask == c->tmask && c->tab_epoch < kEpochs;
    c->tab_ok = false;
    if (tab_reuse) {
        ++c->tab_epoch;
    } else {
        HIP_TRY(launch_fill_u64(c->d_tab, kEmpty, 2 * ((size_t)c->tmask + 1), s));
        c->tab_epoch = 1;
    }
    HIP_TRY(launch_fill_u64(c->d_lb, 0, (size_t)c->lb_words, s));
    c->lb_epoch = 0;

    Run run{c, pb};
    ChunkArgs &ba = run.base;
    ba.store = c->d_store; ba.aux = c->d_aux; ba.tab = c->d_tab; ba.st = c->d_state;
    ba.t_slot = c->t_slot; ba.t_fp2 = c->t_fp2; ba.t_mask = c->t_mask; ba.t_cnt = c->t_cnt; ba.t_off = c->t_off;
    ba.t_list = c->t_list;
    ba.lb_flags = c->d_lb;
    ba.lb_ticket = reinterpret_cast<uint32_t *>(c->d_lb + c->lb_words - 1);
    ba.m_op = c->d_op; ba.m_left = c->d_left; ba.m_right = c->d_right; ba.m_fp2 = c->d_fp2; ba.m_key = c->d_key;
    ba.sum = c->d_sum; ba.saux = c->d_saux; ba.nsw = nsw;
    // Check passes alternate between positive and negative traces (early
    // exit, interleave()): with K words per trace both are in trace order,
    // positives first; packed words and their summary groups mix positive
    // and negative traces already, so they are visited in order.
    if (K > 1) {
        ba.nps = std::min(nsw, (P + 63) / 64);
        ba.npw = P;  // traces
    } else {
        ba.nps = 0;
        ba.npw = 0;
    }
    ba.K = K;
    ba.ptrace = c->d_ptrace;
    if (B == 1) HIP_TRY(upload_init(c->d_ptrace, prep.ptrace.data(), (size_t)W * 8));
    // Packed emits: up to eight waves per store tile, each a word range of
    // whole summary groups of 40 words or more (500+500 traces, per call: one wave
    // per tile 16.1 ms of emits, four 10.1 ms, eight 8.9 ms, sixteen 8.9 ms).
    ba.esplit = 1;
    if (B == 1) {
        HIP_TRY(upload_init(c->d_gword, prep.gword.data(), (size_t)nsw * 8));
        ba.gfirst = c->d_gword;
        ba.glast = c->d_gword + nsw;
        ba.esplit = std::max(1, std::min({8, nsw, c->nw / 40}));
    }
    // Relaxed keys (code/ltli6463.cu:57-141) only above 126 positions; below,
    // every mode is the exact packing, which the exact dedup reproduces.
    ba.W = W;
    ba.kinfo = c->d_kinfo;
    ba.kmode = 0;
    ba.kseed = seed == kDefaultSeed ? 0 : seed;
    {
        int len_sum = 0;
        for (int i = 0; i < W; ++i) len_sum += pb->lens[i];
        if (o.dedup_mode != LTL_DEDUP_EXACT && len_sum > 126) {
            kinfo.assign(W, 0);
            const int stride = len_sum / 126;
            int j = 0;
            for (int i = 0; i < W && j < 126; ++i) {
                const int len = pb->lens[i];
                int take = 0;
                if (o.dedup_mode == LTL_DEDUP_REF1) take = (len + stride - 1) / stride;
                else if (o.dedup_mode == LTL_DEDUP_REF2) take = (126 * len + len_sum) / len_sum;
                const int cnt = std::min(take, 126 - j);
                if (cnt > 0) kinfo[i] = ((uint32_t)j << 8) | (uint32_t)cnt;
                j += cnt;
            }
            ba.kmode = o.dedup_mode;
            ba.t_acc = c->t_succ;  // per candidate: accepted, folded in for dedup survivors only
            ba.kstride = stride;
            ba.kw_end = 0;  // words holding a trace that takes key bits
            for (int i = 0; i < W; ++i)
                if (kinfo[i] & 0xffu) ba.kw_end = std::max(ba.kw_end, c->pk_word[i] + 1);
            HIP_TRY(hipMemcpyAsync(c->d_kinfo, kinfo.data(), (size_t)W * 4, hipMemcpyHostToDevice, s));
        }
    }
    ba.capacity = c->capacity;
    ba.epoch = c->tab_epoch;
    ba.keep_keys = o.hash_slots_log2 ? 0 : 1;
    ba.ws = c->ws;
    ba.nw = c->nw;
    ba.hamming_k = o.hamming_k;
    ba.tmask = c->tmask;
    HIP_TRY(launch_atoms(B, ba, A, s));
    if (c->launch_fn) {
        const ltl_launch_record rec{c1, LTL_ATOM, 1, 1, 0, 0, (uint64_t)A};
        c->launch_fn(&rec, c->launch_user);
    }
    run.all_ltls = all_ltls;

    const int nlev = std::max(maxc, c1) + 2;
    c->lvl_s.assign(nlev, 0);
    c->lvl_e.assign(nlev, 0);
    c->lvl_s[c1] = 0;
    c->lvl_e[c1] = A;
    auto &Ls = c->lvl_s;
    auto &Le = c->lvl_e;

    int cost = c1 + 1, shortage = -1;
    bool last_round = false, found = false;
    int32_t rc = 0;
    DevState &h = *c->h_state;
    auto sync_state = [&]() -> int32_t {
        HIP_TRY(hipMemcpyAsync(&h, c->d_state, sizeof(DevState), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        drain_events(c);
        if (h.collision) return LTL_E_COLLISION;
        if (h.probe_fail) return LTL_E_OOM;
        if (h.lb_fail) {
            report_hip(hipErrorUnknown, "dedup look-back (a wait ran out)", __LINE__);
            return LTL_E_HIP;
        }
        return 0;
    };
    // The blocks of level `cost` in the reference's order
    // (code/ltli6463.cu:1209-1440).
    auto level_blocks = [&](int cost) -> int32_t {
        int32_t r = 0;
        run.cur_level = cost;
        // Generation of a tested level with equal unary costs: ~, X, F and G
        // of the operand level in one pass (regions of 4 * slots within half
        // of the chunk temporaries; the & and | chunks in between use the
        // rest).
        const int lvn = cost - c2;
        const int64_t nn = lvn >= c1 ? Le[lvn] - Ls[lvn] : 0;
        const bool fuse4 = !run.check_only && !run.count_only && !run.otf && run.base.gen_check == 0 &&
                           !run.base.kmode && o.neg_type == 1 && lvn >= c1 && nn > 0 && c2 == c5 &&
                           c5 == c6 && c6 == c7 && 4 * ((nn + 127) / 64 * 64) <= c->CH / 2;
        if (fuse4) {
            uint64_t between = 0;  // candidates of the & and | blocks
            for (int i = c1; 2 * i <= cost - c3; ++i)
                between += (uint64_t)std::max<int64_t>(0, Le[i] - Ls[i]) *
                           (uint64_t)std::max<int64_t>(0, Le[cost - i - c3] - Ls[cost - i - c3]);
            for (int i = c1; 2 * i <= cost - c4; ++i)
                between += (uint64_t)std::max<int64_t>(0, Le[i] - Ls[i]) *
                           (uint64_t)std::max<int64_t>(0, Le[cost - i - c4] - Ls[cost - i - c4]);
            const uint64_t id_xfg = run.all_ltls + (uint64_t)nn + between;
            if ((r = run.unary_nxfg(Ls[lvn], Le[lvn], id_xfg))) return r;
            // & and | of each split in one pass (and U when the level's
            // candidates fit the temporaries); else one pass per block
            // (LTL_GEN_FUSE = 0 or 1 in the environment caps the fusion at
            // none or & and |: tests run every path against the oracle)
            int fused = 0, fuse_max = 2;
            if (const char *e = getenv("LTL_GEN_FUSE")) fuse_max = std::max(0, std::min(2, atoi(e)));
            std::vector<Run::Split2> sv;
            if (c3 == c4 && fuse_max > 0) {
                sv = run.splits_of(cost, c1, c3);
                if (fuse_max == 2 && c8 == c3 && run.fused_slots(sv, true) <= run.ch_cap()) fused = 2;
                else if (run.fused_slots(sv, false) <= run.ch_cap()) fused = 1;
            }
            if (fused) {
                if ((r = run.binary_fused(sv, fused == 2, id_xfg + 3 * (uint64_t)nn))) return r;
                if ((r = run.finish_pending(run.pend_or))) return r;
            } else {
                for (int i = c1; 2 * i <= cost - c3; ++i)
                    if ((r = run.binary(LTL_AND, i, cost - i - c3))) return r;
                for (int i = c1; 2 * i <= cost - c4; ++i)
                    if ((r = run.binary(LTL_OR, i, cost - i - c4))) return r;
            }
            if ((r = run.finish_xfg())) return r;
            if (fused == 2) return run.finish_pending(run.pend_u);
            for (int i = c1; 2 * i <= cost - c8; ++i)
                if ((r = run.binary(LTL_UNTIL, i, cost - i - c8))) return r;
            return 0;
        }
        if ((o.neg_type == 1 && cost - c2 >= c1) || (o.neg_type == 2 && cost - c2 == c1))
            if ((r = run.unary(LTL_NOT, Ls[cost - c2], Le[cost - c2]))) return r;
        for (int i = c1; 2 * i <= cost - c3; ++i)
            if ((r = run.binary(LTL_AND, i, cost - i - c3))) return r;
        for (int i = c1; 2 * i <= cost - c4; ++i)
            if ((r = run.binary(LTL_OR, i, cost - i - c4))) return r;
        if (c5 == c6 && c6 == c7) {
            if (cost - c5 >= c1 && (r = run.unary_xfg(Ls[cost - c5], Le[cost - c5]))) return r;
        } else {
            if (cost - c5 >= c1 && (r = run.unary(LTL_NEXT, Ls[cost - c5], Le[cost - c5]))) return r;
            if (cost - c6 >= c1 && (r = run.unary(LTL_FIN, Ls[cost - c6], Le[cost - c6]))) return r;
            if (cost - c7 >= c1 && (r = run.unary(LTL_GLOB, Ls[cost - c7], Le[cost - c7]))) return r;
        }
        for (int i = c1; 2 * i <= cost - c8; ++i)
            if ((r = run.binary(LTL_UNTIL, i, cost - i - c8))) return r;
        return 0;
    };
    // Learned order and log: tile statistics of the stored levels, the
    // binary blocks of the level, then log rows and / or ranked rounds.
    const bool learned = o.order == LTL_ORDER_LEARNED;
    run.ranked = learned;
    run.logging = c->log_cap > 0 && c->log_per_block > 0;
    c->log_n = 0;
    int tstat_upto = c1 - 1;
    int64_t tstat_used = 0;
    std::vector<int64_t> toff(nlev, 0);
    if (learned || run.logging) {
        const int64_t tiles = c->capacity / 64 + nlev + 1;
        if (tiles > c->tstat_cap) {
            if (c->d_tstat) hipFree(c->d_tstat);
            c->d_tstat = nullptr;
            c->tstat_cap = 0;
            HIP_TRY(hipMalloc((void **)&c->d_tstat, (size_t)tiles * kTileStat * sizeof(float)));
            c->tstat_cap = tiles;
        }
        if (nlev > c->toff_cap) {
            if (c->d_toff) hipFree(c->d_toff);
            c->d_toff = nullptr;
            c->toff_cap = 0;
            HIP_TRY(hipMalloc((void **)&c->d_toff, (size_t)nlev * sizeof(int64_t)));
            c->toff_cap = nlev;
        }
    }
    // Binary blocks of a level recorded in run.lblk; levels up to `upto` are
    // complete in the store.
    auto ensure_tstat = [&](int upto) -> int32_t {
        if (tstat_upto >= upto) return 0;
        for (int lv = tstat_upto + 1; lv <= upto; ++lv) {
            const int64_t n = Le[lv] - Ls[lv];
            toff[lv] = tstat_used;
            if (n > 0) {
                HIP_TRY(timed(c, K_TSTAT, (uint64_t)n * nsw * 16, (uint64_t)n * nsw * 16, (uint64_t)n, [&] {
                    return launch_tile_stats(ba, Ls[lv], n, P, W, c->d_tstat + tstat_used * kTileStat, s);
                }));
            }
            tstat_used += (n + 63) / 64;
        }
        tstat_upto = upto;
        HIP_TRY(upload_pinned(c, c->d_toff, toff.data(), (size_t)nlev * sizeof(int64_t), s));
        return 0;
    };
    auto upload_blocks = [&]() -> int32_t {
        const int nb = (int)run.lblk.size();
        if (nb > c->blk_cap) {
            HIP_TRY(hipStreamSynchronize(s));
            if (c->d_blk) hipFree(c->d_blk);
            c->d_blk = nullptr;
            c->blk_cap = 0;
            HIP_TRY(hipMalloc((void **)&c->d_blk, (size_t)std::max(nb, 64) * sizeof(LBlock)));
            c->blk_cap = std::max(nb, 64);
        }
        HIP_TRY(upload_pinned(c, c->d_blk, run.lblk.data(), (size_t)nb * sizeof(LBlock), s));
        return 0;
    };
    auto feat = [&](int cost) { return FeatCtx{c->d_tstat, c->d_toff, (float)cost, (float)P / (float)W, W, P}; };
    auto log_level = [&](int cost, int upto) -> int32_t {
        if (int32_t e = ensure_tstat(upto)) return e;
        const FeatCtx f = feat(cost);
        for (int b = 0; b < (int)run.lblk.size() && c->log_n < c->log_cap; ++b) {
            const LBlock &K = run.lblk[b];
            const int64_t E = (K.nL + 63) / 64 * K.nTR;
            const int64_t rows = std::min<int64_t>({(int64_t)c->log_per_block, E, c->log_cap - c->log_n});
            HIP_TRY(timed(c, K_LOG, 0, 0, (uint64_t)rows, [&] {
                return launch_log(B, ba, c->d_blk, b, f, rows, c->d_log + c->log_n * (kFeat + 1), s);
            }));
            c->log_n += rows;
        }
        return 0;
    };
    constexpr int64_t kRankMin = 4096;  // about 16M candidates: more work than a sync costs
    // Ranked rounds: the top 1/64 of the tile pairs (at least kRankMin),
    // then three times as many tile pairs as all rounds before (the last
    // round takes the rest when it would be short): round_end(E, r0) is the
    // end of the round that starts at r0.
    auto round_end = [&](int64_t E, int64_t r0) {
        const int64_t next = r0 == 0 ? std::max(kRankMin, E / 64) : 4 * r0;
        return next + next / 4 >= E ? E : next;
    };
    // Tile-pair order of the level in the second half of d_vals: by
    // descending score, or canonical when one round covers the level.
    auto order_level = [&](int cost, int upto) -> int32_t {
        const int64_t E = run.n_ent;
        if (E > ((int64_t)1 << 31)) return LTL_E_LIMIT;
        if (E > c->ent_cap) {
            HIP_TRY(hipStreamSynchronize(s));
            for (void *p : {(void *)c->d_keys, (void *)c->d_vals, c->d_sort})
                if (p) hipFree(p);
            c->d_keys = nullptr; c->d_vals = nullptr; c->d_sort = nullptr;
            c->ent_cap = 0;
            const int64_t cap = std::max<int64_t>(E, 1 << 16);
            c->sort_cap = sort_temp_bytes(cap);
            HIP_TRY(hipMalloc((void **)&c->d_keys, (size_t)cap * 2 * sizeof(float)));
            HIP_TRY(hipMalloc((void **)&c->d_vals, (size_t)cap * 2 * sizeof(uint32_t)));
            HIP_TRY(hipMalloc(&c->d_sort, std::max<size_t>(c->sort_cap, 256)));
            c->ent_cap = cap;
        }
        float *k0 = c->d_keys, *k1 = c->d_keys + c->ent_cap;
        uint32_t *v0 = c->d_vals, *v1 = c->d_vals + c->ent_cap;
        if (round_end(E, 0) >= E) {  // one round
            HIP_TRY(launch_iota(v1, E, s));
            return 0;
        }
        if (int32_t e = ensure_tstat(upto)) return e;
        const FeatCtx f = feat(cost);
        const int nb = (int)run.lblk.size();
        HIP_TRY(timed(c, K_SCORE, 0, 0, (uint64_t)E, [&] {
            return launch_score(c->scorer->m, c->d_blk, nb, f, E, k0, v0, s);
        }));
        HIP_TRY(timed(c, K_SORT, 0, 0, (uint64_t)E, [&] {
            return launch_sort_desc(k0, k1, v0, v1, E, c->d_sort, c->sort_cap, s);
        }));
        return 0;
    };
    auto submit_round = [&](int64_t r0, int64_t n) -> int32_t {
        if (n <= 0) return 0;
        const int nb = (int)run.lblk.size();
        const uint32_t *v1 = c->d_vals + c->ent_cap;
        HIP_TRY(timed(c, K_CHKR, 0, 0, (uint64_t)n * 4096, [&] {
            return launch_check_list(B, run.base, c->d_blk, nb, v1, r0, n, s);
        }));
        run.checked += (uint64_t)n * 4096;
        return 0;
    };
    // Log and / or test (ranked rounds, a sync after each) the recorded blocks.
    auto ranked_level = [&](int cost, int upto) -> int32_t {
        if (int32_t e = upload_blocks()) return e;
        if (run.logging)
            if (int32_t e = log_level(cost, upto)) return e;
        if (!run.ranked) return 0;
        if (int32_t e = order_level(cost, upto)) return e;
        const int64_t E = run.n_ent;
        for (int64_t r0 = 0; r0 < E; r0 = round_end(E, r0)) {
            if (int32_t e = submit_round(r0, round_end(E, r0) - r0)) return e;
            HIP_TRY(hipMemcpyAsync(&h, c->d_state, sizeof(DevState), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (h.found != kEmpty) break;
        }
        return 0;
    };
    auto level_of = [&](uint64_t id) -> int {
        for (const auto &b : run.blocks)
            if (id >= b.id_base && id < b.id_base + b.size) return b.level;
        return -1;
    };

    int64_t n_stored = A;
    // Learned order, before any overflow: the binary blocks of level c + 1
    // only combine levels <= c - 1, so they are tested (ranked, early exit)
    // before level c is generated; an accepted candidate there is minimal,
    // since every level up to c has been tested completely by then.  With
    // neg_type 1 and unit-cost negation the check passes also record
    // candidates whose complement is accepted: once level c has been tested
    // completely without a hit, the negation of such a level-c candidate is
    // an accepted formula of cost c + 1, and level c is never generated.
    const bool anti_ok = learned && o.neg_type == 1 && c2 == 1;
    // plain order: deferred levels and the check of their negation successors
    const bool defer_ok = !learned && !o.dedup_mode;
    const bool anti_plain = defer_ok && o.neg_type == 1 && c2 == 1;
    uint64_t neg_pick = kEmpty;  // smallest id of the level whose complement is accepted
    int plain_ahead = -1;        // level whose binary blocks the lookahead tested
    int looked_ahead = -1;  // level whose binary blocks were tested early
    int found_cost = -1;
    bool neg_answer = false;
    uint64_t neg_id = kEmpty;
    // The binary blocks of level `next` tested before the level below it is
    // generated (plain order): submitted only, a sync follows.
    auto submit_lookahead = [&](int next) -> int32_t {
        if (anti_plain &&
            launch_fill_u64(reinterpret_cast<uint64_t *>(&c->d_state->neg_found[1]), kEmpty, 1, s) != hipSuccess)
            return LTL_E_HIP;
        run.check_only = true;
        run.only_binary = true;
        run.base.anti = anti_plain ? 1 : 0;
        run.base.neg_slot = 1;
        const int32_t r = level_blocks(next);
        run.base.anti = 0;
        run.base.neg_slot = 0;
        run.only_binary = false;
        run.check_only = false;
        return r;
    };
    std::vector<Block> ahead_blocks;  // the lookahead's blocks (ids relative to their first)
    // (an emit's fingerprints overran their buffer: not again this call;
    // LTL_GEN_NX=0 in the environment: never, for the tests of both paths)
    bool nx_off = false, nx_force = false;
    if (const char *e = getenv("LTL_GEN_NX")) {
        nx_off = atoi(e) == 0;
        nx_force = atoi(e) == 2;  // (every level, whatever its size)
    }
    // LTL_GEN_LZW=0 in the environment: levels fingerprinted by their emits
    // still write their words (for the tests of both paths)
    bool lzw_off = false;
    if (const char *e = getenv("LTL_GEN_LZW")) lzw_off = atoi(e) == 0;
    auto materialize_lzw = [&]() -> int32_t {
        if (c->lzw_e <= c->lzw_s) return 0;
        if (launch_materialize(c->lzw_args, c->lzw_s, c->lzw_e, s, false) != hipSuccess) return LTL_E_HIP;
        c->lzw_s = c->lzw_e = 0;
        run.base.lzw_s = run.base.lzw_e = 0;
        return 0;
    };
    constexpr int64_t kNxMinOperands = 1 << 18;
    for (cost = c1 + 1; cost <= maxc; ++cost) {
        run.level_stored = n_stored;
        bool ahead_done = false;          // the lookahead of cost + 1 ran with the check pass
        uint64_t ahead_hit = kEmpty, ahead_neg = kEmpty;
        if (run.otf) {
            const int dif = cost - shortage;  // code/ltli6463.cu:1202-1206
            if (dif == c2 || dif == c1 + c3 || dif == c1 + c4 || dif == c5 || dif == c6 || dif == c7 || dif == c1 + c8)
                last_round = true;
        }
        Ls[cost] = n_stored;
        // Precheck: with exact dedup the first accepted candidate of a level
        // is always the first occurrence of its set (an earlier equal set
        // would have been accepted first), so a check-only pass finds the
        // same result as generation with dedup, without generating the
        // level.  Relaxed keys merge different sets, and the reference tests
        // only candidates whose key is new (code/ltli6463.cu:231-243): those
        // levels are tested in generation, through dedup, except in
        // check-only mode after an overflow, where the reference tests every
        // candidate (code/ltli6463.cu:198-208).
        bool generate = true;
        bool tested = false;  // every candidate of the level went through a check pass
        const bool la = learned && !run.otf;
        if (la) {
            // Level `cost`: unary blocks (binary ones too if they were not
            // tested ahead), then the binary blocks of cost + 1, submitted
            // together; one sync decides.
            const int slot = cost & 1;
            const bool ahead = cost + 1 <= maxc;
            run.check_only = true;
            run.skip_binary = looked_ahead == cost;
            run.base.anti = anti_ok;
            run.base.neg_slot = slot;
            run.lblk.clear();
            run.n_ent = 0;
            rc = level_blocks(cost);
            if (!rc && !run.lblk.empty()) rc = ranked_level(cost, cost - 1);
            // Only unary blocks launched (no binary blocks, or tested ahead):
            // their result first, so that a hit there costs no lookahead
            // (the ranked rounds above end with a sync already).
            if (!rc && run.lblk.empty()) rc = sync_state();
            run.skip_binary = false;
            if (!rc && h.found != kEmpty) {  // found by the pass above
                found = true;
                found_cost = cost;
                Le[cost] = n_stored;
                break;
            }
            int64_t E = 0, r1 = 0;
            if (!rc && ahead) {
                if (launch_fill_u64(reinterpret_cast<uint64_t *>(&c->d_state->neg_found[1 - slot]), kEmpty, 1, s) !=
                    hipSuccess)
                    rc = LTL_E_HIP;
                run.only_binary = true;
                run.base.neg_slot = 1 - slot;
                run.lblk.clear();
                run.n_ent = 0;
                const uint64_t next_base = run.all_ltls;  // first id of level cost + 1
                if (!rc) rc = level_blocks(cost + 1);
                run.only_binary = false;
                E = run.n_ent;
                if (!rc && E > 0) {
                    rc = upload_blocks();
                    if (!rc && run.logging) rc = log_level(cost + 1, cost - 1);
                    if (!rc) rc = order_level(cost + 1, cost - 1);
                    r1 = round_end(E, 0);
                    // (submitted before the level's result is known: no work
                    // once the level has an accepted candidate)
                    run.base.found_below = next_base;
                    if (!rc) rc = submit_round(0, r1);
                    run.base.found_below = 0;
                }
            }
            if (!rc) rc = sync_state();
            if (rc) break;
            if (h.found != kEmpty) {
                found = true;
                found_cost = level_of(h.found);
                Le[cost] = n_stored;
                break;
            }
            if (anti_ok && ahead && h.neg_found[slot] != kEmpty) {
                neg_answer = true;
                neg_id = h.neg_found[slot];
                found_cost = cost + 1;
                Le[cost] = n_stored;
                break;
            }
            for (int64_t r0 = r1; r0 < E && !rc && h.found == kEmpty; r0 = round_end(E, r0))
                if (!(rc = submit_round(r0, round_end(E, r0) - r0))) rc = sync_state();
            if (rc) break;
            if (h.found != kEmpty) {
                found = true;
                found_cost = cost + 1;
                Le[cost] = n_stored;
                break;
            }
            run.check_only = false;
            run.base.anti = 0;
            tested = true;
            if (ahead) looked_ahead = cost + 1;
            else generate = false;  // the last level is tested completely; its sets would serve no later level
        } else if ((o.precheck && !ba.kmode) || run.otf || learned) {
            const uint64_t save = run.all_ltls;
            const size_t nb = run.blocks.size();
            // (plain order: the pass also records the smallest candidate id
            // whose complement is accepted, the negation successor of a
            // deferred level)
            const bool anti = anti_plain && !run.otf;
            // Plain order: the lookahead of cost + 1 (below) is submitted with
            // this pass, one synchronisation for both; its ids follow this
            // level's, so the smallest accepted id tells which of the two it
            // belongs to (once this level has one, the lookahead's launches
            // do no work).  Not while a deferred level is pending: this level
            // then has an accepted candidate.
            const bool ahead_ok = !learned && !run.otf && defer_ok && !run.logging && cost < maxc &&
                                  c->lazy_e == c->lazy_s && run.defer_s < 0;
            for (int pass = 0; pass < 2; ++pass) {
                run.check_only = true;
                run.lblk.clear();
                run.n_ent = 0;
                run.succ_blocks.clear();
                if (anti && launch_fill_u64(reinterpret_cast<uint64_t *>(&c->d_state->neg_found[0]), kEmpty, 1, s) !=
                                hipSuccess) {
                    rc = LTL_E_HIP;
                    break;
                }
                run.base.anti = anti ? 1 : 0;
                run.base.neg_slot = 0;
                run.reuse_binary = pass == 0 && plain_ahead == cost;
                run.bin_seen = 0;
                run.la_found = run.la_neg = kEmpty;
                rc = level_blocks(cost);
                run.reuse_binary = false;
                run.base.anti = 0;
                if (!rc && !run.lblk.empty()) rc = ranked_level(cost, cost - 1);
                uint64_t ahead_base = kEmpty;
                if (!rc && pass == 0 && ahead_ok) {
                    ahead_base = run.all_ltls;
                    const size_t nb2 = run.blocks.size();
                    rc = submit_lookahead(cost + 1);
                    ahead_blocks.assign(run.blocks.begin() + (long)nb2, run.blocks.end());
                }
                if (!rc) rc = sync_state();
                run.check_only = false;
                if (rc) break;
                if (ahead_base != kEmpty) {
                    ahead_done = true;
                    ahead_hit = h.found >= ahead_base ? h.found : kEmpty;
                    ahead_neg = h.neg_found[1];
                    if (h.found >= ahead_base) h.found = kEmpty;
                }
                neg_pick = anti ? std::min<uint64_t>(h.neg_found[0], run.la_neg) : kEmpty;
                h.found = std::min<uint64_t>(h.found, run.la_found);
                // unary blocks over a deferred level: their successors were
                // tested in its generation (smallest store index per operator)
                for (const auto &sb : run.succ_blocks) {
                    const uint64_t m = h.succ_min[sb.second];
                    if (((run.succ_mask >> sb.second) & 1) && m != kEmpty)
                        h.found = std::min<uint64_t>(h.found, sb.first + (m - (uint64_t)run.defer_s));
                }
                if (h.found != kEmpty) break;
                // (a deferred level is always followed by a level that has an
                // accepted candidate; should this one have none, its words and
                // summaries are recomputed, and the pass repeated if it skipped
                // blocks over it)
                if (c->lazy_e > c->lazy_s) {
                    // (its parents first: the level below may lack its words)
                    if ((rc = materialize_lzw())) break;
                    if (launch_materialize(c->lazy_args, c->lazy_s, c->lazy_e, s) != hipSuccess) {
                        rc = LTL_E_HIP;
                        break;
                    }
                    c->lazy_s = c->lazy_e = 0;
                }
                run.defer_s = run.defer_e = -1;
                if (run.succ_blocks.empty()) break;
                run.all_ltls = save;  // again, launching every block
                run.blocks.resize(nb);
            }
            if (rc) break;
            if (h.found != kEmpty) { found = true; found_cost = cost; Le[cost] = n_stored; break; }
            tested = true;
            if (run.otf) generate = false;  // check-only level (overflow)
            else if (!learned) { run.all_ltls = save; run.blocks.resize(nb); }
        }
        // Plain order: a level whose sets no later pass reads in full is
        // deferred: stored with metadata only, no words, no summaries.  That
        // is the level of cost maxc (the last one), and level `cost` when
        // level cost + 1 has an accepted candidate.  The binary blocks of
        // cost + 1 combine levels up to cost - 1, so they are tested here,
        // before the level is generated; an accepted candidate among them
        // makes cost + 1 the last level (levels up to `cost` have none), and
        // the unary blocks of cost + 1 over level `cost` are then tested in
        // its generation instead (the successors ~c, X c, F c, G c of every
        // candidate, at position 0).  Only the unary blocks that come before
        // the hit in canonical order matter: the negation block when the hit
        // is an & or | candidate, all four when it is a U candidate.  The
        // check pass of cost + 1 then takes the smallest id over its binary
        // blocks and these.  Either way its binary blocks are not launched
        // again: their first hit, if any, is the lookahead's (same blocks in
        // the same order; ids shifted by the unary blocks before them).
        bool lazy = false;
        int succ_mask = 0;
        if (generate && tested && defer_ok && !run.otf && !run.logging) {
            if (cost == maxc) {
                lazy = true;
            } else {
                if (!ahead_done) {  // (submitted with the check pass above, or now)
                    const uint64_t save = run.all_ltls;
                    const size_t nb = run.blocks.size();
                    rc = submit_lookahead(cost + 1);
                    if (!rc) rc = sync_state();
                    ahead_blocks.assign(run.blocks.begin() + (long)nb, run.blocks.end());
                    ahead_hit = h.found;
                    ahead_neg = h.neg_found[1];
                    run.all_ltls = save;
                    run.blocks.resize(nb);
                    if (rc) break;
                }
                int hit_op = -1;
                run.la_block = run.la_neg_block = -1;
                const uint64_t la_neg_id = anti_plain ? ahead_neg : kEmpty;
                for (size_t b = 0; b < ahead_blocks.size(); ++b) {
                    const Block &bk = ahead_blocks[b];
                    if (ahead_hit >= bk.id_base && ahead_hit - bk.id_base < bk.size) {
                        hit_op = bk.op;
                        run.la_block = (int64_t)b;
                        run.la_off = ahead_hit - bk.id_base;
                    }
                    if (la_neg_id >= bk.id_base && la_neg_id - bk.id_base < bk.size) {
                        run.la_neg_block = (int64_t)b;
                        run.la_neg_off = la_neg_id - bk.id_base;
                    }
                }
                plain_ahead = cost + 1;
                h.found = ahead_hit;
                if (h.found != kEmpty) {
                    const int avail = (o.neg_type == 1 && c2 == 1 ? 1 : 0) | (c5 == 1 ? 2 : 0) |
                                      (c6 == 1 ? 4 : 0) | (c7 == 1 ? 8 : 0);
                    succ_mask = (hit_op == LTL_AND || hit_op == LTL_OR ? 1 : 15) & avail;
                    // (K words per trace: X, F, G successors are not tested in
                    // generation, so such a level is stored in full)
                    lazy = B == 1 || !(succ_mask & 14);
                    if (!lazy) succ_mask = 0;
                    if (launch_fill_u64(reinterpret_cast<uint64_t *>(&c->d_state->found), kEmpty, 1, s) != hipSuccess) {
                        rc = LTL_E_HIP;
                        break;
                    }
                    h.found = kEmpty;
                }
            }
        }
        if (generate && !o.hash_slots_log2) {
            // Grow the table so that it stays at most half full with every
            // set this level can add and a full chunk pending: a new table,
            // the stored sets re-inserted from their keys.
            run.count_only = true;
            run.counted = 0;
            rc = level_blocks(cost);
            run.count_only = false;
            if (rc) break;
            const uint64_t need =
                2 * (std::min<uint64_t>((uint64_t)c->capacity, (uint64_t)n_stored + run.counted) +
                     std::min<uint64_t>((uint64_t)c->CH, run.counted));
            size_t slots = (size_t)c->tmask + 1;
            if (need > slots && slots < c->tab_max_slots) {
                while (slots < need && slots < c->tab_max_slots) slots <<= 1;
                c->tmask = (uint32_t)(slots - 1);
                ba.tmask = c->tmask;
                if (launch_fill_u64(c->d_tab, kEmpty, 2 * slots, s) != hipSuccess ||
                    launch_rehash(c->d_tab, c->tmask, c->d_key, n_stored, ba.epoch, &c->d_state->probe_fail, s) !=
                        hipSuccess) {
                    rc = LTL_E_HIP;
                    break;
                }
            }
        }
        if (generate) {
            // The level below, emitted without its words: they are written
            // now unless this level is deferred without successor tests (its
            // emits and its ~/X/F/G pass, which then inserts the fingerprints
            // those emits computed, read nothing of it; nor do later passes).
            if (c->lzw_e > c->lzw_s && !(lazy && !(succ_mask & 14)) && (rc = materialize_lzw())) break;
            // a tested level has no accepted candidate: generation skips the check
            run.base.gen_check = tested ? 0 : 1;
            if (lazy) {
                run.base.deferred = 1;
                run.base.succ_mask = succ_mask;
                run.base.t_succ = (succ_mask & 14) ? c->t_succ : nullptr;
                run.base.neg_pick = (succ_mask & 1) ? neg_pick : kEmpty;
                if (succ_mask &&
                    launch_fill_u64(reinterpret_cast<uint64_t *>(c->d_state->succ_min), kEmpty, 4, s) != hipSuccess) {
                    rc = LTL_E_HIP;
                    break;
                }
            }
            // The next level's ~, X, F and G over this one are fingerprinted
            // by this level's emits (plain order, unit unary costs: that
            // level's fused ~/X/F/G pass reads this level), for large levels
            // only: the extra arithmetic lengthens every emit wave, which a
            // level of large chunks hides better than the read it saves
            // costs (500+500 traces of 10-32 positions, per 1000 sets: the next level's pass 0.76 us,
            // the inserts that replace it 0.22 us; the emits' extra 0.30 us
            // on level 9, generated from 0.7 M sets, 0.57 us on level 8,
            // from 0.11 M).
            const int64_t n_prev = Le[cost - 1] - Ls[cost - 1];
            const bool nx = !lazy && !run.base.kmode && !learned && !run.logging && !run.otf && !nx_off &&
                            o.neg_type == 1 && c2 == 1 && c5 == 1 && c6 == 1 && c7 == 1 && cost < maxc &&
                            (n_prev >= kNxMinOperands || nx_force);
            // Such a level's emits also leave its words unwritten: the next
            // level reads it only through those fingerprints (unless it is
            // generated in full or with successor tests: then the words are
            // written first, above), and the levels after that only through
            // the summaries and the until check, which recomputes the words
            // it reads from the parents (k_chkU).
            // (not when the level could overflow the store: check-only levels
            // would then need the words at once)
            bool wskip = nx && !lzw_off;
            if (wskip) {
                run.count_only = true;
                run.counted = 0;
                rc = level_blocks(cost);
                run.count_only = false;
                if (rc) break;
                wskip = (uint64_t)n_stored + run.counted <= (uint64_t)c->capacity;
            }
            if (nx) {
                run.base.nx_fp = c->d_nxfp;
                run.base.nx_ls = n_stored;
                run.base.nx_ush = (int32_t)(n_stored & 63);
                run.base.nx_stride = c->nx_stride;
                run.base.wskip = wskip ? 1 : 0;
            }
            rc = level_blocks(cost);
            run.base.gen_check = 1;
            run.base.deferred = 0;
            run.base.succ_mask = 0;
            run.base.t_succ = nullptr;
            run.base.neg_pick = kEmpty;
            run.base.nx_fp = nullptr;
            run.base.wskip = 0;
            if (rc) break;
            if ((rc = sync_state())) break;
            if (h.nx_over) nx_off = true;
            run.nx_from = nx && !nx_off && h.overflow_chunk == LLONG_MAX ? n_stored : -1;
            if (wskip && h.n_stored > n_stored) {
                c->lzw_s = n_stored;
                c->lzw_e = h.n_stored;
                c->lzw_args = run.base;
                run.base.lzw_s = c->lzw_s;
                run.base.lzw_e = c->lzw_e;
                // (no fingerprints, or a store that overflowed: the next
                // levels read the words as they are)
                if (run.nx_from < 0 && (rc = materialize_lzw())) break;
            }
            if (lazy && h.n_stored > n_stored) {
                c->lazy_s = n_stored;
                c->lazy_e = h.n_stored;
                c->lazy_args = run.base;
                run.defer_s = n_stored;
                run.defer_e = h.n_stored;
                run.succ_mask = succ_mask;
            }
            n_stored = h.n_stored;
            if (h.found != kEmpty) { found = true; found_cost = cost; Le[cost] = n_stored; break; }
            if (h.overflow_chunk != LLONG_MAX) run.otf = true;
        }
        Le[cost] = n_stored;
        if (last_round) break;
        if (run.otf && shortage == -1) shortage = cost;
    }
    if (rc) {
        hipStreamSynchronize(s);
        drain_events(c);
        return res->status = rc;
    }
    if (!noalloc && o.hash_slots_log2) {  // (the table is reusable by the next call)
        c->tab_ok = true;
        c->tab_ptr = c->d_tab;
        c->tab_mask = c->tmask;
    }
    if (cost == maxc + 1) cost--;  // code/ltli6463.cu:1445
    if (found || neg_answer) cost = found_cost;
    res->cost = cost;
    res->n_stored = n_stored;
    res->on_the_fly = run.otf || h.overflow_chunk != LLONG_MAX;
    res->candidates = run.candidates;
    res->checked = run.checked;
    res->words_deferred = c->lazy_e - c->lazy_s;
    res->words_unwritten = c->lzw_e - c->lzw_s;
    res->deferred_succ = run.defer_s >= 0 ? run.succ_mask : 0;
    res->hbm_bytes = run.alg_bytes;
    for (auto &k : c->stats) res->kernel_seconds += k.seconds;

    if (found || neg_answer) {
        // Candidate id -> block -> operator and operand store indices.
        const uint64_t id = neg_answer ? neg_id : h.found;
        const Block *blk = nullptr;
        for (const auto &b : run.blocks)
            if (id >= b.id_base && id < b.id_base + b.size) { blk = &b; break; }
        if (!blk) {
            report_hip(hipErrorUnknown, "block of the found id", __LINE__);
            return res->status = LTL_E_HIP;
        }
        res->all_ltls = blk->id_base + blk->size;
        res->found_id = (int64_t)id;
        const uint64_t rel = id - blk->id_base;
        int32_t l, r;
        if (blk->op == LTL_AND || blk->op == LTL_OR || blk->op == LTL_UNTIL) {
            const uint64_t pair = blk->op == LTL_UNTIL ? rel >> 1 : rel;
            const int32_t sl = (int32_t)(blk->ls + (int64_t)(pair / (uint64_t)blk->nR));
            const int32_t sr = (int32_t)(blk->rs + (int64_t)(pair % (uint64_t)blk->nR));
            const bool swap = blk->op == LTL_UNTIL && (rel & 1);
            l = swap ? sr : sl;
            r = swap ? sl : sr;
        } else {
            l = (int32_t)(blk->ls + (int64_t)rel);
            r = -1;
        }
        std::vector<ltl_node> nodes;
        std::vector<int64_t> src;
        if (int32_t e = build_tree(c, blk->op, l, r, -1, nodes, src, n_stored)) return res->status = e;
        if (neg_answer) {
            // The accepted formula is the negation of the recorded candidate.
            for (auto &nd : nodes) {
                if (nd.left >= 0) ++nd.left;
                if (nd.right >= 0) ++nd.right;
            }
            nodes.insert(nodes.begin(), ltl_node{LTL_NOT, -1, 1, -1});
        }
        res->n_nodes = (int32_t)nodes.size();
        res->nodes = (ltl_node *)malloc(nodes.size() * sizeof(ltl_node));
        memcpy(res->nodes, nodes.data(), nodes.size() * sizeof(ltl_node));
        res->formula = strdup(node_string(nodes, 0, names).c_str());
        res->status = LTL_FOUND;
    } else {
        res->all_ltls = run.all_ltls;
        res->formula = strdup("not_found");
        res->status = LTL_NOT_FOUND;
    }
    res->seconds = now_s() - t0;
    return res->status;
}

// A fingerprint-1 collision between different sets is detected in dedup
// (their fingerprints 2 differ); the call is then repeated with fresh
// fingerprint keys (a new seed, so an input built against the default keys
// does not collide again), up to kMaxAttempts in all.  The result does not
// depend on the keys: dedup is by set equality.  res->retries counts the
// repeats.  With LTL_TEST_WEAK_FP1 in the environment the first attempt
// keys fingerprint 1 with zeros (collisions become common): tests of the
// repeat path.
static constexpr int kMaxAttempts = 4;

static int32_t learn_call(ltl_ctx *c, const ltl_problem *pb, ltl_result *res, bool noalloc) {
    static const bool weak = getenv("LTL_TEST_WEAK_FP1") != nullptr;
    uint64_t seed = kDefaultSeed;
    int32_t rc = 0;
    for (int attempt = 0; attempt < kMaxAttempts; ++attempt) {
        if (attempt) {
            ltl_result_free(res);
            seed = splitmix(seed ^ ((uint64_t)std::chrono::steady_clock::now().time_since_epoch().count() +
                                    (uint64_t)attempt));
        }
        rc = learn_once(c, pb, res, noalloc, seed, weak && attempt == 0);
        res->retries = attempt;
        if (rc != LTL_E_COLLISION) break;
    }
    return rc;
}

extern "C" int32_t ltl_learn(ltl_ctx *c, const ltl_problem *pb, ltl_result *res) {
    return learn_call(c, pb, res, false);
}

// The call's device memory is this device's scratch region of
// LTL_NOALLOC_SCRATCH_BYTES (shared by every context on the device: calls
// that overlap on one GPU race); every call initialises what it reads (state,
// table, look-back words) with kernels before use.
extern "C" int32_t ltl_learn_noalloc(ltl_ctx *c, const ltl_problem *pb, ltl_result *res) {
    return learn_call(c, pb, res, true);
}

// ---------------------------------------------------------------------------
// Batched learning: one persistent launch over many instances.
// ---------------------------------------------------------------------------
static int32_t batch_fill_result(const BatchOut &bo, const AtomNames &names, ltl_result *r) {
    r->cost = bo.cost;
    r->all_ltls = bo.all_ltls;
    r->found_id = bo.found_id;
    r->n_stored = bo.n_stored;
    if (bo.status == BATCH_FOUND) {
        std::vector<ltl_node> nodes(bo.nodes, bo.nodes + bo.n_nodes);
        r->n_nodes = bo.n_nodes;
        r->nodes = (ltl_node *)malloc(nodes.size() * sizeof(ltl_node));
        memcpy(r->nodes, nodes.data(), nodes.size() * sizeof(ltl_node));
        r->formula = strdup(node_string(nodes, 0, names).c_str());
        return r->status = LTL_FOUND;
    }
    r->formula = strdup("not_found");
    return r->status = LTL_NOT_FOUND;
}

// The instances a batched call leaves to the whole-device path (traces longer
// than 64 positions, levels past the persistent kernel's limits) are small
// next to the device: each call's time is mostly its launches and per-level
// synchronisations, with the device idle in between.  Up to kHelpers + 1 of
// them run at once, from host threads, each on a context of its own (stream,
// pool, pinned staging; the calling context is one of them), so that one
// call's launches fill another's gaps (six at once: 0.040 s for 256 instances
// of 20-400 traces, four: 0.042 s).  A helper's pool is an eighth of the calling
// context's budget (a sixteenth of the free memory when that is 0, at most
// 8 GiB); an instance that exhausts it (out of memory, or a store
// that overflowed into check-only levels) is run again on the calling
// context.  The learned order (a shared scorer) and the enumeration log (one
// per context) keep the calls on the calling context, one after another.
static constexpr int kHelpers = 5;

static void run_whole_device(ltl_ctx *c, const ltl_problem *pbs, ltl_result *res, const std::vector<int> &ids) {
    const ltl_options &o = c->opt;
    const bool serial = ids.size() < 2 || o.order == LTL_ORDER_LEARNED || (c->log_cap > 0 && c->log_per_block > 0);
    std::vector<ltl_ctx *> ctxs{c};
    if (!serial) {
        ltl_options ho = o;
        size_t budget = o.hbm_budget_bytes / 8;
        if (!o.hbm_budget_bytes) {
            size_t fr = 0, tot = 0;
            if (hipSetDevice(o.device) == hipSuccess && hipMemGetInfo(&fr, &tot) == hipSuccess)
                budget = std::min<size_t>(fr / 16, (size_t)8 << 30);
        }
        ho.hbm_budget_bytes = budget;
        const int want = (int)std::min<size_t>(kHelpers, ids.size() - 1);
        if (budget >= ((size_t)256 << 20)) {
            while ((int)c->helpers.size() < want) {
                int32_t st = 0;
                ltl_ctx *h = ltl_create(&ho, &st);
                if (!h) break;
                c->helpers.push_back(h);
            }
            for (int k = 0; k < want && k < (int)c->helpers.size(); ++k)
                if (ltl_set_options(c->helpers[k], &ho) == 0) ctxs.push_back(c->helpers[k]);
        }
    }
    if (ctxs.size() == 1) {
        for (int i : ids) ltl_learn(c, &pbs[i], &res[i]);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::vector<int>> again(ctxs.size());
    auto work = [&](size_t t) {
        for (size_t k; (k = next.fetch_add(1)) < ids.size();) {
            const int i = ids[k];
            ltl_learn(ctxs[t], &pbs[i], &res[i]);
            if (t > 0 && (res[i].status == LTL_E_OOM || res[i].on_the_fly)) again[t].push_back(i);
        }
    };
    std::vector<std::thread> th;
    for (size_t t = 1; t < ctxs.size(); ++t) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    for (const auto &v : again)
        for (int i : v) {
            ltl_result_free(&res[i]);
            ltl_learn(c, &pbs[i], &res[i]);
        }
}

extern "C" int32_t ltl_learn_batch(ltl_ctx *c, const ltl_problem *pbs, int32_t n, ltl_result *res) {
    if (!c || n < 0 || (n > 0 && (!pbs || !res))) return LTL_E_ARG;
    const double t0 = now_s();
    const ltl_options &o = c->opt;
    ltl_batch_stats bs{};
    bs.n = n;
    std::vector<AtomNames> names(n);
    std::vector<int> dev;
    for (int i = 0; i < n; ++i) {
        memset(&res[i], 0, sizeof(res[i]));
        res[i].found_id = -1;
        const int32_t pre = precheck_host(o, &pbs[i], &res[i], names[i]);
        if (pre < 0) res[i].status = pre;
        else if (pre == 1) bs.host++;
        else dev.push_back(i);
    }
    std::vector<int> deferred;
    if (!dev.empty() && (o.max_cost + 2 > kBatchLevels || o.dedup_mode != LTL_DEDUP_EXACT)) {
        deferred = dev;
        dev.clear();
    }
    // traces longer than 64 positions: the whole-device path
    for (size_t k = 0; k < dev.size();)
        if (words_per_trace(&pbs[dev[k]]) > 1) {
            deferred.push_back(dev[k]);
            dev.erase(dev.begin() + (long)k);
        } else {
            ++k;
        }
    if (!dev.empty()) {
        if (hipSetDevice(o.device) != hipSuccess) return LTL_E_HIP;
        const int m = (int)dev.size();
        std::vector<BatchInst> inst(m);
        std::vector<uint64_t> in;
        std::vector<int64_t> work(m);
        // the instances' host preparation (packing, per-word records) in
        // parallel; the concatenation below keeps the instance order
        std::vector<Prepared> preps(m);
        {
            const int T = (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 16u);
            std::atomic<int> next{0};
            auto work_fn = [&] {
                for (int k; (k = next.fetch_add(1)) < m;) prepare_instance(&pbs[dev[k]], slot_width(&pbs[dev[k]]), preps[k]);
            };
            if (T <= 1 || m < 16) {
                work_fn();
            } else {
                std::vector<std::thread> th;
                for (int t = 0; t < T; ++t) th.emplace_back(work_fn);
                for (auto &x : th) x.join();
            }
        }
        for (int k = 0; k < m; ++k) {
            const ltl_problem *pb = &pbs[dev[k]];
            const Prepared &q = preps[k];
            BatchInst &I = inst[k];
            memset(&I, 0, sizeof(I));
            I.W = pb->num_pos + pb->num_neg; I.P = pb->num_pos; I.A = pb->num_atoms;
            I.nw = q.nw; I.nsw = q.nsw; I.B = q.B;
            I.max_cost = o.max_cost; I.neg_type = o.neg_type; I.hamming_k = o.hamming_k;
            for (int j = 0; j < 8; ++j) I.cost[j] = o.cost[j];
            auto put = [&](const void *p, size_t bytes) {
                const int64_t off = (int64_t)in.size();
                in.resize(in.size() + (bytes + 7) / 8);
                memcpy(in.data() + off, p, bytes);
                return off;
            };
            I.off_aux = put(q.aux.data(), q.aux.size() * sizeof(WordAux));
            I.off_saux = put(q.saux.data(), q.saux.size() * sizeof(SumAux));
            I.off_atoms = put(q.atom_words.data(), q.atom_words.size() * 8);
            I.off_sums = put(q.atom_sums.data(), q.atom_sums.size() * 8);
            work[k] = (int64_t)I.W * I.A * I.A;
        }
        // largest instances first so that they do not form the tail (one
        // launch per layout code; every instance here is packed)
        std::vector<int32_t> order(m);
        for (int k = 0; k < m; ++k) order[k] = k;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            if (inst[a].B != inst[b].B) return inst[a].B < inst[b].B;
            return work[a] > work[b];
        });

        c->lazy_s = c->lazy_e = 0;
        c->lzw_s = c->lzw_e = 0;
        if (int32_t e = ensure_pool(c)) return e;
        c->tab_ok = false;  // (the regions overwrite the pool)
        const ltl_batch_params &bp = c->batch;
        const int tile = bp.tile == 256 || bp.tile == 1024 ? bp.tile : 512;
        const int wb = bp.word_block == 8 ? 8 : 4;
        int G = bp.workgroups > 0 ? bp.workgroups : (tile == 1024 ? 1 : 2) * std::max(c->n_cu, 1);
        char *p = c->pool;
        auto take = [&](size_t bytes) { char *q = p; p += align_up(bytes, 256); return q; };
        BatchInst *d_inst = (BatchInst *)take(inst.size() * sizeof(BatchInst));
        BatchOut *d_out = (BatchOut *)take((size_t)m * sizeof(BatchOut));
        int32_t *d_order = (int32_t *)take((size_t)m * 4);
        unsigned *d_next = (unsigned *)take(256);
        uint64_t *d_in = (uint64_t *)take(in.size() * 8);
        // the regions start at a multiple of 64 MiB past the inputs, so
        // that calls with inputs of up to 64 MiB share one layout and the
        // regions a previous call left clean need no fill
        const size_t used = align_up((size_t)(p - c->pool), (size_t)64 << 20);
        p = c->pool + std::min(used, c->pool_bytes);
        if (used + (size_t)G * (64ull << 20) > c->pool_bytes)
            G = (int)std::max<size_t>(1, (c->pool_bytes - std::min(used, c->pool_bytes)) / (64ull << 20));
        const uint64_t region = ((c->pool_bytes - std::min(used, c->pool_bytes)) / (size_t)G) / 4096 * 4096;
        uint32_t tslots = 1u << 20;
        while (tslots > 4096 && (uint64_t)tslots * 16 > region / 4) tslots >>= 1;
        // workgroups of the largest launch: regions [0, g_used) are used
        int g_used = 0;
        for (int g0 = 0; g0 < m;) {
            int g1 = g0;
            while (g1 < m && inst[order[g1]].B == inst[order[g0]].B) ++g1;
            g_used = std::max(g_used, std::min(G, g1 - g0));
            g0 = g1;
        }
        if (c->clean_base != p || c->clean_region != region || c->clean_tslots != tslots) c->clean_n = 0;
        const int clean_before = c->clean_n;
        c->clean_n = 0;  // (known again once the launches have finished)
        hipStream_t s = c->stream;
        HIP_TRY(hipMemcpyAsync(d_inst, inst.data(), inst.size() * sizeof(BatchInst), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_order, order.data(), order.size() * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_in, in.data(), in.size() * 8, hipMemcpyHostToDevice, s));
        if (g_used > clean_before)
            HIP_TRY(launch_fill_regions(p + (size_t)clean_before * region, region, tslots, g_used - clean_before, s));
        BatchArgs ba{};
        ba.inst = d_inst; ba.out = d_out; ba.in = d_in;
        ba.regions = p; ba.region_bytes = region; ba.tslots = tslots; ba.next = d_next;
        ba.defer_limit = bp.defer_limit > 0 ? bp.defer_limit : ((int64_t)1 << 22);
        hipEvent_t e0 = get_event(c), e1 = get_event(c);
        HIP_TRY(hipEventRecord(e0, s));
        for (int g0 = 0; g0 < m;) {
            int g1 = g0;
            while (g1 < m && inst[order[g1]].B == inst[order[g0]].B) ++g1;
            ba.order = d_order + g0;
            ba.n_inst = g1 - g0;
            HIP_TRY(hipMemsetAsync(d_next, 0, 4, s));
            HIP_TRY(launch_batch(inst[order[g0]].B, tile, wb, std::min(G, g1 - g0), ba, s));
            g0 = g1;
        }
        HIP_TRY(hipEventRecord(e1, s));
        std::vector<BatchOut> out(m);
        HIP_TRY(hipMemcpyAsync(out.data(), d_out, (size_t)m * sizeof(BatchOut), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        // (every instance leaves its region's table empty)
        c->clean_base = p;
        c->clean_region = region;
        c->clean_tslots = tslots;
        c->clean_n = std::max(clean_before, g_used);
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        c->free_events.push_back(e0);
        c->free_events.push_back(e1);
        bs.kernel_seconds = ms * 1e-3;
        for (int k = 0; k < m; ++k) {
            const int i = dev[k];
            const BatchOut &bo = out[k];
            if (bo.status == BATCH_FOUND || bo.status == BATCH_NOT_FOUND) {
                batch_fill_result(bo, names[i], &res[i]);
                bs.on_device++;
            } else {  // (a fingerprint collision too: ltl_learn repeats with fresh keys)
                deferred.push_back(i);
            }
        }
    }
    c->W = 0;  // the store of the last call is not readable after a batch
    run_whole_device(c, pbs, res, deferred);
    bs.deferred += (int32_t)deferred.size();
    int32_t solved = 0;
    for (int i = 0; i < n; ++i) solved += res[i].status == LTL_FOUND;
    bs.seconds = now_s() - t0;
    c->batch_stats = bs;
    return solved;
}

extern "C" int32_t ltl_set_batch_params(ltl_ctx *c, const ltl_batch_params *p) {
    if (!c || !p) return LTL_E_ARG;
    if (p->tile != 256 && p->tile != 512 && p->tile != 1024) return LTL_E_ARG;
    if (p->word_block != 4 && p->word_block != 8) return LTL_E_ARG;
    if (p->workgroups < 0 || p->defer_limit < 0) return LTL_E_ARG;
    c->batch = *p;
    return 0;
}

extern "C" int32_t ltl_get_batch_params(const ltl_ctx *c, ltl_batch_params *p) {
    if (!c || !p) return LTL_E_ARG;
    *p = c->batch;
    return 0;
}

extern "C" int32_t ltl_batch_last_stats(const ltl_ctx *c, ltl_batch_stats *s) {
    if (!c || !s) return LTL_E_ARG;
    *s = c->batch_stats;
    return 0;
}

extern "C" int32_t ltl_autotune_batch(ltl_ctx *c, const ltl_problem *pbs, int32_t n, ltl_batch_trial *trials,
                                      int32_t max_trials) {
    if (!c || n <= 0 || !pbs) return LTL_E_ARG;
    std::vector<ltl_result> res(n);
    auto run = [&](const ltl_batch_params &p, ltl_batch_trial *t) -> int32_t {
        if (int32_t e = ltl_set_batch_params(c, &p)) return e;
        const int32_t r = ltl_learn_batch(c, pbs, n, res.data());
        for (auto &x : res) ltl_result_free(&x);
        if (r < 0) return r;
        if (t) {
            t->params = p;
            t->seconds = c->batch_stats.seconds;
            t->deferred = c->batch_stats.deferred;
        }
        return 0;
    };
    const ltl_batch_params keep = c->batch;
    if (int32_t e = run(keep, nullptr)) return e;  // loads the code objects
    std::vector<ltl_batch_trial> all;
    for (int tile : {256, 512, 1024})
        for (int wb : {4, 8})
            for (int64_t lim : {(int64_t)1 << 16, (int64_t)1 << 18, (int64_t)1 << 20}) {
                ltl_batch_params p{tile, wb, 0, 0, lim};
                ltl_batch_trial t{};
                if (int32_t e = run(p, &t)) return e;
                all.push_back(t);
            }
    size_t best = 0;
    for (size_t k = 1; k < all.size(); ++k)
        if (all[k].seconds < all[best].seconds) best = k;
    c->batch = all[best].params;
    for (int32_t k = 0; k < max_trials && k < (int32_t)all.size(); ++k) trials[k] = all[k];
    return (int32_t)all.size();
}

extern "C" int32_t ltl_get_kernel_stats(const ltl_ctx *c, ltl_kernel_stat *out, int32_t max) {
    if (!c) return LTL_E_ARG;
    for (int k = 0; k < K_NKIND && k < max; ++k) out[k] = c->stats[k];
    return K_NKIND;
}

extern "C" int32_t ltl_set_scorer(ltl_ctx *c, ltl_scorer *sc) {
    if (!c) return LTL_E_ARG;
    if (sc && sc->device != c->opt.device) return LTL_E_ARG;
    c->scorer = sc;
    return 0;
}

extern "C" int32_t ltl_set_launch_log(ltl_ctx *c, ltl_launch_fn fn, void *user) {
    if (!c) return LTL_E_ARG;
    c->launch_fn = fn;
    c->launch_user = fn ? user : nullptr;
    return 0;
}

extern "C" int32_t ltl_set_log(ltl_ctx *c, int64_t max_rows, int32_t per_block) {
    if (!c || max_rows < 0 || per_block < 0) return LTL_E_ARG;
    HIP_TRY(hipSetDevice(c->opt.device));
    if (max_rows > c->log_cap) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (c->d_log) hipFree(c->d_log);
        c->d_log = nullptr;
        c->log_cap = 0;
        HIP_TRY(hipMalloc((void **)&c->d_log, (size_t)max_rows * (kFeat + 1) * sizeof(float)));
    }
    c->log_cap = max_rows;
    c->log_per_block = per_block;
    c->log_n = 0;
    return 0;
}

extern "C" int64_t ltl_get_log(ltl_ctx *c, float *rows, int64_t max_rows) {
    if (!c || max_rows < 0 || (max_rows && !rows)) return LTL_E_ARG;
    const int64_t n = std::min(max_rows, c->log_n);
    if (n > 0) {
        HIP_TRY(hipSetDevice(c->opt.device));
        HIP_TRY(hipMemcpyAsync(rows, c->d_log, (size_t)n * (kFeat + 1) * sizeof(float), hipMemcpyDeviceToHost,
                               c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return c->log_n;
}

extern "C" int32_t ltl_levels(const ltl_ctx *c, int64_t *start, int64_t *end, int32_t max) {
    if (!c) return LTL_E_ARG;
    const int32_t n = (int32_t)c->lvl_s.size();
    for (int32_t i = 0; i < n && i < max; ++i) {
        if (start) start[i] = c->lvl_s[i];
        if (end) end[i] = c->lvl_e[i];
    }
    return n;
}

// Words of the sets emitted without them: the level below the last
// generated one first (the deferred level's parents are there), then the
// deferred level (words and summaries).
static hipError_t materialize_pending(ltl_ctx *c) {
    if (c->lzw_e > c->lzw_s) {
        const hipError_t e = launch_materialize(c->lzw_args, c->lzw_s, c->lzw_e, c->stream, false);
        if (e != hipSuccess) return e;
        c->lzw_s = c->lzw_e = 0;
    }
    if (c->lazy_e > c->lazy_s) {
        const hipError_t e = launch_materialize(c->lazy_args, c->lazy_s, c->lazy_e, c->stream);
        if (e != hipSuccess) return e;
        c->lazy_s = c->lazy_e = 0;
    }
    return hipSuccess;
}

extern "C" int32_t ltl_store_read(ltl_ctx *c, int64_t first, int64_t count, uint64_t *cs, uint8_t *op,
                                  int32_t *left, int32_t *right) {
    if (!c || first < 0 || count < 0 || !c->W) return LTL_E_ARG;
    if (count == 0) return 0;
    if (first + count > c->capacity) return LTL_E_ARG;
    HIP_TRY(hipSetDevice(c->opt.device));
    HIP_TRY(materialize_pending(c));
    HIP_TRY(hipStreamSynchronize(c->stream));
    // the whole tiles holding [first, first + count), one contiguous copy
    const int64_t t0 = first / 64, t1 = (first + count - 1) / 64 + 1;
    const size_t tile_words = (size_t)c->nw * 64;
    std::vector<uint64_t> buf((size_t)(t1 - t0) * tile_words);
    HIP_TRY(hipMemcpy(buf.data(), c->d_store + (size_t)t0 * tile_words, buf.size() * 8, hipMemcpyDeviceToHost));
    if (cs && c->B == 1)
        for (int64_t k = 0; k < count; ++k) {
            // packed: position p of trace i at bit off + len - 1 - p of its word
            const uint64_t *w0 = buf.data() + set_base(first + k, c->nw) - (size_t)t0 * tile_words;
            for (int i = 0; i < c->W; ++i) {
                const int L = c->lens[i];
                const uint64_t seg = (w0[(size_t)c->pk_word[i] * 64] >> c->pk_off[i]) & len_mask(L);
                cs[(size_t)k * c->W + i] = __builtin_bitreverse64(seg) >> (64 - L);
            }
        }
    else if (cs)
        for (int64_t k = 0; k < count; ++k) {
            // K words per trace: the store order is the caller's
            const uint64_t *w0 = buf.data() + set_base(first + k, c->nw) - (size_t)t0 * tile_words;
            for (int w = 0; w < c->nw; ++w) cs[(size_t)k * c->nw + w] = w0[(size_t)w * 64];
        }
    if (op) HIP_TRY(hipMemcpy(op, c->d_op + first, (size_t)count, hipMemcpyDeviceToHost));
    if (left) HIP_TRY(hipMemcpy(left, c->d_left + first, (size_t)count * 4, hipMemcpyDeviceToHost));
    if (right) HIP_TRY(hipMemcpy(right, c->d_right + first, (size_t)count * 4, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int32_t ltl_store_summaries(ltl_ctx *c, int64_t first, int64_t count, uint64_t *sums) {
    if (!c || !sums || first < 0 || count < 0 || !c->W) return LTL_E_ARG;
    if (count == 0) return 0;
    if (first + count > c->capacity) return LTL_E_ARG;
    HIP_TRY(hipSetDevice(c->opt.device));
    HIP_TRY(materialize_pending(c));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int nsw = c->nsw, W = c->W;
    const int64_t t0 = first / 64, t1 = (first + count - 1) / 64 + 1;
    const size_t tile_words = (size_t)4 * nsw * 64;
    std::vector<uint64_t> buf((size_t)(t1 - t0) * tile_words);
    HIP_TRY(hipMemcpy(buf.data(), c->d_sum + (size_t)t0 * tile_words, buf.size() * 8, hipMemcpyDeviceToHost));
    memset(sums, 0, (size_t)count * 4 * nsw * 8);
    for (int64_t k = 0; k < count; ++k) {
        const uint64_t *z = buf.data() + set_base(first + k, 4 * nsw) - (size_t)t0 * tile_words;
        for (int v = 0; v < 4; ++v)
            for (int t = 0; t < W; ++t)
                if ((z[(size_t)(v * nsw + t / 64) * 64] >> (t % 64)) & 1ull) {
                    const int i = c->sum_order[t];
                    sums[((size_t)k * 4 + v) * nsw + i / 64] |= 1ull << (i % 64);
                }
    }
    return 0;
}

extern "C" const char *ltl_version(void) { return "ltl-amd 0.1 (gfx950)"; }
