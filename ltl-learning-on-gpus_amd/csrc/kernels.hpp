// kernels.hpp -- device data layout and launch interface of the enumeration
// kernels (gfx950, wave64).
//
// Store layout.  A characteristic set (CS) over W traces is held as nw
// words (the packed layout below, or K words per trace).  The store is
// tile-major: sets in tiles of 64,
// word w of set s at ((s / 64) * nw + w) * 64 + s % 64.  One lane handles one
// candidate, so a wave reads or writes 64 consecutive sets of one word as a
// contiguous 512-byte segment, and the words of one tile follow each other
// (a wave streams through 512-byte segments of one region, whatever the
// capacity).  Summaries use the same layout with 4 * nsw words per set.
// Traces longer than 64 positions (layout code B = 0): K words per trace, word
// k (positions 64k..64k+63) of trace i is word i * K + k of the set.
// Packed layout (layout code B = 1, every instance of traces of at most 64
// positions): whole traces packed into words (best fit by decreasing
// length), a trace of length L at bits
// [o, o + L) with position p at bit o + L - 1 - p, so that "the next
// position" is the next lower bit and the until recurrence runs upward like a
// carry.  Summary word g covers the traces 64g .. 64g + 63 of the packing
// order (a group may start or end inside a word).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace ltl {

constexpr uint64_t kEmpty = ~0ull;          // empty table key / value, no result

// First word of set s in a tile-major array of `words` words per set; the
// following words are 64 entries apart.
__host__ __device__ __forceinline__ size_t set_base(int64_t s, int words) {
    return (size_t)(s >> 6) * (size_t)words * 64 + (size_t)(s & 63);
}
// Word stride inside a tile (the layout of set_base): word w of set s at
// set_base(s, words) + w * kTileWS.  ChunkArgs::ws holds the same value.
constexpr int64_t kTileWS = 64;
constexpr uint64_t kPending = 1ull << 59;   // table value: a candidate id (else a store index)
constexpr uint32_t kEpochs = 14;            // table epochs (1..14) between two clears (tab_key)

// Per packed word: dedup hash keys and check masks (wave-uniform).
struct WordAux {
    uint64_t k1, k2;  // multilinear keys of the two fingerprints (two 32-bit halves each)
    uint64_t lm;      // bits inside the traces (length mask)
    uint64_t chk;     // position 0 of every trace in the word
    uint64_t tgt;     // position 0 of every positive trace in the word
    // packed layout only: bits whose next position is in the same trace
    // (every trace bit but the last position's); and, in one word so that it
    // is read with one scalar load, the packing index of the word's first
    // trace as group (bits 0-31) and index within the group (32-39), and the
    // word's trace count (40-47)
    uint64_t inner;
    uint64_t sinfo;
    uint64_t pad_;
};

__host__ __device__ __forceinline__ uint64_t make_sinfo(uint32_t sword, int sbit, int ntr) {
    return (uint64_t)sword | ((uint64_t)sbit << 32) | ((uint64_t)ntr << 40);
}
__host__ __device__ __forceinline__ uint32_t sinfo_word(uint64_t i) { return (uint32_t)i; }
__host__ __device__ __forceinline__ int sinfo_bit(uint64_t i) { return (int)((i >> 32) & 0xff); }
__host__ __device__ __forceinline__ int sinfo_ntr(uint64_t i) { return (int)((i >> 40) & 0xff); }

// Per summary word j (traces 64j..64j+63): positive traces and valid traces.
struct SumAux {
    uint64_t tgt, val;
    uint64_t val2;  // traces of at least two positions
};

// Device-resident run state (one per context).
struct DevState {
    unsigned long long found;     // minimal accepted candidate id, kEmpty = none
    long long n_stored;           // sets in the store
    long long chunk_base;         // store index of the current chunk's first survivor
    long long chunk_total;        // survivors of the current chunk
    long long overflow_chunk;     // first chunk that overflowed the store
    unsigned long long dropped;   // survivors not stored (overflow)
    int collision;                // fingerprint collision detected
    int probe_fail;               // table full
    unsigned long long neg_found[2];  // per level parity: smallest id of a candidate whose
                                      // complement is accepted
    // Deferred level: per successor operator (~, X, F, G), the smallest store
    // index of a set whose successor is accepted, kEmpty = none
    unsigned long long succ_min[4];
    int lb_fail;                  // a look-back wait ran out (the chunk was not stored)
    int nx_over;                  // an emit's next-level fingerprints did not fit ChunkArgs::nx_stride
    // Relaxed keys (dedup modes 1-3): the smallest id of an accepted
    // candidate that survived dedup, folded into `found` after the chunk's
    // dedup launch (a dedup launch that lowered `found` itself would stop
    // its own later workgroups).
    unsigned long long rx_found;
};

// One chunk of one (operator, split) block, in canonical order.  Binary:
// rows a in [a0, a1) of the left level, columns b in [b0, b1) of the right
// level (block-local); unary: operands a in [a0, a1), b0 = b1 = 0.
struct ChunkArgs {
    uint64_t *store;
    const WordAux *aux;
    uint64_t *tab;            // 2 words per slot: key, value
    DevState *st;
    uint32_t *t_slot;         // per candidate: table slot, or kSlotStored | the stored set it duplicates
    uint64_t *t_fp2;          // per candidate: verification fingerprint
    uint64_t *t_mask;         // per tile: survivor masks (2 for U)
    uint32_t *t_cnt;          // per tile: survivors
    uint32_t *t_off;          // per tile: exclusive prefix of t_cnt
    uint32_t *t_list;         // per survivor (chunk order): candidate temp index
    uint8_t *m_op;
    int32_t *m_left, *m_right;
    uint64_t *m_fp2;
    uint64_t *m_key;          // per stored set: its table key (for rehashing a grown table)
    int32_t keep_keys;        // 1: survivors record m_key (the table may grow)
    uint64_t *sum;            // summaries: 4 vectors of nsw words per set, word-major
    const SumAux *saux;
    int64_t capacity;
    int64_t chunk_idx;
    uint64_t id_base;         // canonical id of the block's first candidate
    int64_t ls, rs;           // store index of the left / right level start
    int64_t a0, a1, b0, b1;
    int64_t nR;               // right level size (binary), 1 (unary)
    int64_t ws;               // word stride within a tile of the store and summaries (64)
    int32_t nw;
    int32_t hamming_k;
    uint32_t tmask;           // table slots - 1
    int32_t nbs;              // 64-wide column stripes per row in the chunk
    int32_t n_tiles;          // tiles in the chunk
    int32_t nsw;              // summary words per vector: ceil(W / 64)
    int32_t anti;             // check pass: also record candidates whose complement is accepted
    int32_t npw, nps;         // words of the groups of positive traces (K words per trace: positive
                              // traces); summary words holding positive traces, ceil(P / 64)
    int32_t neg_slot;         // anti: neg_found entry written
    // Reference relaxed uniqueness keys (dedup modes 1-3 above 126 trace
    // positions): kmode 0 = exact (multilinear fingerprints), else the mode; per
    // trace (first key bit << 8) | key bits taken (modes 1, 2); the stride
    // of mode 1; words holding traces that take key bits.
    const uint32_t *kinfo;
    const uint64_t *ptrace;   // packed layout: per packed trace, trace | bit offset << 32 | length << 40
    int32_t kmode, kstride, kw_end, W;
    uint64_t kseed;           // mixed into the relaxed keys' fingerprints (0 = default)
    // Lane shift: unary chunks put candidate a at temp index a - a0 + ush
    // with ush = (ls + a0) % 64; binary chunks put column b of a row at lane
    // offset b - b0 + ush within the row's stripes, ush = (rs + b0) % 64.
    // Either way the 64 lanes of a wave read the 64 sets of one store tile.
    int32_t ush;
    int32_t K;                // words per trace (B = 0), else 1
    // The chunk's kernels stop once an accepted candidate with an id below
    // id_end is known (all ids when id_end is kEmpty; fused X/F/G blocks set
    // each block's end, so that the blocks before an accepted one still
    // store their survivors, as they would one after the other).
    uint64_t id_end = kEmpty;
    uint32_t epoch = 1;       // the table's epoch of this call (tab_key, tab_val)
    // Ranked check round submitted ahead of a sync: its workgroups do no
    // work once an accepted candidate with an id below found_below (a lower
    // level's) is known when they start (0: always work)
    uint64_t found_below = 0;
    // 0: the level was tested completely by a check pass before it is
    // generated, so the gen kernels skip the target check
    int32_t gen_check = 1;
    // Deferred level (packed layout): 1 = the emit stores the metadata only,
    // neither words nor summaries (no later pass reads them), and checks the
    // next level's unary blocks over this level: successor j of a survivor
    // (~c, X c, F c, G c) accepted lowers succ_min[j] to its store index.
    // ~c: the survivor whose candidate id is neg_pick (the smallest id whose
    // complement the level's check pass accepted).  X c, F c, G c (bits 1-3
    // of succ_mask): tested in generation, which writes the accepted ones as
    // bits of t_succ[candidate temp index].
    int32_t deferred = 0;
    int32_t succ_mask = 0;
    uint8_t *t_succ = nullptr;
    uint64_t neg_pick = kEmpty;
    // One-pass dedup of a large chunk (decoupled look-back): per workgroup a
    // status word (epoch << 34 | kind << 32 | count), zeroed once per call,
    // and a ticket counter that the last workgroup returns to 0.
    uint64_t *lb_flags = nullptr;
    uint32_t *lb_ticket = nullptr;
    uint32_t lb_epoch = 0;    // 1, 2, ... per chunk within a call (< 2^30)
    // Relaxed keys (kmode != 0): per candidate temp index, bit 0 = the
    // candidate is accepted.  The reference tests a candidate only when its
    // key is new (code/ltli6463.cu:231-243), so generation records the
    // outcome here and dedup folds it in for the survivors only.
    uint8_t *t_acc = nullptr;
    // Packed layout: the emit's waves per store tile (word ranges at summary
    // group boundaries) and, per summary group, the first and last word that
    // hold its traces.
    int32_t esplit = 1;
    const int32_t *gfirst = nullptr, *glast = nullptr;
    // Next level's unary candidates from an emit (packed layout, exact
    // keys): with nx_fp set, the emit also computes ~c, X c, F c and G c of
    // every survivor c and stores their two fingerprints, value v (f1, f2 of
    // ~, X, F, G in turn) of the survivor at store position p at
    // nx_fp[v * nx_stride + p - nx_ls + nx_ush] -- the temp index that the
    // next level's unary chunk over this level gives it (nx_ls: the level's
    // first store index, nx_ush = nx_ls % 64).  launch_insert_xfg then puts
    // them in the table in place of the ~/X/F/G generation pass.
    uint64_t *nx_fp = nullptr;
    int64_t nx_ls = 0, nx_stride = 0;
    int32_t nx_ush = 0;
    // wskip: such an emit writes the survivors' summaries, metadata and
    // fingerprints but not their words.  The store range [lzw_s, lzw_e) holds
    // sets emitted so (one level): the check kernels that read words
    // recompute a word of such a set from its parents, whose words are
    // written; launch_materialize writes them.
    int32_t wskip = 0;
    int64_t lzw_s = 0, lzw_e = 0;
    // Packed generation, exact keys: the waves that fold one task's words (1,
    // 2 or 4; a task is a tile, or R left rows of a binary stripe).  Set by
    // the launchers for chunks of few tasks, where one wave walking every
    // word of a set is the whole kernel's time; the partial sums meet in LDS.
    int32_t gsplit = 1;
};

// Check passes visit words alternately from the positive and the negative
// traces (sums do not depend on the order): a candidate that holds, or
// fails, at position 0 of almost every trace then fails within a few words
// instead of after all positive traces, and its wave stops early.  j-th
// word of n, the first np of which hold positive traces.
__host__ __device__ __forceinline__ int interleave(int j, int np, int n) {
    const int m = np < n - np ? np : n - np;
    if (j < 2 * m) return (j & 1) ? np + (j >> 1) : (j >> 1);
    j -= 2 * m;
    return np > n - np ? m + j : np + m + j;
}

// Launchers (stream-ordered, asynchronous).  B is the layout code: 1 packed
// words, 0 K words per trace (a.K).
hipError_t launch_gen(int op, int B, const ChunkArgs &a, hipStream_t s);
// X, F and G of one operand block in one pass (packed layout, exact keys):
// block j in {X, F, G} has ids id_base + j * n and its temporaries at offset
// j * slots (t_slot, t_fp2).  neg (a tested level): the negation block too,
// ids from id_base, temporaries at offset 0; then X, F, G have ids
// id_xfg + j * n and temporaries at (j + 1) * slots.
hipError_t launch_gen_xfg(const ChunkArgs &a, int64_t n, int64_t slots, bool neg, uint64_t id_xfg, hipStream_t s);
// The same blocks (~ and X, F, G; neg, a tested level) from fingerprints an
// emit of the operand level computed (ChunkArgs::nx_fp, `fp` with value
// stride `stride`): the table inserts and temporaries of k_gen_xfg, without
// reading the operand level again.
hipError_t launch_insert_xfg(const ChunkArgs &a, int64_t n, int64_t slots, uint64_t id_xfg, const uint64_t *fp,
                             int64_t stride, hipStream_t s);
// The &, | (and, with_until, U) blocks of one split (same operand levels:
// equal costs of the three operators) in one generation pass of a tested
// level, packed layout, exact keys.  The & block is a's (ids from
// a.id_base, temporaries at a's); the | and U blocks' first ids and their
// temporaries' offsets (candidate slots from a's) are f's; each block's
// decide ... emit later runs at its canonical position with its own region.
struct FusedBin {
    uint64_t id_or, id_u;
    int64_t off_or, off_u;
};
hipError_t launch_gen_fused(const ChunkArgs &a, const FusedBin &f, bool with_until, hipStream_t s);
hipError_t launch_decide(int op, const ChunkArgs &a, hipStream_t s);
// A chunk of at most kSmallTiles tiles (4096 candidates) is deduplicated by
// one workgroup: decide, scan, finish and compact in a single launch.  Each of
// its 16 waves decides up to 4 tiles in turn (a table read per candidate and
// tile), so larger chunks take the four-kernel path.
constexpr int64_t kSmallTiles = 64;
hipError_t launch_dedup_small(int op, const ChunkArgs &a, hipStream_t s);
// Chunks of up to kLbMaxTiles tiles (65536 candidates): decide, scan, finish
// and compact in one launch of 4-tile workgroups chained by a decoupled
// look-back (a.lb_flags, a.lb_ticket, a.lb_epoch set).  Above that the chain
// of coherent status reads costs more than the four kernels' launches (config
// 2's chunks of millions of candidates: 42.7 ms per call against 40.7 ms).
constexpr int64_t kLbMaxTiles = 1024;
hipError_t launch_dedup_lb(int op, const ChunkArgs &a, hipStream_t s);
hipError_t launch_finish(const ChunkArgs &a, hipStream_t s);
// Relaxed keys, after a chunk's dedup: `found` lowered to the smallest
// accepted survivor (rx_found); in the chunk that overflowed the store, every
// accepted candidate after the first survivor that did not fit lowers it too
// (from there on the reference tests every candidate, code/ltli6463.cu:198-208,
// 1021-1027).
hipError_t launch_rx_tail(int op, const ChunkArgs &a, hipStream_t s);
// Survivor list (candidate temp index per output position), then emit with
// one lane per output position.
hipError_t launch_compact(int op, const ChunkArgs &a, hipStream_t s);
hipError_t launch_emit(int op, int B, const ChunkArgs &a, hipStream_t s);
// Words and, with sums, summaries of the stored sets [s0, s1) recomputed
// from their parents (operator and operand indices of the metadata): sets
// emitted with deferred = 1 (sums), or with wskip = 1 (words only).
hipError_t launch_materialize(const ChunkArgs &a, int64_t s0, int64_t s1, hipStream_t s, bool sums = true);
hipError_t launch_atoms(int B, const ChunkArgs &a, int n_atoms, hipStream_t s);
// dst[r * ws + c] = src[r * ncols + c] for r < nrows, c < ncols.
hipError_t launch_scatter(uint64_t *dst, int64_t ws, const uint64_t *src, int ncols, int64_t nrows, hipStream_t s);
// The found formula's nodes below its root, as a breadth-first walk from the
// root's operands (l, r) visits them: out[0] = the count (-1: more than cap,
// -2: an index beyond the store), then (op, left, right) of each node.
hipError_t launch_tree(const uint8_t *op, const int32_t *left, const int32_t *right, int64_t n_stored, int32_t l,
                       int32_t r, int32_t *out, int32_t cap, hipStream_t s);
// Check-only pass over a whole block (no dedup, no store): atomicMin of the
// canonical id of every accepted candidate.
hipError_t launch_check(int op, int B, const ChunkArgs &a, hipStream_t s);
hipError_t launch_fill_u64(uint64_t *p, uint64_t v, size_t n, hipStream_t s);
// Insert (keys[s], s) for s < n into an emptied table of tmask + 1 slots
// (keys of table epoch `epoch`).
hipError_t launch_rehash(uint64_t *tab, uint32_t tmask, const uint64_t *keys, int64_t n, uint32_t epoch,
                         int *probe_fail, hipStream_t s);
// The device address of the static scratch region of ltl_learn_noalloc on
// `device` (which must be current): looked up once per device and kept.
hipError_t noalloc_scratch(int device, char **ptr);
// Exclusive scan of t_cnt into t_off (n elements); temp sized by scan_temp_bytes.
size_t scan_temp_bytes(int n_max);
hipError_t launch_scan(const uint32_t *in, uint32_t *out, int n, void *temp, size_t temp_bytes,
                       hipStream_t s);

}  // namespace ltl
