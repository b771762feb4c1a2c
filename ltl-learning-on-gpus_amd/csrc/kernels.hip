// kernels.hip -- characteristic-set enumeration kernels for gfx950 (wave64).
//
// Per chunk of candidates the host enqueues, on one stream:
//   gen     compute every candidate set word by word, fold it into two multilinear
//           fingerprints and the target check, test it (atomicMin of the
//           canonical id of accepted candidates), and intern fingerprint 1 in
//           the open-addressing table (64-bit atomicCAS on the key, atomicMin
//           of the candidate id on the value);
//   decide  a candidate survives iff the table value still holds its own id;
//           duplicates are verified against fingerprint 2 of the winner;
//   scan    survivor counts per 64-candidate tile -> store offsets;
//   finish  advance the store count, detect overflow;
//   compact the survivors' candidate indices, in canonical order;
//   emit    one lane per output position: recompute the survivor, append it
//           to the store (full contiguous segments per wave) with its
//           summaries and parent indices; commit its table value.
// This replaces the reference's per-thread operator kernels + Thrust
// compaction (code/ltli6463.cu:176-873, 994-1035).  Operators follow
// code/ltli6463.cu:203, 295, 385, 468, 551-556, 638-645, 739-771, on words
// of whole packed traces (masks and one carry chain, device_ops.hpp), or on
// K words per trace for traces longer than 64 positions.
#include <hipcub/hipcub.hpp>

#include "../../include/ltl_amd.h"
#include "kernels.hpp"
#include "device_ops.hpp"

namespace ltl {
namespace {

// Word 3 of a raw buffer resource: 32-bit data format (nonzero, so that the
// descriptor is valid), no swizzle, no stride.
constexpr int kBufferWord3 = 0x00020000;

// Workgroups of a binary chunk are laid out so that consecutive workgroups
// dispatched to one XCD (dispatch is round-robin over the 8 XCDs) walk the
// left rows of the same right stripe, which then stays in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
    const int q = nb / 8, r = nb % 8, x = bid % 8, j = bid / 8;
    return x < r ? x * (q + 1) + j : r * (q + 1) + (x - r) * q + j;
}

__device__ __forceinline__ bool run_stopped(const ChunkArgs &A) {
    return __hip_atomic_load(&A.st->found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < A.id_end;
}
// A look-back wait of the dedup ran out: the survivor lists of the call are
// not trusted (the call fails at its next synchronisation).
__device__ __forceinline__ bool look_back_failed(const ChunkArgs &A) {
    return __hip_atomic_load(&A.st->lb_fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}
__device__ __forceinline__ bool chunk_otf(const ChunkArgs &A) {
    return __hip_atomic_load(&A.st->overflow_chunk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < A.chunk_idx;
}

struct Tile {
    int64_t a, b;      // binary: left row, right column (block-local); unary: operand
    int64_t ti;        // tile index in the chunk
    bool valid;
};

// Word-split generation (A.gsplit = S, 1, 2 or 4): the 4 waves of a
// workgroup take 4 / S tasks, each folded by S waves over word ranges
// [w0, w1) of an even split; S = 1 is one wave per task over every word.
struct Split {
    int S, r, tw, per;  // waves per task, the wave's range, its task in the workgroup, tasks per workgroup
    int w0, w1;
};
__device__ __forceinline__ Split split_of(const ChunkArgs &A) {
    Split s;
    s.S = A.gsplit;
    const int v = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    s.r = v & (s.S - 1);
    s.tw = v / s.S;
    s.per = 4 / s.S;
    s.w0 = (int)((int64_t)s.r * A.nw / s.S);
    s.w1 = (int)((int64_t)(s.r + 1) * A.nw / s.S);
    return s;
}

// Binary chunks: tasks of R consecutive left rows of one 64-column right
// stripe, numbered stripe by stripe (task q: stripe q / nAg, rows R (q % nAg)
// .. + R - 1 with nAg = ceil(rows / R)); a workgroup takes per consecutive
// tasks, so no wave idles when the left operand has few rows (5 atoms: the
// rows of one stripe, then the next stripe's).  t.a is the task's first row,
// t.ti its tile; row r's tile is t.ti + r * nbs.  Without a task (false) t is
// an invalid lane of the chunk's first row.
template <int R = 1>
__device__ __forceinline__ bool tile_binary(const ChunkArgs &A, Tile &t, const Split &sp) {
    const int64_t nAg = (A.a1 - A.a0 + R - 1) / R;
    const int64_t q = (int64_t)xcd_remap((int)blockIdx.x, (int)gridDim.x) * sp.per + sp.tw;
    const int64_t bs = q / nAg;
    t.a = A.a0 + (q % nAg) * R;
    if (bs >= A.nbs) {
        t.a = A.a0;
        t.b = A.b0;
        t.ti = 0;
        t.valid = false;
        return false;
    }
    t.b = A.b0 - A.ush + 64 * bs + (threadIdx.x & 63);
    t.valid = t.b >= A.b0 && t.b < A.b1;
    t.ti = (t.a - A.a0) * A.nbs + bs;
    return true;
}
template <int R = 1>
__device__ __forceinline__ bool tile_binary(const ChunkArgs &A, Tile &t) {
    return tile_binary<R>(A, t, split_of(A));
}

__device__ __forceinline__ bool tile_unary(const ChunkArgs &A, Tile &t, const Split &sp) {
    t.ti = (int64_t)blockIdx.x * sp.per + sp.tw;
    if (t.ti >= A.n_tiles) {
        t.ti = 0;
        t.a = A.a0;
        t.b = 0;
        t.valid = false;
        return false;
    }
    t.a = A.a0 - A.ush + 64 * t.ti + (threadIdx.x & 63);
    t.b = 0;
    t.valid = t.a >= A.a0 && t.a < A.a1;
    return true;
}
__device__ __forceinline__ bool tile_unary(const ChunkArgs &A, Tile &t) { return tile_unary(A, t, split_of(A)); }

// The S waves of a task meet before its tail: waves 1 .. S-1 leave their
// sums (fingerprints mod 2^64, mismatch counts) in LDS, wave 0 adds them.
// Every wave of the workgroup reaches the barrier (S is uniform); waves
// without a task take part with zeros.  LDS: (4 - 4 / S) * N * 512 bytes.
template <int N>
__device__ __forceinline__ void split_add(uint64_t (&v)[N], const Split &sp) {
    extern __shared__ uint64_t split_lds[];
    const int lane = (int)(threadIdx.x & 63);
    uint64_t *b = split_lds + (size_t)sp.tw * (sp.S - 1) * N * 64;
    if (sp.r > 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) b[((size_t)(sp.r - 1) * N + i) * 64 + lane] = v[i];
    }
    __syncthreads();
    if (sp.r == 0) {
        for (int j = 1; j < sp.S; ++j) {
#pragma unroll
            for (int i = 0; i < N; ++i) v[i] += b[((size_t)(j - 1) * N + i) * 64 + lane];
        }
    }
}
__device__ __forceinline__ void acc_put(uint64_t *v, const Acc &a) {
    v[0] = a.f1;
    v[1] = a.f2;
    v[2] = a.mism;
}
__device__ __forceinline__ void acc_get(const uint64_t *v, Acc &a) {
    a.f1 = v[0];
    a.f2 = v[1];
    a.mism = (uint32_t)v[2];
}

// ---------------------------------------------------------------------------
// gen
// ---------------------------------------------------------------------------
// Successors X c, F c, G c of a candidate (bits 1-3 of A.succ_mask) tested
// at position 0 of every trace, word by word: their mismatch counts.  (~c,
// bit 0, comes from the level's check pass: the complements it accepts.)
struct SuccAcc {
    uint32_t m[3] = {0, 0, 0};
};

template <bool SUCC>
__device__ __forceinline__ void succ_fold(SuccAcc &s, uint64_t c, const WordAux *x, int mask) {
    if constexpr (SUCC) {
        const uint64_t tgt = x->tgt, chk = x->chk;
        if (mask & 2) s.m[0] += (uint32_t)__popcll((pk_unary<LTL_NEXT>(c, x) ^ tgt) & chk);
        if (mask & 4) s.m[1] += (uint32_t)__popcll((pk_unary<LTL_FIN>(c, x) ^ tgt) & chk);
        if (mask & 8) s.m[2] += (uint32_t)__popcll((pk_unary<LTL_GLOB>(c, x) ^ tgt) & chk);
    }
}

// Accepted successors X c, F c, G c of a candidate as bits 1-3.
__device__ __forceinline__ uint8_t succ_flags(const ChunkArgs &A, const SuccAcc &s) {
    const uint32_t k = (uint32_t)A.hamming_k;
    uint8_t f = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
        if ((A.succ_mask & (2 << j)) && s.m[j] <= k) f |= (uint8_t)(2 << j);
    return f;
}

// After the last word: test the candidate(s) of this lane (CHK: the level
// was not tested by a check pass beforehand), then intern fingerprint 1 with
// the candidate id as pending value (unless the chunk is check-only), and
// (SUCC) record the candidates' accepted successors.
// Relaxed keys (RX): outside check-only chunks a candidate counts only if
// it survives dedup (the reference tests a candidate only when its key is
// new, code/ltli6463.cu:231-243), so its outcome goes to t_acc and dedup
// folds it in; in a check-only chunk every candidate is tested
// (code/ltli6463.cu:198-208).
template <int OP, bool CHK, bool SUCC, bool RX = false>
__device__ __forceinline__ void gen2_tail(const ChunkArgs &A, const Tile &t, const Acc &c0, const Acc &c1,
                                          const SuccAcc &s0, const SuccAcc &s1, bool otf) {
    const uint64_t pair = (uint64_t)t.a * (uint64_t)A.nR + (uint64_t)t.b;
    const uint64_t id0 = A.id_base + (OP == LTL_UNTIL ? 2 * pair : pair);
    const bool acc0 = CHK && c0.mism <= (uint32_t)A.hamming_k;
    const bool acc1 = CHK && OP == LTL_UNTIL && c1.mism <= (uint32_t)A.hamming_k;
    if (!RX || otf) {
        if (acc0) atomicMin(&A.st->found, (unsigned long long)id0);
        if (acc1) atomicMin(&A.st->found, (unsigned long long)(id0 + 1));
    }
    if (otf) return;
    const size_t ix = (size_t)t.ti * 64 + (threadIdx.x & 63);
    if constexpr (RX) {
        if constexpr (OP == LTL_UNTIL) {
            A.t_acc[2 * ix] = (uint8_t)acc0;
            A.t_acc[2 * ix + 1] = (uint8_t)acc1;
        } else {
            A.t_acc[ix] = (uint8_t)acc0;
        }
    }
    if constexpr (OP == LTL_UNTIL) {
        A.t_slot[2 * ix] = tab_insert<true>(A.tab, A.tmask, tab_key(c0.f1, A.epoch), tab_val(kPending | id0, A.epoch), A.epoch, &A.st->probe_fail);
        A.t_fp2[2 * ix] = c0.f2;
        A.t_slot[2 * ix + 1] = tab_insert<true>(A.tab, A.tmask, tab_key(c1.f1, A.epoch), tab_val(kPending | (id0 + 1), A.epoch), A.epoch,
                                          &A.st->probe_fail);
        A.t_fp2[2 * ix + 1] = c1.f2;
        if constexpr (SUCC) {
            A.t_succ[2 * ix] = succ_flags(A, s0);
            A.t_succ[2 * ix + 1] = succ_flags(A, s1);
        }
    } else {
        A.t_slot[ix] = tab_insert<true>(A.tab, A.tmask, tab_key(c0.f1, A.epoch), tab_val(kPending | id0, A.epoch), A.epoch, &A.st->probe_fail);
        A.t_fp2[ix] = c0.f2;
        if constexpr (SUCC) A.t_succ[ix] = succ_flags(A, s0);
    }
}

template <bool CHK, bool SUCC, bool RX = false>
__device__ __forceinline__ void gen1_tail(const ChunkArgs &A, const Tile &t, const Acc &c0, const SuccAcc &s0,
                                          bool otf) {
    const uint64_t id0 = A.id_base + (uint64_t)t.a;
    const bool acc0 = CHK && c0.mism <= (uint32_t)A.hamming_k;
    if ((!RX || otf) && acc0) atomicMin(&A.st->found, (unsigned long long)id0);
    if (otf) return;
    const size_t ix = (size_t)t.ti * 64 + (threadIdx.x & 63);
    if constexpr (RX) A.t_acc[ix] = (uint8_t)acc0;
    A.t_slot[ix] = tab_insert<true>(A.tab, A.tmask, tab_key(c0.f1, A.epoch), tab_val(kPending | id0, A.epoch), A.epoch, &A.st->probe_fail);
    A.t_fp2[ix] = c0.f2;
    if constexpr (SUCC) A.t_succ[ix] = succ_flags(A, s0);
}

// R left rows per wave: each right word is loaded once for R candidates
// (the left words are wave-uniform scalar loads), so the right operand's
// reads from L2 drop R-fold; the fingerprint arithmetic per candidate stays.
template <int OP, bool RX, bool CHK, bool SUCC, int R>
__global__ __launch_bounds__(256) void k_gen2(ChunkArgs A) {
    // (split: every wave reaches split_add; the stop test moves to the tail)
    const Split sp = split_of(A);
    const bool split = !RX && !SUCC && sp.S > 1;
    if (!split && run_stopped(A)) return;
    Tile t;
    const bool live = tile_binary<R>(A, t, sp);
    if (!split && !live) return;
    const bool otf = chunk_otf(A);
    const int nrow = (int)(A.a1 - t.a < R ? A.a1 - t.a : R);  // wave-uniform
    const int64_t sr = A.rs + (t.valid ? t.b : A.b0);
    const uint64_t *pr = A.store + set_base(sr, A.nw);
    const uint64_t *pl[R];
#pragma unroll
    for (int r = 0; r < R; ++r) pl[r] = A.store + set_base(A.ls + t.a + (r < nrow ? r : 0), A.nw);
    Acc c0[R], c1[R];
    KeyAcc k0[R], k1[R];
    SuccAcc s0[R], s1[R];
    const int sm = A.succ_mask;
    auto step = [&](int w, const uint64_t (&l)[R], uint64_t rv) {
        const WordAux *x = A.aux + w;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if constexpr (OP == LTL_UNTIL) {
                const uint64_t u0 = pk_until(l[r], rv, x->inner), u1 = pk_until(rv, l[r], x->inner);
                fold<CHK>(c0[r], u0, x);
                fold<CHK>(c1[r], u1, x);
                succ_fold<SUCC>(s0[r], u0, x, sm);
                succ_fold<SUCC>(s1[r], u1, x, sm);
                if constexpr (RX) {
                    key_fold_pk(k0[r], u0, w, x, A);
                    key_fold_pk(k1[r], u1, w, x, A);
                }
            } else {
                const uint64_t v = OP == LTL_AND ? (l[r] & rv) : (l[r] | rv);
                fold<CHK>(c0[r], v, x);
                succ_fold<SUCC>(s0[r], v, x, sm);
                if constexpr (RX) key_fold_pk(k0[r], v, w, x, A);
            }
        }
    };
    // Batches of UB words: all right-operand loads of a batch are issued
    // before use, so that each lane keeps UB loads in flight.  Measured on
    // 500+500 traces of 10-32 positions (gen_and_or per call): R = 1, UB = 16: 6.4 ms; R = 2, UB = 8:
    // 4.9 ms; R = 4, UB = 4 or 8: 5.5-5.6 ms.  U: R = 1, UB = 4: 5.45 ms;
    // UB = 8 or R = 2: 5.7-6.0 ms (two carry chains and four fingerprints
    // per pair keep it VALU-bound).
    constexpr int UB = R == 1 ? (OP == LTL_UNTIL ? 4 : 16) : 8;
    const int we = live ? sp.w1 : sp.w0;
    int w = sp.w0;
    for (; w + UB <= we; w += UB) {
        uint64_t rv[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) rv[u] = pr[(size_t)(w + u) * kTileWS];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            uint64_t l[R];
#pragma unroll
            for (int r = 0; r < R; ++r) l[r] = pl[r][(size_t)(w + u) * kTileWS];
            step(w + u, l, rv[u]);
        }
    }
    for (; w < we; ++w) {
        uint64_t l[R];
#pragma unroll
        for (int r = 0; r < R; ++r) l[r] = pl[r][(size_t)w * kTileWS];
        step(w, l, pr[(size_t)w * kTileWS]);
    }
    if constexpr (!RX && !SUCC) {
        if (split) {
            constexpr int NO = OP == LTL_UNTIL ? 2 : 1;
            uint64_t v[3 * NO * R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                acc_put(v + 3 * NO * r, c0[r]);
                if constexpr (NO == 2) acc_put(v + 3 * NO * r + 3, c1[r]);
            }
            split_add(v, sp);
            if (sp.r != 0 || !live || run_stopped(A)) return;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                acc_get(v + 3 * NO * r, c0[r]);
                if constexpr (NO == 2) acc_get(v + 3 * NO * r + 3, c1[r]);
            }
        }
    }
    if (!t.valid) return;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (r >= nrow) break;
        if constexpr (RX) {
            key_finish(c0[r], k0[r], A.kseed);
            if constexpr (OP == LTL_UNTIL) key_finish(c1[r], k1[r], A.kseed);
        }
        Tile tr = t;
        tr.a = t.a + r;
        tr.ti = t.ti + (int64_t)r * A.nbs;
        gen2_tail<OP, CHK, SUCC, RX>(A, tr, c0[r], c1[r], s0[r], s1[r], otf);
    }
}

template <int OP, bool RX, bool CHK, bool SUCC>
__global__ __launch_bounds__(256) void k_gen1(ChunkArgs A) {
    const Split sp = split_of(A);
    const bool split = !RX && !SUCC && sp.S > 1;
    if (!split && run_stopped(A)) return;
    Tile t;
    const bool live = tile_unary(A, t, sp);
    if (!split && !live) return;
    const bool otf = chunk_otf(A);
    const int64_t s = A.ls + (t.valid ? t.a : A.a0);
    const uint64_t *p = A.store + set_base(s, A.nw);
    Acc c0;
    KeyAcc k0;
    SuccAcc s0;
    const int sm = A.succ_mask;
    auto step = [&](int w, uint64_t v) {
        const WordAux *x = A.aux + w;
        const uint64_t y = pk_unary<OP>(v, x);
        fold<CHK>(c0, y, x);
        succ_fold<SUCC>(s0, y, x, sm);
        if constexpr (RX) key_fold_pk(k0, y, w, x, A);
    };
    constexpr int UB = 8;
    const int we = live ? sp.w1 : sp.w0;
    int w = sp.w0;
    for (; w + UB <= we; w += UB) {
        uint64_t v[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) v[u] = p[(size_t)(w + u) * kTileWS];
#pragma unroll
        for (int u = 0; u < UB; ++u) step(w + u, v[u]);
    }
    for (; w < we; ++w) step(w, p[(size_t)w * kTileWS]);
    if constexpr (!RX && !SUCC) {
        if (split) {
            uint64_t v[3];
            acc_put(v, c0);
            split_add(v, sp);
            if (sp.r != 0 || !live || run_stopped(A)) return;
            acc_get(v, c0);
        }
    }
    if (!t.valid) return;
    if constexpr (RX) key_finish(c0, k0, A.kseed);
    gen1_tail<CHK, SUCC, RX>(A, t, c0, s0, otf);
}

// X, F and G of one operand block in one pass over the operand words (the
// three blocks read the same level when their costs are equal): three
// candidates per operand, three fingerprint pairs, three inserts.  The table
// keeps the smallest id, and X ids < F ids < G ids, so the dedup outcome is
// the one of three consecutive passes.  NEG: the negation block of the same
// operand level too (four candidates per operand; the negation block's ids
// start at A.id_base, the X block's at id_xfg, after the & and | blocks that
// come between them in canonical order).  Temporaries of candidate j at
// j * slots.  Packed layout, exact keys.
template <bool CHK, bool SUCC, bool NEG>
__global__ __launch_bounds__(256) void k_gen_xfg(ChunkArgs A, int64_t n, int64_t slots, uint64_t id_xfg) {
    const Split sp = split_of(A);
    const bool split = !SUCC && sp.S > 1;
    if (!split && run_stopped(A)) return;
    Tile t;
    const bool live = tile_unary(A, t, sp);
    if (!split && !live) return;
    const bool otf = chunk_otf(A);
    const int64_t s = A.ls + (t.valid ? t.a : A.a0);
    const uint64_t *p = A.store + set_base(s, A.nw);
    constexpr int NC = NEG ? 4 : 3;
    Acc c[NC];
    SuccAcc sc[NC];
    const int sm = A.succ_mask;
    auto step = [&](int w, uint64_t v) {
        const WordAux *x = A.aux + w;
        uint64_t y[NC];
        if constexpr (NEG) y[0] = pk_unary<LTL_NOT>(v, x);
        y[NC - 3] = pk_unary<LTL_NEXT>(v, x);
        y[NC - 2] = pk_unary<LTL_FIN>(v, x);
        y[NC - 1] = pk_unary<LTL_GLOB>(v, x);
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            fold<CHK>(c[j], y[j], x);
            succ_fold<SUCC>(sc[j], y[j], x, sm);
        }
    };
    constexpr int UB = 8;
    const int we = live ? sp.w1 : sp.w0;
    int w = sp.w0;
    for (; w + UB <= we; w += UB) {
        uint64_t v[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) v[u] = p[(size_t)(w + u) * kTileWS];
#pragma unroll
        for (int u = 0; u < UB; ++u) step(w + u, v[u]);
    }
    for (; w < we; ++w) step(w, p[(size_t)w * kTileWS]);
    if constexpr (!SUCC) {
        if (split) {
            uint64_t v[3 * NC];
#pragma unroll
            for (int j = 0; j < NC; ++j) acc_put(v + 3 * j, c[j]);
            split_add(v, sp);
            if (sp.r != 0 || !live || run_stopped(A)) return;
#pragma unroll
            for (int j = 0; j < NC; ++j) acc_get(v + 3 * j, c[j]);
        }
    }
    if (!t.valid) return;
    const size_t ix = (size_t)t.ti * 64 + (threadIdx.x & 63);
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        const uint64_t id = NEG ? (j == 0 ? A.id_base : id_xfg + (uint64_t)(j - 1) * (uint64_t)n) + (uint64_t)t.a
                                : A.id_base + (uint64_t)j * (uint64_t)n + (uint64_t)t.a;
        if (CHK && c[j].mism <= (uint32_t)A.hamming_k) atomicMin(&A.st->found, (unsigned long long)id);
        if (otf) continue;
        const size_t k = (size_t)j * (size_t)slots + ix;
        A.t_slot[k] = tab_insert<true>(A.tab, A.tmask, tab_key(c[j].f1, A.epoch), tab_val(kPending | id, A.epoch), A.epoch,
                                &A.st->probe_fail);
        A.t_fp2[k] = c[j].f2;
        if constexpr (SUCC) A.t_succ[k] = succ_flags(A, sc[j]);
    }
}

// The binary blocks of one split in one pass (tested levels, packed
// layout, exact keys): per pair (l, r) of rows a of the left level and
// columns b of the right level, l & r (the & block, ids from A.id_base,
// temporaries at A's), l | r (the | block, F.id_or, temporaries F.off_or
// candidate slots further) and, with U, l U r and r U l (the U block,
// F.id_u, F.off_u).  The operand words are read once for the three blocks;
// the table keeps the smallest id, so inserting candidates of later blocks
// early changes no outcome, provided each block is decided in canonical
// order (with every earlier block committed).  R left rows per wave (& |:
// 2; with U: 1).
template <bool U, bool SUCC, int R>
__device__ __forceinline__ void gen2f_body(const ChunkArgs &A, const FusedBin &F) {
    constexpr int NC = U ? 4 : 2;
    const Split sp = split_of(A);
    const bool split = !SUCC && sp.S > 1;
    if (!split && run_stopped(A)) return;
    Tile t;
    const bool live = tile_binary<R>(A, t, sp);
    if (!split && !live) return;
    const bool otf = chunk_otf(A);
    const int nrow = (int)(A.a1 - t.a < R ? A.a1 - t.a : R);  // wave-uniform
    // The right operand is the stripe's tile (lane l: set 64 bs + l - ush of
    // the level, the first lane at a tile start), read with buffer loads
    // bounded to that tile: word w of lane l at byte 512 w + 8 l.  Lanes past
    // the block's columns read their own sets of the tile (unused).  Loads of
    // the next batch of words are issued before this batch is folded.
    const int64_t bs = t.ti - (t.a - A.a0) * A.nbs;
    const int64_t s_first = A.rs + A.b0 - A.ush + 64 * bs;
    const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(A.store + set_base(s_first, A.nw)), (short)0, A.nw * 512, kBufferWord3);
    const int lane8 = (int)(threadIdx.x & 63) * 8;
    auto rload = [&](int w) -> uint64_t {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(rr, lane8, w * 512, 0);
        return (uint64_t)v[0] | ((uint64_t)v[1] << 32);
    };
    const uint64_t *pl[R];
#pragma unroll
    for (int r = 0; r < R; ++r) pl[r] = A.store + set_base(A.ls + t.a + (r < nrow ? r : 0), A.nw);
    Acc c[R][NC];
    SuccAcc sc[R][NC];
    const int sm = A.succ_mask;
    // per word: the left rows' words and the word's keys (wave-uniform,
    // scalar loads), loaded for a whole batch of words before any is used so
    // that the batch waits for its scalar loads once
    struct Word {
        uint64_t l[R], k1, k2, inner;
    };
    auto load = [&](int w, Word &d) {
#pragma unroll
        for (int r = 0; r < R; ++r) d.l[r] = pl[r][(size_t)w * kTileWS];
        d.k1 = A.aux[w].k1;
        d.k2 = A.aux[w].k2;
        if constexpr (U) d.inner = A.aux[w].inner;
    };
    auto step = [&](int w, const Word &d, uint64_t rv) {
        const WordAux *x = A.aux + w;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            uint64_t y[NC];
            y[0] = d.l[r] & rv;
            y[1] = d.l[r] | rv;
            if constexpr (U) {
                y[2] = pk_until(d.l[r], rv, d.inner);
                y[3] = pk_until(rv, d.l[r], d.inner);
            }
#pragma unroll
            for (int j = 0; j < NC; ++j) {
                ml_fold(c[r][j].f1, y[j], d.k1);
                ml_fold(c[r][j].f2, y[j], d.k2);
                succ_fold<SUCC>(sc[r][j], y[j], x, sm);
            }
        }
    };
    constexpr int UB = 4;
    const int we = live ? sp.w1 : sp.w0;
    int w = sp.w0;
    // Two batches per iteration in alternate registers (no copies, which
    // would wait for the loads in flight); past the last batch the last word
    // is loaded again, unused.
    auto batch = [&](int w0, const uint64_t (&rv)[UB]) {
        Word d[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) load(w0 + u, d[u]);
#pragma unroll
        for (int u = 0; u < UB; ++u) step(w0 + u, d[u], rv[u]);
    };
    uint64_t ra[UB], rb[UB];
    if (w + UB <= we) {
#pragma unroll
        for (int u = 0; u < UB; ++u) ra[u] = rload(w + u);
    }
    for (; w + 2 * UB <= we; w += 2 * UB) {
#pragma unroll
        for (int u = 0; u < UB; ++u) rb[u] = rload(w + UB + u);
        batch(w, ra);
#pragma unroll
        for (int u = 0; u < UB; ++u) ra[u] = rload(w + 2 * UB + u < we ? w + 2 * UB + u : we - 1);
        batch(w + UB, rb);
    }
    if (w + UB <= we) {
        batch(w, ra);
        w += UB;
    }
    for (; w < we; ++w) {
        Word d;
        load(w, d);
        step(w, d, rload(w));
    }
    if constexpr (!SUCC) {
        if (split) {
            uint64_t v[3 * NC * R];
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int j = 0; j < NC; ++j) acc_put(v + 3 * (NC * r + j), c[r][j]);
            split_add(v, sp);
            if (sp.r != 0 || !live || run_stopped(A)) return;
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int j = 0; j < NC; ++j) acc_get(v + 3 * (NC * r + j), c[r][j]);
        }
    }
    if (!t.valid || otf) return;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (r >= nrow) break;
        const uint64_t pair = (uint64_t)(t.a + r) * (uint64_t)A.nR + (uint64_t)t.b;
        const size_t ix = (size_t)(t.ti + (int64_t)r * A.nbs) * 64 + (threadIdx.x & 63);
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            const uint64_t id = j == 0 ? A.id_base + pair : j == 1 ? F.id_or + pair : F.id_u + 2 * pair + (uint64_t)(j - 2);
            const size_t k = j == 0 ? ix : j == 1 ? (size_t)F.off_or + ix : (size_t)F.off_u + 2 * ix + (size_t)(j - 2);
            A.t_slot[k] = tab_insert<true>(A.tab, A.tmask, tab_key(c[r][j].f1, A.epoch), tab_val(kPending | id, A.epoch),
                                     A.epoch, &A.st->probe_fail);
            A.t_fp2[k] = c[r][j].f2;
            if constexpr (SUCC) A.t_succ[k] = succ_flags(A, sc[r][j]);
        }
    }
}

// Without successor tests the pass fits 64 registers: 8 waves per SIMD
// (6 at the compiler's own choice of 80) hide more of the operand loads'
// latency (7-10% faster in a standalone microbenchmark of the pass); the
// successor tests' counters need more registers (k_gen2fs: forced to 64
// they spilled to scratch).
template <bool U, int R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_gen2f(ChunkArgs A, FusedBin F) {
    gen2f_body<U, false, R>(A, F);
}
template <bool U, int R>
__global__ __launch_bounds__(256) void k_gen2fs(ChunkArgs A, FusedBin F) {
    gen2f_body<U, true, R>(A, F);
}

// Multi-word traces (K words per trace): the fused binary pass of a tested
// level, as k_gen2f -- per pair (l, r) l & r, l | r and, with U, l U r and
// r U l (the until over a trace's K words from the top word down), each
// fingerprinted and inserted into its block's region at its canonical id.
// One left row per wave, one wave per task; no successor tests (multi-word
// levels are not deferred with them).
template <bool U, int K>
__global__ __launch_bounds__(256) void k_gen2fm(ChunkArgs A, FusedBin F) {
    if (run_stopped(A)) return;
    Tile t;
    if (!tile_binary(A, t)) return;
    const bool otf = chunk_otf(A);
    const uint64_t *pl = A.store + set_base(A.ls + t.a, A.nw);
    // The right operand through buffer loads bounded to the stripe's tile,
    // one trace ahead in alternate registers (as in k_gen2f).
    const int64_t bs = t.ti - (t.a - A.a0) * A.nbs;
    const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(A.store + set_base(A.rs + A.b0 - A.ush + 64 * bs, A.nw)), (short)0, A.nw * 512, kBufferWord3);
    const int lane8 = (int)(threadIdx.x & 63) * 8;
    auto rload = [&](int i, uint64_t (&r)[K]) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b64(rr, lane8, (i * K + k) * 512, 0);
            r[k] = (uint64_t)v[0] | ((uint64_t)v[1] << 32);
        }
    };
    constexpr int NC = U ? 4 : 2;
    Acc c[NC];
    auto trace = [&](int i, const uint64_t (&r)[K]) {
        uint64_t l[K], u0[K], u1[K];
#pragma unroll
        for (int k = 0; k < K; ++k) l[k] = pl[(size_t)(i * K + k) * kTileWS];
        if constexpr (U) {
            trace_until<K>(l, r, u0);
            trace_until<K>(r, l, u1);
        }
        const WordAux *aux = A.aux + i * K;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            fold<false>(c[0], l[k] & r[k], aux + k);
            fold<false>(c[1], l[k] | r[k], aux + k);
            if constexpr (U) {
                fold<false>(c[2], u0[k], aux + k);
                fold<false>(c[3], u1[k], aux + k);
            }
        }
    };
    const int T = A.nw / K;
    uint64_t ra[K], rb[K];
    rload(0, ra);
    int i = 0;
    for (; i + 2 <= T; i += 2) {
        rload(i + 1, rb);
        trace(i, ra);
        rload(i + 2 < T ? i + 2 : T - 1, ra);
        trace(i + 1, rb);
    }
    if (i < T) trace(i, ra);
    if (!t.valid || otf) return;
    const uint64_t pair = (uint64_t)t.a * (uint64_t)A.nR + (uint64_t)t.b;
    const size_t ix = (size_t)t.ti * 64 + (threadIdx.x & 63);
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        const uint64_t id = j == 0 ? A.id_base + pair : j == 1 ? F.id_or + pair : F.id_u + 2 * pair + (uint64_t)(j - 2);
        const size_t k = j == 0 ? ix : j == 1 ? (size_t)F.off_or + ix : (size_t)F.off_u + 2 * ix + (size_t)(j - 2);
        A.t_slot[k] = tab_insert<true>(A.tab, A.tmask, tab_key(c[j].f1, A.epoch), tab_val(kPending | id, A.epoch), A.epoch,
                                 &A.st->probe_fail);
        A.t_fp2[k] = c[j].f2;
    }
}

// Multi-word traces (B = 0): K words per trace, combined from the top word
// down.  & and | are word-wise and use k_gen2<OP, false>.
template <int OP, int K>
__global__ __launch_bounds__(256) void k_gen1m(ChunkArgs A) {
    if (run_stopped(A)) return;
    Tile t;
    if (!tile_unary(A, t)) return;
    const bool otf = chunk_otf(A);
    const int64_t s = A.ls + (t.valid ? t.a : A.a0);
    const uint64_t *p = A.store + set_base(s, A.nw);
    const int T = A.nw / K;
    Acc c0;
#pragma unroll 2
    for (int i = 0; i < T; ++i) {
        uint64_t x[K], y[K];
#pragma unroll
        for (int k = 0; k < K; ++k) x[k] = p[(size_t)(i * K + k) * kTileWS];
        const WordAux *aux = A.aux + i * K;
        trace_unary<OP, K>(x, aux, y);
#pragma unroll
        for (int k = 0; k < K; ++k) fold(c0, y[k], aux + k);
    }
    if (!t.valid) return;
    gen1_tail<true, false>(A, t, c0, SuccAcc{}, otf);
}

// k_gen_xfg for K words per trace (no successor tests): X, F, G (NEG: and
// ~) of each operand, trace by trace.
template <int K, bool CHK, bool NEG>
__global__ __launch_bounds__(256) void k_gen_xfgm(ChunkArgs A, int64_t n, int64_t slots, uint64_t id_xfg) {
    if (run_stopped(A)) return;
    Tile t;
    if (!tile_unary(A, t)) return;
    const bool otf = chunk_otf(A);
    const int64_t s = A.ls + (t.valid ? t.a : A.a0);
    const uint64_t *p = A.store + set_base(s, A.nw);
    const int T = A.nw / K;
    constexpr int NC = NEG ? 4 : 3;
    Acc c[NC];
    for (int i = 0; i < T; ++i) {
        uint64_t x[K], y[K];
#pragma unroll
        for (int k = 0; k < K; ++k) x[k] = p[(size_t)(i * K + k) * kTileWS];
        const WordAux *aux = A.aux + i * K;
        if constexpr (NEG) {
            trace_unary<LTL_NOT, K>(x, aux, y);
#pragma unroll
            for (int k = 0; k < K; ++k) fold<CHK>(c[0], y[k], aux + k);
        }
        trace_unary<LTL_NEXT, K>(x, aux, y);
#pragma unroll
        for (int k = 0; k < K; ++k) fold<CHK>(c[NC - 3], y[k], aux + k);
        trace_unary<LTL_FIN, K>(x, aux, y);
#pragma unroll
        for (int k = 0; k < K; ++k) fold<CHK>(c[NC - 2], y[k], aux + k);
        trace_unary<LTL_GLOB, K>(x, aux, y);
#pragma unroll
        for (int k = 0; k < K; ++k) fold<CHK>(c[NC - 1], y[k], aux + k);
    }
    if (!t.valid) return;
    const size_t ix = (size_t)t.ti * 64 + (threadIdx.x & 63);
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        const uint64_t id = NEG ? (j == 0 ? A.id_base : id_xfg + (uint64_t)(j - 1) * (uint64_t)n) + (uint64_t)t.a
                                : A.id_base + (uint64_t)j * (uint64_t)n + (uint64_t)t.a;
        if (CHK && c[j].mism <= (uint32_t)A.hamming_k) atomicMin(&A.st->found, (unsigned long long)id);
        if (otf) continue;
        const size_t k = (size_t)j * (size_t)slots + ix;
        A.t_slot[k] = tab_insert<true>(A.tab, A.tmask, tab_key(c[j].f1, A.epoch), tab_val(kPending | id, A.epoch), A.epoch,
                                &A.st->probe_fail);
        A.t_fp2[k] = c[j].f2;
    }
}

template <int K>
__global__ __launch_bounds__(256) void k_gen2m(ChunkArgs A) {  // U
    if (run_stopped(A)) return;
    Tile t;
    if (!tile_binary(A, t)) return;
    const bool otf = chunk_otf(A);
    const int64_t sl = A.ls + t.a, sr = A.rs + (t.valid ? t.b : A.b0);
    const uint64_t *pl = A.store + set_base(sl, A.nw);
    const uint64_t *pr = A.store + set_base(sr, A.nw);
    const int T = A.nw / K;
    Acc c0, c1;
    for (int i = 0; i < T; ++i) {
        uint64_t l[K], r[K], u0[K], u1[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            l[k] = pl[(size_t)(i * K + k) * kTileWS];
            r[k] = pr[(size_t)(i * K + k) * kTileWS];
        }
        trace_until<K>(l, r, u0);
        trace_until<K>(r, l, u1);
        const WordAux *aux = A.aux + i * K;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            fold(c0, u0[k], aux + k);
            fold(c1, u1[k], aux + k);
        }
    }
    if (!t.valid) return;
    gen2_tail<LTL_UNTIL, true, false>(A, t, c0, c1, SuccAcc{}, SuccAcc{}, otf);
}

// Atoms: interned with committed values (store index = atom index), so that
// the first of equal atoms wins (code/ltli6463.cu:144-173).
// One wave per atom, lanes over the words; the hash sums (mod 2^64) are
// reduced across the wave.
__global__ __launch_bounds__(64) void k_atoms(ChunkArgs A, int n_atoms) {
    const int a = (int)blockIdx.x, l = (int)threadIdx.x;
    const uint64_t *p = A.store + set_base(a, A.nw);
    Acc c0;
    KeyAcc k0;
    for (int w = l; w < A.nw; w += 64) {
        const uint64_t v = p[(size_t)w * kTileWS];
        fold(c0, v, A.aux + w);
        if (A.kmode) key_fold_pk(k0, v, w, A.aux + w, A);  // relaxed keys: packed layout only
    }
    for (int o = 1; o < 64; o <<= 1) {
        c0.f1 += __shfl_xor(c0.f1, o);
        c0.f2 += __shfl_xor(c0.f2, o);
        k0.h ^= __shfl_xor(k0.h, o);  // contributions of different traces do not overlap (modes 1, 2)
        k0.l ^= __shfl_xor(k0.l, o);
    }
    if (A.kmode) key_finish(c0, k0, A.kseed);
    if (l != 0) return;
    const uint64_t key = tab_key(c0.f1, A.epoch);
    tab_insert(A.tab, A.tmask, key, tab_val((uint64_t)a, A.epoch), A.epoch, &A.st->probe_fail);
    A.m_key[a] = key;
    A.m_fp2[a] = c0.f2;
    A.m_op[a] = LTL_ATOM;
    A.m_left[a] = -1;
    A.m_right[a] = -1;
}

// dst[r * ws + c] = src[r * ncols + c]: rows of a contiguous staging copy
// into the word-major store.
__global__ void k_scatter(uint64_t *dst, int64_t ws, const uint64_t *src, int ncols, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[(i / ncols) * ws + i % ncols] = src[i];
}

// ---------------------------------------------------------------------------
// decide
// ---------------------------------------------------------------------------
// Temp index of a pending candidate id of this chunk (its tile and lane).
__device__ __forceinline__ int64_t temp_index_of(const ChunkArgs &A, int op, uint64_t id) {
    const uint64_t rel = id - A.id_base;
    if (op == LTL_AND || op == LTL_OR || op == LTL_UNTIL) {
        const uint64_t pair = op == LTL_UNTIL ? rel >> 1 : rel;
        const int64_t a = (int64_t)(pair / (uint64_t)A.nR), b = (int64_t)(pair % (uint64_t)A.nR);
        if (a < A.a0 || a >= A.a1 || b < A.b0 || b >= A.b1) return -1;
        const int64_t off = b - A.b0 + A.ush;
        const int64_t ti = (a - A.a0) * A.nbs + off / 64;
        const int64_t ix = ti * 64 + off % 64;
        return op == LTL_UNTIL ? 2 * ix + (int64_t)(rel & 1) : ix;
    }
    const int64_t a = (int64_t)rel;
    if (a < A.a0 || a >= A.a1) return -1;
    return a - A.a0 + A.ush;
}

// Survivors of tile ti (64 candidates, two per pair for U): a candidate
// survives when its table slot still holds its own pending id (it is the
// smallest id of its fingerprint); a slot taken by a stored set or by another
// candidate with a different second fingerprint is a collision.  One wave per
// tile; m[d] is the survivor mask of the tile's d-th candidate of each pair.
template <int OP>
__device__ __forceinline__ uint32_t decide_tile(const ChunkArgs &A, int64_t ti, int lane, uint64_t m[2]) {
    constexpr bool bin = OP == LTL_AND || OP == LTL_OR || OP == LTL_UNTIL;
    constexpr int ND = OP == LTL_UNTIL ? 2 : 1;
    bool valid;
    uint64_t id0;
    if (bin) {
        const int64_t a = A.a0 + ti / A.nbs, b = A.b0 - A.ush + (ti % A.nbs) * 64 + lane;
        valid = b >= A.b0 && b < A.b1;
        const uint64_t pair = (uint64_t)a * (uint64_t)A.nR + (uint64_t)b;
        id0 = A.id_base + ND * pair;
    } else {
        const int64_t a = A.a0 - A.ush + ti * 64 + lane;
        valid = a >= A.a0 && a < A.a1;
        id0 = A.id_base + (uint64_t)a;
    }
    const size_t ix = (size_t)ti * 64 + lane;
    uint32_t cnt = 0;
    m[1] = 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        bool surv = false;
        if (valid) {
            const size_t k = ND * ix + d;
            const uint32_t slot = A.t_slot[k];
            const uint64_t f2 = A.t_fp2[k];
            const uint64_t id = id0 + d;
            // (found stored at the insert: a duplicate, no table read)
            const uint64_t v = (slot & kSlotStored) ? tab_val(slot & ~kSlotStored, A.epoch)
                                                    : A.tab[2 * (size_t)slot + 1];
            if (v == tab_val(kPending | id, A.epoch)) {
                surv = true;
                // relaxed keys: the survivor's own test (set in generation)
                if (A.t_acc && A.t_acc[k]) atomicMin(&A.st->rx_found, (unsigned long long)id);
            } else if (tab_committed(v, A.epoch)) {
                if (A.m_fp2[tab_payload(v)] != f2) A.st->collision = 1;
            } else {
                const int64_t o = temp_index_of(A, OP, tab_payload(v));
                if (o < 0 || A.t_fp2[o] != f2) A.st->collision = 1;
            }
        }
        m[d] = __ballot(surv);
        cnt += (uint32_t)__popcll(m[d]);
    }
    return cnt;
}

template <int OP>
__global__ __launch_bounds__(256) void k_decide(ChunkArgs A) {
    if (run_stopped(A) || chunk_otf(A)) return;
    const int64_t ti = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (ti >= A.n_tiles) return;
    const int lane = threadIdx.x & 63;
    constexpr int ND = OP == LTL_UNTIL ? 2 : 1;
    uint64_t m[2];
    const uint32_t cnt = decide_tile<OP>(A, ti, lane, m);
    if (lane == 0) {
        for (int d = 0; d < ND; ++d) A.t_mask[ND * ti + d] = m[d];
        A.t_cnt[ti] = cnt;
    }
}

__global__ void k_finish(ChunkArgs A) {
    if (run_stopped(A) || chunk_otf(A)) return;
    const int64_t n = A.n_tiles;
    const long long total = (long long)A.t_off[n - 1] + (long long)A.t_cnt[n - 1];
    const long long base = A.st->n_stored;
    A.st->chunk_base = base;
    A.st->chunk_total = total;
    const long long room = A.capacity - base;
    if (total > room) {
        if (A.st->overflow_chunk > A.chunk_idx) A.st->overflow_chunk = A.chunk_idx;
        A.st->dropped += (unsigned long long)(total - room);
        A.st->n_stored = A.capacity;
    } else {
        A.st->n_stored = base + total;
    }
}

// Canonical id of the candidate at temp index k of a chunk (the inverse of
// temp_index_of); false for lanes outside the chunk's rows and columns.
template <int OP>
__device__ __forceinline__ bool id_of_temp(const ChunkArgs &A, int64_t k, uint64_t &id) {
    if constexpr (OP == LTL_AND || OP == LTL_OR || OP == LTL_UNTIL) {
        constexpr int ND = OP == LTL_UNTIL ? 2 : 1;
        const int64_t ix = k / ND, ti = ix / 64;
        const int64_t a = A.a0 + ti / A.nbs, b = A.b0 - A.ush + (ti % A.nbs) * 64 + ix % 64;
        if (a >= A.a1 || b < A.b0 || b >= A.b1) return false;
        id = A.id_base + ND * ((uint64_t)a * (uint64_t)A.nR + (uint64_t)b) + (uint64_t)(k % ND);
        return true;
    } else {
        const int64_t a = A.a0 - A.ush + k;
        if (a < A.a0 || a >= A.a1) return false;
        id = A.id_base + (uint64_t)a;
        return true;
    }
}

// Relaxed keys, after the chunk's dedup (one launch of at most 1024
// workgroups): the smallest accepted survivor becomes `found`.  In the chunk
// that overflowed the store, the first survivor that did not fit switches the
// reference to testing every candidate (code/ltli6463.cu:198-208, 1021-1027):
// every accepted candidate after it lowers `found` too.
template <int OP>
__global__ __launch_bounds__(256) void k_rx_tail(ChunkArgs A) {
    if (run_stopped(A) || look_back_failed(A)) return;
    const unsigned long long rx = A.st->rx_found;
    if (blockIdx.x == 0 && threadIdx.x == 0 && rx != kEmpty) atomicMin(&A.st->found, rx);
    if (A.st->overflow_chunk != A.chunk_idx) return;
    const int64_t room = A.capacity - A.st->chunk_base;
    if (room < 0 || room >= A.st->chunk_total) return;
    uint64_t over;
    if (!id_of_temp<OP>(A, (int64_t)A.t_list[room], over)) return;
    constexpr int ND = OP == LTL_UNTIL ? 2 : 1;
    const int64_t n = (int64_t)A.n_tiles * 64 * ND;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        uint64_t id;
        if (A.t_acc[k] && id_of_temp<OP>(A, k, id) && id > over) atomicMin(&A.st->found, (unsigned long long)id);
    }
}

// ---------------------------------------------------------------------------
// emit
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t lanes_below() {
    const int lane = threadIdx.x & 63;
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

__device__ __forceinline__ void commit(const ChunkArgs &A, int64_t pos, size_t k, int op, int32_t l, int32_t r) {
    A.m_op[pos] = (uint8_t)op;
    A.m_left[pos] = l;
    A.m_right[pos] = r;
    A.m_fp2[pos] = A.t_fp2[k];
    // (the key is kept for re-inserting the stored sets into a grown table;
    // a table of fixed size never grows, and the read would pull a random
    // table line per survivor)
    if (A.keep_keys) A.m_key[pos] = A.tab[2 * (size_t)A.t_slot[k]];
    A.tab[2 * (size_t)A.t_slot[k] + 1] = tab_val((uint64_t)pos, A.epoch);
}

template <int OP>
__global__ __launch_bounds__(256) void k_compact(ChunkArgs A) {
    if (run_stopped(A) || chunk_otf(A)) return;
    const int64_t ti = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (ti >= A.n_tiles || A.t_cnt[ti] == 0) return;
    constexpr int ND = OP == LTL_UNTIL ? 2 : 1;
    const uint64_t below = lanes_below(), bit = 1ull << (threadIdx.x & 63);
    const uint64_t m0 = A.t_mask[ND * ti], m1 = ND == 2 ? A.t_mask[ND * ti + 1] : 0;
    const uint32_t p0 = A.t_off[ti] + (uint32_t)(__popcll(m0 & below) + __popcll(m1 & below));
    const uint32_t ix = (uint32_t)(ti * 64 + (threadIdx.x & 63));
    if (m0 & bit) A.t_list[p0] = ND * ix;
    if (m1 & bit) A.t_list[p0 + ((m0 & bit) ? 1 : 0)] = ND * ix + 1;
}

// A chunk of at most kSmallTiles tiles: decide, scan, finish and compact in
// one workgroup of 16 waves (one launch instead of a decide, a device scan, a
// finish and a compact).  The tile counts, offsets and survivor masks stay in
// LDS; t_list, the chunk's base and total and the stored count come out as
// from the four kernels.
template <int OP>
__global__ __launch_bounds__(1024) void k_dedup_small(ChunkArgs A) {
    if (run_stopped(A) || chunk_otf(A)) return;
    constexpr int ND = OP == LTL_UNTIL ? 2 : 1;
    __shared__ uint32_t s_cnt[kSmallTiles];  // survivors per tile, then the exclusive offset
    __shared__ uint64_t s_mask[2 * kSmallTiles];
    __shared__ uint32_t s_wsum[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = (int)A.n_tiles;  // <= kSmallTiles (checked at launch)
    for (int ti = wave; ti < n; ti += 16) {
        uint64_t m[2];
        const uint32_t cnt = decide_tile<OP>(A, ti, lane, m);
        if (lane == 0) {
            s_mask[2 * ti] = m[0];
            s_mask[2 * ti + 1] = m[1];
            s_cnt[ti] = cnt;
        }
    }
    __syncthreads();
    // exclusive scan over the tile counts: thread t holds tile t
    const int t = (int)threadIdx.x;
    const uint32_t v = t < n ? s_cnt[t] : 0u;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wsum[wave] = x;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < 16; ++w) {
        const uint32_t ws = s_wsum[w];
        before += w < wave ? ws : 0u;
        total += ws;
    }
    if (t < n) s_cnt[t] = before + x - v;
    if (t == 0) {
        const long long base = A.st->n_stored;
        A.st->chunk_base = base;
        A.st->chunk_total = (long long)total;
        const long long room = A.capacity - base;
        if ((long long)total > room) {
            if (A.st->overflow_chunk > A.chunk_idx) A.st->overflow_chunk = A.chunk_idx;
            A.st->dropped += (unsigned long long)((long long)total - room);
            A.st->n_stored = A.capacity;
        } else {
            A.st->n_stored = base + (long long)total;
        }
    }
    __syncthreads();
    const uint64_t below = lanes_below(), bit = 1ull << lane;
    for (int ti = wave; ti < n; ti += 16) {
        const uint64_t m0 = s_mask[2 * ti], m1 = ND == 2 ? s_mask[2 * ti + 1] : 0;
        if (!(m0 | m1)) continue;
        const uint32_t p0 = s_cnt[ti] + (uint32_t)(__popcll(m0 & below) + __popcll(m1 & below));
        const uint32_t ix = (uint32_t)(ti * 64 + lane);
        if (m0 & bit) A.t_list[p0] = ND * ix;
        if (m1 & bit) A.t_list[p0 + ((m0 & bit) ? 1 : 0)] = ND * ix + 1;
    }
}

// Chunks of more than kSmallTiles tiles: one launch of 4-tile workgroups
// (one wave per tile) in place of decide, device scan, finish and compact.
// A workgroup takes a ticket (its place in tile order), decides its tiles,
// publishes its survivor count, then adds the counts of the workgroups before
// it, walking back until one that has published its inclusive prefix
// (decoupled look-back; every workgroup it waits for took its ticket earlier
// and is running).  Status words carry the chunk's epoch, so words of earlier
// chunks read as not yet published.  The waits are bounded: one that runs
// out marks the status word poisoned and sets lb_fail; a poisoned workgroup
// writes no survivor list entries, the kernels that read the list do nothing
// once lb_fail is set, and the call fails with a HIP error.  The last workgroup does the
// finish and returns the ticket counter to 0.  The status words are
// relaxed agent-scope atomics (coherent across the XCDs' L2s): a count
// travels inside its status word, and what else the workgroups write (t_list,
// the chunk's base and total) is read by later launches, so no release or
// acquire fence (an L2 write-back or invalidate per wait) is needed.
constexpr uint32_t kLbSpinMax = 1u << 20;

template <int OP>
__global__ __launch_bounds__(256) void k_dedup_lb(ChunkArgs A) {
    if (run_stopped(A) || chunk_otf(A)) return;
    constexpr int ND = OP == LTL_UNTIL ? 2 : 1;
    __shared__ uint32_t s_cnt[4];
    __shared__ uint32_t s_blk, s_excl, s_poison;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t nblk = (uint32_t)((A.n_tiles + 3) / 4);
    if (threadIdx.x == 0) s_poison = 0;
    if (threadIdx.x == 0) s_blk = atomicAdd(A.lb_ticket, 1u);
    __syncthreads();
    const uint32_t blk = s_blk;
    const int64_t ti = (int64_t)blk * 4 + wave;
    uint64_t m[2] = {0, 0};
    uint32_t cnt = 0;
    if (blk < nblk && ti < A.n_tiles) cnt = decide_tile<OP>(A, ti, lane, m);
    if (lane == 0) s_cnt[wave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0 && blk < nblk) {
        const uint32_t agg = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
        const uint64_t tag = (uint64_t)A.lb_epoch << 2;
        uint32_t excl = 0;
        bool poison = false;
        if (blk > 0) {
            __hip_atomic_store(&A.lb_flags[blk], ((tag | 1) << 32) | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int64_t p = (int64_t)blk - 1;
            uint32_t spins = 0;
            while (p >= 0) {
                const uint64_t v = __hip_atomic_load(&A.lb_flags[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t hi = (uint32_t)(v >> 32);
                if ((hi >> 2) != A.lb_epoch) {
                    if (++spins > kLbSpinMax) { poison = true; break; }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                if ((hi & 3) == 3) { poison = true; break; }
                excl += (uint32_t)v;
                if ((hi & 3) == 2) break;
                --p;
            }
        }
        __hip_atomic_store(&A.lb_flags[blk], ((tag | (poison ? 3u : 2u)) << 32) | (uint32_t)(excl + agg),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (poison) __hip_atomic_store(&A.st->lb_fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_excl = excl;
        s_poison = poison ? 1u : 0u;
        if (blk == nblk - 1) {
            // every workgroup has taken its ticket
            __hip_atomic_store(A.lb_ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const long long total = poison ? 0 : (long long)excl + (long long)agg;
            const long long base = A.st->n_stored;
            A.st->chunk_base = base;
            A.st->chunk_total = total;
            const long long room = A.capacity - base;
            if (total > room) {
                if (A.st->overflow_chunk > A.chunk_idx) A.st->overflow_chunk = A.chunk_idx;
                A.st->dropped += (unsigned long long)(total - room);
                A.st->n_stored = A.capacity;
            } else {
                A.st->n_stored = base + total;
            }
        }
    }
    __syncthreads();
    // (a poisoned workgroup's prefix is unknown: it writes no list entries)
    if (blk >= nblk || cnt == 0 || s_poison) return;
    uint32_t before = s_excl;
    for (int w = 0; w < wave; ++w) before += s_cnt[w];
    const uint64_t below = lanes_below(), bit = 1ull << lane;
    const uint64_t m0 = m[0], m1 = ND == 2 ? m[1] : 0;
    const uint32_t p0 = before + (uint32_t)(__popcll(m0 & below) + __popcll(m1 & below));
    const uint32_t ix = (uint32_t)(ti * 64 + lane);
    if (m0 & bit) A.t_list[p0] = ND * ix;
    if (m1 & bit) A.t_list[p0 + ((m0 & bit) ? 1 : 0)] = ND * ix + 1;
}

// `aux` and `saux` are A.aux and A.saux as separate restrict parameters: the
// per-word records are then known not to be written through the store
// pointer, and are read with scalar loads instead of vector loads that wait
// behind the data loads.
//
// Summary vectors that follow from the operands' summaries are taken from
// there (one load per vector and summary word instead of a bit per trace and
// word): for ~, all four (b0 and b1 negated, nz = not fl, fl = not nz); for &,
// b0, b1 and fl (ANDed); for |, b0, b1 and nz (ORed); for X, b0 (the
// operand's b1); for F, b0 and nz (the operand's nz); for G, b0 and fl (the
// operand's fl); for U, nz (the right operand's).  The rest come from the
// words.
//
// Word ranges: the A.esplit waves of a store tile each take the summary
// groups [g0, g1) of an even split of the nsw groups, and the words that hold
// their traces, [gfirst[g0], glast[g1 - 1]] (a word shared by two groups at a
// range boundary is written by both ranges, with the same value).  More waves
// stream each tile's words in parallel: a chunk's survivors alone are too few
// waves to hide the memory latency.
//
// NX (ChunkArgs::nx_fp): the next level's ~, X, F and G of each survivor are
// fingerprinted from the words computed here (each range hashes the words
// it owns: a word shared with the previous range at a group boundary is that
// range's); a workgroup is the esplit (<= 8) waves of one store tile, and
// wave 0 adds the ranges' sums in LDS and stores them.
template <int OP, bool NX = false>
__global__ __launch_bounds__(NX ? 512 : 256) void k_emit(ChunkArgs A, const WordAux *__restrict__ aux,
                                                         const SumAux *__restrict__ saux) {
    if (run_stopped(A) || chunk_otf(A) || look_back_failed(A)) return;
    // Waves cover whole store tiles: lane l of the waves of tile v writes
    // position 64 * (chunk_base / 64 + v) + l, so every store instruction of a
    // wave fills one 512-byte segment of one tile.
    const int64_t base = A.st->chunk_base;
    const int K = A.esplit;
    // (wave-uniform: the range loops below run on the scalar unit)
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t q = NX ? (int64_t)blockIdx.x * K + wv : (int64_t)blockIdx.x * 4 + wv;
    const int rg = (int)(q % K);
    const int64_t pos = (base & ~63ll) + (q / K) * 64 + (threadIdx.x & 63);
    const int64_t i = pos - base;
    // (NX: every lane stays until the workgroup's ranges are added)
    const bool live = i >= 0 && i < A.st->chunk_total && pos < A.capacity;
    if (!NX && !live) return;
    constexpr bool bin = OP == LTL_AND || OP == LTL_OR || OP == LTL_UNTIL;
    const uint32_t k = live ? A.t_list[i] : 0;  // temp index of the candidate
    const uint32_t ix = OP == LTL_UNTIL ? k >> 1 : k;
    const bool swap = OP == LTL_UNTIL && (k & 1);
    int64_t sl, sr = -1;
    if (bin) {
        const int64_t ti = ix / 64;
        sl = A.ls + A.a0 + ti / A.nbs;
        sr = A.rs + A.b0 - A.ush + (ti % A.nbs) * 64 + (ix % 64);
    } else {
        sl = A.ls + A.a0 - A.ush + ix;
    }
    const uint64_t *pl = A.store + set_base(sl, A.nw);
    const uint64_t *pr = A.store + set_base(bin ? sr : sl, A.nw);
    uint64_t *qs = A.store + set_base(pos, A.nw);
    uint64_t *z = A.sum + set_base(pos, 4 * A.nsw);
    const uint64_t *zl = A.sum + set_base(sl, 4 * A.nsw), *zr = A.sum + set_base(bin ? sr : sl, 4 * A.nsw);
    // vectors computed from the words (bit v for vector v); the others follow from the operands
    constexpr int VM = OP == LTL_NOT ? 0 : OP == LTL_AND ? 4 : OP == LTL_OR ? 8 : OP == LTL_NEXT ? 14
                     : OP == LTL_FIN ? 10 : OP == LTL_GLOB ? 6 : 11;
    uint64_t acc[4] = {0, 0, 0, 0};
    const int g0 = rg * A.nsw / K, g1 = (rg + 1) * A.nsw / K;
    const int w0 = K > 1 ? A.gfirst[g0] : 0, w1 = K > 1 ? A.glast[g1 - 1] + 1 : A.nw;
    // Summary word j is complete: the vectors that follow from the operands'
    // summaries, the accumulated ones, stored; the accumulators cleared.
    auto flush = [&](int j) {
        uint64_t out[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) out[v] = acc[v];
        if constexpr (OP == LTL_NOT) {
            const SumAux sa = saux[j];
            const uint64_t o0 = zl[(size_t)(0 * A.nsw + j) * kTileWS], o1 = zl[(size_t)(1 * A.nsw + j) * kTileWS];
            const uint64_t o2 = zl[(size_t)(2 * A.nsw + j) * kTileWS], o3 = zl[(size_t)(3 * A.nsw + j) * kTileWS];
            out[0] = ~o0 & sa.val;
            out[1] = ~o1 & sa.val2;
            out[2] = ~o3 & sa.val;
            out[3] = ~o2 & sa.val;
        } else if constexpr (OP == LTL_NEXT) {
            out[0] = zl[(size_t)(1 * A.nsw + j) * kTileWS];                   // X f at 0 = f at 1
        } else if constexpr (OP == LTL_FIN) {
            out[0] = out[2] = zl[(size_t)(2 * A.nsw + j) * kTileWS];          // F f at 0 / anywhere = f anywhere
        } else if constexpr (OP == LTL_GLOB) {
            out[0] = out[3] = zl[(size_t)(3 * A.nsw + j) * kTileWS];          // G f at 0 / everywhere = f everywhere
        } else if constexpr (OP == LTL_UNTIL) {
            // l U r holds somewhere iff r does (the right operand: sl when swapped)
            out[2] = (swap ? zl : zr)[(size_t)(2 * A.nsw + j) * kTileWS];
        } else {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if (v == (OP == LTL_AND ? 2 : 3)) continue;  // from the words
                const uint64_t a = zl[(size_t)(v * A.nsw + j) * kTileWS], b = zr[(size_t)(v * A.nsw + j) * kTileWS];
                out[v] = OP == LTL_AND ? (a & b) : (a | b);
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            __builtin_nontemporal_store(out[v], z + (size_t)(v * A.nsw + j) * kTileWS);
            acc[v] = 0;
        }
    };
    // NX: the next level's candidates ~c, X c, F c, G c, hashed from word hw0 on
    Acc nx[4];
    const int hw0 = NX && K > 1 && rg > 0 && A.glast[g0 - 1] == w0 ? w0 + 1 : w0;
    auto step = [&](int w, uint64_t l, uint64_t r) {
        const WordAux *x = aux + w;
        uint64_t c;
        if constexpr (OP == LTL_AND) c = l & r;
        else if constexpr (OP == LTL_OR) c = l | r;
        else if constexpr (OP == LTL_UNTIL) c = swap ? pk_until(r, l, x->inner) : pk_until(l, r, x->inner);
        else c = pk_unary<OP>(l, x);
        // streamed: not re-read in this chunk
        if (!NX || !A.wskip) __builtin_nontemporal_store(c, qs + (size_t)w * kTileWS);
        summarize_pk<VM>(c, x, acc, A.W, flush, g0, g1);
        if constexpr (NX) {
            if (w >= hw0) {
                fold<false>(nx[0], pk_unary<LTL_NOT>(c, x), x);
                fold<false>(nx[1], pk_unary<LTL_NEXT>(c, x), x);
                fold<false>(nx[2], pk_unary<LTL_FIN>(c, x), x);
                fold<false>(nx[3], pk_unary<LTL_GLOB>(c, x), x);
            }
        }
    };
    constexpr int UB = bin ? 4 : 8;
    if (live) {
        int w = w0;
        for (; w + UB <= w1; w += UB) {
            uint64_t l[UB], r[UB];
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                l[u] = pl[(size_t)(w + u) * kTileWS];
                r[u] = bin ? pr[(size_t)(w + u) * kTileWS] : 0;
            }
#pragma unroll
            for (int u = 0; u < UB; ++u) step(w + u, l[u], r[u]);
        }
        for (; w < w1; ++w) step(w, pl[(size_t)w * kTileWS], bin ? pr[(size_t)w * kTileWS] : 0);
    }
    if constexpr (NX) {
        __shared__ uint64_t nx_lds[7][8][64];
        const int lane = (int)(threadIdx.x & 63);
        uint64_t v[8];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            v[2 * j] = nx[j].f1;
            v[2 * j + 1] = nx[j].f2;
        }
        if (rg > 0) {
#pragma unroll
            for (int j = 0; j < 8; ++j) nx_lds[rg - 1][j][lane] = v[j];
        }
        __syncthreads();
        if (rg != 0 || !live) return;
        for (int r = 1; r < K; ++r) {
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] += nx_lds[r - 1][j][lane];
        }
        const int64_t nxi = pos - A.nx_ls + A.nx_ush;
        if (nxi < 0 || nxi >= A.nx_stride) {
            A.st->nx_over = 1;  // (the level is not taken from these fingerprints)
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) A.nx_fp[(size_t)j * (size_t)A.nx_stride + (size_t)nxi] = v[j];
        }
    }
    if (rg != 0 || !live) return;
    if (bin) commit(A, pos, k, OP, (int32_t)(swap ? sr : sl), (int32_t)(swap ? sl : sr));
    else commit(A, pos, k, OP, (int32_t)sl, -1);
}

// ~, X, F and G blocks of a tested level (k_gen_xfg<false, false, true>'s
// tail) from the fingerprints an emit of their operand level stored: one
// lane per operand; ids and temporaries as there.
__global__ __launch_bounds__(256) void k_insert_xfg(ChunkArgs A, int64_t n, int64_t slots, uint64_t id_xfg,
                                                    const uint64_t *__restrict__ fp, int64_t stride) {
    if (run_stopped(A) || chunk_otf(A)) return;
    const int64_t ix = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (ix >= (int64_t)A.n_tiles * 64) return;
    const int64_t a = A.a0 - A.ush + ix;
    if (a < A.a0 || a >= A.a1) return;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint64_t id = (j == 0 ? A.id_base : id_xfg + (uint64_t)(j - 1) * (uint64_t)n) + (uint64_t)a;
        const uint64_t f1 = fp[(size_t)(2 * j) * (size_t)stride + ix], f2 = fp[(size_t)(2 * j + 1) * (size_t)stride + ix];
        const size_t k = (size_t)j * (size_t)slots + ix;
        A.t_slot[k] = tab_insert<true>(A.tab, A.tmask, tab_key(f1, A.epoch), tab_val(kPending | id, A.epoch), A.epoch,
                                 &A.st->probe_fail);
        A.t_fp2[k] = f2;
    }
}

// Deferred level: survivors get their metadata and table commit only (no
// words, no summaries), and each survivor whose successor j was accepted in
// generation lowers succ_min[j] to its store index.
__global__ __launch_bounds__(256) void k_emit_meta(ChunkArgs A, int op) {
    if (run_stopped(A) || chunk_otf(A) || look_back_failed(A)) return;
    const int64_t base = A.st->chunk_base;
    const int64_t pos = (base & ~63ll) + (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t i = pos - base;
    if (i < 0 || i >= A.st->chunk_total) return;
    if (pos >= A.capacity) return;
    const bool bin = op == LTL_AND || op == LTL_OR || op == LTL_UNTIL;
    const uint32_t k = A.t_list[i];
    const uint32_t ix = op == LTL_UNTIL ? k >> 1 : k;
    const bool swap = op == LTL_UNTIL && (k & 1);
    uint64_t id;  // canonical candidate id
    if (bin) {
        const int64_t ti = ix / 64;
        const int64_t a = A.a0 + ti / A.nbs, b = A.b0 - A.ush + (ti % A.nbs) * 64 + (ix % 64);
        const uint64_t pair = (uint64_t)a * (uint64_t)A.nR + (uint64_t)b;
        id = A.id_base + (op == LTL_UNTIL ? 2 * pair + (swap ? 1 : 0) : pair);
        const int64_t sl = A.ls + a, sr = A.rs + b;
        commit(A, pos, k, op, (int32_t)(swap ? sr : sl), (int32_t)(swap ? sl : sr));
    } else {
        const int64_t a = A.a0 - A.ush + ix;
        id = A.id_base + (uint64_t)a;
        commit(A, pos, k, op, (int32_t)(A.ls + a), -1);
    }
    // ~c: the smallest candidate id whose complement the check pass of this
    // level accepted is the first occurrence of its set, so it is stored
    if (id == A.neg_pick) atomicMin(&A.st->succ_min[0], (unsigned long long)pos);
    if (A.t_succ) {
        const uint32_t f = A.t_succ[k];
#pragma unroll
        for (int j = 1; j < 4; ++j)
            if ((f >> j) & 1u) atomicMin(&A.st->succ_min[j], (unsigned long long)pos);
    }
}

// Words and summaries of the stored sets [s0, s1), recomputed from their
// parents as the metadata records them (m_op, m_left, m_right; until with
// the operands in formula order): the sets of a deferred level.  One lane
// per set; packed layout.
// SUMS: the summaries too (a deferred level); else the words only (a level
// stored with its summaries but without words).
template <bool SUMS>
__global__ __launch_bounds__(256) void k_materialize(ChunkArgs A, int64_t s0, int64_t s1) {
    const int64_t s = (s0 & ~63ll) + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s < s0 || s >= s1) return;
    const int op = A.m_op[s];
    const int64_t l = A.m_left[s], r = A.m_right[s] >= 0 ? A.m_right[s] : l;
    const uint64_t *pl = A.store + set_base(l, A.nw), *pr = A.store + set_base(r, A.nw);
    uint64_t *q = A.store + set_base(s, A.nw);
    uint64_t *z = A.sum + set_base(s, 4 * A.nsw);
    uint64_t acc[4] = {0, 0, 0, 0};
    auto flush = [&](int j) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            z[(size_t)(v * A.nsw + j) * kTileWS] = acc[v];
            acc[v] = 0;
        }
    };
    for (int w = 0; w < A.nw; ++w) {
        const WordAux *x = A.aux + w;
        const uint64_t a = pl[(size_t)w * kTileWS], b = pr[(size_t)w * kTileWS];
        uint64_t v;
        switch (op) {
        case LTL_NOT: v = pk_unary<LTL_NOT>(a, x); break;
        case LTL_NEXT: v = pk_unary<LTL_NEXT>(a, x); break;
        case LTL_FIN: v = pk_unary<LTL_FIN>(a, x); break;
        case LTL_GLOB: v = pk_unary<LTL_GLOB>(a, x); break;
        case LTL_AND: v = a & b; break;
        case LTL_OR: v = a | b; break;
        case LTL_UNTIL: v = pk_until(a, b, x->inner); break;
        default: v = a; break;  // atoms are stored with their words
        }
        q[(size_t)w * kTileWS] = v;
        if constexpr (SUMS) summarize_pk<15>(v, x, acc, A.W, flush);
    }
}

// Multi-word traces (B = 0): k_materialize for K words per trace, one trace
// per step, its summary bits from all K words.
template <int K, bool SUMS>
__global__ __launch_bounds__(256) void k_materialize_m(ChunkArgs A, int64_t s0, int64_t s1) {
    const int64_t s = (s0 & ~63ll) + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s < s0 || s >= s1) return;
    const int op = A.m_op[s];
    const int64_t l = A.m_left[s], r = A.m_right[s] >= 0 ? A.m_right[s] : l;
    const uint64_t *pl = A.store + set_base(l, A.nw), *pr = A.store + set_base(r, A.nw);
    uint64_t *q = A.store + set_base(s, A.nw);
    uint64_t *z = A.sum + set_base(s, 4 * A.nsw);
    const int T = A.nw / K;
    uint64_t v[4] = {0, 0, 0, 0};
    for (int i = 0; i < T; ++i) {
        uint64_t x[K], y[K], w[K];
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            x[kk] = pl[(size_t)(i * K + kk) * kTileWS];
            y[kk] = pr[(size_t)(i * K + kk) * kTileWS];
        }
        const WordAux *aux = A.aux + i * K;
        switch (op) {
        case LTL_NOT: trace_unary<LTL_NOT, K>(x, aux, w); break;
        case LTL_NEXT: trace_unary<LTL_NEXT, K>(x, aux, w); break;
        case LTL_FIN: trace_unary<LTL_FIN, K>(x, aux, w); break;
        case LTL_GLOB: trace_unary<LTL_GLOB, K>(x, aux, w); break;
        case LTL_UNTIL: trace_until<K>(x, y, w); break;
        default:
#pragma unroll
            for (int kk = 0; kk < K; ++kk)
                w[kk] = op == LTL_AND ? (x[kk] & y[kk]) : op == LTL_OR ? (x[kk] | y[kk]) : x[kk];
            break;
        }
#pragma unroll
        for (int kk = 0; kk < K; ++kk) q[(size_t)(i * K + kk) * kTileWS] = w[kk];
        if constexpr (SUMS) {
            summarize_trace<K>(w, aux, i & 63, v);
            flush_summary(z, i, T, A.nsw, kTileWS, v);
        }
    }
}

// Multi-word traces (B = 0): one trace (K words) per step; its summary bits
// come from all K words.
// One workgroup of E = A.esplit waves per store tile: wave r writes the
// traces of summary groups [r nsw / E, (r + 1) nsw / E) of the tile's 64
// survivors (lane l: position 64 (chunk_base / 64 + tile) + l), so that E
// waves share each survivor's walk over its nw words.  NX: also the
// fingerprints of the survivor's ~, X, F and G (the next level's unary
// candidates, as k_emit<OP, true>), the waves' partial sums added in LDS;
// with A.wskip no words.
template <int OP, int K, bool NX = false>
__global__ __launch_bounds__(512) void k_emitm(ChunkArgs A) {
    if (run_stopped(A) || chunk_otf(A) || look_back_failed(A)) return;
    const int64_t base = A.st->chunk_base;
    const int E = A.esplit;
    const int rg = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // (wave-uniform)
    const int lane = (int)(threadIdx.x & 63);
    const int64_t pos = (base & ~63ll) + (int64_t)blockIdx.x * 64 + lane;
    const int64_t rel = pos - base;
    // (NX: every lane stays until the workgroup's ranges are added)
    const bool live = rel >= 0 && rel < A.st->chunk_total && pos < A.capacity;
    if (!NX && !live) return;
    constexpr bool bin = OP == LTL_AND || OP == LTL_OR || OP == LTL_UNTIL;
    const uint32_t k = live ? A.t_list[rel] : 0;  // temp index of the candidate
    const uint32_t ix = OP == LTL_UNTIL ? k >> 1 : k;
    const bool swap = OP == LTL_UNTIL && (k & 1);
    int64_t sl, sr = -1;
    if (bin) {
        const int64_t ti = ix / 64;
        sl = A.ls + A.a0 + ti / A.nbs;
        sr = A.rs + A.b0 - A.ush + (ti % A.nbs) * 64 + (ix % 64);
    } else {
        sl = A.ls + A.a0 - A.ush + ix;
    }
    const uint64_t *pl = A.store + set_base(swap ? sr : sl, A.nw);
    const uint64_t *pr = A.store + set_base(bin ? (swap ? sl : sr) : sl, A.nw);
    uint64_t *q = A.store + set_base(pos, A.nw);
    uint64_t *z = A.sum + set_base(pos, 4 * A.nsw);
    const int T = A.nw / K;
    const int g0 = rg * A.nsw / E, g1 = (rg + 1) * A.nsw / E;
    const int i0 = 64 * g0, i1 = T < 64 * g1 ? T : 64 * g1;
    uint64_t v[4] = {0, 0, 0, 0};
    Acc nx[4];
    for (int i = live ? i0 : i1; i < i1; ++i) {
        uint64_t x[K], r[K], y[K];
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            x[kk] = pl[(size_t)(i * K + kk) * kTileWS];
            r[kk] = bin ? pr[(size_t)(i * K + kk) * kTileWS] : 0ull;
        }
        const WordAux *aux = A.aux + i * K;
        if constexpr (OP == LTL_AND) {
#pragma unroll
            for (int kk = 0; kk < K; ++kk) y[kk] = x[kk] & r[kk];
        } else if constexpr (OP == LTL_OR) {
#pragma unroll
            for (int kk = 0; kk < K; ++kk) y[kk] = x[kk] | r[kk];
        } else if constexpr (OP == LTL_UNTIL) {
            trace_until<K>(x, r, y);
        } else {
            trace_unary<OP, K>(x, aux, y);
        }
        if (!NX || !A.wskip) {
#pragma unroll
            for (int kk = 0; kk < K; ++kk) __builtin_nontemporal_store(y[kk], q + (size_t)(i * K + kk) * kTileWS);
        }
        summarize_trace<K>(y, aux, i & 63, v);
        flush_summary(z, i, T, A.nsw, kTileWS, v);
        if constexpr (NX) {
            uint64_t u[K];
            trace_unary<LTL_NOT, K>(y, aux, u);
#pragma unroll
            for (int kk = 0; kk < K; ++kk) fold<false>(nx[0], u[kk], aux + kk);
            trace_unary<LTL_NEXT, K>(y, aux, u);
#pragma unroll
            for (int kk = 0; kk < K; ++kk) fold<false>(nx[1], u[kk], aux + kk);
            trace_unary<LTL_FIN, K>(y, aux, u);
#pragma unroll
            for (int kk = 0; kk < K; ++kk) fold<false>(nx[2], u[kk], aux + kk);
            trace_unary<LTL_GLOB, K>(y, aux, u);
#pragma unroll
            for (int kk = 0; kk < K; ++kk) fold<false>(nx[3], u[kk], aux + kk);
        }
    }
    if constexpr (NX) {
        __shared__ uint64_t nx_lds[7][8][64];
        uint64_t f[8];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            f[2 * j] = nx[j].f1;
            f[2 * j + 1] = nx[j].f2;
        }
        if (rg > 0) {
#pragma unroll
            for (int j = 0; j < 8; ++j) nx_lds[rg - 1][j][lane] = f[j];
        }
        __syncthreads();
        if (rg != 0 || !live) return;
        for (int w = 1; w < E; ++w) {
#pragma unroll
            for (int j = 0; j < 8; ++j) f[j] += nx_lds[w - 1][j][lane];
        }
        const int64_t nxi = pos - A.nx_ls + A.nx_ush;
        if (nxi < 0 || nxi >= A.nx_stride) {
            A.st->nx_over = 1;  // (the level is not taken from these fingerprints)
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) A.nx_fp[(size_t)j * (size_t)A.nx_stride + (size_t)nxi] = f[j];
        }
    }
    if (rg != 0 || !live) return;
    if (bin) commit(A, pos, k, OP, (int32_t)(swap ? sr : sl), (int32_t)(swap ? sl : sr));
    else commit(A, pos, k, OP, (int32_t)sl, -1);
}

// ---------------------------------------------------------------------------
// check-only pass (no dedup, no store)
// ---------------------------------------------------------------------------
// ~, X, F, G: one summary vector of the operand (b0 negated, b1, nz, fl).
// Words of a stored set, which may lie in the range whose words were not
// written: then word w is recomputed from the set's parents (whose words
// are written), as k_materialize does.
struct SetWords {
    const uint64_t *p, *pa, *pb;
    int op;  // -1: the words are at p
};
__device__ __forceinline__ bool in_lzw(const ChunkArgs &A, int64_t s) { return s >= A.lzw_s && s < A.lzw_e; }
__device__ __forceinline__ SetWords set_words(const ChunkArgs &A, int64_t s) {
    SetWords d;
    d.p = d.pa = d.pb = A.store + set_base(s, A.nw);
    d.op = -1;
    if (in_lzw(A, s)) {
        d.op = A.m_op[s];
        const int64_t l = A.m_left[s], r = A.m_right[s] >= 0 ? A.m_right[s] : l;
        d.pa = A.store + set_base(l, A.nw);
        d.pb = A.store + set_base(r, A.nw);
    }
    return d;
}
__device__ __forceinline__ uint64_t word_at(const ChunkArgs &A, const SetWords &d, int w) {
    if (d.op < 0) return d.p[(size_t)w * kTileWS];
    const WordAux *x = A.aux + w;
    const uint64_t a = d.pa[(size_t)w * kTileWS], b = d.pb[(size_t)w * kTileWS];
    switch (d.op) {
    case LTL_NOT: return pk_unary<LTL_NOT>(a, x);
    case LTL_NEXT: return pk_unary<LTL_NEXT>(a, x);
    case LTL_FIN: return pk_unary<LTL_FIN>(a, x);
    case LTL_GLOB: return pk_unary<LTL_GLOB>(a, x);
    case LTL_AND: return a & b;
    case LTL_OR: return a | b;
    case LTL_UNTIL: return pk_until(a, b, x->inner);
    default: return a;
    }
}

template <int OP>
__global__ __launch_bounds__(256) void k_chk1(ChunkArgs A) {
    if (run_stopped(A)) return;
    Tile t;
    if (!tile_unary(A, t)) return;
    constexpr int VEC = OP == LTL_NOT ? 0 : OP == LTL_NEXT ? 1 : OP == LTL_FIN ? 2 : 3;
    const uint64_t *p = A.sum + set_base(A.ls + (t.valid ? t.a : A.a0), 4 * A.nsw) + (size_t)(VEC * A.nsw) * kTileWS;
    const uint32_t k = (uint32_t)A.hamming_k;
    const bool anti = A.anti && t.valid;
    uint32_t m = t.valid ? 0u : 0x7fffffffu;  // invalid lanes never accept
    uint32_t seen = 0;                         // traces tested so far
    for (int i = 0; i < A.nsw; ++i) {
        const int j = interleave(i, A.nps, A.nsw);
        uint64_t v = p[(size_t)j * kTileWS];
        if (OP == LTL_NOT) v = ~v;
        m += (uint32_t)__popcll((v ^ A.saux[j].tgt) & A.saux[j].val);
        seen += (uint32_t)__popcll(A.saux[j].val);
        // Every lane of the wave already fails (and, with anti, classifies
        // more than k traces correctly, so its complement fails too).
        if (!__any(m <= k || (anti && seen - m <= k))) return;
    }
    const uint64_t id = A.id_base + (uint64_t)t.a;
    if (m <= k) atomicMin(&A.st->found, (unsigned long long)id);
    else if (anti && seen - m <= k) atomicMin(&A.st->neg_found[A.neg_slot], (unsigned long long)id);
}

// & and |: position 0 of the result is the AND / OR of the operands' b0.
template <int OP>
__global__ __launch_bounds__(256) void k_chk2(ChunkArgs A) {
    if (run_stopped(A)) return;
    Tile t;
    if (!tile_binary(A, t)) return;
    const int64_t sl = A.ls + t.a, sr = A.rs + (t.valid ? t.b : A.b0);
    const uint64_t *pl = A.sum + set_base(sl, 4 * A.nsw);
    const uint64_t *pr = A.sum + set_base(sr, 4 * A.nsw);
    const uint32_t k = (uint32_t)A.hamming_k;
    const bool anti = A.anti && t.valid;
    uint32_t m = t.valid ? 0u : 0x7fffffffu;  // invalid lanes never accept
    uint32_t seen = 0;                         // traces tested so far
    for (int i = 0; i < A.nsw; ++i) {
        const int j = interleave(i, A.nps, A.nsw);
        const uint64_t l = pl[(size_t)j * kTileWS], r = pr[(size_t)j * kTileWS];
        const uint64_t c = OP == LTL_AND ? (l & r) : (l | r);
        m += (uint32_t)__popcll((c ^ A.saux[j].tgt) & A.saux[j].val);
        seen += (uint32_t)__popcll(A.saux[j].val);
        // every lane of the wave already fails (with anti: its complement too)
        if (!__any(m <= k || (anti && seen - m <= k))) return;
    }
    const uint64_t id = A.id_base + (uint64_t)t.a * (uint64_t)A.nR + (uint64_t)t.b;
    if (m <= k) atomicMin(&A.st->found, (unsigned long long)id);
    else if (anti && seen - m <= k) atomicMin(&A.st->neg_found[A.neg_slot], (unsigned long long)id);
}

__global__ __launch_bounds__(256) void k_chkU(ChunkArgs A) {
    if (run_stopped(A)) return;
    Tile t;
    if (!tile_binary(A, t)) return;
    const int64_t sl = A.ls + t.a, sr = A.rs + (t.valid ? t.b : A.b0);
    const uint64_t *pl = A.store + set_base(sl, A.nw);
    const uint64_t *pr = A.store + set_base(sr, A.nw);
    const uint32_t k = (uint32_t)A.hamming_k;
    const bool anti = A.anti && t.valid;
    uint32_t m0 = t.valid ? 0u : 0x7fffffffu, m1 = m0;  // invalid lanes never accept
    uint32_t seen = 0;                                  // traces tested so far
    auto step = [&](int w, uint64_t l, uint64_t r) {
        const WordAux *x = A.aux + w;
        m0 += (uint32_t)__popcll((pk_until(l, r, x->inner) ^ x->tgt) & x->chk);
        m1 += (uint32_t)__popcll((pk_until(r, l, x->inner) ^ x->tgt) & x->chk);
        seen += (uint32_t)__popcll(x->chk);
    };
    auto done = [&]() {
        return !__any(m0 <= k || m1 <= k || (anti && (seen - m0 <= k || seen - m1 <= k)));
    };
    // An operand level whose words were not written (a level is in the range
    // as a whole, so the test is uniform): its words recomputed from the
    // parents, word by word, until the wave is done.
    if (in_lzw(A, A.ls) || in_lzw(A, A.rs)) {
        const SetWords L = set_words(A, sl), Rw = set_words(A, sr);
        for (int i = 0; i < A.nw; ++i) {
            const int w = interleave(i, A.npw, A.nw);
            step(w, word_at(A, L, w), word_at(A, Rw, w));
            if ((i & 3) == 3 && done()) return;
        }
    } else {
        constexpr int UB = 4;
        int i = 0;
        for (; i + UB <= A.nw; i += UB) {
            uint64_t l[UB], r[UB];
            int wu[UB];
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                wu[u] = interleave(i + u, A.npw, A.nw);
                l[u] = pl[(size_t)wu[u] * kTileWS];
                r[u] = pr[(size_t)wu[u] * kTileWS];
            }
#pragma unroll
            for (int u = 0; u < UB; ++u) step(wu[u], l[u], r[u]);
            // Mismatch and match counts only grow: stop once every lane of the
            // wave fails (with anti: the complements too).
            if (done()) return;
        }
        for (; i < A.nw; ++i) {
            const int w = interleave(i, A.npw, A.nw);
            step(w, pl[(size_t)w * kTileWS], pr[(size_t)w * kTileWS]);
        }
    }
    if (!t.valid) return;
    const uint64_t id0 = A.id_base + 2 * ((uint64_t)t.a * (uint64_t)A.nR + (uint64_t)t.b);
    if (m0 <= k) atomicMin(&A.st->found, (unsigned long long)id0);
    else if (m1 <= k) atomicMin(&A.st->found, (unsigned long long)(id0 + 1));
    else if (anti && seen - m0 <= k) atomicMin(&A.st->neg_found[A.neg_slot], (unsigned long long)id0);
    else if (anti && seen - m1 <= k) atomicMin(&A.st->neg_found[A.neg_slot], (unsigned long long)(id0 + 1));
}

// Multi-word traces (B = 0): position 0 of U from the first word of each
// trace that decides it, traces in interleaved order.
// The K words of trace i of a set whose words may be unwritten (SetWords):
// loaded, or recomputed from the parents' words of that trace.
template <int K>
__device__ __forceinline__ void trace_words(const ChunkArgs &A, const SetWords &d, int i, uint64_t (&y)[K]) {
    if (d.op < 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) y[k] = d.p[(size_t)(i * K + k) * kTileWS];
        return;
    }
    uint64_t a[K], b[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        a[k] = d.pa[(size_t)(i * K + k) * kTileWS];
        b[k] = d.pb[(size_t)(i * K + k) * kTileWS];
    }
    const WordAux *aux = A.aux + i * K;
    switch (d.op) {
    case LTL_NOT: trace_unary<LTL_NOT, K>(a, aux, y); break;
    case LTL_NEXT: trace_unary<LTL_NEXT, K>(a, aux, y); break;
    case LTL_FIN: trace_unary<LTL_FIN, K>(a, aux, y); break;
    case LTL_GLOB: trace_unary<LTL_GLOB, K>(a, aux, y); break;
    case LTL_AND:
#pragma unroll
        for (int k = 0; k < K; ++k) y[k] = a[k] & b[k];
        break;
    case LTL_OR:
#pragma unroll
        for (int k = 0; k < K; ++k) y[k] = a[k] | b[k];
        break;
    case LTL_UNTIL: trace_until<K>(a, b, y); break;
    default:
#pragma unroll
        for (int k = 0; k < K; ++k) y[k] = a[k];
    }
}

// k_chkUm over an operand level stored without words: each trace's words
// recomputed from the parents.  false: the wave is done (no lane can still
// be accepted).
template <int K>
__device__ bool chkUm_unwritten(const ChunkArgs &A, int64_t sl, int64_t sr, bool anti, uint32_t &m0, uint32_t &m1) {
    const uint32_t k = (uint32_t)A.hamming_k;
    const SetWords L = set_words(A, sl), Rw = set_words(A, sr);
    const int T = A.nw / K;
    for (int j = 0; j < T; ++j) {
        const int i = interleave(j, A.npw, T);
        uint64_t l[K], r[K];
        trace_words<K>(A, L, i, l);
        trace_words<K>(A, Rw, i, r);
        uint32_t h0, h1;
        until_at0_trace(l, r, 1, K, h0, h1);
        const uint32_t tg = (uint32_t)(A.aux[i * K].tgt & 1ull);
        m0 += h0 ^ tg;
        m1 += h1 ^ tg;
        const uint32_t seen = (uint32_t)j + 1;
        if ((j & 3) == 3 && !__any(m0 <= k || m1 <= k || (anti && (seen - m0 <= k || seen - m1 <= k)))) return false;
    }
    return true;
}

__global__ __launch_bounds__(256) void k_chkUm(ChunkArgs A) {
    if (run_stopped(A)) return;
    Tile t;
    if (!tile_binary(A, t)) return;
    const int64_t sl = A.ls + t.a, sr = A.rs + (t.valid ? t.b : A.b0);
    const uint64_t *pl = A.store + set_base(sl, A.nw);
    const uint64_t *pr = A.store + set_base(sr, A.nw);
    const uint32_t k = (uint32_t)A.hamming_k;
    const bool anti = A.anti && t.valid;
    const int K = A.K, T = A.nw / K;
    uint32_t m0 = t.valid ? 0u : 0x7fffffffu, m1 = m0;  // invalid lanes never accept
    // (an operand level stored without words: in the range as a whole)
    if (in_lzw(A, A.ls) || in_lzw(A, A.rs)) {
        bool more = false;
        switch (K) {
        case 2: more = chkUm_unwritten<2>(A, sl, sr, anti, m0, m1); break;
        case 3: more = chkUm_unwritten<3>(A, sl, sr, anti, m0, m1); break;
        case 4: more = chkUm_unwritten<4>(A, sl, sr, anti, m0, m1); break;
        default: return;
        }
        if (!more) return;
    } else {
        for (int j = 0; j < T; ++j) {
            const int i = interleave(j, A.npw, T);
            uint32_t h0, h1;
            until_at0_trace(pl + (size_t)i * K * kTileWS, pr + (size_t)i * K * kTileWS, kTileWS, K, h0, h1);
            const uint32_t tg = (uint32_t)(A.aux[i * K].tgt & 1ull);
            m0 += h0 ^ tg;
            m1 += h1 ^ tg;
            // Mismatch and match counts only grow: stop once every lane of the
            // wave fails (with anti: the complements too).
            const uint32_t seen = (uint32_t)j + 1;
            if ((j & 3) == 3 && !__any(m0 <= k || m1 <= k || (anti && (seen - m0 <= k || seen - m1 <= k)))) return;
        }
    }
    if (!t.valid) return;
    const uint32_t seen = (uint32_t)T;
    const uint64_t id0 = A.id_base + 2 * ((uint64_t)t.a * (uint64_t)A.nR + (uint64_t)t.b);
    if (m0 <= k) atomicMin(&A.st->found, (unsigned long long)id0);
    else if (m1 <= k) atomicMin(&A.st->found, (unsigned long long)(id0 + 1));
    else if (anti && seen - m0 <= k) atomicMin(&A.st->neg_found[A.neg_slot], (unsigned long long)id0);
    else if (anti && seen - m1 <= k) atomicMin(&A.st->neg_found[A.neg_slot], (unsigned long long)(id0 + 1));
}

// (the stored keys carry the call's epoch)
__global__ void k_rehash(uint64_t *tab, uint32_t tmask, const uint64_t *keys, int64_t n, uint32_t e, int *probe_fail) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        tab_insert(tab, tmask, keys[i], tab_val((uint64_t)i, e), e, probe_fail);
}

__global__ void k_fill(uint64_t *p, uint64_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
// R: left rows per wave (tile_binary<R>); a.gsplit waves per task.
static dim3 grid_binary(const ChunkArgs &a, int R = 1) {
    const int64_t per = 4 / a.gsplit, tasks = (a.a1 - a.a0 + R - 1) / R * a.nbs;
    return dim3((unsigned)((tasks + per - 1) / per));
}
// Left rows per wave of the binary generation kernels (& and |; U keeps one
// row: its two carry chains per word and candidate pair are the larger
// cost); relaxed keys fold per-row key accumulators and keep one row.
template <int OP>
constexpr int gen_rows() { return OP == LTL_UNTIL ? 1 : 2; }
constexpr int kGenRows = 2;
static dim3 grid_tiles(const ChunkArgs &a) {
    const int per = 4 / a.gsplit;
    return dim3((unsigned)((a.n_tiles + per - 1) / per));
}

// Waves per task of a packed generation launch (exact keys, no successor
// tests): one while the chunk has enough tasks to fill the device; a chunk
// of few tasks folds each task's words with 2 or 4 waves, since one wave
// walking every word of its sets is then the whole launch (a fused
// ~/X/F/G pass over 81 sets of 341 words took 55 us).  tasks: the waves of
// the launch with one wave per task.
static int gen_split(int64_t tasks, int nw) {
    // (LTL_GEN_SPLIT = 1, 2 or 4 in the environment forces the choice: tests
    // run every mode against the oracle)
    if (const char *e = getenv("LTL_GEN_SPLIT")) {
        const int f = atoi(e);
        if (f == 1 || f == 2 || f == 4) return f;
    }
    if (nw >= 64 && tasks <= 2048) return 4;
    if (nw >= 32 && tasks <= 4096) return 2;
    return 1;
}
// LDS of split_add for n sums per lane.
static size_t split_lds(const ChunkArgs &a, int n) { return a.gsplit > 1 ? (size_t)(4 - 4 / a.gsplit) * n * 512 : 0; }
static int64_t binary_tasks(const ChunkArgs &a, int R) { return (int64_t)a.nbs * ((a.a1 - a.a0 + R - 1) / R); }

// The chunk kernels run on the packed layout (B = 1); K words per trace
// (B = 0) has kernels of its own.
// Successor checks (a.t_succ) run with exact keys on a tested level only.
// Exact keys without successor checks: split over a.gsplit waves per task.
#define LTL_DISPATCH_BX(KER, OP, B, RX, S)                                                      \
    if ((B) != 1) return hipErrorInvalidValue;                                                  \
    if (a.t_succ) {                                                                             \
        if (RX || a.gen_check) return hipErrorInvalidValue;                                     \
        hipLaunchKernelGGL((KER<OP, false, false, true>), grid_tiles(a), dim3(256), 0, S, a);   \
    } else {                                                                                    \
        ChunkArgs b = a;                                                                        \
        if (!RX) b.gsplit = gen_split(a.n_tiles, a.nw);                                         \
        if (a.gen_check)                                                                        \
            hipLaunchKernelGGL((KER<OP, RX, true, false>), grid_tiles(b), dim3(256), split_lds(b, 3), S, b); \
        else                                                                                    \
            hipLaunchKernelGGL((KER<OP, RX, false, false>), grid_tiles(b), dim3(256), split_lds(b, 3), S, b); \
    }

// The same for k_gen2 with R left rows per wave (R = 1 with relaxed keys).
#define LTL_DISPATCH_GEN2(OP, B, RX, S)                                                                  \
    if ((B) != 1) return hipErrorInvalidValue;                                                           \
    if (a.t_succ) {                                                                               \
        if (RX || a.gen_check) return hipErrorInvalidValue;                                              \
        hipLaunchKernelGGL((k_gen2<OP, false, false, true, gen_rows<OP>()>), grid_binary(a, gen_rows<OP>()), dim3(256), 0, S, a); \
    } else if (RX) {                                                                                     \
        if (a.gen_check) hipLaunchKernelGGL((k_gen2<OP, RX, true, false, 1>), grid_binary(a), dim3(256), 0, S, a); \
        else hipLaunchKernelGGL((k_gen2<OP, RX, false, false, 1>), grid_binary(a), dim3(256), 0, S, a);  \
    } else {                                                                                             \
        constexpr int R_ = gen_rows<OP>(), N_ = 3 * R_ * (OP == LTL_UNTIL ? 2 : 1);                       \
        ChunkArgs b = a;                                                                                 \
        b.gsplit = gen_split(binary_tasks(a, R_), a.nw);                                                 \
        if (a.gen_check)                                                                                 \
            hipLaunchKernelGGL((k_gen2<OP, false, true, false, R_>), grid_binary(b, R_), dim3(256), split_lds(b, N_), S, b); \
        else                                                                                             \
            hipLaunchKernelGGL((k_gen2<OP, false, false, false, R_>), grid_binary(b, R_), dim3(256), split_lds(b, N_), S, b); \
    }

#define LTL_DISPATCH_B(KER, OP, B, GRID, BLOCK, S, ...)                               \
    if ((B) != 1) return hipErrorInvalidValue;                                          \
    hipLaunchKernelGGL((KER<OP>), GRID, BLOCK, 0, S, __VA_ARGS__);

#define LTL_DISPATCH_K(KER, OP, K, GRID, S, ...)                                            \
    switch (K) {                                                                            \
    case 2: hipLaunchKernelGGL((KER<OP, 2>), GRID, dim3(256), 0, S, __VA_ARGS__); break;    \
    case 3: hipLaunchKernelGGL((KER<OP, 3>), GRID, dim3(256), 0, S, __VA_ARGS__); break;    \
    case 4: hipLaunchKernelGGL((KER<OP, 4>), GRID, dim3(256), 0, S, __VA_ARGS__); break;    \
    default: return hipErrorInvalidValue;                                                   \
    }
#define LTL_DISPATCH_KNX(KER, OP, K, GRID, BLOCK, S, ...)                                     \
    switch (K) {                                                                              \
    case 2: hipLaunchKernelGGL((KER<OP, 2, true>), GRID, BLOCK, 0, S, __VA_ARGS__); break;    \
    case 3: hipLaunchKernelGGL((KER<OP, 3, true>), GRID, BLOCK, 0, S, __VA_ARGS__); break;    \
    case 4: hipLaunchKernelGGL((KER<OP, 4, true>), GRID, BLOCK, 0, S, __VA_ARGS__); break;    \
    default: return hipErrorInvalidValue;                                                     \
    }
#define LTL_DISPATCH_KB(KER, OP, K, GRID, BLOCK, S, ...)                                      \
    switch (K) {                                                                              \
    case 2: hipLaunchKernelGGL((KER<OP, 2>), GRID, BLOCK, 0, S, __VA_ARGS__); break;          \
    case 3: hipLaunchKernelGGL((KER<OP, 3>), GRID, BLOCK, 0, S, __VA_ARGS__); break;          \
    case 4: hipLaunchKernelGGL((KER<OP, 4>), GRID, BLOCK, 0, S, __VA_ARGS__); break;          \
    default: return hipErrorInvalidValue;                                                     \
    }

// Multi-word traces (B = 0, a.K words per trace).
static hipError_t launch_gen_mw(int op, const ChunkArgs &a, hipStream_t s) {
    switch (op) {
    case LTL_NOT: LTL_DISPATCH_K(k_gen1m, LTL_NOT, a.K, grid_tiles(a), s, a); break;
    case LTL_NEXT: LTL_DISPATCH_K(k_gen1m, LTL_NEXT, a.K, grid_tiles(a), s, a); break;
    case LTL_FIN: LTL_DISPATCH_K(k_gen1m, LTL_FIN, a.K, grid_tiles(a), s, a); break;
    case LTL_GLOB: LTL_DISPATCH_K(k_gen1m, LTL_GLOB, a.K, grid_tiles(a), s, a); break;
    case LTL_AND:
        hipLaunchKernelGGL((k_gen2<LTL_AND, false, true, false, kGenRows>), grid_binary(a, kGenRows), dim3(256), 0, s, a);
        break;
    case LTL_OR:
        hipLaunchKernelGGL((k_gen2<LTL_OR, false, true, false, kGenRows>), grid_binary(a, kGenRows), dim3(256), 0, s, a);
        break;
    case LTL_UNTIL:
        switch (a.K) {
        case 2: hipLaunchKernelGGL(k_gen2m<2>, grid_binary(a), dim3(256), 0, s, a); break;
        case 3: hipLaunchKernelGGL(k_gen2m<3>, grid_binary(a), dim3(256), 0, s, a); break;
        case 4: hipLaunchKernelGGL(k_gen2m<4>, grid_binary(a), dim3(256), 0, s, a); break;
        default: return hipErrorInvalidValue;
        }
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <bool RX>
static hipError_t launch_gen_x(int op, int B, const ChunkArgs &a, hipStream_t s) {
    switch (op) {
    case LTL_NOT: LTL_DISPATCH_BX(k_gen1, LTL_NOT, B, RX, s); break;
    case LTL_NEXT: LTL_DISPATCH_BX(k_gen1, LTL_NEXT, B, RX, s); break;
    case LTL_FIN: LTL_DISPATCH_BX(k_gen1, LTL_FIN, B, RX, s); break;
    case LTL_GLOB: LTL_DISPATCH_BX(k_gen1, LTL_GLOB, B, RX, s); break;
    case LTL_AND: LTL_DISPATCH_GEN2(LTL_AND, B, RX, s); break;
    case LTL_OR: LTL_DISPATCH_GEN2(LTL_OR, B, RX, s); break;
    case LTL_UNTIL: LTL_DISPATCH_GEN2(LTL_UNTIL, B, RX, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int K>
static void launch_gen_xfgm(const ChunkArgs &a, int64_t n, int64_t slots, bool neg, uint64_t id_xfg, hipStream_t s) {
    const dim3 g = grid_tiles(a), b(256);
    if (neg) hipLaunchKernelGGL((k_gen_xfgm<K, false, true>), g, b, 0, s, a, n, slots, id_xfg);
    else if (a.gen_check) hipLaunchKernelGGL((k_gen_xfgm<K, true, false>), g, b, 0, s, a, n, slots, a.id_base);
    else hipLaunchKernelGGL((k_gen_xfgm<K, false, false>), g, b, 0, s, a, n, slots, a.id_base);
}

hipError_t launch_gen_xfg(const ChunkArgs &a, int64_t n, int64_t slots, bool neg, uint64_t id_xfg, hipStream_t s) {
    if (a.kmode || a.ws != kTileWS) return hipErrorInvalidValue;
    if (a.K > 1) {  // K words per trace
        if (a.t_succ || (neg && a.gen_check)) return hipErrorInvalidValue;
        switch (a.K) {
        case 2: launch_gen_xfgm<2>(a, n, slots, neg, id_xfg, s); break;
        case 3: launch_gen_xfgm<3>(a, n, slots, neg, id_xfg, s); break;
        case 4: launch_gen_xfgm<4>(a, n, slots, neg, id_xfg, s); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    const dim3 g = grid_tiles(a), b(256);
    // without successor tests: split over gsplit waves per tile
    ChunkArgs c = a;
    c.gsplit = a.t_succ ? 1 : gen_split(a.n_tiles, a.nw);
    const dim3 gc = grid_tiles(c);
    if (neg) {
        if (a.gen_check) return hipErrorInvalidValue;  // tested levels only
        if (a.t_succ) hipLaunchKernelGGL((k_gen_xfg<false, true, true>), g, b, 0, s, a, n, slots, id_xfg);
        else hipLaunchKernelGGL((k_gen_xfg<false, false, true>), gc, b, split_lds(c, 12), s, c, n, slots, id_xfg);
    } else if (a.t_succ) {
        if (a.gen_check) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_gen_xfg<false, true, false>), g, b, 0, s, a, n, slots, a.id_base);
    } else if (a.gen_check) {
        hipLaunchKernelGGL((k_gen_xfg<true, false, false>), gc, b, split_lds(c, 9), s, c, n, slots, a.id_base);
    } else {
        hipLaunchKernelGGL((k_gen_xfg<false, false, false>), gc, b, split_lds(c, 9), s, c, n, slots, a.id_base);
    }
    return hipGetLastError();
}

hipError_t launch_gen_fused(const ChunkArgs &a, const FusedBin &f, bool with_until, hipStream_t s) {
    if (a.kmode || a.ws != kTileWS || a.gen_check || a.t_acc) return hipErrorInvalidValue;
    if (a.K > 1) {
        if (a.t_succ || a.gsplit != 1) return hipErrorInvalidValue;
        const dim3 g = grid_binary(a, 1);
        switch (a.K * 2 + (with_until ? 1 : 0)) {
        case 4: hipLaunchKernelGGL((k_gen2fm<false, 2>), g, dim3(256), 0, s, a, f); break;
        case 5: hipLaunchKernelGGL((k_gen2fm<true, 2>), g, dim3(256), 0, s, a, f); break;
        case 6: hipLaunchKernelGGL((k_gen2fm<false, 3>), g, dim3(256), 0, s, a, f); break;
        case 7: hipLaunchKernelGGL((k_gen2fm<true, 3>), g, dim3(256), 0, s, a, f); break;
        case 8: hipLaunchKernelGGL((k_gen2fm<false, 4>), g, dim3(256), 0, s, a, f); break;
        case 9: hipLaunchKernelGGL((k_gen2fm<true, 4>), g, dim3(256), 0, s, a, f); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    ChunkArgs b = a;
    if (with_until) {
        constexpr int R = 1, N = 3 * 4 * R;
        if (a.t_succ) {
            hipLaunchKernelGGL((k_gen2fs<true, R>), grid_binary(a, R), dim3(256), 0, s, a, f);
        } else {
            b.gsplit = gen_split(binary_tasks(a, R), a.nw);
            hipLaunchKernelGGL((k_gen2f<true, R>), grid_binary(b, R), dim3(256), split_lds(b, N), s, b, f);
        }
    } else {
        constexpr int R = 2, N = 3 * 2 * R;
        if (a.t_succ) {
            hipLaunchKernelGGL((k_gen2fs<false, R>), grid_binary(a, R), dim3(256), 0, s, a, f);
        } else {
            b.gsplit = gen_split(binary_tasks(a, R), a.nw);
            hipLaunchKernelGGL((k_gen2f<false, R>), grid_binary(b, R), dim3(256), split_lds(b, N), s, b, f);
        }
    }
    return hipGetLastError();
}

hipError_t launch_insert_xfg(const ChunkArgs &a, int64_t n, int64_t slots, uint64_t id_xfg, const uint64_t *fp,
                             int64_t stride, hipStream_t s) {
    if (a.kmode || a.ws != kTileWS || a.t_succ || a.gen_check || !fp) return hipErrorInvalidValue;
    if ((int64_t)a.n_tiles * 64 > stride) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_insert_xfg, dim3((unsigned)(((int64_t)a.n_tiles * 64 + 255) / 256)), dim3(256), 0, s, a, n,
                       slots, id_xfg, fp, stride);
    return hipGetLastError();
}

hipError_t launch_gen(int op, int B, const ChunkArgs &a, hipStream_t s) {
    if (a.ws != kTileWS) return hipErrorInvalidValue;
    // relaxed keys: generation tests every candidate into t_acc
    if (a.kmode && (!a.t_acc || !a.gen_check)) return hipErrorInvalidValue;
    if (B == 0) return a.kmode || a.t_succ ? hipErrorInvalidValue : launch_gen_mw(op, a, s);
    return a.kmode ? launch_gen_x<true>(op, B, a, s) : launch_gen_x<false>(op, B, a, s);
}

hipError_t launch_emit(int op, int B, const ChunkArgs &a, hipStream_t s) {
    if (a.ws != kTileWS || (B != 1 && a.t_succ)) return hipErrorInvalidValue;
    // one lane per survivor; the survivor count is on the device, so the
    // grid covers every candidate of the chunk and idle lanes exit (the packed
    // emit: a.esplit waves per store tile)
    const int nd = op == LTL_UNTIL ? 2 : 1;
    const int64_t pos_tiles = ((int64_t)a.n_tiles * 64 * nd + 63 + 63) / 64;
    int K = B == 1 && !a.deferred ? a.esplit : 1;
    if (K < 1 || (K > 1 && (!a.gfirst || !a.glast || K > a.nsw))) return hipErrorInvalidValue;
    // A chunk of few store tiles: up to 16 word ranges per tile (ranges of at
    // least 16 words), since each wave's walk over its range sets the time.
    if (K > 1 && pos_tiles <= 256) K = std::max(K, std::min({16, a.nsw, a.nw / 16}));
    const bool nx = B == 1 && !a.deferred && a.nx_fp;
    if (nx) K = std::min(K, 8);  // (one workgroup of K waves per store tile)
    if (B == 0 && !a.deferred) {
        // one workgroup of E waves per store tile, each over whole summary
        // groups of the traces (E <= 8 for the LDS sums of NX)
        ChunkArgs m = a;
        m.esplit = std::max(1, std::min(8, a.nsw));
        const dim3 gm((unsigned)pos_tiles), bm((unsigned)(64 * m.esplit));
        if (a.nx_fp) {
            switch (op) {
            case LTL_NOT: LTL_DISPATCH_KNX(k_emitm, LTL_NOT, a.K, gm, bm, s, m); break;
            case LTL_NEXT: LTL_DISPATCH_KNX(k_emitm, LTL_NEXT, a.K, gm, bm, s, m); break;
            case LTL_FIN: LTL_DISPATCH_KNX(k_emitm, LTL_FIN, a.K, gm, bm, s, m); break;
            case LTL_GLOB: LTL_DISPATCH_KNX(k_emitm, LTL_GLOB, a.K, gm, bm, s, m); break;
            case LTL_AND: LTL_DISPATCH_KNX(k_emitm, LTL_AND, a.K, gm, bm, s, m); break;
            case LTL_OR: LTL_DISPATCH_KNX(k_emitm, LTL_OR, a.K, gm, bm, s, m); break;
            case LTL_UNTIL: LTL_DISPATCH_KNX(k_emitm, LTL_UNTIL, a.K, gm, bm, s, m); break;
            default: return hipErrorInvalidValue;
            }
        } else {
            switch (op) {
            case LTL_NOT: LTL_DISPATCH_KB(k_emitm, LTL_NOT, a.K, gm, bm, s, m); break;
            case LTL_NEXT: LTL_DISPATCH_KB(k_emitm, LTL_NEXT, a.K, gm, bm, s, m); break;
            case LTL_FIN: LTL_DISPATCH_KB(k_emitm, LTL_FIN, a.K, gm, bm, s, m); break;
            case LTL_GLOB: LTL_DISPATCH_KB(k_emitm, LTL_GLOB, a.K, gm, bm, s, m); break;
            case LTL_AND: LTL_DISPATCH_KB(k_emitm, LTL_AND, a.K, gm, bm, s, m); break;
            case LTL_OR: LTL_DISPATCH_KB(k_emitm, LTL_OR, a.K, gm, bm, s, m); break;
            case LTL_UNTIL: LTL_DISPATCH_KB(k_emitm, LTL_UNTIL, a.K, gm, bm, s, m); break;
            default: return hipErrorInvalidValue;
            }
        }
        return hipGetLastError();
    }
    ChunkArgs e = a;
    e.esplit = K;
    const dim3 g((unsigned)((pos_tiles * K + 3) / 4)), b(256);
    if (nx) {
        const dim3 gx((unsigned)pos_tiles), bx((unsigned)(64 * K));
        switch (op) {
        case LTL_NOT: hipLaunchKernelGGL((k_emit<LTL_NOT, true>), gx, bx, 0, s, e, e.aux, e.saux); break;
        case LTL_NEXT: hipLaunchKernelGGL((k_emit<LTL_NEXT, true>), gx, bx, 0, s, e, e.aux, e.saux); break;
        case LTL_FIN: hipLaunchKernelGGL((k_emit<LTL_FIN, true>), gx, bx, 0, s, e, e.aux, e.saux); break;
        case LTL_GLOB: hipLaunchKernelGGL((k_emit<LTL_GLOB, true>), gx, bx, 0, s, e, e.aux, e.saux); break;
        case LTL_AND: hipLaunchKernelGGL((k_emit<LTL_AND, true>), gx, bx, 0, s, e, e.aux, e.saux); break;
        case LTL_OR: hipLaunchKernelGGL((k_emit<LTL_OR, true>), gx, bx, 0, s, e, e.aux, e.saux); break;
        case LTL_UNTIL: hipLaunchKernelGGL((k_emit<LTL_UNTIL, true>), gx, bx, 0, s, e, e.aux, e.saux); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    if (a.deferred) {
        if (op < LTL_NOT || op > LTL_UNTIL) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_emit_meta, g, b, 0, s, a, op);
        return hipGetLastError();
    }
    switch (op) {
    case LTL_NOT: LTL_DISPATCH_B(k_emit, LTL_NOT, B, g, b, s, e, e.aux, e.saux); break;
    case LTL_NEXT: LTL_DISPATCH_B(k_emit, LTL_NEXT, B, g, b, s, e, e.aux, e.saux); break;
    case LTL_FIN: LTL_DISPATCH_B(k_emit, LTL_FIN, B, g, b, s, e, e.aux, e.saux); break;
    case LTL_GLOB: LTL_DISPATCH_B(k_emit, LTL_GLOB, B, g, b, s, e, e.aux, e.saux); break;
    case LTL_AND: LTL_DISPATCH_B(k_emit, LTL_AND, B, g, b, s, e, e.aux, e.saux); break;
    case LTL_OR: LTL_DISPATCH_B(k_emit, LTL_OR, B, g, b, s, e, e.aux, e.saux); break;
    case LTL_UNTIL: LTL_DISPATCH_B(k_emit, LTL_UNTIL, B, g, b, s, e, e.aux, e.saux); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_materialize(const ChunkArgs &a, int64_t s0, int64_t s1, hipStream_t s, bool sums) {
    if (a.ws != kTileWS || s0 < 0 || s1 > a.capacity) return hipErrorInvalidValue;
    if (s1 <= s0) return hipSuccess;
    const int64_t n = s1 - (s0 & ~63ll);
    const dim3 g((unsigned)((n + 255) / 256)), b(256);
    switch (a.K * 2 + (sums ? 1 : 0)) {
    case 2: hipLaunchKernelGGL(k_materialize<false>, g, b, 0, s, a, s0, s1); break;
    case 3: hipLaunchKernelGGL(k_materialize<true>, g, b, 0, s, a, s0, s1); break;
    case 4: hipLaunchKernelGGL((k_materialize_m<2, false>), g, b, 0, s, a, s0, s1); break;
    case 5: hipLaunchKernelGGL((k_materialize_m<2, true>), g, b, 0, s, a, s0, s1); break;
    case 6: hipLaunchKernelGGL((k_materialize_m<3, false>), g, b, 0, s, a, s0, s1); break;
    case 7: hipLaunchKernelGGL((k_materialize_m<3, true>), g, b, 0, s, a, s0, s1); break;
    case 8: hipLaunchKernelGGL((k_materialize_m<4, false>), g, b, 0, s, a, s0, s1); break;
    case 9: hipLaunchKernelGGL((k_materialize_m<4, true>), g, b, 0, s, a, s0, s1); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_compact(int op, const ChunkArgs &a, hipStream_t s) {
    const dim3 g = grid_tiles(a), b(256);
    switch (op) {
    case LTL_NOT: case LTL_NEXT: case LTL_FIN: case LTL_GLOB:
        hipLaunchKernelGGL(k_compact<LTL_NOT>, g, b, 0, s, a); break;
    case LTL_AND: case LTL_OR:
        hipLaunchKernelGGL(k_compact<LTL_AND>, g, b, 0, s, a); break;
    case LTL_UNTIL: hipLaunchKernelGGL(k_compact<LTL_UNTIL>, g, b, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_decide(int op, const ChunkArgs &a, hipStream_t s) {
    const dim3 g = grid_tiles(a), b(256);
    switch (op) {
    case LTL_NOT: hipLaunchKernelGGL(k_decide<LTL_NOT>, g, b, 0, s, a); break;
    case LTL_NEXT: hipLaunchKernelGGL(k_decide<LTL_NEXT>, g, b, 0, s, a); break;
    case LTL_FIN: hipLaunchKernelGGL(k_decide<LTL_FIN>, g, b, 0, s, a); break;
    case LTL_GLOB: hipLaunchKernelGGL(k_decide<LTL_GLOB>, g, b, 0, s, a); break;
    case LTL_AND: hipLaunchKernelGGL(k_decide<LTL_AND>, g, b, 0, s, a); break;
    case LTL_OR: hipLaunchKernelGGL(k_decide<LTL_OR>, g, b, 0, s, a); break;
    case LTL_UNTIL: hipLaunchKernelGGL(k_decide<LTL_UNTIL>, g, b, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_dedup_small(int op, const ChunkArgs &a, hipStream_t s) {
    if (a.n_tiles < 1 || a.n_tiles > kSmallTiles) return hipErrorInvalidValue;
    const dim3 g(1), b(1024);
    switch (op) {
    case LTL_NOT: hipLaunchKernelGGL(k_dedup_small<LTL_NOT>, g, b, 0, s, a); break;
    case LTL_NEXT: hipLaunchKernelGGL(k_dedup_small<LTL_NEXT>, g, b, 0, s, a); break;
    case LTL_FIN: hipLaunchKernelGGL(k_dedup_small<LTL_FIN>, g, b, 0, s, a); break;
    case LTL_GLOB: hipLaunchKernelGGL(k_dedup_small<LTL_GLOB>, g, b, 0, s, a); break;
    case LTL_AND: hipLaunchKernelGGL(k_dedup_small<LTL_AND>, g, b, 0, s, a); break;
    case LTL_OR: hipLaunchKernelGGL(k_dedup_small<LTL_OR>, g, b, 0, s, a); break;
    case LTL_UNTIL: hipLaunchKernelGGL(k_dedup_small<LTL_UNTIL>, g, b, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_dedup_lb(int op, const ChunkArgs &a, hipStream_t s) {
    if (a.n_tiles < 1 || !a.lb_flags || !a.lb_ticket || a.lb_epoch == 0 || a.lb_epoch >= (1u << 30))
        return hipErrorInvalidValue;
    const dim3 g = grid_tiles(a), b(256);
    switch (op) {
    case LTL_NOT: hipLaunchKernelGGL(k_dedup_lb<LTL_NOT>, g, b, 0, s, a); break;
    case LTL_NEXT: hipLaunchKernelGGL(k_dedup_lb<LTL_NEXT>, g, b, 0, s, a); break;
    case LTL_FIN: hipLaunchKernelGGL(k_dedup_lb<LTL_FIN>, g, b, 0, s, a); break;
    case LTL_GLOB: hipLaunchKernelGGL(k_dedup_lb<LTL_GLOB>, g, b, 0, s, a); break;
    case LTL_AND: hipLaunchKernelGGL(k_dedup_lb<LTL_AND>, g, b, 0, s, a); break;
    case LTL_OR: hipLaunchKernelGGL(k_dedup_lb<LTL_OR>, g, b, 0, s, a); break;
    case LTL_UNTIL: hipLaunchKernelGGL(k_dedup_lb<LTL_UNTIL>, g, b, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_rx_tail(int op, const ChunkArgs &a, hipStream_t s) {
    if (!a.t_acc || a.n_tiles < 1) return hipErrorInvalidValue;
    const int nd = op == LTL_UNTIL ? 2 : 1;
    const dim3 g((unsigned)std::min<int64_t>(((int64_t)a.n_tiles * 64 * nd + 255) / 256, 1024)), b(256);
    switch (op) {
    case LTL_NOT: case LTL_NEXT: case LTL_FIN: case LTL_GLOB:
        hipLaunchKernelGGL(k_rx_tail<LTL_NOT>, g, b, 0, s, a); break;
    case LTL_AND: case LTL_OR: hipLaunchKernelGGL(k_rx_tail<LTL_AND>, g, b, 0, s, a); break;
    case LTL_UNTIL: hipLaunchKernelGGL(k_rx_tail<LTL_UNTIL>, g, b, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_finish(const ChunkArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(1), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_atoms(int B, const ChunkArgs &a, int n_atoms, hipStream_t s) {
    if (n_atoms <= 0) return hipSuccess;
    if (a.ws != kTileWS) return hipErrorInvalidValue;
    (void)B;  // one kernel for both layouts: the fingerprints are per word
    hipLaunchKernelGGL(k_atoms, dim3((unsigned)n_atoms), dim3(64), 0, s, a, n_atoms);
    return hipGetLastError();
}

// One thread walks the formula from the store's metadata (a formula has a
// few dozen nodes at most; the walk replaces a host copy per node).  The
// nodes written so far are the walk's queue: the next node to expand is the
// first one at or after qi that is not an atom.
__global__ void k_tree(const uint8_t *op, const int32_t *left, const int32_t *right, int64_t n_stored, int32_t l,
                       int32_t r, int32_t *out, int32_t cap) {
    if (threadIdx.x || blockIdx.x) return;
    int32_t n = 0, qi = -1;
    for (;;) {
        int32_t kids[2];
        if (qi < 0) {
            kids[0] = l;
            kids[1] = r;
            qi = 0;
        } else {
            while (qi < n && out[1 + 3 * qi] == LTL_ATOM) ++qi;
            if (qi >= n) break;
            kids[0] = out[2 + 3 * qi];
            kids[1] = out[3 + 3 * qi];
            ++qi;
        }
        for (int s = 0; s < 2; ++s) {
            const int32_t k = kids[s];
            if (k < 0) continue;
            if (k >= n_stored) { out[0] = -2; return; }
            if (n >= cap) { out[0] = -1; return; }
            out[1 + 3 * n] = op[k];
            out[2 + 3 * n] = left[k];
            out[3 + 3 * n] = right[k];
            ++n;
        }
    }
    out[0] = n;
}

hipError_t launch_tree(const uint8_t *op, const int32_t *left, const int32_t *right, int64_t n_stored, int32_t l,
                       int32_t r, int32_t *out, int32_t cap, hipStream_t s) {
    hipLaunchKernelGGL(k_tree, dim3(1), dim3(64), 0, s, op, left, right, n_stored, l, r, out, cap);
    return hipGetLastError();
}

hipError_t launch_scatter(uint64_t *dst, int64_t ws, const uint64_t *src, int ncols, int64_t nrows, hipStream_t s) {
    const int64_t n = nrows * ncols;
    if (n <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_scatter, dim3(blocks), dim3(256), 0, s, dst, ws, src, ncols, n);
    return hipGetLastError();
}

hipError_t launch_check(int op, int B, const ChunkArgs &a, hipStream_t s) {
    if (a.ws != kTileWS) return hipErrorInvalidValue;
    const dim3 b(256);
    switch (op) {
    case LTL_NOT: hipLaunchKernelGGL(k_chk1<LTL_NOT>, grid_tiles(a), b, 0, s, a); break;
    case LTL_NEXT: hipLaunchKernelGGL(k_chk1<LTL_NEXT>, grid_tiles(a), b, 0, s, a); break;
    case LTL_FIN: hipLaunchKernelGGL(k_chk1<LTL_FIN>, grid_tiles(a), b, 0, s, a); break;
    case LTL_GLOB: hipLaunchKernelGGL(k_chk1<LTL_GLOB>, grid_tiles(a), b, 0, s, a); break;
    case LTL_AND: hipLaunchKernelGGL(k_chk2<LTL_AND>, grid_binary(a), b, 0, s, a); break;
    case LTL_OR: hipLaunchKernelGGL(k_chk2<LTL_OR>, grid_binary(a), b, 0, s, a); break;
    case LTL_UNTIL:
        switch (B) {
        case 0: hipLaunchKernelGGL(k_chkUm, grid_binary(a), b, 0, s, a); break;
        case 1: hipLaunchKernelGGL(k_chkU, grid_binary(a), b, 0, s, a); break;
        default: return hipErrorInvalidValue;
        }
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_rehash(uint64_t *tab, uint32_t tmask, const uint64_t *keys, int64_t n, uint32_t epoch,
                         int *probe_fail, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_rehash, dim3(blocks), dim3(256), 0, s, tab, tmask, keys, n, epoch, probe_fail);
    return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t *p, uint64_t v, size_t n, hipStream_t s) {
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill, dim3(blocks ? blocks : 1), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

size_t scan_temp_bytes(int n_max) {
    size_t bytes = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr, n_max);
    return bytes;
}

hipError_t launch_scan(const uint32_t *in, uint32_t *out, int n, void *temp, size_t temp_bytes, hipStream_t s) {
    return hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, in, out, n, s);
}

}  // namespace ltl
