// device_ops.hpp -- per-word operators, fingerprints, summaries and the
// dedup-table insert shared by the chunk kernels and the batched kernel.
//
// Operators follow code/ltli6463.cu:203, 295, 385, 468, 551-556, 638-645,
// 739-771, restated for the two layouts: traces packed whole into words (a
// mask per word keeps every shift and carry inside its trace; bits outside
// the traces stay zero), and K words per trace for traces longer than 64
// positions.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/ltl_amd.h"
#include "kernels.hpp"

namespace ltl {


// Suffix OR of a 64-bit word: bit p set iff some bit at or above p is.
__device__ __forceinline__ uint64_t suffix_or64(uint64_t x) {
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) x |= x >> s;
    return x;
}

// Strong until, y_p = r_p | (l_p & y_{p+1}), as one carry chain: with the
// bits reversed (position p -> bit 63 - p) the recurrence runs upward,
// Y_q = R_q | (L_q & Y_{q-1}), which is exactly the carry out of bit q of
// a + R + cin with a = L | R (generate R_q, propagate L_q); cin is the result
// just above the word (multi-word traces).  The carry out of bit q of a sum
// a + b is (a_q & b_q) | ((a_q | b_q) & ~s_q); as R is a subset of a, that is
// R_q where s_q is set and a_q where it is not -- one bit select, no shift
// and no carry out of the top bit (its own bit gives it).
__device__ __forceinline__ uint64_t until_carry64(uint64_t l, uint64_t r, uint64_t cin) {
    const uint64_t L = __builtin_bitreverse64(l), R = __builtin_bitreverse64(r), a = L | R;
    const uint64_t s = a + R + cin;
    return __builtin_bitreverse64((s & R) | (~s & a));
}

// Fingerprints are multilinear hashes (Carter and Wegman's multilinear
// family, as in MMH, Halevi and Krawczyk, FSE 1997, without its prime
// reduction): the sum mod 2^64 of every 32-bit half of every word times its
// own 32-bit key, each product exact in 64 bits.  Two different sets differ
// in some half, by d with 0 < |d| < 2^32; for the other keys fixed, two
// values k != k' of that half's key give products that differ by
// (k - k') d, nonzero and below 2^64 in magnitude, so at most one key value
// makes the sums equal.  The host draws odd keys (2^31 values per half):
// Pr[collision] <= 2^-31 per fingerprint, independent keys for the two, and
// sets that differ in one half only never collide.  One multiply-add per
// half (v_mad_u64_u32).
__device__ __forceinline__ void ml_fold(uint64_t &f, uint64_t c, uint64_t k) {
    f += (uint64_t)(uint32_t)c * (uint64_t)(uint32_t)k;
    f += (uint64_t)(uint32_t)(c >> 32) * (uint64_t)(uint32_t)(k >> 32);
}

struct Acc {
    uint64_t f1 = 0, f2 = 0;
    uint32_t mism = 0;
};

// Fold word c of a candidate: both fingerprints and (CHK) the mismatch
// count.  Fields are read one by one from the (wave-uniform) per-word
// record, so that only what is used occupies scalar registers.
template <bool CHK = true>
__device__ __forceinline__ void fold(Acc &a, uint64_t c, const WordAux *x) {
    ml_fold(a.f1, c, x->k1);
    ml_fold(a.f2, c, x->k2);
    if constexpr (CHK) a.mism += (uint32_t)__popcll((c ^ x->tgt) & x->chk);
}

// The reference's relaxed uniqueness keys (code/ltli6463.cu:57-141), used
// as dedup keys in modes 1-3 when the traces hold more than 126 positions:
// key bit j of the reference is bit j of l (j < 63) or bit j - 63 of h.
// Mode 1 takes every stride-th position of each trace, mode 2 the first
// (126 len_i + lenSum) / lenSum positions, both until 126 bits are taken;
// mode 3 XORs a 64-bit mixer of each trace into h (first half of the
// traces) or l.  Folded word by word like the fingerprints.
struct KeyAcc {
    uint64_t h = 0, l = 0;
};

__device__ __forceinline__ void key_put(KeyAcc &k, uint64_t v, int j0, int cnt) {
    constexpr uint64_t M63 = (1ull << 63) - 1ull;
    if (j0 < 63) {
        k.l |= (v << j0) & M63;
        if (j0 + cnt > 63) k.h |= v >> (63 - j0);
    } else {
        k.h |= v << (j0 - 63);
    }
}

// Fingerprints of a relaxed key: equal keys give equal fingerprints; two
// different keys share fingerprint 1 with probability 2^-64 (detected as a
// collision through fingerprint 2).
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ void key_finish(Acc &a, const KeyAcc &k, uint64_t seed = 0) {
    a.f1 = mix64(k.h ^ mix64(k.l + (0x9e3779b97f4a7c15ull ^ seed)));
    a.f2 = mix64(k.l ^ mix64(k.h + (0xd1b54a32d192ed03ull ^ seed)));
}

// Flush the summary words of a set when trace i (of n) closes summary word
// i / 64, and clear them.
__device__ __forceinline__ void flush_summary(uint64_t *z, int i, int n, int nsw, int64_t ws, uint64_t (&v)[4]) {
    if ((i % 64) != 63 && i != n - 1) return;
    const int j = i / 64;
#pragma unroll
    for (int k = 0; k < 4; ++k) __builtin_nontemporal_store(v[k], z + (size_t)(k * nsw + j) * ws);
    v[0] = v[1] = v[2] = v[3] = 0;
}

// Table key from fingerprint 1: a 64-bit finalizer spreads the hash sum over
// all bits (slot index = low bits); the empty key is never produced.
__device__ __forceinline__ uint64_t table_key(uint64_t f) {
    f ^= f >> 33; f *= 0xff51afd7ed558ccdull;
    f ^= f >> 33; f *= 0xc4ceb9fe1a85ec53ull;
    f ^= f >> 33;
    return f == kEmpty ? kEmpty - 2 : f;
}

// Table epochs: a table of fixed size is cleared once and then serves up
// to kEpochs calls.  A call's epoch e (1..kEpochs) is the top 4 bits of
// every key it inserts, so a key of another epoch (or kEmpty, whose top bits
// are 15) marks a free slot.  Values carry 15 - e in their top 4 bits: a
// value of the current epoch is smaller than any earlier epoch's, so the
// first insert's atomicMin replaces a stale value; below them, kPending
// (bit 59) marks a candidate id, else the value is a store index.
constexpr uint64_t kKeyBody = (1ull << 60) - 1;
__device__ __forceinline__ uint64_t tab_key(uint64_t f1, uint32_t e) {
    return (table_key(f1) & kKeyBody) | ((uint64_t)e << 60);
}
__device__ __forceinline__ uint64_t tab_val(uint64_t v, uint32_t e) { return ((uint64_t)(15u - e) << 60) | v; }
// a store index of epoch e (not a candidate id)
__device__ __forceinline__ bool tab_committed(uint64_t v, uint32_t e) { return (v & ~(kPending - 1)) == tab_val(0, e); }
__device__ __forceinline__ uint64_t tab_payload(uint64_t v) { return v & (kPending - 1); }

// A candidate insert's result with this bit set: the key was found holding a
// stored set of the epoch (a committed value, which is final: values only
// decrease), whose store index the low 31 bits give; the candidate is a
// duplicate of it (the table's slot indices and the store's indices stay
// below 2^31).
constexpr uint32_t kSlotStored = 1u << 31;

// Linear probing; key CAS over a free slot (another epoch's key), then
// atomicMin of the value.  key and val carry the epoch e (tab_key, tab_val).
// STORED: a key found with a committed value returns kSlotStored | its store
// index instead of the slot.
template <bool STORED = false>
__device__ inline uint32_t tab_insert(uint64_t *tab, uint32_t tmask, uint64_t key, uint64_t val, uint32_t e,
                                      int *probe_fail) {
    uint32_t p = (uint32_t)key & tmask;
    for (uint32_t n = 0; n <= tmask; ++n) {
        unsigned long long *kp = reinterpret_cast<unsigned long long *>(tab + 2 * (size_t)p);
        unsigned long long k = __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // (within an epoch a slot only goes from free to a key of the epoch:
        // a failed CAS has seen such a key, so this runs at most twice)
        while ((uint32_t)(k >> 60) != e) {
            const unsigned long long seen = atomicCAS(kp, k, (unsigned long long)key);
            k = seen == k ? (unsigned long long)key : seen;
        }
        if (k == key) {
            // Values only decrease; a read that is already <= val makes the
            // atomic a no-op.  Popular sets (the empty set under G, the full
            // set under F, ...) are committed early, so their many duplicates
            // become cached loads instead of serialized atomics on one slot.
            unsigned long long cur = __hip_atomic_load(kp + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (STORED && tab_committed(cur, e)) return kSlotStored | (uint32_t)tab_payload(cur);
            if ((unsigned long long)val < cur) atomicMin(kp + 1, (unsigned long long)val);
            return p;
        }
        p = (p + 1) & tmask;
    }
    *probe_fail = 1;
    return 0;
}

// ---------------------------------------------------------------------------
// Packed layout (B = 1): whole traces packed into words, position p of a
// trace of length L at bit o + L - 1 - p.  Per word, lm covers the traces'
// bits and inner every bit but each trace's lowest (its last position).
// The operators become masks and one carry chain (restated from the
// per-position definitions; code/ltli6463.cu:203, 468, 551-556, 638-645,
// 739-771 for one trace per slot):
//   X f = (f << 1) & inner               (the next position is the next lower bit)
//   l U r: Y_b = r_b | (l_b & inner_b & Y_{b-1}), the carry out of bit b of
//          ((l & inner) | r) + r -- a trace's lowest bit has inner = 0, so no
//          carry crosses from one trace into the next
//   F f = true U f,  G f = ~F ~f
// ---------------------------------------------------------------------------
// Y_b = g_b | (p_b & Y_{b-1}) is the carry out of bit b of s = a + g with
// a = p | g: (a_b & g_b) | ((a_b | g_b) & ~s_b), which, g being a subset of
// a, selects g_b where s_b is set and a_b where it is not (v_bfi_b32 per
// half; with the add and the or, 6 instructions per word).
__device__ __forceinline__ uint64_t pk_chain(uint64_t p, uint64_t g) {
    const uint64_t a = p | g, s = a + g;
    return (s & g) | (~s & a);
}

__device__ __forceinline__ uint64_t pk_until(uint64_t l, uint64_t r, uint64_t inner) {
    return pk_chain(l & inner, r);
}

template <int OP>
__device__ __forceinline__ uint64_t pk_unary(uint64_t x, const WordAux *a) {
    if constexpr (OP == LTL_NOT) return ~x & a->lm;
    else if constexpr (OP == LTL_NEXT) return (x << 1) & a->inner;
    else if constexpr (OP == LTL_FIN) return pk_chain(a->inner, x);
    else return ~pk_chain(a->inner, ~x & a->lm) & a->lm;
}

// Summary bits of the traces of one packed word: positions 0 (b0) and 1
// (b1), any position (nz = F at 0) and every position (fl = G at 0), taken
// at the traces' position-0 bits (chk, wave-uniform) from the word's top bit
// down and shifted into the four accumulators (acc = acc << 1 | bit).  Trace
// t in packing order belongs to group t / 64 (summary word t / 64); after
// the group's last trace, flush(t / 64) stores and clears the accumulators,
// so the i-th of the n traces of a group ends at bit n - 1 - i.  VMASK
// selects the vectors (bit v for vector v; 0: only the flushes).  Bit
// positions are wave-uniform: the loops run on the scalar unit, one per
// 32-bit half, and each bit costs one extract and one shift-or per vector.
// The traces of a half are taken in runs that end at a group's end, so the
// end-of-group test runs once per run, not once per trace.  [g_lo, g_hi):
// only the traces of these groups are taken (a word range of a split emit).
template <int VMASK, class Flush>
__device__ __forceinline__ void summarize_pk(uint64_t c, const WordAux *x, uint64_t (&acc)[4], int W,
                                             Flush &&flush, int g_lo = 0, int g_hi = 0x7fffffff) {
    const uint64_t inner = x->inner, lm = x->lm;
    uint64_t u[4] = {c, 0, 0, 0};
    if constexpr ((VMASK & 2) != 0) u[1] = (c << 1) & inner;
    if constexpr ((VMASK & 4) != 0) u[2] = pk_chain(inner, c);
    if constexpr ((VMASK & 8) != 0) u[3] = ~pk_chain(inner, ~c & lm) & lm;
    uint64_t chk = x->chk;
    const uint64_t si = x->sinfo;
    int g = (int)sinfo_word(si);                       // group of the word's first trace
    int t = g * 64 + sinfo_bit(si);                    // its packed index
    // Groups below g_lo belong to an earlier word range: their traces, the
    // word's first ones, are dropped from the top of chk.
    if (g < g_lo) {
        for (int k = g_lo * 64 - t, n = __popcll(chk); k > 0 && n > 0; --k, --n)
            chk &= ~(1ull << (63 - __builtin_clzll(chk)));
        if (!chk) return;
        g = g_lo;
        t = g_lo * 64;
    }
    if (g >= g_hi) return;
    int left = (W - 1 < (t | 63) ? W - 1 : (t | 63)) - t + 1;  // traces until the group is complete
    if constexpr (VMASK == 0) {  // only the flushes
        for (int n = __popcll(chk); left > 0 && n >= left;) {
            n -= left;
            flush(g);
            if (++g >= g_hi) return;
            left = W - 64 * g < 64 ? W - 64 * g : 64;
        }
        return;
    }
#pragma unroll
    for (int h = 1; h >= 0; --h) {
        uint32_t m = (uint32_t)(chk >> (32 * h));
        uint32_t uh[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) uh[j] = (uint32_t)(u[j] >> (32 * h));
        while (m && left > 0) {  // left <= 0: every trace taken
            const int pc = __popc(m), n = pc < left ? pc : left;
            for (int k = 0; k < n; ++k) {
                const int b = 31 - __builtin_clz(m);
                m ^= 1u << b;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if ((VMASK >> j) & 1) acc[j] = (acc[j] << 1) | __builtin_amdgcn_ubfe(uh[j], (uint32_t)b, 1u);
            }
            left -= n;
            if (left == 0) {
                flush(g);
                if (++g >= g_hi) return;
                left = W - 64 * g < 64 ? W - 64 * g : 64;
            }
        }
    }
}

// Relaxed keys over the traces of one packed word (each trace back in the
// reference's bit order, bit k = position k).
__device__ __forceinline__ void key_fold_pk(KeyAcc &k, uint64_t c, int w, const WordAux *x, const ChunkArgs &A) {
    if (A.kmode != 3 && w >= A.kw_end) return;
    const uint64_t si = x->sinfo;
    const int t0 = (int)sinfo_word(si) * 64 + sinfo_bit(si), n = sinfo_ntr(si);
    for (int q = 0; q < n; ++q) {
        const uint64_t rec = A.ptrace[t0 + q];
        const int i = (int)(uint32_t)rec, off = (int)((rec >> 32) & 63), len = (int)((rec >> 40) & 127);
        const uint64_t seg = (c >> off) & (len >= 64 ? ~0ull : ((1ull << len) - 1ull));
        uint64_t v = __builtin_bitreverse64(seg) >> (64 - len);
        if (A.kmode == 3) {
            v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ull;
            v = (v ^ (v >> 27)) * 0x94d049bb133111ebull;
            v ^= v >> 31;
            if (i < A.W / 2) k.h ^= v; else k.l ^= v;
            continue;
        }
        const uint32_t info = A.kinfo[i];
        const int cnt = (int)(info & 0xffu), j0 = (int)(info >> 8);
        if (cnt == 0) continue;
        if (A.kmode == 1) {
            for (int m = 0; m < cnt; ++m)
                if ((v >> (m * A.kstride)) & 1ull) key_put(k, 1ull, j0 + m, 1);
        } else {
            key_put(k, cnt >= 64 ? v : (v & ((1ull << cnt) - 1ull)), j0, cnt);
        }
    }
}

// ---------------------------------------------------------------------------
// Traces longer than 64 positions: K words per trace, word k (positions
// 64k..64k+63) of trace i is word i*K + k of a set.  The 64-bit operators
// apply per word; what crosses a word boundary comes from the next higher
// word, so a trace's words are combined from the top down: X moves bit 0 of
// word k+1 into bit 63 of word k; F holds on all of word k once a higher word
// has a set bit; G holds in word k only while every higher word is full; U
// holds on the top run of l's ones in word k when it holds at position 0 of
// word k+1.  Restated from the per-position definitions
// (code/ltli6463.cu:468, 551-556, 638-645, 739-771 for one word).
// ---------------------------------------------------------------------------
template <int OP, int K>
__device__ __forceinline__ void trace_unary(const uint64_t (&x)[K], const WordAux *aux, uint64_t (&y)[K]) {
    if constexpr (OP == LTL_NOT) {
#pragma unroll
        for (int k = 0; k < K; ++k) y[k] = ~x[k] & aux[k].lm;
    } else if constexpr (OP == LTL_NEXT) {
#pragma unroll
        for (int k = 0; k < K; ++k) y[k] = (x[k] >> 1) | (k + 1 < K ? x[k + 1] << 63 : 0ull);
    } else if constexpr (OP == LTL_FIN) {
        uint64_t any = 0;  // all ones once a higher word has a set bit
#pragma unroll
        for (int k = K - 1; k >= 0; --k) {
            y[k] = suffix_or64(x[k]) | (any & aux[k].lm);
            any |= 0ull - (uint64_t)(x[k] != 0);
        }
    } else {
        uint64_t all = ~0ull;  // all ones while every higher word is full
#pragma unroll
        for (int k = K - 1; k >= 0; --k) {
            const uint64_t lm = aux[k].lm;
            y[k] = all & x[k] & ~suffix_or64(~x[k] & lm);
            all &= 0ull - (uint64_t)(x[k] == lm);
        }
    }
}

template <int K>
__device__ __forceinline__ void trace_until(const uint64_t (&l)[K], const uint64_t (&r)[K], uint64_t (&y)[K]) {
    uint64_t c = 0;  // the result at the next word's first position
#pragma unroll
    for (int k = K - 1; k >= 0; --k) {
        y[k] = until_carry64(l[k], r[k], c);
        c = y[k] & 1ull;
    }
}

// Summary bits of trace `at` (of its summary word) from its K result words.
template <int K>
__device__ __forceinline__ void summarize_trace(const uint64_t (&y)[K], const WordAux *aux, int at,
                                                uint64_t (&v)[4]) {
    bool nz = false, fl = true;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        nz |= y[k] != 0;
        fl &= y[k] == aux[k].lm;
    }
    v[0] |= (y[0] & 1ull) << at;
    v[1] |= ((y[0] >> 1) & 1ull) << at;
    v[2] |= (uint64_t)nz << at;
    v[3] |= (uint64_t)fl << at;
}

// Position 0 of l U r (h0) and of r U l (h1) on one trace of K words at pl,
// pr (word stride ws): decided in the first word where the left operand has
// a zero or the right operand a one at or before that zero.
__device__ __forceinline__ void until_at0_trace(const uint64_t *pl, const uint64_t *pr, int64_t ws, int K,
                                                uint32_t &h0, uint32_t &h1) {
    bool d0 = false, d1 = false;
    h0 = h1 = 0;
    for (int k = 0; k < K && !(d0 && d1); ++k) {
        const uint64_t l = pl[(size_t)k * ws], r = pr[(size_t)k * ws], nl = ~l, nr = ~r;
        if (!d0) {
            if (r & (nl ^ (nl - 1))) { h0 = 1; d0 = true; }
            else if (nl) d0 = true;
        }
        if (!d1) {
            if (l & (nr ^ (nr - 1))) { h1 = 1; d1 = true; }
            else if (nr) d1 = true;
        }
    }
}

}  // namespace ltl
