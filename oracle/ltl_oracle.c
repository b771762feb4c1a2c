/*
 * ltl_oracle.c -- CPU oracle for the LTL characteristic-set enumeration.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See
 * ltl_oracle.h for what it restates and the canonical candidate order.
 * Every function cites the reference lines it follows; paths are relative to
 * the reference repository root.
 */
#include "ltl_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Operators (code/ltli6463.cu:203-204, 295, 385, 468, 551-556, 638-645,
 * 739-752).  A trace of length len occupies bits [0, len); bits at or above
 * len are zero in every CS.  len == 64 is an extension: the reference caps
 * lengths at 63 (code/ltli6463.cu:1078) and its mask (1<<len)-1 would be
 * undefined there, so the mask is written out for 64.                      */
/* ------------------------------------------------------------------------ */
static inline uint64_t len_mask(int len) { return len >= 64 ? ~0ull : ((1ull << len) - 1ull); }

uint64_t orc_op_not(uint64_t x, int len) { return ~x & len_mask(len); }
uint64_t orc_op_next(uint64_t x) { return x >> 1; }
uint64_t orc_op_fin(uint64_t x) {
    x |= x >> 1;  x |= x >> 2;  x |= x >> 4;
    x |= x >> 8;  x |= x >> 16; x |= x >> 32;
    return x;
}
uint64_t orc_op_glob(uint64_t x, int len) {
    uint64_t c = orc_op_fin(~x & len_mask(len));
    return x & ~c;
}
uint64_t orc_op_until(uint64_t l, uint64_t r) {
    /* y_k = r_k | (l_k & y_{k+1}) solved as a parallel prefix over 6
     * doubling steps (code/ltli6463.cu:739-752). */
    uint64_t run = l, y = r;
    for (int s = 1; s < 64; s <<= 1) {
        y |= run & (y >> s);
        run &= run >> s;
    }
    return y;
}

/* Multi-word traces (an extension: the reference stops at 63 positions,
 * code/ltli6463.cu:1078).  Word k of a trace holds positions 64k..64k+63;
 * the single-word operators apply per word, and what crosses a word boundary
 * comes from the next higher word: X moves bit 0 of word k+1 into bit 63 of
 * word k; F holds at every position of word k when a higher word has a set
 * bit; G holds in word k only where every higher word is full; U holds on
 * the top run of l's ones in word k when it holds at position 0 of word k+1. */
static uint64_t word_mask(int len, int k) {
    const int lo = 64 * k;
    if (len <= lo) return 0;
    return len - lo >= 64 ? ~0ull : ((1ull << (len - lo)) - 1ull);
}

void orc_trace_op(int32_t op, const uint64_t *x, const uint64_t *y, int32_t K, int32_t len, uint64_t *out) {
    int carry = 0;
    switch (op) {
    case ORC_NOT:
        for (int k = 0; k < K; ++k) out[k] = ~x[k] & word_mask(len, k);
        return;
    case ORC_AND:
        for (int k = 0; k < K; ++k) out[k] = x[k] & y[k];
        return;
    case ORC_OR:
        for (int k = 0; k < K; ++k) out[k] = x[k] | y[k];
        return;
    case ORC_NEXT:
        for (int k = 0; k < K; ++k) out[k] = (x[k] >> 1) | (k + 1 < K ? x[k + 1] << 63 : 0ull);
        return;
    case ORC_FIN:
        for (int k = K - 1; k >= 0; --k) {
            out[k] = orc_op_fin(x[k]) | (carry ? word_mask(len, k) : 0ull);
            carry |= x[k] != 0;
        }
        return;
    case ORC_GLOB:
        carry = 1;
        for (int k = K - 1; k >= 0; --k) {
            const uint64_t m = word_mask(len, k);
            out[k] = carry ? x[k] & ~orc_op_fin(~x[k] & m) : 0ull;
            carry = carry && x[k] == m;
        }
        return;
    default:
        for (int k = K - 1; k >= 0; --k) {
            uint64_t u = orc_op_until(x[k], y[k]);
            if (carry) u |= ~orc_op_fin(~x[k]);
            out[k] = u;
            carry = (int)(u & 1ull);
        }
    }
}

/* Target check (code/ltli6463.cu:240-242): positives must hold at position
 * 0, negatives must not.  Returned as a count so that a Hamming tolerance
 * can be applied (the exact reference check is count == 0). */
int32_t orc_mismatches(const uint64_t *cs, int32_t num_pos, int32_t num_neg) {
    int32_t m = 0;
    for (int i = 0; i < num_pos; ++i) m += (int32_t)(~cs[i] & 1ull);
    for (int i = num_pos; i < num_pos + num_neg; ++i) m += (int32_t)(cs[i] & 1ull);
    return m;
}

/* Reference uniqueness key (code/ltli6463.cu:57-141).  For len_sum <= 126
 * every position maps to its own key bit, which is exact; above that the
 * three relaxed modes sample (1), take length-proportional prefixes (2) or
 * XOR-fold a per-trace 64-bit finalizer (3; the SplitMix64 finalizer of
 * Vigna's public-domain generator). */
static void put_bit(uint64_t *h, uint64_t *l, int j, int bit) {
    if (!bit) return;
    if (j < 63) *l |= 1ull << j;
    else *h |= 1ull << (j - 63);
}
static uint64_t splitmix_final(uint64_t x) {
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}
void orc_relaxed_key(const uint64_t *cs, const uint8_t *lens, int32_t W,
                     int32_t mode, int32_t len_sum, uint64_t *h, uint64_t *l) {
    *h = 0; *l = 0;
    if (len_sum <= 126) {
        int j = 0;
        for (int i = 0; i < W; ++i)
            for (int k = 0; k < lens[i] && j < 126; ++k, ++j)
                put_bit(h, l, j, (int)((cs[i] >> k) & 1ull));
        return;
    }
    if (mode == 1) {
        int stride = len_sum / 126, j = 0;
        for (int i = 0; i < W; ++i)
            for (int k = 0; k < lens[i] && j < 126; k += stride, ++j)
                put_bit(h, l, j, (int)((cs[i] >> k) & 1ull));
    } else if (mode == 2) {
        int j = 0;
        for (int i = 0; i < W; ++i) {
            int take = (126 * lens[i] + len_sum) / len_sum;
            for (int k = 0; k < take && j < 126; ++k, ++j)
                put_bit(h, l, j, k < 64 ? (int)((cs[i] >> k) & 1ull) : 0);
        }
    } else {
        for (int i = 0; i < W; ++i) {
            uint64_t x = splitmix_final(cs[i]);
            if (i < W / 2) *h ^= x; else *l ^= x;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Enumerator (code/ltli6463.cu:1043-1461).                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    const orc_problem *pb;
    const orc_options *opt;
    int W, P, A;   /* W = words per set = T * K */
    int T, K;      /* traces, words per trace */
    int len_sum;
    /* append-only store (code/ltli6463.cu:1166-1178) */
    uint64_t *words; uint8_t *op; int32_t *left, *right;
    int64_t n, alloc;
    /* dedup set: exact mode keys by full CS (slot -> store index),
     * reference modes by the (h, l) key pair */
    int64_t *slot_idx; uint64_t *slot_h, *slot_k0, *slot_k1;
    int64_t tcap, tn;
    int on_the_fly;
    uint64_t all_ltls;
    int found; int64_t found_id; int f_op; int32_t f_l, f_r;
    uint64_t *cand;
    int64_t *lvl_s, *lvl_e;
} ectx;

static uint64_t hash_words(const uint64_t *w, int n) {
    uint64_t h = 0x6a09e667f3bcc909ull ^ (uint64_t)n;
    for (int i = 0; i < n; ++i) {
        h ^= splitmix_final(w[i] + 0x9e3779b97f4a7c15ull * (uint64_t)(i + 1));
        h = (h << 27) | (h >> 37);
        h *= 0xff51afd7ed558ccdull;
    }
    return splitmix_final(h);
}

static int store_push(ectx *e, const uint64_t *cs, int op, int32_t l, int32_t r) {
    if (e->n == e->alloc) {
        int64_t na = e->alloc ? e->alloc * 2 : 1024;
        uint64_t *nw = (uint64_t *)realloc(e->words, (size_t)na * e->W * sizeof(uint64_t));
        uint8_t *no = (uint8_t *)realloc(e->op, (size_t)na);
        int32_t *nl = (int32_t *)realloc(e->left, (size_t)na * sizeof(int32_t));
        int32_t *nr = (int32_t *)realloc(e->right, (size_t)na * sizeof(int32_t));
        if (nw) e->words = nw;
        if (no) e->op = no;
        if (nl) e->left = nl;
        if (nr) e->right = nr;
        if (!nw || !no || !nl || !nr) return -1;
        e->alloc = na;
    }
    memcpy(e->words + (size_t)e->n * e->W, cs, (size_t)e->W * sizeof(uint64_t));
    e->op[e->n] = (uint8_t)op; e->left[e->n] = l; e->right[e->n] = r;
    e->n++;
    return 0;
}

static void key_of(ectx *e, const uint64_t *cs, uint64_t *hv, uint64_t *k0, uint64_t *k1) {
    if (e->opt->dedup_mode == 0) {
        *hv = hash_words(cs, e->W); *k0 = 0; *k1 = 0;
    } else {
        orc_relaxed_key(cs, e->pb->lens, e->T, e->opt->dedup_mode, e->len_sum, k0, k1);
        uint64_t t[2] = {*k0, *k1};
        *hv = hash_words(t, 2);
    }
}

static int table_grow(ectx *e) {
    int64_t nc = e->tcap ? e->tcap * 2 : (1 << 12);
    int64_t *si = (int64_t *)malloc((size_t)nc * sizeof(int64_t));
    uint64_t *sh = (uint64_t *)malloc((size_t)nc * sizeof(uint64_t));
    uint64_t *s0 = (uint64_t *)malloc((size_t)nc * sizeof(uint64_t));
    uint64_t *s1 = (uint64_t *)malloc((size_t)nc * sizeof(uint64_t));
    if (!si || !sh || !s0 || !s1) { free(si); free(sh); free(s0); free(s1); return -1; }
    for (int64_t i = 0; i < nc; ++i) si[i] = -1;
    for (int64_t i = 0; i < e->tcap; ++i) {
        if (e->slot_idx[i] < 0) continue;
        int64_t p = (int64_t)(e->slot_h[i] & (uint64_t)(nc - 1));
        while (si[p] >= 0) p = (p + 1) & (nc - 1);
        si[p] = e->slot_idx[i]; sh[p] = e->slot_h[i]; s0[p] = e->slot_k0[i]; s1[p] = e->slot_k1[i];
    }
    free(e->slot_idx); free(e->slot_h); free(e->slot_k0); free(e->slot_k1);
    e->slot_idx = si; e->slot_h = sh; e->slot_k0 = s0; e->slot_k1 = s1; e->tcap = nc;
    return 0;
}

/* Returns 1 if a CS equivalent to cs is already interned, 0 otherwise; in
 * the latter case *slot is where it would go. */
static int table_find(ectx *e, const uint64_t *cs, uint64_t hv, uint64_t k0, uint64_t k1, int64_t *slot) {
    int64_t p = (int64_t)(hv & (uint64_t)(e->tcap - 1));
    for (;;) {
        int64_t si = e->slot_idx[p];
        if (si < 0) { *slot = p; return 0; }
        if (e->slot_h[p] == hv) {
            if (e->opt->dedup_mode == 0) {
                if (!memcmp(e->words + (size_t)si * e->W, cs, (size_t)e->W * sizeof(uint64_t))) return 1;
            } else if (e->slot_k0[p] == k0 && e->slot_k1[p] == k1) {
                return 1;
            }
        }
        p = (p + 1) & (e->tcap - 1);
    }
}

static int accept(ectx *e, const uint64_t *cs) {
    if (e->K == 1) return orc_mismatches(cs, e->P, e->T - e->P) <= e->opt->hamming_k;
    int32_t m = 0;
    for (int i = 0; i < e->T; ++i) m += (int32_t)(((cs[(size_t)i * e->K] & 1ull) != 0) != (i < e->P));
    return m <= e->opt->hamming_k;
}

/* Operator op on every trace of a K-word set. */
static void apply_set(const ectx *e, int op, const uint64_t *x, const uint64_t *y, uint64_t *out) {
    const int K = e->K;
    for (int i = 0; i < e->T; ++i)
        orc_trace_op(op, x + (size_t)i * K, y ? y + (size_t)i * K : NULL, K, e->pb->lens[i], out + (size_t)i * K);
}

/* One candidate, as one thread of any operator kernel does it
 * (code/ltli6463.cu:212-253): dedup, store if unique, check.  In overflow
 * ("onTheFly") mode only the check runs (code/ltli6463.cu:198-208). */
static int consider(ectx *e, const uint64_t *cs, int op, int32_t l, int32_t r, uint64_t id) {
    if (!e->on_the_fly) {
        uint64_t hv, k0, k1; int64_t slot;
        key_of(e, cs, &hv, &k0, &k1);
        if (table_find(e, cs, hv, k0, k1, &slot)) return 0;
        if (e->opt->capacity > 0 && e->n >= e->opt->capacity) {
            /* code/ltli6463.cu:1021-1027: a unique that no longer fits
             * switches the run to check-only mode. */
            e->on_the_fly = 1;
        } else {
            if (store_push(e, cs, op, l, r)) return -1;
            e->slot_idx[slot] = e->n - 1; e->slot_h[slot] = hv;
            e->slot_k0[slot] = k0; e->slot_k1[slot] = k1;
            if (++e->tn * 2 > e->tcap && table_grow(e)) return -1;
        }
    }
    if (accept(e, cs)) {
        e->found = 1; e->found_id = (int64_t)id; e->f_op = op; e->f_l = l; e->f_r = r;
        return 1;
    }
    return 0;
}

static int unary_block(ectx *e, int op, int64_t s, int64_t t) {
    const int W = e->W;
    const uint8_t *lens = e->pb->lens;
    const uint64_t base = e->all_ltls;
    int rc = 0;
    for (int64_t k = s; k < t && !rc; ++k) {
        const uint64_t *x = e->words + (size_t)k * W;
        if (e->K > 1) apply_set(e, op, x, NULL, e->cand);
        else
            for (int i = 0; i < W; ++i) {
                switch (op) {
                case ORC_NOT:  e->cand[i] = orc_op_not(x[i], lens[i]); break;
                case ORC_NEXT: e->cand[i] = orc_op_next(x[i]); break;
                case ORC_FIN:  e->cand[i] = orc_op_fin(x[i]); break;
                default:       e->cand[i] = orc_op_glob(x[i], lens[i]); break;
                }
            }
        rc = consider(e, e->cand, op, (int32_t)k, -1, base + (uint64_t)(k - s));
    }
    e->all_ltls += (uint64_t)(t - s);
    return rc;
}

static int binary_block(ectx *e, int op, int64_t ls, int64_t le, int64_t rs, int64_t re) {
    const int W = e->W;
    const int64_t nR = re - rs;
    const uint64_t base = e->all_ltls;
    int rc = 0;
    for (int64_t a = ls; a < le && !rc; ++a) {
        for (int64_t b = rs; b < re && !rc; ++b) {
            const uint64_t pair = (uint64_t)((a - ls) * nR + (b - rs));
            if (op == ORC_UNTIL) {
                const uint64_t *x = e->words + (size_t)a * W, *y = e->words + (size_t)b * W;
                if (e->K > 1) apply_set(e, ORC_UNTIL, x, y, e->cand);
                else for (int i = 0; i < W; ++i) e->cand[i] = orc_op_until(x[i], y[i]);
                rc = consider(e, e->cand, op, (int32_t)a, (int32_t)b, base + 2 * pair);
                if (rc) break;
                x = e->words + (size_t)a * W; y = e->words + (size_t)b * W;
                if (e->K > 1) apply_set(e, ORC_UNTIL, y, x, e->cand);
                else for (int i = 0; i < W; ++i) e->cand[i] = orc_op_until(y[i], x[i]);
                rc = consider(e, e->cand, op, (int32_t)b, (int32_t)a, base + 2 * pair + 1);
            } else {
                const uint64_t *x = e->words + (size_t)a * W, *y = e->words + (size_t)b * W;
                if (op == ORC_AND) for (int i = 0; i < W; ++i) e->cand[i] = x[i] & y[i];
                else               for (int i = 0; i < W; ++i) e->cand[i] = x[i] | y[i];
                rc = consider(e, e->cand, op, (int32_t)a, (int32_t)b, base + pair);
            }
        }
    }
    e->all_ltls += (uint64_t)((le - ls) * nR) * (op == ORC_UNTIL ? 2u : 1u);
    return rc;
}

/* ----- OpenMP block processing (CPU baseline with threads > 1) ----------
 * Same semantics as the sequential blocks: a batch of candidates is hashed
 * and checked in parallel; dedup is sharded by hash so that each shard sees
 * its candidates in canonical order (first occurrence wins, as in the
 * sequential loop); survivors are appended in canonical order. */
#ifdef _OPENMP
static void cand_compute(ectx *e, int op, int64_t ls, int64_t nR, int64_t rs, int64_t idx, uint64_t *out,
                         int32_t *l, int32_t *r) {
    const int W = e->W;
    const uint8_t *lens = e->pb->lens;
    if (op == ORC_AND || op == ORC_OR || op == ORC_UNTIL) {
        int64_t pair = op == ORC_UNTIL ? idx >> 1 : idx;
        int dir = op == ORC_UNTIL ? (int)(idx & 1) : 0;
        int64_t a = ls + pair / nR, b = rs + pair % nR;
        const uint64_t *x = e->words + (size_t)a * W, *y = e->words + (size_t)b * W;
        if (op == ORC_AND) for (int i = 0; i < W; ++i) out[i] = x[i] & y[i];
        else if (op == ORC_OR) for (int i = 0; i < W; ++i) out[i] = x[i] | y[i];
        else if (e->K > 1) apply_set(e, ORC_UNTIL, dir ? y : x, dir ? x : y, out);
        else if (!dir) for (int i = 0; i < W; ++i) out[i] = orc_op_until(x[i], y[i]);
        else for (int i = 0; i < W; ++i) out[i] = orc_op_until(y[i], x[i]);
        *l = (int32_t)(dir ? b : a); *r = (int32_t)(dir ? a : b);
    } else {
        const uint64_t *x = e->words + (size_t)(ls + idx) * W;
        if (e->K > 1) { apply_set(e, op, x, NULL, out); *l = (int32_t)(ls + idx); *r = -1; return; }
        for (int i = 0; i < W; ++i) {
            switch (op) {
            case ORC_NOT:  out[i] = orc_op_not(x[i], lens[i]); break;
            case ORC_NEXT: out[i] = orc_op_next(x[i]); break;
            case ORC_FIN:  out[i] = orc_op_fin(x[i]); break;
            default:       out[i] = orc_op_glob(x[i], lens[i]); break;
            }
        }
        *l = (int32_t)(ls + idx); *r = -1;
    }
}

static int par_block(ectx *e, int op, int64_t ls, int64_t le, int64_t rs, int64_t re) {
    const int W = e->W;
    const int binary = (op == ORC_AND || op == ORC_OR || op == ORC_UNTIL);
    const int64_t nR = binary ? re - rs : 1;
    const int64_t total = binary ? (le - ls) * nR * (op == ORC_UNTIL ? 2 : 1) : (le - ls);
    const uint64_t base = e->all_ltls;
    const int T = e->opt->threads;
    const int64_t B = 1 << 17;
    int64_t lcap = 1; while (lcap < 2 * B) lcap <<= 1;
    uint64_t *hv = (uint64_t *)malloc((size_t)B * 8), *k0 = (uint64_t *)malloc((size_t)B * 8);
    uint64_t *k1 = (uint64_t *)malloc((size_t)B * 8);
    uint8_t *acc = (uint8_t *)malloc((size_t)B), *uniq = (uint8_t *)malloc((size_t)B);
    int64_t *pos = (int64_t *)malloc((size_t)B * 8);
    int64_t *ltab = (int64_t *)malloc((size_t)T * lcap * 8);
    uint64_t *scratch = (uint64_t *)malloc((size_t)T * 2 * W * 8);
    int rc = 0;
    for (int64_t off = 0; off < total && !rc; off += B) {
        const int64_t nb = total - off < B ? total - off : B;
        const int otf = e->on_the_fly;
        /* 1: candidate CS, key, check, and read-only lookup in the store */
        #pragma omp parallel for num_threads(T) schedule(static)
        for (int64_t c = 0; c < nb; ++c) {
            uint64_t *buf = scratch + (size_t)omp_get_thread_num() * 2 * W;
            int32_t l, r; int64_t slot;
            cand_compute(e, op, ls, nR, rs, off + c, buf, &l, &r);
            acc[c] = (uint8_t)accept(e, buf);
            uniq[c] = 0;
            if (otf) continue;
            key_of(e, buf, &hv[c], &k0[c], &k1[c]);
            uniq[c] = (uint8_t)!table_find(e, buf, hv[c], k0[c], k1[c], &slot);
        }
        /* 2: first occurrence within the batch; shard s owns the hashes with
         * (hv >> 48) % T == s and walks the batch in canonical order */
        if (!otf) {
            #pragma omp parallel for num_threads(T) schedule(static, 1)
            for (int s = 0; s < T; ++s) {
                int64_t *lt = ltab + (size_t)s * lcap;
                uint64_t *b0 = scratch + (size_t)s * 2 * W, *b1 = b0 + W;
                for (int64_t i = 0; i < lcap; ++i) lt[i] = -1;
                for (int64_t c = 0; c < nb; ++c) {
                    if (!uniq[c] || (int)((hv[c] >> 48) % (uint64_t)T) != s) continue;
                    int64_t p = (int64_t)(hv[c] & (uint64_t)(lcap - 1));
                    int dup = 0, have = 0;
                    while (lt[p] >= 0) {
                        int64_t o = lt[p];
                        if (hv[o] == hv[c]) {
                            if (e->opt->dedup_mode) dup = k0[o] == k0[c] && k1[o] == k1[c];
                            else {
                                int32_t l, r;
                                if (!have) { cand_compute(e, op, ls, nR, rs, off + c, b0, &l, &r); have = 1; }
                                cand_compute(e, op, ls, nR, rs, off + o, b1, &l, &r);
                                dup = !memcmp(b0, b1, (size_t)W * 8);
                            }
                            if (dup) break;
                        }
                        p = (p + 1) & (lcap - 1);
                    }
                    if (dup) uniq[c] = 0; else lt[p] = c;
                }
            }
        }
        /* 3: canonical-order append up to capacity (code/ltli6463.cu:1021-1027);
         * the first unique that does not fit starts check-only mode */
        int64_t over = nb, k = 0;
        if (!otf) {
            for (int64_t c = 0; c < nb; ++c) {
                pos[c] = -1;
                if (!uniq[c]) continue;
                if (e->opt->capacity > 0 && e->n + k >= e->opt->capacity) {
                    if (over == nb) over = c;
                    continue;
                }
                pos[c] = e->n + k++;
            }
            if (e->n + k > e->alloc) {
                int64_t na = e->alloc ? e->alloc : 1024;
                while (na < e->n + k) na *= 2;
                uint64_t *nw = (uint64_t *)realloc(e->words, (size_t)na * W * 8);
                uint8_t *no = (uint8_t *)realloc(e->op, (size_t)na);
                int32_t *nl = (int32_t *)realloc(e->left, (size_t)na * 4);
                int32_t *nr = (int32_t *)realloc(e->right, (size_t)na * 4);
                if (nw) e->words = nw;
                if (no) e->op = no;
                if (nl) e->left = nl;
                if (nr) e->right = nr;
                if (!nw || !no || !nl || !nr) { rc = -1; break; }
                e->alloc = na;
            }
            #pragma omp parallel for num_threads(T) schedule(static)
            for (int64_t c = 0; c < nb; ++c) {
                if (pos[c] < 0) continue;
                int32_t l, r;
                cand_compute(e, op, ls, nR, rs, off + c, e->words + (size_t)pos[c] * W, &l, &r);
                e->op[pos[c]] = (uint8_t)op; e->left[pos[c]] = l; e->right[pos[c]] = r;
            }
            for (int64_t c = 0; c < nb; ++c) {
                if (pos[c] < 0) continue;
                int64_t slot = 0;
                table_find(e, e->words + (size_t)pos[c] * W, hv[c], k0[c], k1[c], &slot);
                e->slot_idx[slot] = pos[c]; e->slot_h[slot] = hv[c];
                e->slot_k0[slot] = k0[c]; e->slot_k1[slot] = k1[c];
                if (++e->tn * 2 > e->tcap && table_grow(e)) { rc = -1; break; }
            }
            if (rc) break;
            e->n += k;
            if (over < nb) e->on_the_fly = 1;
        }
        /* 4: the first checked candidate that is accepted.  A candidate is
         * checked when it is unique or when check-only mode was on at its
         * turn (code/ltli6463.cu:198-208 vs 236-243). */
        for (int64_t c = 0; c < nb; ++c) {
            if (!acc[c] || !(otf || uniq[c] || c > over)) continue;
            int32_t l, r;
            cand_compute(e, op, ls, nR, rs, off + c, e->cand, &l, &r);
            e->found = 1; e->found_id = (int64_t)(base + off + c);
            e->f_op = op; e->f_l = l; e->f_r = r; rc = 1;
            break;
        }
    }
    free(hv); free(k0); free(k1); free(acc); free(uniq); free(pos); free(ltab); free(scratch);
    e->all_ltls += (uint64_t)total;
    return rc;
}
#endif

static int run_unary(ectx *e, int op, int64_t s, int64_t t) {
#ifdef _OPENMP
    if (e->opt->threads > 1 && t - s > 4096) return par_block(e, op, s, t, 0, 0);
#endif
    return unary_block(e, op, s, t);
}
static int run_binary(ectx *e, int op, int64_t ls, int64_t le, int64_t rs, int64_t re) {
#ifdef _OPENMP
    if (e->opt->threads > 1 && (le - ls) * (re - rs) > 4096) return par_block(e, op, ls, le, rs, re);
#endif
    return binary_block(e, op, ls, le, rs, re);
}

/* Formula string (code/ltli6463.cu:877-928). */
typedef struct { char *s; size_t n, cap; } sbuf;
static void sb_put(sbuf *b, const char *t) {
    size_t k = strlen(t);
    if (b->n + k + 1 > b->cap) {
        size_t nc = (b->cap ? b->cap * 2 : 64);
        while (nc < b->n + k + 1) nc *= 2;
        b->s = (char *)realloc(b->s, nc); b->cap = nc;
    }
    memcpy(b->s + b->n, t, k + 1); b->n += k;
}
static char atom_name(const orc_problem *pb, int a) { return pb->names ? pb->names[a] : (char)('a' + a); }
static void emit(const orc_problem *pb, const uint8_t *op, const int32_t *L, const int32_t *R,
                 int o, int32_t l, int32_t r, int atom, sbuf *b);
static void emit_idx(const orc_problem *pb, const uint8_t *op, const int32_t *L, const int32_t *R,
                     int32_t k, sbuf *b) {
    emit(pb, op, L, R, op[k], L[k], R[k], k, b);
}
static void emit(const orc_problem *pb, const uint8_t *op, const int32_t *L, const int32_t *R,
                 int o, int32_t l, int32_t r, int atom, sbuf *b) {
    switch (o) {
    case ORC_ATOM: { char t[2] = {atom_name(pb, atom), 0}; sb_put(b, t); return; }
    case ORC_NOT:  sb_put(b, "~(");  emit_idx(pb, op, L, R, l, b); sb_put(b, ")"); return;
    case ORC_NEXT: sb_put(b, "X(");  emit_idx(pb, op, L, R, l, b); sb_put(b, ")"); return;
    case ORC_FIN:  sb_put(b, "F(");  emit_idx(pb, op, L, R, l, b); sb_put(b, ")"); return;
    case ORC_GLOB: sb_put(b, "G(");  emit_idx(pb, op, L, R, l, b); sb_put(b, ")"); return;
    default: {
        const char *mid = o == ORC_AND ? ")&(" : o == ORC_OR ? ")|(" : ")U(";
        sb_put(b, "("); emit_idx(pb, op, L, R, l, b); sb_put(b, mid);
        emit_idx(pb, op, L, R, r, b); sb_put(b, ")");
    }
    }
}

static double now_s(void) {
    struct timespec ts; clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int32_t orc_enumerate(const orc_problem *pb, const orc_options *opt, orc_result *res) {
    memset(res, 0, sizeof(*res));
    res->found_id = -1; res->root_op = -1; res->root_left = -1; res->root_right = -1;
    const int T = pb->num_pos + pb->num_neg;
    const int K = pb->words_per_trace > 1 ? pb->words_per_trace : 1;
    const int W = T * K;
    if (T <= 0 || pb->num_atoms <= 0 || K > 4) return res->status = -1;
    for (int i = 0; i < T; ++i) if (pb->lens[i] < 1 || pb->lens[i] > 64 * K) return res->status = -1;
    for (int i = 0; i < 8; ++i) if (opt->cost[i] <= 0) return res->status = -1;
    /* the reference's relaxed keys are defined on one word per trace */
    if (K > 1 && opt->dedup_mode != 0) return res->status = -1;
    const int c1 = opt->cost[0], c2 = opt->cost[1], c3 = opt->cost[2], c4 = opt->cost[3];
    const int c5 = opt->cost[4], c6 = opt->cost[5], c7 = opt->cost[6], c8 = opt->cost[7];
    const int maxc = opt->max_cost;

    double t0 = now_s();
    ectx e; memset(&e, 0, sizeof(e));
    e.pb = pb; e.opt = opt; e.W = W; e.P = pb->num_pos; e.A = pb->num_atoms;
    e.T = T; e.K = K;
    for (int i = 0; i < T; ++i) e.len_sum += pb->lens[i];
    e.cand = (uint64_t *)malloc((size_t)W * sizeof(uint64_t));
    res->n_levels = (maxc > c1 ? maxc : c1) + 2;
    e.lvl_s = (int64_t *)calloc((size_t)res->n_levels, sizeof(int64_t));
    e.lvl_e = (int64_t *)calloc((size_t)res->n_levels, sizeof(int64_t));
    table_grow(&e);
    res->cost = c1;

    /* Atoms (code/ltli6463.cu:1100-1136): stored in alphabet order without
     * dedup among themselves, each checked before the next; all are
     * interned (code/ltli6463.cu:144-173). */
    int rc = 0;
    for (int a = 0; a < e.A; ++a) {
        const uint64_t *cs = pb->atoms + (size_t)a * W;
        store_push(&e, cs, ORC_ATOM, -1, -1);
        uint64_t hv, k0, k1; int64_t slot;
        key_of(&e, cs, &hv, &k0, &k1);
        if (!table_find(&e, cs, hv, k0, k1, &slot)) {
            e.slot_idx[slot] = e.n - 1; e.slot_h[slot] = hv; e.slot_k0[slot] = k0; e.slot_k1[slot] = k1;
            if (++e.tn * 2 > e.tcap) table_grow(&e);
        }
        e.all_ltls++;
        if (accept(&e, cs)) {
            e.found = 1; e.found_id = a; e.f_op = ORC_ATOM; e.f_l = -1; e.f_r = -1;
            res->cost = c1;
            rc = 1;
            break;
        }
    }
    if (c1 < res->n_levels) { e.lvl_s[c1] = 0; e.lvl_e[c1] = e.n; }

    int cost = c1 + 1, last_round = 0, shortage = -1;
    if (!rc) {
        for (cost = c1 + 1; cost <= maxc; ++cost) {
            if (e.on_the_fly) {
                /* code/ltli6463.cu:1202-1206 */
                int dif = cost - shortage;
                if (dif == c2 || dif == c1 + c3 || dif == c1 + c4 || dif == c5 || dif == c6 ||
                    dif == c7 || dif == c1 + c8) last_round = 1;
            }
            e.lvl_s[cost] = e.n;
            /* ~ (code/ltli6463.cu:1209-1242) */
            if ((opt->neg_type == 1 && cost - c2 >= c1) || (opt->neg_type == 2 && cost - c2 == c1))
                if ((rc = run_unary(&e, ORC_NOT, e.lvl_s[cost - c2], e.lvl_e[cost - c2]))) break;
            /* & and | (code/ltli6463.cu:1244-1310) */
            for (int i = c1; 2 * i <= cost - c3 && !rc; ++i)
                rc = run_binary(&e, ORC_AND, e.lvl_s[i], e.lvl_e[i], e.lvl_s[cost - i - c3], e.lvl_e[cost - i - c3]);
            if (rc) break;
            for (int i = c1; 2 * i <= cost - c4 && !rc; ++i)
                rc = run_binary(&e, ORC_OR, e.lvl_s[i], e.lvl_e[i], e.lvl_s[cost - i - c4], e.lvl_e[cost - i - c4]);
            if (rc) break;
            /* X, F, G (code/ltli6463.cu:1312-1406) */
            if (cost - c5 >= c1 && (rc = run_unary(&e, ORC_NEXT, e.lvl_s[cost - c5], e.lvl_e[cost - c5]))) break;
            if (cost - c6 >= c1 && (rc = run_unary(&e, ORC_FIN, e.lvl_s[cost - c6], e.lvl_e[cost - c6]))) break;
            if (cost - c7 >= c1 && (rc = run_unary(&e, ORC_GLOB, e.lvl_s[cost - c7], e.lvl_e[cost - c7]))) break;
            /* U (code/ltli6463.cu:1408-1440) */
            for (int i = c1; 2 * i <= cost - c8 && !rc; ++i)
                rc = run_binary(&e, ORC_UNTIL, e.lvl_s[i], e.lvl_e[i], e.lvl_s[cost - i - c8], e.lvl_e[cost - i - c8]);
            if (rc) break;
            e.lvl_e[cost] = e.n;
            if (last_round) break;
            if (e.on_the_fly && shortage == -1) shortage = cost;
        }
        if (cost == maxc + 1) cost--;   /* code/ltli6463.cu:1445 */
        if (rc == 1) e.lvl_e[cost] = e.n;
        res->cost = cost;
    }

    res->seconds = now_s() - t0;
    res->all_ltls = e.all_ltls;
    res->on_the_fly = e.on_the_fly;
    res->n_stored = e.n;
    res->level_start = e.lvl_s; res->level_end = e.lvl_e;
    if (rc < 0) {
        res->status = -2;
    } else if (e.found) {
        res->status = 1;
        res->found_id = e.found_id;
        res->root_op = e.f_op; res->root_left = e.f_l; res->root_right = e.f_r;
        sbuf b = {0, 0, 0};
        emit(pb, e.op, e.left, e.right, e.f_op, e.f_l, e.f_r, (int)e.found_id, &b);
        res->formula = b.s;
    } else {
        res->status = 0;
        res->formula = strdup("not_found");
    }
    if (opt->keep_store) {
        res->words = e.words; res->op = e.op; res->left = e.left; res->right = e.right;
    } else {
        free(e.words); free(e.op); free(e.left); free(e.right);
    }
    free(e.slot_idx); free(e.slot_h); free(e.slot_k0); free(e.slot_k1); free(e.cand);
    return res->status;
}

void orc_result_free(orc_result *res) {
    free(res->level_start); free(res->level_end); free(res->words); free(res->op);
    free(res->left); free(res->right); free(res->formula);
    memset(res, 0, sizeof(*res));
}

/* ------------------------------------------------------------------------ */
/* Textbook evaluator and parser for the output grammar (code/ltli6463.cu:
 * 877-928; the same grammar is parsed by code/dc.py:136-183).  Semantics per
 * position k of a trace of length n (strong next, finite-trace F/G/U):
 *   X f @k  = k+1 < n and f @k+1
 *   F f @k  = exists j in [k, n): f @j
 *   G f @k  = forall j in [k, n): f @j
 *   f U g @k = exists j in [k, n): g @j and forall m in [k, j): f @m        */
/* ------------------------------------------------------------------------ */
typedef struct node { int op; int atom; struct node *a, *b; } node;

static void free_node(node *n) { if (!n) return; free_node(n->a); free_node(n->b); free(n); }

static node *parse_f(const char **p, const orc_problem *pb);
static node *mk(int op) { node *n = (node *)calloc(1, sizeof(node)); n->op = op; return n; }
static int expect(const char **p, char c) { if (**p != c) return 0; ++*p; return 1; }

static node *parse_f(const char **p, const orc_problem *pb) {
    char c = **p;
    if (c == '~' || c == 'X' || c == 'F' || c == 'G') {
        ++*p;
        if (!expect(p, '(')) return NULL;
        node *k = parse_f(p, pb);
        if (!k || !expect(p, ')')) { free_node(k); return NULL; }
        node *n = mk(c == '~' ? ORC_NOT : c == 'X' ? ORC_NEXT : c == 'F' ? ORC_FIN : ORC_GLOB);
        n->a = k; return n;
    }
    if (c == '(') {
        ++*p;
        node *l = parse_f(p, pb);
        if (!l || !expect(p, ')')) { free_node(l); return NULL; }
        char o = **p;
        int op = o == '&' ? ORC_AND : o == '|' ? ORC_OR : o == 'U' ? ORC_UNTIL : -1;
        if (op < 0) { free_node(l); return NULL; }
        ++*p;
        if (!expect(p, '(')) { free_node(l); return NULL; }
        node *r = parse_f(p, pb);
        if (!r || !expect(p, ')')) { free_node(l); free_node(r); return NULL; }
        node *n = mk(op); n->a = l; n->b = r; return n;
    }
    for (int a = 0; a < pb->num_atoms; ++a) {
        if (c == atom_name(pb, a)) { ++*p; node *n = mk(ORC_ATOM); n->atom = a; return n; }
    }
    return NULL;
}

static uint64_t eval_pos(const node *n, const orc_problem *pb, int t, int len) {
    const int W = pb->num_pos + pb->num_neg;
    if (n->op == ORC_ATOM) return pb->atoms[(size_t)n->atom * W + t] & len_mask(len);
    uint64_t a = eval_pos(n->a, pb, t, len), b = n->b ? eval_pos(n->b, pb, t, len) : 0, out = 0;
    for (int k = 0; k < len; ++k) {
        int v = 0;
        switch (n->op) {
        case ORC_NOT:  v = !((a >> k) & 1); break;
        case ORC_AND:  v = ((a >> k) & 1) && ((b >> k) & 1); break;
        case ORC_OR:   v = ((a >> k) & 1) || ((b >> k) & 1); break;
        case ORC_NEXT: v = k + 1 < len && ((a >> (k + 1)) & 1); break;
        case ORC_FIN:  for (int j = k; j < len && !v; ++j) v = (a >> j) & 1; break;
        case ORC_GLOB: v = 1; for (int j = k; j < len && v; ++j) v = (a >> j) & 1; break;
        default:
            for (int j = k; j < len && !v; ++j) {
                if (!((b >> j) & 1)) continue;
                int ok = 1;
                for (int m = k; m < j && ok; ++m) ok = (a >> m) & 1;
                v = ok;
            }
        }
        if (v) out |= 1ull << k;
    }
    return out;
}

/* Value at position k of a trace of length len, from the operands' values
 * per position (the definitions above, for any length). */
static uint8_t pos_value(int op, const uint8_t *a, const uint8_t *b, int k, int len) {
    switch (op) {
    case ORC_NOT: return (uint8_t)!a[k];
    case ORC_AND: return (uint8_t)(a[k] && b[k]);
    case ORC_OR: return (uint8_t)(a[k] || b[k]);
    case ORC_NEXT: return (uint8_t)(k + 1 < len && a[k + 1]);
    case ORC_FIN:
        for (int j = k; j < len; ++j) if (a[j]) return 1;
        return 0;
    case ORC_GLOB:
        for (int j = k; j < len; ++j) if (!a[j]) return 0;
        return 1;
    default:
        /* the first j >= k with b@j decides, unless a fails before it */
        for (int j = k; j < len; ++j) {
            if (b[j]) return 1;
            if (!a[j]) return 0;
        }
        return 0;
    }
}

static void unpack_bits(const uint64_t *w, int len, uint8_t *out) {
    for (int k = 0; k < len; ++k) out[k] = (uint8_t)((w[k >> 6] >> (k & 63)) & 1ull);
}
static void pack_bits(const uint8_t *v, int len, int K, uint64_t *w) {
    for (int k = 0; k < K; ++k) w[k] = 0;
    for (int k = 0; k < len; ++k) if (v[k]) w[k >> 6] |= 1ull << (k & 63);
}

/* Per-position values of n on trace t of a K-word problem (len <= 255). */
static void eval_bits(const node *n, const orc_problem *pb, int t, int len, int K, uint8_t *out) {
    const int T = pb->num_pos + pb->num_neg;
    if (n->op == ORC_ATOM) {
        unpack_bits(pb->atoms + ((size_t)n->atom * T + t) * K, len, out);
        return;
    }
    uint8_t a[256], b[256] = {0};
    eval_bits(n->a, pb, t, len, K, a);
    if (n->b) eval_bits(n->b, pb, t, len, K, b);
    for (int k = 0; k < len; ++k) out[k] = pos_value(n->op, a, b, k, len);
}

int32_t orc_eval_formula(const char *formula, const orc_problem *pb, uint64_t *out_cs) {
    const char *p = formula;
    node *n = parse_f(&p, pb);
    if (!n || *p) { free_node(n); return -1; }
    const int W = pb->num_pos + pb->num_neg;
    const int K = pb->words_per_trace > 1 ? pb->words_per_trace : 1;
    if (K == 1) {
        for (int t = 0; t < W; ++t) out_cs[t] = eval_pos(n, pb, t, pb->lens[t]);
    } else {
        uint8_t v[256];
        for (int t = 0; t < W; ++t) {
            eval_bits(n, pb, t, pb->lens[t], K, v);
            pack_bits(v, pb->lens[t], K, out_cs + (size_t)t * K);
        }
    }
    free_node(n);
    return 0;
}

static int32_t node_cost(const node *n, const int32_t *cost) {
    if (!n) return 0;
    return cost[n->op] + node_cost(n->a, cost) + node_cost(n->b, cost);
}

int32_t orc_formula_cost(const char *formula, const int32_t *cost) {
    /* parse against a 26-letter alphabet */
    static const char letters[] = "abcdefghijklmnopqrstuvwxyz";
    uint8_t lens[1] = {1};
    uint64_t atoms[26] = {0};
    orc_problem pb = {1, 0, 26, lens, atoms, letters, 1, 0};
    const char *p = formula;
    node *n = parse_f(&p, &pb);
    if (!n || *p) { free_node(n); return -1; }
    int32_t c = node_cost(n, cost);
    free_node(n);
    return c;
}

/* ------------------------------------------------------------------------ */
/* Brute force: every syntactic formula by cost, no dedup, per-position
 * semantics as in eval_pos.  Pins the minimal cost independently of the
 * bit-parallel operators and of observational-equivalence pruning.        */
/* ------------------------------------------------------------------------ */
static uint64_t bf_unary(int op, uint64_t a, int len) {
    uint64_t out = 0;
    for (int k = 0; k < len; ++k) {
        int v = 0;
        if (op == ORC_NOT) v = !((a >> k) & 1);
        else if (op == ORC_NEXT) v = k + 1 < len && ((a >> (k + 1)) & 1);
        else if (op == ORC_FIN) { for (int j = k; j < len && !v; ++j) v = (a >> j) & 1; }
        else { v = 1; for (int j = k; j < len && v; ++j) v = (a >> j) & 1; }
        if (v) out |= 1ull << k;
    }
    return out;
}
static uint64_t bf_binary(int op, uint64_t a, uint64_t b, int len) {
    uint64_t out = 0;
    for (int k = 0; k < len; ++k) {
        int v = 0;
        if (op == ORC_AND) v = ((a >> k) & 1) && ((b >> k) & 1);
        else if (op == ORC_OR) v = ((a >> k) & 1) || ((b >> k) & 1);
        else {
            for (int j = k; j < len && !v; ++j) {
                if (!((b >> j) & 1)) continue;
                int ok = 1;
                for (int m = k; m < j && ok; ++m) ok = (a >> m) & 1;
                v = ok;
            }
        }
        if (v) out |= 1ull << k;
    }
    return out;
}

/* One trace of K words through an operator, per position (b = NULL for a
 * unary operator). */
static void bf_trace(int op, const uint64_t *a, const uint64_t *b, int len, int K, uint64_t *out) {
    if (K == 1) {
        out[0] = b ? bf_binary(op, a[0], b[0], len) : bf_unary(op, a[0], len);
        return;
    }
    uint8_t x[256], y[256] = {0}, v[256];
    unpack_bits(a, len, x);
    if (b) unpack_bits(b, len, y);
    for (int k = 0; k < len; ++k) v[k] = pos_value(op, x, y, k, len);
    pack_bits(v, len, K, out);
}

static int32_t mism_stride(const uint64_t *cs, int P, int N, int K) {
    int32_t m = 0;
    for (int i = 0; i < P + N; ++i) m += (int32_t)(((cs[(size_t)i * K] & 1ull) != 0) != (i < P));
    return m;
}

int32_t orc_bruteforce_min_cost(const orc_problem *pb, const orc_options *opt, int64_t max_formulas) {
    const int T = pb->num_pos + pb->num_neg;
    const int K = pb->words_per_trace > 1 ? pb->words_per_trace : 1;
    const int W = T * K;   /* words per formula */
    const int c1 = opt->cost[0], maxc = opt->max_cost;
    if (maxc < c1) return -1;
    uint64_t **lv = (uint64_t **)calloc((size_t)maxc + 1, sizeof(uint64_t *));
    int64_t *cnt = (int64_t *)calloc((size_t)maxc + 1, sizeof(int64_t));
    int64_t total = 0;
    int32_t result = -1;
    uint64_t *tmp = (uint64_t *)malloc((size_t)W * 8);
    for (int c = c1; c <= maxc && result < 0; ++c) {
        /* count */
        int64_t n = 0;
        if (c == c1) n += pb->num_atoms;
        const int un[4] = {ORC_NOT, ORC_NEXT, ORC_FIN, ORC_GLOB};
        for (int u = 0; u < 4; ++u) {
            int sub = c - opt->cost[un[u]];
            if (sub < c1) continue;
            if (un[u] == ORC_NOT && opt->neg_type == 2) { if (sub == c1) n += pb->num_atoms; }
            else n += cnt[sub];
        }
        const int bn[3] = {ORC_AND, ORC_OR, ORC_UNTIL};
        for (int u = 0; u < 3; ++u)
            for (int i = c1; c - opt->cost[bn[u]] - i >= c1; ++i) n += cnt[i] * cnt[c - opt->cost[bn[u]] - i];
        total += n;
        if (total > max_formulas) { result = -2; break; }
        lv[c] = (uint64_t *)malloc((size_t)(n ? n : 1) * W * 8);
        int64_t k = 0;
#define BF_EMIT() do { memcpy(lv[c] + (size_t)k * W, tmp, (size_t)W * 8); ++k; \
            if (mism_stride(tmp, pb->num_pos, pb->num_neg, K) <= opt->hamming_k) result = c; } while (0)
        if (c == c1)
            for (int a = 0; a < pb->num_atoms; ++a) {
                for (int t = 0; t < W; ++t) tmp[t] = pb->atoms[(size_t)a * W + t];
                BF_EMIT();
            }
        for (int u = 0; u < 4; ++u) {
            int sub = c - opt->cost[un[u]];
            if (sub < c1) continue;
            int64_t m = cnt[sub];
            if (un[u] == ORC_NOT && opt->neg_type == 2) { if (sub != c1) continue; m = pb->num_atoms; }
            for (int64_t f = 0; f < m; ++f) {
                for (int t = 0; t < T; ++t)
                    bf_trace(un[u], lv[sub] + (size_t)f * W + (size_t)t * K, NULL, pb->lens[t], K,
                             tmp + (size_t)t * K);
                BF_EMIT();
            }
        }
        for (int u = 0; u < 3; ++u)
            for (int i = c1; c - opt->cost[bn[u]] - i >= c1; ++i) {
                int j = c - opt->cost[bn[u]] - i;
                for (int64_t f = 0; f < cnt[i]; ++f)
                    for (int64_t g = 0; g < cnt[j]; ++g) {
                        for (int t = 0; t < T; ++t)
                            bf_trace(bn[u], lv[i] + (size_t)f * W + (size_t)t * K,
                                     lv[j] + (size_t)g * W + (size_t)t * K, pb->lens[t], K, tmp + (size_t)t * K);
                        BF_EMIT();
                    }
            }
#undef BF_EMIT
        cnt[c] = k;
    }
    for (int c = 0; c <= maxc; ++c) free(lv[c]);
    free(lv); free(cnt); free(tmp);
    return result;
}

/* ------------------------------------------------------------------------ */
/* Trace file (code/ltli6463.cu:1500-1573).  The alphabet is the last line's
 * a-z letters, columns in file order; atoms are the sorted letters.        */
/* ------------------------------------------------------------------------ */
int32_t orc_read_trace_file(const char *path, int32_t max_traces,
                            int32_t *num_pos, int32_t *num_neg, int32_t *num_atoms,
                            uint8_t *lens, uint64_t *atoms, char *names) {
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = (char *)malloc((size_t)sz + 2);
    size_t got = fread(buf, 1, (size_t)sz, f);
    fclose(f);
    buf[got] = 0;
    /* last line: skip one trailing newline */
    long end = (long)got;
    if (end > 0 && buf[end - 1] == '\n') end--;
    long st = end;
    while (st > 0 && buf[st - 1] != '\n') st--;
    char alpha[64]; int na = 0;
    int inset[26] = {0};
    for (long i = st; i < end; ++i)
        if (buf[i] >= 'a' && buf[i] <= 'z') {
            if (na < 63) alpha[na++] = buf[i];
            inset[buf[i] - 'a'] = 1;
        }
    int A = 0; int col_of[26];
    for (int c = 0; c < 26; ++c) if (inset[c]) { names[A] = (char)('a' + c); col_of[c] = A++; }
    names[A] = 0;
    *num_atoms = A;
    /* traces: tokens per position; membership of each sorted letter */
    int W = 0, P = -1, section = 0;
    char *line = buf;
    uint64_t bits[26];
    while (line && *line && section < 2) {
        char *nl = strchr(line, '\n');
        if (nl) *nl = 0;
        if (!strcmp(line, "---")) {
            if (section == 0) P = W;
            section++;
        } else {
            if (W >= max_traces) { free(buf); return -2; }
            memset(bits, 0, sizeof(bits));
            int pos = 0, j = 0;
            for (char *c = line; *c; ++c) {
                if (*c == ';') { pos++; j = 0; }
                else if (*c == ',') continue;
                else {
                    if (*c == '1' && j < na && pos < 64) bits[col_of[alpha[j] - 'a']] |= 1ull << pos;
                    j++;
                }
            }
            int len = pos + 1;
            if (len > 64) { free(buf); return -3; }
            lens[W] = (uint8_t)len;
            for (int a = 0; a < A; ++a) atoms[(size_t)a * max_traces + W] = bits[a];
            W++;
        }
        line = nl ? nl + 1 : NULL;
    }
    if (P < 0) P = W;
    *num_pos = P; *num_neg = W - P;
    /* compact atom-major rows from stride max_traces to stride W */
    for (int a = 1; a < A; ++a)
        memmove(atoms + (size_t)a * W, atoms + (size_t)a * max_traces, (size_t)W * 8);
    free(buf);
    return 0;
}
