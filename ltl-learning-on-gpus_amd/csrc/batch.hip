// batch.hip -- persistent kernel: one workgroup runs one learning instance at
// a time, pulling instances from a device work queue.
//
// Per instance, everything the host scheduler and the chunk kernels do for
// ltl_learn happens inside the workgroup, with workgroup barriers in place of
// kernel boundaries:
//   atoms -> for each cost level (code/ltli6463.cu:1199-1440 order):
//     precheck every candidate of the level on summaries (first accepted id
//     wins) -> generate the level in steps of `tile` candidates: fingerprints
//     and table insert (atomicCAS key, atomicMin id), barrier, decide
//     (own id survived; duplicates verified on fingerprint 2), workgroup scan
//     of survivors, emit in canonical order, commit, barrier.
// The instance's sets, summaries, parent indices and dedup table live in the
// workgroup's region of the pool; used table slots are cleared afterwards.
#include "batch.hpp"
#include "device_ops.hpp"

namespace ltl {
namespace {

// (every instance leaves its region's table empty: one table epoch)
constexpr uint32_t kBatchEpoch = 1;

template <int T>
struct Shared {
    unsigned long long found;
    long long n;                      // sets stored
    int coll, pfail;
    int inst;
    int total;
    int32_t lvl_s[kBatchLevels], lvl_e[kBatchLevels];
    int32_t kl[kBatchNodes], kr[kBatchNodes];   // formula gather queue
    uint64_t fp2[2 * T];
    int wsum[T / 64];
};

struct Inst {
    int W, P, A, nw, nsw;
    int c[8];
    int max_cost, neg_type, hk;
    const WordAux *aux;
    const SumAux *saux;
    uint64_t *tab;
    uint32_t tmask;
    uint64_t *store, *sum;
    int64_t ws, cap;
    uint64_t *m_fp2;
    int32_t *m_left, *m_right;
    uint32_t *m_slot;
    uint8_t *m_op;
};

__device__ __forceinline__ uint64_t aload(const uint64_t *p) {
    return __hip_atomic_load(reinterpret_cast<const unsigned long long *>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of x over the workgroup; *total receives the sum.
template <int T>
__device__ int wg_scan(Shared<T> &S, int x, int *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int v = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    if (lane == 63) S.wsum[wave] = v;
    __syncthreads();
    if (wave == 0) {
        const int w = lane < T / 64 ? S.wsum[lane] : 0;
        int u = w;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(u, o, 64);
            if (lane >= o) u += y;
        }
        if (lane < T / 64) S.wsum[lane] = u - w;
        if (lane == T / 64 - 1) S.total = u;
    }
    __syncthreads();
    *total = S.total;
    return S.wsum[wave] + v - x;
}

// Blocks of level `cost` in the reference's order; f returns true to stop.
template <class F>
__device__ bool level_blocks(const Inst &X, const int32_t *Ls, const int32_t *Le, int cost, F &&f) {
    const int c1 = X.c[0], c2 = X.c[1], c3 = X.c[2], c4 = X.c[3], c5 = X.c[4], c6 = X.c[5], c7 = X.c[6],
              c8 = X.c[7];
    if ((X.neg_type == 1 && cost - c2 >= c1) || (X.neg_type == 2 && cost - c2 == c1))
        if (f(LTL_NOT, Ls[cost - c2], Le[cost - c2], 0, 0)) return true;
    for (int i = c1; 2 * i <= cost - c3; ++i)
        if (f(LTL_AND, Ls[i], Le[i], Ls[cost - i - c3], Le[cost - i - c3])) return true;
    for (int i = c1; 2 * i <= cost - c4; ++i)
        if (f(LTL_OR, Ls[i], Le[i], Ls[cost - i - c4], Le[cost - i - c4])) return true;
    if (cost - c5 >= c1 && f(LTL_NEXT, Ls[cost - c5], Le[cost - c5], 0, 0)) return true;
    if (cost - c6 >= c1 && f(LTL_FIN, Ls[cost - c6], Le[cost - c6], 0, 0)) return true;
    if (cost - c7 >= c1 && f(LTL_GLOB, Ls[cost - c7], Le[cost - c7], 0, 0)) return true;
    for (int i = c1; 2 * i <= cost - c8; ++i)
        if (f(LTL_UNTIL, Ls[i], Le[i], Ls[cost - i - c8], Le[cost - i - c8])) return true;
    return false;
}

__device__ __forceinline__ bool is_bin(int op) { return op == LTL_AND || op == LTL_OR || op == LTL_UNTIL; }

__device__ __forceinline__ uint64_t block_size(int op, int ls, int le, int rs, int re) {
    const int64_t nL = le - ls, nR = is_bin(op) ? re - rs : 1;
    if (nL <= 0 || nR <= 0) return 0;
    return (uint64_t)(nL * nR) * (op == LTL_UNTIL ? 2u : 1u);
}

// Check-only test of one candidate pair / operand.  Returns mismatches of
// the first (m0) and, for U, the swapped (m1) candidate.
__device__ void check_one(const Inst &X, int op, int64_t sl, int64_t sr, uint32_t &m0, uint32_t &m1) {
    m0 = m1 = 0;
    if (op == LTL_UNTIL) {
        const uint64_t *pl = X.store + sl, *pr = X.store + sr;
        for (int w = 0; w < X.nw; ++w) {
            const WordAux *x = X.aux + w;
            const uint64_t l = pl[(size_t)w * X.ws], r = pr[(size_t)w * X.ws];
            m0 += (uint32_t)__popcll((pk_until(l, r, x->inner) ^ x->tgt) & x->chk);
            m1 += (uint32_t)__popcll((pk_until(r, l, x->inner) ^ x->tgt) & x->chk);
        }
    } else if (op == LTL_AND || op == LTL_OR) {
        const uint64_t *pl = X.sum + sl, *pr = X.sum + sr;
        for (int j = 0; j < X.nsw; ++j) {
            const uint64_t l = pl[(size_t)j * X.ws], r = pr[(size_t)j * X.ws];
            const uint64_t c = op == LTL_AND ? (l & r) : (l | r);
            m0 += (uint32_t)__popcll((c ^ X.saux[j].tgt) & X.saux[j].val);
        }
    } else {
        const int vec = op == LTL_NOT ? 0 : op == LTL_NEXT ? 1 : op == LTL_FIN ? 2 : 3;
        const uint64_t *p = X.sum + sl + (size_t)(vec * X.nsw) * X.ws;
        for (int j = 0; j < X.nsw; ++j) {
            uint64_t v = p[(size_t)j * X.ws];
            if (op == LTL_NOT) v = ~v;
            m0 += (uint32_t)__popcll((v ^ X.saux[j].tgt) & X.saux[j].val);
        }
    }
}

template <int OP>
__device__ __forceinline__ uint64_t combine(uint64_t l, uint64_t r, const WordAux *x, bool swap) {
    if constexpr (OP == LTL_AND) return l & r;
    else if constexpr (OP == LTL_OR) return l | r;
    else if constexpr (OP == LTL_UNTIL) return swap ? pk_until(r, l, x->inner) : pk_until(l, r, x->inner);
    else return pk_unary<OP>(l, x);
}

// Fingerprints of the candidate(s) of one lane: c0 = op(l, r), c1 = r U l.
template <int OP, int UB>
__device__ void fingerprint(const Inst &X, int64_t sl, int64_t sr, Acc &c0, Acc &c1) {
    constexpr bool bin = OP == LTL_AND || OP == LTL_OR || OP == LTL_UNTIL;
    const uint64_t *pl = X.store + sl, *pr = X.store + sr;
    auto step = [&](int w, uint64_t l, uint64_t r) {
        const WordAux *x = X.aux + w;
        fold(c0, combine<OP>(l, r, x, false), x);
        if constexpr (OP == LTL_UNTIL) fold(c1, pk_until(r, l, x->inner), x);
    };
    int w = 0;
    for (; w + UB <= X.nw; w += UB) {
        uint64_t l[UB], r[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            l[u] = pl[(size_t)(w + u) * X.ws];
            r[u] = bin ? pr[(size_t)(w + u) * X.ws] : 0;
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) step(w + u, l[u], r[u]);
    }
    for (; w < X.nw; ++w) step(w, pl[(size_t)w * X.ws], bin ? pr[(size_t)w * X.ws] : 0);
}

// Write the survivor at pos: set words, summaries, parent indices,
// fingerprint 2, table slot; commit the table value to pos.
template <int OP, int UB>
__device__ void emit_one(const Inst &X, int64_t pos, int64_t sl, int64_t sr, bool swap, uint32_t slot,
                         uint64_t f2) {
    constexpr bool bin = OP == LTL_AND || OP == LTL_OR || OP == LTL_UNTIL;
    const uint64_t *pl = X.store + sl, *pr = X.store + sr;
    uint64_t *q = X.store + pos, *z = X.sum + pos;
    uint64_t acc[4] = {0, 0, 0, 0};
    auto flush = [&](int j) {  // summary word j is complete
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            z[(size_t)(v * X.nsw + j) * X.ws] = acc[v];
            acc[v] = 0;
        }
    };
    auto step = [&](int w, uint64_t l, uint64_t r) {
        const WordAux *x = X.aux + w;
        const uint64_t c = combine<OP>(l, r, x, swap);
        q[(size_t)w * X.ws] = c;
        summarize_pk<15>(c, x, acc, X.W, flush);
    };
    int w = 0;
    for (; w + UB <= X.nw; w += UB) {
        uint64_t l[UB], r[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            l[u] = pl[(size_t)(w + u) * X.ws];
            r[u] = bin ? pr[(size_t)(w + u) * X.ws] : 0;
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) step(w + u, l[u], r[u]);
    }
    for (; w < X.nw; ++w) step(w, pl[(size_t)w * X.ws], bin ? pr[(size_t)w * X.ws] : 0);
    X.m_op[pos] = (uint8_t)OP;
    X.m_left[pos] = (int32_t)(swap ? sr : sl);
    X.m_right[pos] = bin ? (int32_t)(swap ? sl : sr) : -1;
    X.m_fp2[pos] = f2;
    X.m_slot[pos] = slot;
    X.tab[2 * (size_t)slot + 1] = tab_val((uint64_t)pos, kBatchEpoch);
}

// Generate one block; returns true if the instance must stop (deferred or
// error; S.coll / S.pfail / overflow).
template <int OP, int T, int UB>
__device__ bool gen_block(const Inst &X, Shared<T> &S, int ls, int le, int rs, int re, uint64_t base_id,
                          bool *overflow) {
    constexpr bool bin = OP == LTL_AND || OP == LTL_OR || OP == LTL_UNTIL;
    constexpr int ND = OP == LTL_UNTIL ? 2 : 1;
    const int64_t nL = le - ls, nR = bin ? re - rs : 1;
    if (nL <= 0 || nR <= 0) return false;
    const int64_t npairs = nL * nR;
    for (int64_t b0 = 0; b0 < npairs; b0 += T) {
        const int64_t t = b0 + threadIdx.x;
        const bool valid = t < npairs;
        const int64_t a = valid ? t / nR : 0, b = valid ? t % nR : 0;
        const int64_t sl = ls + a, sr = bin ? rs + b : sl;
        const uint64_t id0 = base_id + (uint64_t)(ND * t);
        uint32_t slot0 = 0, slot1 = 0;
        uint64_t f20 = 0, f21 = 0;
        if (valid) {
            Acc c0, c1;
            fingerprint<OP, UB>(X, sl, sr, c0, c1);
            slot0 = tab_insert(X.tab, X.tmask, tab_key(c0.f1, kBatchEpoch), tab_val(kPending | id0, kBatchEpoch),
                               kBatchEpoch, &S.pfail);
            f20 = c0.f2;
            S.fp2[2 * threadIdx.x] = f20;
            if constexpr (ND == 2) {
                slot1 = tab_insert(X.tab, X.tmask, tab_key(c1.f1, kBatchEpoch),
                                   tab_val(kPending | (id0 + 1), kBatchEpoch), kBatchEpoch, &S.pfail);
                f21 = c1.f2;
                S.fp2[2 * threadIdx.x + 1] = f21;
            }
        }
        __syncthreads();
        bool s0 = false, s1 = false;
        if (valid) {
            const uint64_t first = base_id + (uint64_t)(ND * b0);
            auto decide = [&](uint32_t slot, uint64_t id, uint64_t f2) -> bool {
                const uint64_t v = aload(X.tab + 2 * (size_t)slot + 1);
                if (v == tab_val(kPending | id, kBatchEpoch)) return true;
                if (tab_committed(v, kBatchEpoch)) {
                    if (X.m_fp2[tab_payload(v)] != f2) S.coll = 1;
                } else {
                    const uint64_t rel = tab_payload(v) - first;
                    const uint64_t k = ND == 2 ? rel : 2 * rel;   // index into S.fp2
                    if (rel >= (uint64_t)(ND * T) || S.fp2[k] != f2) S.coll = 1;
                }
                return false;
            };
            s0 = decide(slot0, id0, f20);
            if constexpr (ND == 2) s1 = decide(slot1, id0 + 1, f21);
        }
        int total = 0;
        const int off = wg_scan<T>(S, (int)s0 + (int)s1, &total);
        if (S.n + total > X.cap) { *overflow = true; return true; }
        if (s0) emit_one<OP, UB>(X, S.n + off, sl, sr, false, slot0, f20);
        if (s1) emit_one<OP, UB>(X, S.n + off + (s0 ? 1 : 0), sl, sr, true, slot1, f21);
        __syncthreads();
        if (threadIdx.x == 0) S.n += total;
        __syncthreads();
        if (S.coll || S.pfail) return true;
    }
    return false;
}

template <int T, int UB>
__device__ bool gen_any(int op, const Inst &X, Shared<T> &S, int ls, int le, int rs, int re, uint64_t base_id,
                        bool *overflow) {
    switch (op) {
    case LTL_NOT: return gen_block<LTL_NOT, T, UB>(X, S, ls, le, rs, re, base_id, overflow);
    case LTL_AND: return gen_block<LTL_AND, T, UB>(X, S, ls, le, rs, re, base_id, overflow);
    case LTL_OR: return gen_block<LTL_OR, T, UB>(X, S, ls, le, rs, re, base_id, overflow);
    case LTL_NEXT: return gen_block<LTL_NEXT, T, UB>(X, S, ls, le, rs, re, base_id, overflow);
    case LTL_FIN: return gen_block<LTL_FIN, T, UB>(X, S, ls, le, rs, re, base_id, overflow);
    case LTL_GLOB: return gen_block<LTL_GLOB, T, UB>(X, S, ls, le, rs, re, base_id, overflow);
    default: return gen_block<LTL_UNTIL, T, UB>(X, S, ls, le, rs, re, base_id, overflow);
    }
}

template <int T, int UB>
__device__ void run_instance(const BatchArgs &G, const BatchInst &I, BatchOut &O, Shared<T> &S, char *region) {
    Inst X;
    X.W = I.W; X.P = I.P; X.A = I.A; X.nw = I.nw; X.nsw = I.nsw;
    for (int k = 0; k < 8; ++k) X.c[k] = I.cost[k];
    X.max_cost = I.max_cost; X.neg_type = I.neg_type; X.hk = I.hamming_k;
    X.aux = reinterpret_cast<const WordAux *>(G.in + I.off_aux);
    X.saux = reinterpret_cast<const SumAux *>(G.in + I.off_saux);
    X.tab = reinterpret_cast<uint64_t *>(region);
    X.tmask = G.tslots - 1;
    const uint64_t per_cs = (uint64_t)X.nw * 8 + (uint64_t)X.nsw * 32 + 21;
    const uint64_t avail = G.region_bytes - (uint64_t)G.tslots * 16 - 4096;
    int64_t cap = (int64_t)(avail / per_cs);
    cap = min(cap, (int64_t)(G.tslots / 2) - 2 * T);
    cap = cap / 64 * 64;
    X.cap = cap;
    X.ws = cap;
    char *p = region + (size_t)G.tslots * 16;
    X.store = reinterpret_cast<uint64_t *>(p); p += (size_t)cap * X.nw * 8;
    X.sum = reinterpret_cast<uint64_t *>(p); p += (size_t)cap * 4 * X.nsw * 8;
    X.m_fp2 = reinterpret_cast<uint64_t *>(p); p += (size_t)cap * 8;
    X.m_left = reinterpret_cast<int32_t *>(p); p += (size_t)cap * 4;
    X.m_right = reinterpret_cast<int32_t *>(p); p += (size_t)cap * 4;
    X.m_slot = reinterpret_cast<uint32_t *>(p); p += (size_t)cap * 4;
    X.m_op = reinterpret_cast<uint8_t *>(p);

    int status = BATCH_RUNNING, error = 0;
    if (cap < X.A + 64 || X.max_cost + 2 > kBatchLevels) status = BATCH_DEFERRED;

    if (threadIdx.x == 0) {
        S.found = kEmpty;
        S.n = X.A;
        S.coll = S.pfail = 0;
        for (int c = 0; c < kBatchLevels; ++c) S.lvl_s[c] = S.lvl_e[c] = 0;
        S.lvl_e[X.c[0]] = X.A;
    }
    // atoms: words and summaries into positions 0..A-1, then interned with
    // committed values (the first of equal atoms wins)
    if (status == BATCH_RUNNING) {
        const uint64_t *aw = G.in + I.off_atoms, *as = G.in + I.off_sums;
        for (int k = threadIdx.x; k < X.nw * X.A; k += T) X.store[(size_t)(k / X.A) * X.ws + k % X.A] = aw[k];
        for (int k = threadIdx.x; k < 4 * X.nsw * X.A; k += T) X.sum[(size_t)(k / X.A) * X.ws + k % X.A] = as[k];
    }
    __syncthreads();
    if (status == BATCH_RUNNING && (int)threadIdx.x < X.A) {
        const int a = threadIdx.x;
        Acc c0;
        for (int w = 0; w < X.nw; ++w) fold(c0, X.store[(size_t)w * X.ws + a], X.aux + w);
        X.m_slot[a] = tab_insert(X.tab, X.tmask, tab_key(c0.f1, kBatchEpoch), tab_val((uint64_t)a, kBatchEpoch),
                                 kBatchEpoch, &S.pfail);
        X.m_fp2[a] = c0.f2;
        X.m_op[a] = LTL_ATOM;
        X.m_left[a] = X.m_right[a] = -1;
    }
    __syncthreads();

    uint64_t all_ltls = (uint64_t)X.A;
    int cost = X.c[0] + 1;
    int f_op = -1;
    int32_t f_l = -1, f_r = -1;
    uint64_t f_all = 0;
    if (status == BATCH_RUNNING) {
        for (; cost <= X.max_cost; ++cost) {
            if (threadIdx.x == 0) S.lvl_s[cost] = (int32_t)S.n;
            __syncthreads();
            // defer levels that are too large for one workgroup
            uint64_t tot = 0;
            level_blocks(X, S.lvl_s, S.lvl_e, cost, [&](int op, int ls, int le, int rs, int re) {
                tot += block_size(op, ls, le, rs, re);
                return false;
            });
            if ((int64_t)tot > G.defer_limit) { status = BATCH_DEFERRED; break; }
            // precheck
            uint64_t base = all_ltls;
            const bool found = level_blocks(X, S.lvl_s, S.lvl_e, cost, [&](int op, int ls, int le, int rs, int re) {
                const uint64_t sz = block_size(op, ls, le, rs, re);
                if (!sz) return false;
                const bool bin = is_bin(op);
                const int64_t nR = bin ? re - rs : 1, npairs = (int64_t)(le - ls) * nR;
                for (int64_t b0 = 0; b0 < npairs; b0 += T) {
                    const int64_t t = b0 + threadIdx.x;
                    if (t < npairs) {
                        const int64_t sl = ls + t / nR, sr = bin ? rs + t % nR : sl;
                        uint32_t m0, m1;
                        check_one(X, op, sl, sr, m0, m1);
                        const uint64_t id = base + (uint64_t)(op == LTL_UNTIL ? 2 * t : t);
                        if (m0 <= (uint32_t)X.hk) atomicMin(&S.found, (unsigned long long)id);
                        else if (op == LTL_UNTIL && m1 <= (uint32_t)X.hk) atomicMin(&S.found, (unsigned long long)(id + 1));
                    }
                    __syncthreads();
                    if (S.found != kEmpty) {
                        const uint64_t rel = S.found - base;
                        if (bin) {
                            const uint64_t pair = op == LTL_UNTIL ? rel >> 1 : rel;
                            const int32_t a = (int32_t)(ls + (int64_t)(pair / (uint64_t)nR));
                            const int32_t b = (int32_t)(rs + (int64_t)(pair % (uint64_t)nR));
                            const bool swap = op == LTL_UNTIL && (rel & 1);
                            f_l = swap ? b : a;
                            f_r = swap ? a : b;
                        } else {
                            f_l = (int32_t)(ls + (int64_t)rel);
                            f_r = -1;
                        }
                        f_op = op;
                        f_all = base + sz;
                        return true;
                    }
                }
                base += sz;
                return false;
            });
            if (found) { status = BATCH_FOUND; break; }
            // generate the level
            bool overflow = false;
            const bool stop = level_blocks(X, S.lvl_s, S.lvl_e, cost, [&](int op, int ls, int le, int rs, int re) {
                const bool r = gen_any<T, UB>(op, X, S, ls, le, rs, re, all_ltls, &overflow);
                all_ltls += block_size(op, ls, le, rs, re);
                return r;
            });
            if (stop) {
                if (S.coll) { status = BATCH_ERROR; error = 1; }
                else if (S.pfail) { status = BATCH_ERROR; error = 2; }
                else status = BATCH_DEFERRED;
                break;
            }
            if (threadIdx.x == 0) S.lvl_e[cost] = (int32_t)S.n;
            __syncthreads();
        }
        if (status == BATCH_RUNNING) {
            status = BATCH_NOT_FOUND;
            if (cost == X.max_cost + 1) cost--;  // code/ltli6463.cu:1445
        }
    }

    // result, formula gather (breadth first over the parent indices)
    if (threadIdx.x == 0) {
        O.status = status;
        O.error = error;
        O.cost = cost;
        O.n_stored = S.n;
        O.found_id = status == BATCH_FOUND ? (int64_t)S.found : -1;
        O.all_ltls = status == BATCH_FOUND ? f_all : all_ltls;
        O.n_nodes = 0;
        if (status == BATCH_FOUND) {
            int n = 1, h = 0;
            int32_t *kl = S.kl, *kr = S.kr;
            O.nodes[0] = ltl_node{f_op, -1, -1, -1};
            kl[0] = f_l; kr[0] = f_r;
            bool ok = true;
            for (; h < n && ok; ++h) {
                const int32_t kids[2] = {kl[h], kr[h]};
                for (int sd = 0; sd < 2 && ok; ++sd) {
                    if (kids[sd] < 0) continue;
                    if (n >= kBatchNodes) { ok = false; break; }
                    const int32_t k = kids[sd];
                    const int op = X.m_op[k];
                    O.nodes[n] = ltl_node{op, op == LTL_ATOM ? k : -1, -1, -1};
                    kl[n] = op == LTL_ATOM ? -1 : X.m_left[k];
                    kr[n] = op == LTL_ATOM ? -1 : X.m_right[k];
                    if (sd == 0) O.nodes[h].left = n; else O.nodes[h].right = n;
                    ++n;
                }
            }
            if (!ok) { O.status = BATCH_DEFERRED; O.error = 3; }
            O.n_nodes = n;
        }
    }
    // leave the table empty for the next instance: the slots of stored sets
    // are the only ones used on a clean finish; otherwise clear everything
    if (status == BATCH_FOUND || status == BATCH_NOT_FOUND) {
        for (long long s = threadIdx.x; s < S.n; s += T) {
            const uint32_t slot = X.m_slot[s];
            X.tab[2 * (size_t)slot] = kEmpty;
            X.tab[2 * (size_t)slot + 1] = kEmpty;
        }
    } else {
        for (size_t k = threadIdx.x; k < 2 * (size_t)G.tslots; k += T) X.tab[k] = kEmpty;
    }
    __syncthreads();
}

template <int T, int UB>
__global__ __launch_bounds__(T) void k_batch(BatchArgs G) {
    __shared__ Shared<T> S;
    char *region = G.regions + (size_t)blockIdx.x * G.region_bytes;
    for (;;) {
        if (threadIdx.x == 0) S.inst = (int)atomicAdd(G.next, 1u);
        __syncthreads();
        const int q = S.inst;
        __syncthreads();
        if (q >= G.n_inst) return;
        const int i = G.order[q];
        run_instance<T, UB>(G, G.inst[i], G.out[i], S, region);
    }
}

__global__ void k_fill_regions(char *regions, uint64_t region_bytes, uint32_t tslots, int n_regions) {
    const size_t per = 2 * (size_t)tslots, total = per * (size_t)n_regions;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (size_t)gridDim.x * blockDim.x) {
        uint64_t *t = reinterpret_cast<uint64_t *>(regions + (k / per) * region_bytes);
        t[k % per] = kEmpty;
    }
}

}  // namespace

static void launch_b(int tile, int word_block, const dim3 &g, const BatchArgs &a, hipStream_t s) {
#define LTL_BATCH(T, UB) hipLaunchKernelGGL((k_batch<T, UB>), g, dim3(T), 0, s, a)
    if (word_block == 8) {
        if (tile == 256) LTL_BATCH(256, 8);
        else if (tile == 512) LTL_BATCH(512, 8);
        else LTL_BATCH(1024, 8);
    } else {
        if (tile == 256) LTL_BATCH(256, 4);
        else if (tile == 512) LTL_BATCH(512, 4);
        else LTL_BATCH(1024, 4);
    }
#undef LTL_BATCH
}

hipError_t launch_batch(int B, int tile, int word_block, int workgroups, const BatchArgs &a, hipStream_t s) {
    const dim3 g((unsigned)workgroups);
    // instances in the packed layout (traces of at most 64 positions)
    if (B != 1) return hipErrorInvalidValue;
    launch_b(tile, word_block, g, a, s);
    return hipGetLastError();
}

hipError_t launch_fill_regions(char *regions, uint64_t region_bytes, uint32_t tslots, int n_regions, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_regions, dim3(4096), dim3(256), 0, s, regions, region_bytes, tslots, n_regions);
    return hipGetLastError();
}

}  // namespace ltl
