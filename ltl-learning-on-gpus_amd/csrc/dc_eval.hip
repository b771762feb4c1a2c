// dc_eval.hip -- formula matches on the device (see dc_eval.hpp).
#include "dc_eval.hpp"

#include <algorithm>
#include <cstring>

namespace ltl {
namespace {

struct EvalJobDev {
    const uint64_t *atoms;
    const uint8_t *lens;
    uint64_t *out;  // [ceil(W / 64)]: bit t % 64 of word t / 64 = value on trace t
    int32_t W, prog, n_ops, pad;
};

__device__ __forceinline__ uint64_t wmask(int len, int k) {
    const int lo = 64 * k;
    if (len <= lo) return 0ull;
    return len - lo >= 64 ? ~0ull : ((1ull << (len - lo)) - 1ull);
}
// bit p set iff some bit at or above p is
__device__ __forceinline__ uint64_t suffix_or(uint64_t x) {
    x |= x >> 1; x |= x >> 2; x |= x >> 4; x |= x >> 8; x |= x >> 16; x |= x >> 32;
    return x;
}
// y_p = r_p | (l_p & y_{p+1}) within one word, as a doubling prefix
__device__ __forceinline__ uint64_t until_word(uint64_t l, uint64_t r) {
    uint64_t run = l, y = r;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        y |= run & (y >> s);
        run &= run >> s;
    }
    return y;
}

// One thread per trace of job blockIdx.y (a program over a stack of K-word
// trace values); each wave covers the 64 traces of one output word.
template <int K>
__global__ __launch_bounds__(256) void k_match(const EvalJobDev *jobs, const EvalOp *ops) {
    const EvalJobDev J = jobs[blockIdx.y];
    const int t0 = (int)blockIdx.x * 256;
    if (t0 >= J.W) return;
    const int t = t0 + (int)threadIdx.x;
    const bool ok = t < J.W;
    const int len = ok ? J.lens[t] : 0;
    uint64_t st[kEvalDepth][K];
    int sp = 0;
    for (int i = 0; i < J.n_ops; ++i) {
        const EvalOp o = ops[J.prog + i];
        if (o.op == LTL_ATOM) {
            const uint64_t *src = J.atoms + ((size_t)o.arg * J.W + (ok ? t : 0)) * K;
#pragma unroll
            for (int k = 0; k < K; ++k) st[sp][k] = ok ? src[k] & wmask(len, k) : 0ull;
            ++sp;
            continue;
        }
        if (o.op == LTL_NOT || o.op == LTL_NEXT || o.op == LTL_FIN || o.op == LTL_GLOB) {
            uint64_t *x = st[sp - 1];
            if (o.op == LTL_NOT) {
#pragma unroll
                for (int k = 0; k < K; ++k) x[k] = ~x[k] & wmask(len, k);
            } else if (o.op == LTL_NEXT) {
#pragma unroll
                for (int k = 0; k < K; ++k) x[k] = (x[k] >> 1) | (k + 1 < K ? x[k + 1] << 63 : 0ull);
            } else if (o.op == LTL_FIN) {
                uint64_t c = 0;
#pragma unroll
                for (int k = K - 1; k >= 0; --k) {
                    const uint64_t xk = x[k];
                    x[k] = suffix_or(xk) | (c ? wmask(len, k) : 0ull);
                    c |= xk;
                }
            } else {
                bool c = true;
#pragma unroll
                for (int k = K - 1; k >= 0; --k) {
                    const uint64_t xk = x[k], m = wmask(len, k);
                    x[k] = c ? xk & ~suffix_or(~xk & m) : 0ull;
                    c = c && xk == m;
                }
            }
            continue;
        }
        // binary: the first evaluated operand lies below the second
        uint64_t *a = st[sp - 2];
        const uint64_t *b = st[sp - 1];
        if (o.op == LTL_AND) {
#pragma unroll
            for (int k = 0; k < K; ++k) a[k] &= b[k];
        } else if (o.op == LTL_OR) {
#pragma unroll
            for (int k = 0; k < K; ++k) a[k] |= b[k];
        } else {
            const uint64_t *l = o.arg ? b : a, *r = o.arg ? a : b;
            uint64_t c = 0;
#pragma unroll
            for (int k = K - 1; k >= 0; --k) {
                const uint64_t lk = l[k], rk = r[k];
                const uint64_t y = until_word(lk, rk) | (c ? ~suffix_or(~lk) : 0ull);
                a[k] = y;
                c = y & 1ull;
            }
        }
        --sp;
    }
    const unsigned long long word = __ballot(ok && (st[0][0] & 1ull));
    if ((threadIdx.x & 63) == 0 && ok) J.out[t >> 6] = word;
}

int need(const Node &n) {
    if (n.op == LTL_ATOM) return 1;
    if (!n.b) return need(*n.a);
    const int x = need(*n.a), y = need(*n.b);
    return x == y ? x + 1 : std::max(x, y);
}

void emit(const Node &n, std::vector<EvalOp> &ops) {
    if (n.op == LTL_ATOM) {
        ops.push_back({(int16_t)LTL_ATOM, (int16_t)n.atom});
        return;
    }
    if (!n.b) {
        emit(*n.a, ops);
        ops.push_back({(int16_t)n.op, 0});
        return;
    }
    const bool right_first = need(*n.b) > need(*n.a);
    emit(right_first ? *n.b : *n.a, ops);
    emit(right_first ? *n.a : *n.b, ops);
    ops.push_back({(int16_t)n.op, (int16_t)(right_first ? 1 : 0)});
}

}  // namespace

DevTraces::~DevTraces() {
    if (atoms) hipFree(atoms);
    if (lens) hipFree(lens);
}

int32_t dev_traces_upload(const ltl_problem *pb, hipStream_t s, DevTraces &dt) {
    dt.W = pb->num_pos + pb->num_neg;
    dt.K = pb->words_per_trace > 1 ? pb->words_per_trace : 1;
    dt.A = pb->num_atoms;
    if (dt.W < 1 || dt.K > 4 || dt.A < 1) return LTL_E_LIMIT;
    const size_t nb = (size_t)dt.A * dt.W * dt.K * sizeof(uint64_t);
    if (hipMalloc((void **)&dt.atoms, nb) != hipSuccess || hipMalloc((void **)&dt.lens, (size_t)dt.W) != hipSuccess ||
        hipMemcpyAsync(dt.atoms, pb->atoms, nb, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dt.lens, pb->lens, (size_t)dt.W, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return LTL_E_HIP;
    return 0;
}

int compile_formula(const Node &f, std::vector<EvalOp> &ops) {
    const int d = need(f);
    if (d > kEvalDepth) return -1;
    emit(f, ops);
    return d;
}

static bool atoms_below(const Node &n, int A) {
    if (n.op == LTL_ATOM) return n.atom >= 0 && n.atom < A;
    return atoms_below(*n.a, A) && (!n.b || atoms_below(*n.b, A));
}

MatchBuf::~MatchBuf() {
    if (d) hipFree(d);
}

int32_t dev_match(const std::vector<MatchJob> &jobs, hipStream_t s, MatchBuf &buf,
                  std::vector<std::vector<uint8_t>> &out) {
    const int n = (int)jobs.size();
    out.assign(n, {});
    if (n == 0) return 0;
    // jobs grouped by words per trace (one launch each)
    std::vector<int> order(n);
    for (int j = 0; j < n; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return jobs[x].traces->K < jobs[y].traces->K; });
    std::vector<EvalOp> ops;
    std::vector<EvalJobDev> jd(n);
    std::vector<size_t> out_off(n);
    size_t out_words = 0;
    for (int q = 0; q < n; ++q) {
        const MatchJob &J = jobs[order[q]];
        if (!atoms_below(*J.formula, J.traces->A)) return LTL_E_LIMIT;
        const int32_t prog = (int32_t)ops.size();
        if (compile_formula(*J.formula, ops) < 0) return LTL_E_LIMIT;
        jd[q] = {J.traces->atoms, J.traces->lens, nullptr, J.traces->W, prog, (int32_t)(ops.size() - prog), 0};
        out_off[q] = out_words;
        out_words += ((size_t)J.traces->W + 63) / 64;
    }
    // one device buffer: jobs, outputs, program
    const size_t b_jobs = (size_t)n * sizeof(EvalJobDev), b_out = out_words * 8, b_ops = ops.size() * sizeof(EvalOp);
    if (buf.cap < b_jobs + b_out + b_ops) {
        if (buf.d) hipFree(buf.d);
        buf.d = nullptr;
        buf.cap = 0;
        const size_t want = std::max<size_t>(b_jobs + b_out + b_ops, (size_t)1 << 20);
        if (hipMalloc((void **)&buf.d, want) != hipSuccess) return LTL_E_HIP;
        buf.cap = want;
    }
    char *d = buf.d;
    uint64_t *d_out = reinterpret_cast<uint64_t *>(d + b_jobs);
    EvalOp *d_ops = reinterpret_cast<EvalOp *>(d + b_jobs + b_out);
    for (int q = 0; q < n; ++q) jd[q].out = d_out + out_off[q];
    std::vector<uint64_t> h_out(out_words);
    hipError_t e = hipMemcpyAsync(d, jd.data(), b_jobs, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_ops, ops.data(), b_ops, hipMemcpyHostToDevice, s);
    for (int q0 = 0; e == hipSuccess && q0 < n;) {
        const int K = jobs[order[q0]].traces->K;
        int q1 = q0, maxW = 0;
        while (q1 < n && q1 - q0 < 65535 && jobs[order[q1]].traces->K == K) maxW = std::max(maxW, jd[q1++].W);
        const dim3 grid((unsigned)((maxW + 255) / 256), (unsigned)(q1 - q0));
        const EvalJobDev *dj = reinterpret_cast<const EvalJobDev *>(d) + q0;
        switch (K) {
        case 1: hipLaunchKernelGGL(k_match<1>, grid, dim3(256), 0, s, dj, d_ops); break;
        case 2: hipLaunchKernelGGL(k_match<2>, grid, dim3(256), 0, s, dj, d_ops); break;
        case 3: hipLaunchKernelGGL(k_match<3>, grid, dim3(256), 0, s, dj, d_ops); break;
        default: hipLaunchKernelGGL(k_match<4>, grid, dim3(256), 0, s, dj, d_ops); break;
        }
        e = hipGetLastError();
        q0 = q1;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h_out.data(), d_out, b_out, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return LTL_E_HIP;
    for (int q = 0; q < n; ++q) {
        const int W = jd[q].W;
        std::vector<uint8_t> &m = out[order[q]];
        m.resize(W);
        for (int t = 0; t < W; ++t) m[t] = (uint8_t)((h_out[out_off[q] + t / 64] >> (t % 64)) & 1ull);
    }
    return 0;
}

}  // namespace ltl
