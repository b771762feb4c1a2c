// scorer.hip -- tile statistics, the BF16 MLP (forward, backward, Adam) on
// the gfx950 matrix cores, tile-pair scoring, ordered check rounds and the
// rows of the enumeration log the scorer is trained on.
//
// Matrix-core tiles: v_mfma_f32_16x16x32_bf16 on one wave computes
// D[16][16] += A[16][32] * B[32][16].  Lane l (g = l / 16, c = l % 16) holds
// A[c][8g .. 8g+7], B[8g .. 8g+7][c] (eight BF16 values each) and
// D[4g .. 4g+3][c] (four FP32 values).  Activations are staged in LDS in the
// row-major form each product reads as 16-byte vectors: sample-major rows
// for the forward A operands, feature-major rows for the weight gradients,
// whose reduction dimension is the sample.
#include <hipcub/hipcub.hpp>

#include "../../include/ltl_amd.h"
#include "scorer.hpp"
#include "device_ops.hpp"

namespace ltl {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// LDS row strides in BF16 elements: 16-byte multiples whose rows start on
// different banks (80, 272, 144 bytes).
constexpr int XS = 40;   // 32-wide rows
constexpr int HS = 136;  // 128-wide rows
constexpr int TS = 72;   // 64-wide rows (one per feature, 64 samples)

// FP32 -> BF16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t bf16r(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ bf16x8 ld8(const uint16_t *p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(p));
}
__device__ __forceinline__ f32x4 mma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// Orders the LDS accesses of one wave (writes by some lanes, reads by others).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct Fwd {
    float h2[8][4];   // second hidden layer (FP32, after ReLU): sample 4g+i, unit 16t+c
    uint32_t m1;      // bit 4t+i: first hidden pre-activation > 0
    float y[4];       // outputs of samples 4g+i (every lane of group g)
};

// Forward pass of the 16 samples in xs (16 rows, stride XS, BF16 bits).  The
// first hidden layer goes to h1s (16 rows, stride HS) and, if h1t is given,
// to its unit-major copy at sample column s0 (stride TS).
__device__ __forceinline__ void fwd16(const MlpDev &M, const uint16_t *xs, uint16_t *h1s, uint16_t *h1t, int s0,
                                      Fwd &f) {
    const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
    const float *p = M.p;
    wave_sync();
    const bf16x8 a = ld8(xs + c * XS + 8 * g);
    f.m1 = 0;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        f32x4 z = {0.f, 0.f, 0.f, 0.f};
        z = mma(a, ld8(M.w1b + (16 * t + c) * kIn + 8 * g), z);
        const float bias = p[P_B1 + 16 * t + c];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float v = z[i] + bias;
            const bool on = v > 0.f;
            f.m1 |= (uint32_t)on << (4 * t + i);
            const uint16_t hb = bf16r(on ? v : 0.f);
            h1s[(4 * g + i) * HS + 16 * t + c] = hb;
            if (h1t) h1t[(16 * t + c) * TS + s0 + 4 * g + i] = hb;
        }
    }
    wave_sync();
    f32x4 acc[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        const bf16x8 a2 = ld8(h1s + c * HS + 32 * kb + 8 * g);
#pragma unroll
        for (int t = 0; t < 8; ++t) acc[t] = mma(a2, ld8(M.w2b + (16 * t + c) * kHid + 32 * kb + 8 * g), acc[t]);
    }
    float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const float bias = p[P_B2 + 16 * t + c], w3 = p[P_W3 + 16 * t + c];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float v = acc[t][i] + bias;
            v = v > 0.f ? v : 0.f;
            f.h2[t][i] = v;
            part[i] += v * w3;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float s = part[i];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        s += __shfl_xor(s, 4);
        s += __shfl_xor(s, 8);
        f.y[i] = s + p[P_B3];
    }
}

// Forward pass of 64 samples (rows of xs) by one wave: four 16-row groups
// share every weight fragment, so each B load feeds four products.  y[q][i]
// is the output of sample 16q + 4g + i (every lane of group g).
__device__ __forceinline__ void fwd64(const MlpDev &M, const uint16_t *xs, uint16_t *h1s, float (&y)[4][4]) {
    const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
    const float *p = M.p;
    wave_sync();
    bf16x8 a[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) a[q] = ld8(xs + (16 * q + c) * XS + 8 * g);
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const bf16x8 b = ld8(M.w1b + (16 * t + c) * kIn + 8 * g);
        const float bias = p[P_B1 + 16 * t + c];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            f32x4 z = {0.f, 0.f, 0.f, 0.f};
            z = mma(a[q], b, z);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float v = z[i] + bias;
                h1s[(16 * q + 4 * g + i) * HS + 16 * t + c] = bf16r(v > 0.f ? v : 0.f);
            }
        }
    }
    wave_sync();
    f32x4 acc[4][8];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int t = 0; t < 8; ++t) acc[q][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        bf16x8 a2[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) a2[q] = ld8(h1s + (16 * q + c) * HS + 32 * kb + 8 * g);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const bf16x8 b = ld8(M.w2b + (16 * t + c) * kHid + 32 * kb + 8 * g);
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q][t] = mma(a2[q], b, acc[q][t]);
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const float bias = p[P_B2 + 16 * t + c], w3 = p[P_W3 + 16 * t + c];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v = acc[q][t][i] + bias;
                v = v > 0.f ? v : 0.f;
                part[i] += v * w3;
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float s = part[i];
            s += __shfl_xor(s, 1);
            s += __shfl_xor(s, 2);
            s += __shfl_xor(s, 4);
            s += __shfl_xor(s, 8);
            y[q][i] = s + p[P_B3];
        }
    }
}

// ---------------------------------------------------------------------------
// training
// ---------------------------------------------------------------------------
// Gradients of sum_s inv_n * (y_s - t_s)^2, added into M.g (and the loss into
// M.loss).  Each workgroup takes batches of 64 samples; wave w runs samples
// 16w .. 16w+15 forward and backward, then the four waves form the weight
// gradients of the batch on the matrix cores (wave w: rows 32w .. 32w+31 of
// dW2 and dW1), accumulated in registers over the workgroup's batches.
__global__ __launch_bounds__(256) void k_grad(MlpDev M, const float *X, int xsd, const float *T, int tsd, int64_t n,
                                              float inv_n) {
    __shared__ __align__(16) uint16_t xs[64 * XS];
    __shared__ __align__(16) uint16_t xt[32 * TS];
    __shared__ __align__(16) uint16_t hs[64 * HS];   // first hidden layer, then dL/dz2 (sample-major)
    __shared__ __align__(16) uint16_t h1t[128 * TS];
    __shared__ __align__(16) uint16_t d2t[128 * TS];
    __shared__ __align__(16) uint16_t d1t[128 * TS];
    __shared__ float tg[64];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, c = l & 15;
    const float *p = M.p;
    f32x4 gW2[2][8], gW1[2][2];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi) {
#pragma unroll
        for (int j = 0; j < 8; ++j) gW2[mi][j] = f32x4{0.f, 0.f, 0.f, 0.f};
        gW1[mi][0] = gW1[mi][1] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    float gb1[8], gb2[8], gw3[8], gb3 = 0.f, lsum = 0.f;
#pragma unroll
    for (int t = 0; t < 8; ++t) gb1[t] = gb2[t] = gw3[t] = 0.f;

    for (int64_t base = (int64_t)blockIdx.x * 64; base < n; base += (int64_t)gridDim.x * 64) {
        __syncthreads();
        {
            const int s = threadIdx.x >> 2, f0 = 8 * (threadIdx.x & 3);
            const int64_t r = base + s;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int f = f0 + j;
                const uint16_t b = bf16r(r < n && f < kFeat ? X[r * xsd + f] : 0.f);
                xs[s * XS + f] = b;
                xt[f * TS + s] = b;
            }
            if ((threadIdx.x & 3) == 0) tg[s] = r < n ? T[r * tsd] : 0.f;
        }
        __syncthreads();
        Fwd f;
        fwd16(M, xs + 16 * w * XS, hs + 16 * w * HS, h1t, 16 * w, f);
        float dy[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int s = 16 * w + 4 * g + i;
            const float d = base + s < n ? f.y[i] - tg[s] : 0.f;
            dy[i] = 2.f * inv_n * d;
            if (c == 0) {
                lsum += d * d * inv_n;
                gb3 += dy[i];
            }
        }
        wave_sync();  // layer-2 reads of the wave's rows of hs are done
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const float w3 = p[P_W3 + 16 * t + c];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                gw3[t] += dy[i] * f.h2[t][i];
                const float d = f.h2[t][i] > 0.f ? dy[i] * w3 : 0.f;
                gb2[t] += d;
                const uint16_t b = bf16r(d);
                hs[(16 * w + 4 * g + i) * HS + 16 * t + c] = b;
                d2t[(16 * t + c) * TS + 16 * w + 4 * g + i] = b;
            }
        }
        wave_sync();
        f32x4 acc[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            const bf16x8 a = ld8(hs + (16 * w + c) * HS + 32 * kb + 8 * g);
#pragma unroll
            for (int t = 0; t < 8; ++t) acc[t] = mma(a, ld8(M.w2tb + (16 * t + c) * kHid + 32 * kb + 8 * g), acc[t]);
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float d = (f.m1 >> (4 * t + i)) & 1u ? acc[t][i] : 0.f;
                gb1[t] += d;
                d1t[(16 * t + c) * TS + 16 * w + 4 * g + i] = bf16r(d);
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
            for (int mi = 0; mi < 2; ++mi) {
                const int m = 2 * w + mi;
                const bf16x8 a2 = ld8(d2t + (16 * m + c) * TS + 32 * ks + 8 * g);
#pragma unroll
                for (int j = 0; j < 8; ++j) gW2[mi][j] = mma(a2, ld8(h1t + (16 * j + c) * TS + 32 * ks + 8 * g), gW2[mi][j]);
                const bf16x8 a1 = ld8(d1t + (16 * m + c) * TS + 32 * ks + 8 * g);
#pragma unroll
                for (int j = 0; j < 2; ++j) gW1[mi][j] = mma(a1, ld8(xt + (16 * j + c) * TS + 32 * ks + 8 * g), gW1[mi][j]);
            }
        }
    }
#pragma unroll
    for (int mi = 0; mi < 2; ++mi) {
        const int m = 2 * w + mi;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int j = 0; j < 8; ++j) atomicAdd(&M.g[P_W2 + (16 * m + 4 * g + i) * kHid + 16 * j + c], gW2[mi][j][i]);
#pragma unroll
            for (int j = 0; j < 2; ++j) atomicAdd(&M.g[P_W1 + (16 * m + 4 * g + i) * kIn + 16 * j + c], gW1[mi][j][i]);
        }
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        float v1 = gb1[t], v2 = gb2[t], v3 = gw3[t];
        v1 += __shfl_xor(v1, 16); v1 += __shfl_xor(v1, 32);
        v2 += __shfl_xor(v2, 16); v2 += __shfl_xor(v2, 32);
        v3 += __shfl_xor(v3, 16); v3 += __shfl_xor(v3, 32);
        if (g == 0) {
            atomicAdd(&M.g[P_B1 + 16 * t + c], v1);
            atomicAdd(&M.g[P_B2 + 16 * t + c], v2);
            atomicAdd(&M.g[P_W3 + 16 * t + c], v3);
        }
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        gb3 += __shfl_xor(gb3, o);
        lsum += __shfl_xor(lsum, o);
    }
    if (l == 0) {
        atomicAdd(&M.g[P_B3], gb3);
        atomicAdd(M.loss, lsum);
    }
}

// Adam step (bias-corrected: c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)),
// clearing the gradient and refreshing the BF16 weight copies.
__device__ __forceinline__ void store_bf16(const MlpDev &M, int i, float v) {
    if (i < P_B1) {
        M.w1b[i] = bf16r(v);
    } else if (i >= P_W2 && i < P_B2) {
        const int k = i - P_W2, r = k / kHid, q = k % kHid;
        const uint16_t b = bf16r(v);
        M.w2b[k] = b;
        M.w2tb[q * kHid + r] = b;
    }
}

__global__ void k_adam(MlpDev M, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P_N) return;
    const float gr = M.g[i];
    const float m = b1 * M.m[i] + (1.f - b1) * gr;
    const float v = b2 * M.v[i] + (1.f - b2) * gr * gr;
    M.m[i] = m;
    M.v[i] = v;
    const float pv = M.p[i] - lr * (m * c1) / (sqrtf(v * c2) + eps);
    M.p[i] = pv;
    M.g[i] = 0.f;
    store_bf16(M, i, pv);
}

__global__ void k_refresh(MlpDev M) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P_N) store_bf16(M, i, M.p[i]);
}

__global__ void k_iota(uint32_t *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

__global__ void k_zero(float *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.f;
}

// Forward pass of feature rows: one wave (workgroup) per 64 rows.
__global__ __launch_bounds__(64) void k_forward(MlpDev M, const float *X, int xsd, int64_t n, float *Y) {
    __shared__ __align__(16) uint16_t xs[64 * XS];
    __shared__ __align__(16) uint16_t hs[64 * HS];
    const int l = threadIdx.x;
    const int64_t s0 = (int64_t)blockIdx.x * 64;
    {
        const int64_t r = s0 + l;
        for (int j = 0; j < 32; ++j) xs[l * XS + j] = bf16r(r < n && j < kFeat ? X[r * xsd + j] : 0.f);
    }
    float y[4][4];
    fwd64(M, xs, hs, y);
    const int g = l >> 4, c = l & 15;
    if (c == 0)
        for (int q = 0; q < 4; ++q)
            for (int i = 0; i < 4; ++i)
                if (s0 + 16 * q + 4 * g + i < n) Y[s0 + 16 * q + 4 * g + i] = y[q][i];
}

// ---------------------------------------------------------------------------
// tile statistics, features, scoring
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wmax(float v) {
    for (int o = 1; o < 64; o <<= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wmin(float v) {
    for (int o = 1; o < 64; o <<= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wsum(float v) {
    for (int o = 1; o < 64; o <<= 1) v += __shfl_xor(v, o);
    return v;
}

// One wave per tile of 64 sets, one lane per set, from the summaries
// (vector 0: position 0, vector 2: some position).
__global__ __launch_bounds__(256) void k_tile_stats(ChunkArgs A, int64_t first, int64_t n, float invP, float invN,
                                                    float *out) {
    const int64_t tile = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile * 64 >= n) return;
    const int l = threadIdx.x & 63;
    const int64_t k = tile * 64 + l;
    const bool ok = k < n;
    const int64_t s = first + (ok ? k : tile * 64);
    uint32_t p0 = 0, n0 = 0, pz = 0, nz = 0;
    const uint64_t *zs = A.sum + set_base(s, 4 * A.nsw);
    for (int j = 0; j < A.nsw; ++j) {
        const uint64_t tg = A.saux[j].tgt, ng = A.saux[j].val & ~tg;
        const uint64_t b0 = zs[(size_t)j * A.ws];
        const uint64_t z = zs[(size_t)(2 * A.nsw + j) * A.ws];
        p0 += __popcll(b0 & tg);
        n0 += __popcll(b0 & ng);
        pz += __popcll(z & tg);
        nz += __popcll(z & ng);
    }
    const float fp0 = p0 * invP, fn0 = n0 * invN;
    const float mx = wmax(ok ? fp0 : -1.f), mn = wmin(ok ? fn0 : 2.f);
    const float s0 = wsum(ok ? fp0 : 0.f), s1 = wsum(ok ? fn0 : 0.f);
    const float s2 = wsum(ok ? pz * invP : 0.f), s3 = wsum(ok ? nz * invN : 0.f);
    if (l == 0) {
        const float inv = 1.f / (float)min<int64_t>(64, n - tile * 64);
        float *o = out + tile * kTileStat;
        o[0] = mx; o[1] = s0 * inv; o[2] = mn; o[3] = s1 * inv; o[4] = s2 * inv; o[5] = s3 * inv;
        o[6] = 0.f; o[7] = 0.f;
    }
}

__device__ __forceinline__ int find_block(const LBlock *blk, int nblk, int64_t e) {
    int lo = 0, hi = nblk - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (blk[mid].e_off <= e) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ void features(const LBlock &K, const FeatCtx &F, int64_t loc, float *x) {
    const int64_t ta = loc / K.nTR, tb = loc - ta * K.nTR;
    const float *L = F.tstat + (F.toff[K.lvl_l] + ta) * kTileStat;
    const float *R = F.tstat + (F.toff[K.lvl_r] + tb) * kTileStat;
    for (int j = 0; j < 6; ++j) {
        x[j] = L[j];
        x[6 + j] = R[j];
    }
    for (int j = 0; j < 7; ++j) x[12 + j] = K.op - 1 == j ? 1.f : 0.f;
    x[19] = F.cost * (1.f / 32.f);
    x[20] = (float)K.lvl_l / F.cost;
    x[21] = (float)K.lvl_r / F.cost;
    x[22] = F.pos_share;
    x[23] = 1.f;
}

// One wave (workgroup) per 64 tile pairs; each lane builds one entry's
// features.
__global__ __launch_bounds__(64) void k_score(MlpDev M, const LBlock *blk, int nblk, FeatCtx F, int64_t n_ent,
                                              float *keys, uint32_t *vals) {
    __shared__ __align__(16) uint16_t xs[64 * XS];
    __shared__ __align__(16) uint16_t hs[64 * HS];
    const int l = threadIdx.x;
    const int64_t e0 = (int64_t)blockIdx.x * 64;
    {
        const int64_t e = e0 + l;
        float x[kFeat];
        if (e < n_ent) {
            const int b = find_block(blk, nblk, e);
            features(blk[b], F, e - blk[b].e_off, x);
        } else {
            for (int j = 0; j < kFeat; ++j) x[j] = 0.f;
        }
        for (int j = 0; j < 32; ++j) xs[l * XS + j] = j < kFeat ? bf16r(x[j]) : (uint16_t)0;
    }
    float y[4][4];
    fwd64(M, xs, hs, y);
    const int g = l >> 4, c = l & 15;
    if (c == 0)
        for (int q = 0; q < 4; ++q)
            for (int i = 0; i < 4; ++i) {
                const int64_t e = e0 + 16 * q + 4 * g + i;
                if (e < n_ent) {
                    keys[e] = y[q][i];
                    vals[e] = (uint32_t)e;
                }
            }
}

// ---------------------------------------------------------------------------
// ordered check rounds and log rows
// ---------------------------------------------------------------------------
// Mismatches of one operand pair (U: both directions) and the number of
// traces tested.  EARLY: stop once no lane of the wave can still be
// accepted (m <= stop) or, with anti, have its complement accepted
// (seen - m <= stop).
template <int B, bool EARLY>
__device__ __forceinline__ void pair_mism(const ChunkArgs &A, int op, int64_t sl, int64_t sr, uint32_t stop,
                                          bool anti, uint32_t &m0, uint32_t &m1, uint32_t &seen) {
    if (B == 0 && op == LTL_UNTIL) {
        // multi-word traces: one trace at a time, decided in its first words
        const uint64_t *pl = A.store + set_base(sl, A.nw), *pr = A.store + set_base(sr, A.nw);
        const int K = A.K, T = A.nw / K;
        for (int j = 0; j < T; ++j) {
            const int i = EARLY ? interleave(j, A.npw, T) : j;
            uint32_t h0, h1;
            until_at0_trace(pl + (size_t)i * K * A.ws, pr + (size_t)i * K * A.ws, A.ws, K, h0, h1);
            const uint32_t tg = (uint32_t)(A.aux[i * K].tgt & 1ull);
            m0 += h0 ^ tg;
            m1 += h1 ^ tg;
            ++seen;
            if (EARLY && (j & 3) == 3 &&
                !__any(m0 <= stop || m1 <= stop || (anti && (seen - m0 <= stop || seen - m1 <= stop))))
                return;
        }
    } else if (op == LTL_UNTIL) {
        const uint64_t *pl = A.store + set_base(sl, A.nw), *pr = A.store + set_base(sr, A.nw);
        constexpr int UB = 4;
        int i = 0;
        for (; i + UB <= A.nw; i += UB) {
            uint64_t lv[UB], rv[UB];
            int wu[UB];
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                wu[u] = EARLY ? interleave(i + u, A.npw, A.nw) : i + u;
                lv[u] = pl[(size_t)wu[u] * A.ws];
                rv[u] = pr[(size_t)wu[u] * A.ws];
            }
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                const WordAux *x = A.aux + wu[u];
                m0 += (uint32_t)__popcll((pk_until(lv[u], rv[u], x->inner) ^ x->tgt) & x->chk);
                m1 += (uint32_t)__popcll((pk_until(rv[u], lv[u], x->inner) ^ x->tgt) & x->chk);
                seen += (uint32_t)__popcll(x->chk);
            }
            if (EARLY && !__any(m0 <= stop || m1 <= stop || (anti && (seen - m0 <= stop || seen - m1 <= stop))))
                return;
        }
        for (; i < A.nw; ++i) {
            const int w = EARLY ? interleave(i, A.npw, A.nw) : i;
            const WordAux *x = A.aux + w;
            const uint64_t lv = pl[(size_t)w * A.ws], rv = pr[(size_t)w * A.ws];
            m0 += (uint32_t)__popcll((pk_until(lv, rv, x->inner) ^ x->tgt) & x->chk);
            m1 += (uint32_t)__popcll((pk_until(rv, lv, x->inner) ^ x->tgt) & x->chk);
            seen += (uint32_t)__popcll(x->chk);
        }
    } else {
        const uint64_t *pl = A.sum + set_base(sl, 4 * A.nsw), *pr = A.sum + set_base(sr, 4 * A.nsw);
        for (int i = 0; i < A.nsw; ++i) {
            const int j = EARLY ? interleave(i, A.nps, A.nsw) : i;
            const uint64_t a = pl[(size_t)j * A.ws], b = pr[(size_t)j * A.ws];
            const uint64_t v = op == LTL_AND ? (a & b) : (a | b);
            m0 += (uint32_t)__popcll((v ^ A.saux[j].tgt) & A.saux[j].val);
            seen += (uint32_t)__popcll(A.saux[j].val);
            if (EARLY && !__any(m0 <= stop || (anti && seen - m0 <= stop))) return;
        }
        m1 = 0x7fffffffu;
    }
}

// Sixteen workgroups per tile pair, one wave per left row of the tile
// (row 4 * (blockIdx % 16) + wave), lane = right column.
template <int B>
__global__ __launch_bounds__(256) void k_chk_list(ChunkArgs A, const LBlock *blk, int nblk, const uint32_t *ent,
                                                  int64_t r0) {
    if (A.found_below && *(volatile const unsigned long long *)&A.st->found < A.found_below) return;
    const int64_t e = ent[r0 + (int64_t)(blockIdx.x >> 4)];
    const LBlock &K = blk[find_block(blk, nblk, e)];
    const int64_t loc = e - K.e_off, ta = loc / K.nTR, tb = loc - ta * K.nTR;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int64_t row = ta * 64 + 4 * (int64_t)(blockIdx.x & 15) + w;
    if (row >= K.nL) return;
    const int64_t col = tb * 64 + l;
    const bool valid = col < K.nR;
    const uint32_t k = (uint32_t)A.hamming_k;
    const bool anti = A.anti && valid;
    uint32_t m0 = valid ? 0u : 0x7fffffffu, m1 = m0, seen = 0;
    pair_mism<B, true>(A, K.op, K.ls + row, K.rs + (valid ? col : tb * 64), k, anti, m0, m1, seen);
    if (!valid) return;
    const uint64_t pair = (uint64_t)row * (uint64_t)K.nR + (uint64_t)col;
    const uint64_t id = K.id_base + (K.op == LTL_UNTIL ? 2 * pair : pair);
    if (m0 <= k) atomicMin(&A.st->found, (unsigned long long)id);
    else if (K.op == LTL_UNTIL && m1 <= k) atomicMin(&A.st->found, (unsigned long long)(id + 1));
    else if (anti && seen - m0 <= k) atomicMin(&A.st->neg_found[A.neg_slot], (unsigned long long)id);
    else if (anti && K.op == LTL_UNTIL && seen - m1 <= k)
        atomicMin(&A.st->neg_found[A.neg_slot], (unsigned long long)(id + 1));
}

// One workgroup per log row: the fewest mismatches over the tile pair's
// candidates (all words, no early exit) -> label 1 - min / W.
template <int B>
__global__ __launch_bounds__(256) void k_log(ChunkArgs A, const LBlock *blk, int b, FeatCtx F, int64_t n_rows,
                                             float *rows) {
    __shared__ uint32_t red[4];
    const LBlock &K = blk[b];
    const int64_t E = (K.nL + 63) / 64 * K.nTR;
    const int64_t loc = (int64_t)blockIdx.x * E / n_rows;
    const int64_t ta = loc / K.nTR, tb = loc - ta * K.nTR;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int64_t col = tb * 64 + l;
    const bool valid = col < K.nR;
    const int64_t sr = K.rs + (valid ? col : tb * 64);
    uint32_t best = 0xffffffffu;
    for (int r = 0; r < 16; ++r) {
        const int64_t row = ta * 64 + 16 * w + r;
        if (row >= K.nL) break;
        uint32_t m0 = 0, m1 = 0, seen = 0;
        pair_mism<B, false>(A, K.op, K.ls + row, sr, 0, false, m0, m1, seen);
        if (valid) best = min(best, min(m0, m1));
    }
    for (int o = 1; o < 64; o <<= 1) best = min(best, (uint32_t)__shfl_xor((int)best, o));
    if (l == 0) red[w] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t mm = min(min(red[0], red[1]), min(red[2], red[3]));
        float x[kFeat];
        features(K, F, loc, x);
        float *o = rows + (int64_t)blockIdx.x * (kFeat + 1);
        for (int j = 0; j < kFeat; ++j) o[j] = x[j];
        o[kFeat] = 1.f - (float)mm / (float)F.W;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_tile_stats(const ChunkArgs &a, int64_t first, int64_t n, int32_t P, int32_t W, float *tstat,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t tiles = (n + 63) / 64;
    const float invP = 1.f / (float)(P > 0 ? P : 1), invN = 1.f / (float)(W - P > 0 ? W - P : 1);
    hipLaunchKernelGGL(k_tile_stats, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, s, a, first, n, invP, invN,
                       tstat);
    return hipGetLastError();
}

hipError_t launch_score(const MlpDev &m, const LBlock *blk, int nblk, const FeatCtx &f, int64_t n_ent, float *keys,
                        uint32_t *vals, hipStream_t s) {
    if (n_ent <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_score, dim3((unsigned)((n_ent + 63) / 64)), dim3(64), 0, s, m, blk, nblk, f, n_ent, keys,
                       vals);
    return hipGetLastError();
}

size_t sort_temp_bytes(int64_t n) {
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, (const float *)nullptr, (float *)nullptr,
                                                     (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                     (int)n) != hipSuccess)
        return 0;
    return bytes;
}

hipError_t launch_sort_desc(const float *kin, float *kout, const uint32_t *vin, uint32_t *vout, int64_t n, void *temp,
                            size_t temp_bytes, hipStream_t s) {
    return hipcub::DeviceRadixSort::SortPairsDescending(temp, temp_bytes, kin, kout, vin, vout, (int)n, 0, 32, s);
}

#define LTL_SCORER_B(KER, B, GRID, S, ...)                                          \
    switch (B) {                                                                    \
    case 0: hipLaunchKernelGGL(KER<0>, GRID, dim3(256), 0, S, __VA_ARGS__); break;   \
    case 1: hipLaunchKernelGGL(KER<1>, GRID, dim3(256), 0, S, __VA_ARGS__); break;   \
    default: return hipErrorInvalidValue;                                           \
    }

hipError_t launch_check_list(int B, const ChunkArgs &a, const LBlock *blk, int nblk, const uint32_t *ent, int64_t r0,
                             int64_t n, hipStream_t s) {
    for (int64_t d = 0; d < n; d += (int64_t)1 << 26) {
        const unsigned grid = (unsigned)(std::min<int64_t>(n - d, (int64_t)1 << 26) * 16);
        LTL_SCORER_B(k_chk_list, B, dim3(grid), s, a, blk, nblk, ent, r0 + d);
    }
    return hipGetLastError();
}

hipError_t launch_log(int B, const ChunkArgs &a, const LBlock *blk, int b, const FeatCtx &f, int64_t n_rows,
                      float *rows, hipStream_t s) {
    if (n_rows <= 0) return hipSuccess;
    LTL_SCORER_B(k_log, B, dim3((unsigned)n_rows), s, a, blk, b, f, n_rows, rows);
    return hipGetLastError();
}

hipError_t launch_iota(uint32_t *p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_iota, dim3(blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_zero(float *p, int64_t n, hipStream_t s) {
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_zero, dim3(blocks ? blocks : 1), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_grad(const MlpDev &m, const float *x, int xs, const float *t, int ts, int64_t n, float inv_n,
                       hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const unsigned grid = (unsigned)std::min<int64_t>((n + 63) / 64, 512);
    hipLaunchKernelGGL(k_grad, dim3(grid), dim3(256), 0, s, m, x, xs, t, ts, n, inv_n);
    return hipGetLastError();
}

hipError_t launch_adam(const MlpDev &m, float lr, float b1, float b2, float eps, int step, hipStream_t s) {
    const float c1 = 1.f / (1.f - powf(b1, (float)step)), c2 = 1.f / (1.f - powf(b2, (float)step));
    hipLaunchKernelGGL(k_adam, dim3((P_N + 255) / 256), dim3(256), 0, s, m, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_refresh(const MlpDev &m, hipStream_t s) {
    hipLaunchKernelGGL(k_refresh, dim3((P_N + 255) / 256), dim3(256), 0, s, m);
    return hipGetLastError();
}

hipError_t launch_forward(const MlpDev &m, const float *x, int xs, int64_t n, float *y, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_forward, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, m, x, xs, n, y);
    return hipGetLastError();
}

}  // namespace ltl
