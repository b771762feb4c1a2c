// scorer.hpp -- learned operand ordering: a small BF16 MLP on MFMA that ranks
// (left tile, right tile) work of the binary operators within a cost level.
//
// A tile is 64 consecutive sets of one cost level; a tile pair of a binary
// block covers the 64 x 64 operand pairs (x 2 directions for U) of one left
// and one right tile.  The scorer predicts, per tile pair, the best accuracy
// (1 - misclassified / traces) of any candidate in it; the check pass of a
// level then tests tile pairs in descending score and stops at the first
// round with an accepted candidate.  Every level below has been tested
// completely, so any accepted candidate of this level has minimal cost.
//
// Network: x (24 features, padded to 32) -> 128 (ReLU) -> 128 (ReLU) -> 1.
// Inputs and weights of the two hidden layers enter the matrix cores as BF16
// (v_mfma_f32_16x16x32_bf16, round to nearest even), accumulation is FP32;
// biases, the second hidden activation and the output layer stay FP32.
// Backward: the output-layer gradient in FP32, the hidden-layer gradients
// rounded to BF16 for the matrix cores.  Features of a tile pair:
//   0-5   left tile, over its sets: max and mean of the fraction of positive
//         traces whose position 0 holds; min and mean of that fraction over
//         negative traces; mean fraction of positive, then of negative,
//         traces where the set holds at some position
//   6-11  the same for the right tile
//   12-18 operator one-hot (~ & | X F G U)
//   19    cost of the level / 32
//   20,21 cost of the left / right operand's level / cost of the level
//   22    share of positive traces
//   23    1
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "kernels.hpp"

namespace ltl {

constexpr int kFeat = 24, kIn = 32, kHid = 128;
constexpr int kTileStat = 8;  // floats per tile (6 used)
// parameter layout (floats)
constexpr int P_W1 = 0;                       // [128][32]
constexpr int P_B1 = P_W1 + kHid * kIn;       // [128]
constexpr int P_W2 = P_B1 + kHid;             // [128][128]
constexpr int P_B2 = P_W2 + kHid * kHid;      // [128]
constexpr int P_W3 = P_B2 + kHid;             // [128]
constexpr int P_B3 = P_W3 + kHid;             // [1]
constexpr int P_N = P_B3 + 1;

struct MlpDev {
    float *p;         // fp32 master parameters [P_N]
    float *g;         // gradients [P_N]
    float *m, *v;     // Adam moments [P_N]
    uint16_t *w1b;    // BF16 bits [128][32]
    uint16_t *w2b;    // [128][128]
    uint16_t *w2tb;   // [128][128], transposed
    float *loss;      // [1]
};

// One binary block of a level, as the tile-pair ranking sees it.
struct LBlock {
    int32_t op, lvl_l, lvl_r, pad;
    int64_t ls, nL, rs, nR;   // store index of the operand ranges and their sizes
    uint64_t id_base;         // canonical id of the block's first candidate
    int64_t nTR;              // right tiles
    int64_t e_off;            // first tile-pair entry of the block in the level
};

// What the feature builder needs besides the blocks.
struct FeatCtx {
    const float *tstat;       // per tile: kTileStat floats
    const int64_t *toff;      // first tile of each level in tstat
    float cost, pos_share;
    int32_t W, P;
};

// Tile statistics of the sets [first, first + n) (one level), written to
// tstat[t0 ..].
hipError_t launch_tile_stats(const ChunkArgs &a, int64_t first, int64_t n, int32_t P, int32_t W, float *tstat,
                             hipStream_t s);
// Score entries [0, n_ent) of the level's blocks; keys = scores, vals = entry.
hipError_t launch_score(const MlpDev &m, const LBlock *blk, int nblk, const FeatCtx &f, int64_t n_ent,
                        float *keys, uint32_t *vals, hipStream_t s);
size_t sort_temp_bytes(int64_t n);
hipError_t launch_sort_desc(const float *kin, float *kout, const uint32_t *vin, uint32_t *vout, int64_t n, void *temp,
                            size_t temp_bytes, hipStream_t s);
// Check the tile pairs ent[r0 .. r0 + n): atomicMin of the id of every
// accepted candidate.
hipError_t launch_check_list(int B, const ChunkArgs &a, const LBlock *blk, int nblk, const uint32_t *ent, int64_t r0,
                             int64_t n, hipStream_t s);
// Log rows (features + label) of n_rows tile pairs of block b: row k covers
// block-local entry floor(k * entries / n_rows).
hipError_t launch_log(int B, const ChunkArgs &a, const LBlock *blk, int b, const FeatCtx &f, int64_t n_rows,
                      float *rows, hipStream_t s);

// Training and inference on feature rows (x: stride xs floats, t: stride ts).
hipError_t launch_zero(float *p, int64_t n, hipStream_t s);
hipError_t launch_iota(uint32_t *p, int64_t n, hipStream_t s);
hipError_t launch_grad(const MlpDev &m, const float *x, int xs, const float *t, int ts, int64_t n, float inv_n,
                       hipStream_t s);
hipError_t launch_adam(const MlpDev &m, float lr, float b1, float b2, float eps, int step, hipStream_t s);
hipError_t launch_refresh(const MlpDev &m, hipStream_t s);
hipError_t launch_forward(const MlpDev &m, const float *x, int xs, int64_t n, float *y, hipStream_t s);

}  // namespace ltl

// Host side of a scorer (include/ltl_amd.h: ltl_scorer_*).
struct ltl_scorer {
    int device = 0;
    hipStream_t stream = nullptr;
    ltl::MlpDev m{};
    float *d_f = nullptr;        // p, g, m, v, loss
    uint16_t *d_b = nullptr;     // w1b, w2b, w2tb
    float *d_x = nullptr;        // staging: rows / features
    float *d_y = nullptr;        // staging: scores / targets
    int64_t cap_x = 0, cap_y = 0;
    int step = 0;
};
