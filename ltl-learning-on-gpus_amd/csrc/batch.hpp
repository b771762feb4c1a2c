// batch.hpp -- many independent learning instances in one persistent launch.
//
// Each workgroup takes instances from a device work queue and runs the whole
// enumeration of one instance by itself (cost levels, precheck, fingerprint
// dedup, ordered compaction through a workgroup scan, formula gather), in
// the same canonical order as the chunk kernels, so results are identical to
// ltl_learn.  Instances whose next level would exceed `defer_limit`
// candidates, or the capacity of the workgroup's memory region, are deferred
// to the whole-device path.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/ltl_amd.h"
#include "kernels.hpp"

namespace ltl {

constexpr int kBatchNodes = 64;     // formula nodes gathered on the device
constexpr int kBatchLevels = 64;    // cost levels tracked per instance

enum { BATCH_RUNNING = 0, BATCH_FOUND = 1, BATCH_NOT_FOUND = 2, BATCH_DEFERRED = 3, BATCH_ERROR = 4 };

struct BatchInst {
    int32_t W, P, A, nw, nsw, B;
    int32_t max_cost, neg_type, hamming_k, pad;
    int32_t cost[8];
    int64_t off_aux, off_saux, off_atoms, off_sums;  // u64-word offsets into the input buffer
};

struct BatchOut {
    int32_t status;    // BATCH_*
    int32_t cost;
    uint64_t all_ltls;
    int64_t found_id;
    int64_t n_stored;
    int32_t n_nodes;
    int32_t error;     // 1 collision, 2 table full, 3 too many nodes
    ltl_node nodes[kBatchNodes];
};

struct BatchArgs {
    const BatchInst *inst;
    BatchOut *out;
    const uint64_t *in;
    const int32_t *order;      // queue position -> instance index
    int32_t n_inst;
    int32_t pad;
    char *regions;             // one region per workgroup
    uint64_t region_bytes;
    uint32_t tslots;           // table slots per region (power of two)
    uint32_t pad2;
    unsigned int *next;        // work-queue counter
    int64_t defer_limit;       // candidates per level
};

// tile = threads per workgroup (candidates per batch step): 256, 512, 1024;
// word_block = loads in flight per lane in the word loops: 4 or 8.
// B = 1: the packed layout of the chunk kernels (G.order lists the instances).
hipError_t launch_batch(int B, int tile, int word_block, int workgroups, const BatchArgs &a, hipStream_t s);
hipError_t launch_fill_regions(char *regions, uint64_t region_bytes, uint32_t tslots, int n_regions,
                               hipStream_t s);

}  // namespace ltl
