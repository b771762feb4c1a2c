// dc_eval.hpp -- formula matches on the device (the reference's LTLmatch,
// code/dc.py:136-186: the value at position 0 of a formula on every trace).
//
// A problem's traces are copied to the device once (DevTraces); formulas are
// compiled to stack programs (postfix, the child that needs the deeper stack
// first, so that a formula of n nodes needs at most log2(n) + 1 entries) and
// evaluated one thread per (formula, trace), many formulas per launch, with
// the operator semantics of the enumeration kernels (positions past the
// trace length are false; what crosses a 64-position word comes from the
// next higher word).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "../../include/ltl_amd.h"
#include "formula.hpp"

namespace ltl {

constexpr int kEvalDepth = 16;  // stack entries of a program

struct EvalOp {
    int16_t op;   // LTL_* code
    int16_t arg;  // atom index (LTL_ATOM); binary: 1 = right operand evaluated first
};

// A problem's traces on the device.
struct DevTraces {
    uint64_t *atoms = nullptr;  // [A][W][K], as ltl_problem.atoms
    uint8_t *lens = nullptr;    // [W]
    int32_t W = 0, K = 1, A = 0;
    DevTraces() = default;
    DevTraces(const DevTraces &) = delete;
    DevTraces &operator=(const DevTraces &) = delete;
    ~DevTraces();
};

// Copy pb's traces to the current device (hipMalloc; synchronous on s).
int32_t dev_traces_upload(const ltl_problem *pb, hipStream_t s, DevTraces &dt);

// Stack program of a formula; returns its stack depth, or -1 if it exceeds
// kEvalDepth.
int compile_formula(const Node &f, std::vector<EvalOp> &ops);

// One formula on one trace set.
struct MatchJob {
    const DevTraces *traces;
    const Node *formula;
};

// Device staging of dev_match calls (grows, kept between calls).
struct MatchBuf {
    char *d = nullptr;
    size_t cap = 0;
    MatchBuf() = default;
    MatchBuf(const MatchBuf &) = delete;
    MatchBuf &operator=(const MatchBuf &) = delete;
    ~MatchBuf();
};

// Position-0 values: out[j][t] = 1 iff job j's formula holds at position 0
// of trace t.  One upload, one launch per words-per-trace width, one
// download; synchronous on s.  Returns 0, LTL_E_LIMIT (a formula deeper than
// the stack, or an atom outside the trace set) or LTL_E_HIP.
int32_t dev_match(const std::vector<MatchJob> &jobs, hipStream_t s, MatchBuf &buf,
                  std::vector<std::vector<uint8_t>> &out);

}  // namespace ltl
