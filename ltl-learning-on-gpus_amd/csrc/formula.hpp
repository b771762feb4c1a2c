// formula.hpp -- formula trees in the reference output grammar, host side.
//
// Grammar (code/ltli6463.cu:877-928): a letter, ~(f), X(f), F(f), G(f),
// (f)&(g), (f)|(g), (f)U(g).  Evaluation is bit-parallel on one uint64 per
// trace, with the same operator semantics as the device kernels.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace ltl {

// Trace word operators.  Bits at or above the trace length stay zero.
inline uint64_t len_mask(int len) { return len >= 64 ? ~0ull : ((1ull << len) - 1ull); }
inline uint64_t op_not(uint64_t x, int len) { return ~x & len_mask(len); }
inline uint64_t op_next(uint64_t x) { return x >> 1; }
inline uint64_t suffix_or(uint64_t x) {
    x |= x >> 1; x |= x >> 2; x |= x >> 4; x |= x >> 8; x |= x >> 16; x |= x >> 32;
    return x;
}
inline uint64_t op_glob(uint64_t x, int len) { return x & ~suffix_or(~x & len_mask(len)); }
// Strong until: y_k = r_k | (l_k & y_{k+1}), as a doubling prefix.
inline uint64_t op_until(uint64_t l, uint64_t r) {
    uint64_t run = l, y = r;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) { y |= run & (y >> s); run &= run >> s; }
    return y;
}

struct Node {
    int op = 0;     // LTL_* code
    int atom = -1;
    std::unique_ptr<Node> a, b;
};

// Parse; `names` maps letters to atom indices.  Returns nullptr on error.
std::unique_ptr<Node> parse_formula(const std::string &s, const std::string &names);
std::string print_formula(const Node &n, const std::string &names);
int formula_size(const Node &n, const uint16_t *cost);
// cs[t] for every trace t; atoms is atom-major [A][W].
void eval_formula(const Node &n, int W, const uint8_t *lens, const uint64_t *atoms, uint64_t *cs);

}  // namespace ltl
