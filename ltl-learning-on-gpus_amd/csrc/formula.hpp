// formula.hpp -- formula trees in the reference output grammar, host side.
//
// Grammar (code/ltli6463.cu:877-928): a letter, ~(f), X(f), F(f), G(f),
// (f)&(g), (f)|(g), (f)U(g).  Evaluation is bit-parallel on one uint64 per
// trace, or K per trace for traces longer than 64 positions, with the same
// operator semantics as the device kernels.
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

// Positions 64k..64k+63 of a trace of length len.
inline uint64_t word_mask(int len, int k) {
    const int lo = 64 * k;
    if (len <= lo) return 0;
    return len - lo >= 64 ? ~0ull : ((1ull << (len - lo)) - 1ull);
}

// Operator op on one trace of K words (word k = positions 64k..64k+63); y
// is the right operand of & | U.  What crosses a word boundary comes from
// the next higher word: X takes its bit 0, F holds on all of word k below a
// set bit there, G only while the higher words are full, U on the top run
// of l's ones when it holds at the higher word's first position.
inline void trace_op(int op, const uint64_t *x, const uint64_t *y, int K, int len, uint64_t *out) {
    uint64_t c = 0;
    switch (op) {
    case 1:  // ~
        for (int k = 0; k < K; ++k) out[k] = ~x[k] & word_mask(len, k);
        break;
    case 2:  // &
        for (int k = 0; k < K; ++k) out[k] = x[k] & y[k];
        break;
    case 3:  // |
        for (int k = 0; k < K; ++k) out[k] = x[k] | y[k];
        break;
    case 4:  // X
        for (int k = 0; k < K; ++k) out[k] = (x[k] >> 1) | (k + 1 < K ? x[k + 1] << 63 : 0ull);
        break;
    case 5:  // F
        for (int k = K - 1; k >= 0; --k) {
            out[k] = suffix_or(x[k]) | (c ? word_mask(len, k) : 0ull);
            c |= x[k];
        }
        break;
    case 6:  // G
        c = 1;
        for (int k = K - 1; k >= 0; --k) {
            const uint64_t m = word_mask(len, k);
            out[k] = c ? x[k] & ~suffix_or(~x[k] & m) : 0ull;
            c = c && x[k] == m;
        }
        break;
    default:  // U
        for (int k = K - 1; k >= 0; --k) {
            out[k] = op_until(x[k], y[k]) | (c ? ~suffix_or(~x[k]) : 0ull);
            c = out[k] & 1ull;
        }
    }
}

struct Node {
    int op = 0;     // LTL_* code
    int atom = -1;
    std::unique_ptr<Node> a, b;
};

// Atom names: ltl_problem.names is NULL (a, b, c, ... for up to 26 atoms,
// else p0, p1, ...), num_atoms single characters, or a comma-separated list
// of num_atoms names, each a lowercase letter or _ followed by lowercase
// letters, digits or _ (more than 26 propositions).  Returns false if the
// list is malformed.
using AtomNames = std::vector<std::string>;
bool atom_names(const char *names, int num_atoms, AtomNames &out);
// Parse in the reference grammar; an atom is one character when every name
// is one character, else a maximal run of lowercase letters, digits and _.
// Returns nullptr on error.
std::unique_ptr<Node> parse_formula(const std::string &s, const AtomNames &names);
std::string print_formula(const Node &n, const AtomNames &names);
int formula_size(const Node &n, const uint16_t *cost);
// cs[t] for every trace t; atoms is atom-major [A][W].
void eval_formula(const Node &n, int W, const uint8_t *lens, const uint64_t *atoms, uint64_t *cs);
// K words per trace: cs[t*K + k]; atoms [A][W][K].
void eval_formula(const Node &n, int W, int K, const uint8_t *lens, const uint64_t *atoms, uint64_t *cs);

}  // namespace ltl
