// host_util.cpp -- host-side helpers around the boundary: formula trees,
// trace-file reading, the seeded workload generator.  No GPU code.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ltl_amd.h"
#include "../../include/ltl_gen.h"
#include "formula.hpp"

namespace ltl {

static bool is_name_char(char c) { return (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9') || c == '_'; }

bool atom_names(const char *names, int A, AtomNames &out) {
    out.clear();
    if (A < 0) return false;
    if (!names) {
        for (int a = 0; a < A; ++a)
            out.push_back(A <= 26 ? std::string(1, (char)('a' + a)) : "p" + std::to_string(a));
        return true;
    }
    if (!strchr(names, ',')) {
        if ((int)strlen(names) < A) return false;
        for (int a = 0; a < A; ++a) out.push_back(std::string(1, names[a]));
        return true;
    }
    std::string cur;
    for (const char *p = names;; ++p) {
        if (*p == ',' || *p == 0) {
            if (cur.empty() || !((cur[0] >= 'a' && cur[0] <= 'z') || cur[0] == '_')) return false;
            for (char c : cur)
                if (!is_name_char(c)) return false;
            for (const auto &o : out)
                if (o == cur) return false;
            out.push_back(cur);
            cur.clear();
            if (*p == 0) break;
        } else {
            cur += *p;
        }
    }
    return (int)out.size() == A;
}

namespace {
struct Parser {
    const std::string &s;
    const AtomNames &names;
    bool single;  // every name is one character
    size_t p = 0;
    bool eat(char c) {
        if (p < s.size() && s[p] == c) { ++p; return true; }
        return false;
    }
    std::unique_ptr<Node> f() {
        if (p >= s.size()) return nullptr;
        char c = s[p];
        if (c == '~' || c == 'X' || c == 'F' || c == 'G') {
            ++p;
            if (!eat('(')) return nullptr;
            auto k = f();
            if (!k || !eat(')')) return nullptr;
            auto n = std::make_unique<Node>();
            n->op = c == '~' ? LTL_NOT : c == 'X' ? LTL_NEXT : c == 'F' ? LTL_FIN : LTL_GLOB;
            n->a = std::move(k);
            return n;
        }
        if (c == '(') {
            ++p;
            auto l = f();
            if (!l || !eat(')') || p >= s.size()) return nullptr;
            char o = s[p++];
            int op = o == '&' ? LTL_AND : o == '|' ? LTL_OR : o == 'U' ? LTL_UNTIL : -1;
            if (op < 0 || !eat('(')) return nullptr;
            auto r = f();
            if (!r || !eat(')')) return nullptr;
            auto n = std::make_unique<Node>();
            n->op = op; n->a = std::move(l); n->b = std::move(r);
            return n;
        }
        std::string tok;
        if (single) {
            tok = std::string(1, c);
        } else {
            for (size_t q = p; q < s.size() && is_name_char(s[q]); ++q) tok += s[q];
        }
        size_t k = 0;
        while (k < names.size() && names[k] != tok) ++k;
        if (tok.empty() || k == names.size()) return nullptr;
        p += tok.size();
        auto n = std::make_unique<Node>();
        n->op = LTL_ATOM; n->atom = (int)k;
        return n;
    }
};
}  // namespace

std::unique_ptr<Node> parse_formula(const std::string &s, const AtomNames &names) {
    bool single = true;
    for (const auto &n : names) single = single && n.size() == 1;
    Parser ps{s, names, single};
    auto n = ps.f();
    if (!n || ps.p != s.size()) return nullptr;
    return n;
}

std::string print_formula(const Node &n, const AtomNames &names) {
    switch (n.op) {
    case LTL_ATOM: return names[n.atom];
    case LTL_NOT: return "~(" + print_formula(*n.a, names) + ")";
    case LTL_NEXT: return "X(" + print_formula(*n.a, names) + ")";
    case LTL_FIN: return "F(" + print_formula(*n.a, names) + ")";
    case LTL_GLOB: return "G(" + print_formula(*n.a, names) + ")";
    default: {
        const char *mid = n.op == LTL_AND ? ")&(" : n.op == LTL_OR ? ")|(" : ")U(";
        return "(" + print_formula(*n.a, names) + mid + print_formula(*n.b, names) + ")";
    }
    }
}

int formula_size(const Node &n, const uint16_t *cost) {
    int c = cost[n.op];
    if (n.a) c += formula_size(*n.a, cost);
    if (n.b) c += formula_size(*n.b, cost);
    return c;
}

void eval_formula(const Node &n, int W, int K, const uint8_t *lens, const uint64_t *atoms, uint64_t *cs) {
    if (K <= 1) {
        eval_formula(n, W, lens, atoms, cs);
        return;
    }
    const size_t nw = (size_t)W * K;
    if (n.op == LTL_ATOM) {
        for (int t = 0; t < W; ++t)
            for (int k = 0; k < K; ++k)
                cs[(size_t)t * K + k] = atoms[(size_t)n.atom * nw + (size_t)t * K + k] & word_mask(lens[t], k);
        return;
    }
    std::vector<uint64_t> a(nw), b;
    eval_formula(*n.a, W, K, lens, atoms, a.data());
    if (n.b) { b.resize(nw); eval_formula(*n.b, W, K, lens, atoms, b.data()); }
    for (int t = 0; t < W; ++t)
        trace_op(n.op, a.data() + (size_t)t * K, n.b ? b.data() + (size_t)t * K : nullptr, K, lens[t],
                 cs + (size_t)t * K);
}

void eval_formula(const Node &n, int W, const uint8_t *lens, const uint64_t *atoms, uint64_t *cs) {
    if (n.op == LTL_ATOM) {
        for (int t = 0; t < W; ++t) cs[t] = atoms[(size_t)n.atom * W + t] & len_mask(lens[t]);
        return;
    }
    std::vector<uint64_t> a(W), b;
    eval_formula(*n.a, W, lens, atoms, a.data());
    if (n.b) { b.resize(W); eval_formula(*n.b, W, lens, atoms, b.data()); }
    for (int t = 0; t < W; ++t) {
        switch (n.op) {
        case LTL_NOT: cs[t] = op_not(a[t], lens[t]); break;
        case LTL_AND: cs[t] = a[t] & b[t]; break;
        case LTL_OR: cs[t] = a[t] | b[t]; break;
        case LTL_NEXT: cs[t] = op_next(a[t]); break;
        case LTL_FIN: cs[t] = suffix_or(a[t]); break;
        case LTL_GLOB: cs[t] = op_glob(a[t], lens[t]); break;
        default: cs[t] = op_until(a[t], b[t]); break;
        }
    }
}

// ---------------------------------------------------------------------------
// Random numbers: SplitMix64 seeding xoshiro256** (Vigna, public domain).
// ---------------------------------------------------------------------------
struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x) {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) { for (auto &w : s) w = splitmix(seed); }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    int range(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }  // inclusive
};

namespace {

std::unique_ptr<Node> random_formula(Rng &g, int size, int A) {
    auto n = std::make_unique<Node>();
    if (size <= 1) { n->op = LTL_ATOM; n->atom = g.range(0, A - 1); return n; }
    static const int un[4] = {LTL_NOT, LTL_NEXT, LTL_FIN, LTL_GLOB};
    static const int bn[3] = {LTL_AND, LTL_OR, LTL_UNTIL};
    if (size == 2 || g.uniform() < 0.4) {
        n->op = un[g.range(0, 3)];
        n->a = random_formula(g, size - 1, A);
    } else {
        int ls = g.range(1, size - 2);
        n->op = bn[g.range(0, 2)];
        n->a = random_formula(g, ls, A);
        n->b = random_formula(g, size - 1 - ls, A);
    }
    return n;
}

struct Draw {
    int P, N, A, len_lo, len_hi;
    std::vector<double> rate;  // per atom
};

// Draw traces until both classes are full.  Returns false if `budget`
// draws did not fill them (the formula is then too skewed).
bool fill_classes(Rng &g, const Draw &d, const Node &f, std::vector<uint8_t> &lens,
                  std::vector<uint64_t> &atoms, long budget) {
    const int W = d.P + d.N;
    std::vector<std::vector<uint64_t>> pos, neg;  // per trace: [A] words
    std::vector<uint8_t> plen, nlen;
    std::vector<uint64_t> tr(d.A);
    uint64_t cs;
    uint8_t len8;
    for (long it = 0; it < budget && ((int)pos.size() < d.P || (int)neg.size() < d.N); ++it) {
        int len = g.range(d.len_lo, d.len_hi);
        for (int a = 0; a < d.A; ++a) {
            uint64_t w = 0;
            for (int k = 0; k < len; ++k)
                if (g.uniform() < d.rate[a]) w |= 1ull << k;
            tr[a] = w;
        }
        len8 = (uint8_t)len;
        eval_formula(f, 1, &len8, tr.data(), &cs);
        if (cs & 1) {
            if ((int)pos.size() < d.P) { pos.push_back(tr); plen.push_back(len8); }
        } else if ((int)neg.size() < d.N) {
            neg.push_back(tr); nlen.push_back(len8);
        }
    }
    if ((int)pos.size() < d.P || (int)neg.size() < d.N) return false;
    lens.assign(W, 0);
    atoms.assign((size_t)d.A * W, 0);
    for (int t = 0; t < W; ++t) {
        const auto &src = t < d.P ? pos[t] : neg[t - d.P];
        lens[t] = t < d.P ? plen[t] : nlen[t - d.P];
        for (int a = 0; a < d.A; ++a) atoms[(size_t)a * W + t] = src[a];
    }
    return true;
}

void emit(ltl_gen_instance *out, const Draw &d, const std::vector<uint8_t> &lens,
          const std::vector<uint64_t> &atoms, const std::string &planted) {
    const int W = d.P + d.N;
    out->num_pos = d.P; out->num_neg = d.N; out->num_atoms = d.A;
    out->lens = (uint8_t *)malloc(W);
    out->atoms = (uint64_t *)malloc((size_t)d.A * W * 8);
    memcpy(out->lens, lens.data(), W);
    memcpy(out->atoms, atoms.data(), (size_t)d.A * W * 8);
    for (int a = 0; a < d.A; ++a) out->names[a] = (char)('a' + a);
    out->names[d.A] = 0;
    snprintf(out->planted, sizeof(out->planted), "%s", planted.c_str());
}

// Letter names a, b, c, ... of n <= 26 atoms.
AtomNames letters(int n) {
    AtomNames v;
    for (int a = 0; a < n; ++a) v.push_back(std::string(1, (char)('a' + a)));
    return v;
}

}  // namespace
}  // namespace ltl

using namespace ltl;

extern "C" int32_t ltl_gen_custom(uint64_t seed, int32_t num_pos, int32_t num_neg, int32_t num_atoms,
                                  int32_t len_min, int32_t len_max, double bit_rate,
                                  const char *planted, int32_t planted_size, ltl_gen_instance *out) {
    memset(out, 0, sizeof(*out));
    if (num_pos < 1 || num_neg < 1 || num_atoms < 1 || num_atoms > 26 || len_min < 1 || len_max > 64 ||
        len_min > len_max)
        return LTL_E_ARG;
    Rng g(seed);
    Draw d{num_pos, num_neg, num_atoms, len_min, len_max, std::vector<double>(num_atoms, bit_rate)};
    const AtomNames names = letters(num_atoms);
    std::vector<uint8_t> lens;
    std::vector<uint64_t> atoms;
    for (int attempt = 0; attempt < 1000; ++attempt) {
        std::unique_ptr<Node> f = planted ? parse_formula(planted, names)
                                          : random_formula(g, planted_size, num_atoms);
        if (!f) return LTL_E_ARG;
        if (fill_classes(g, d, *f, lens, atoms, 200L * (num_pos + num_neg) + 20000)) {
            emit(out, d, lens, atoms, print_formula(*f, names));
            uint16_t unit[8] = {1, 1, 1, 1, 1, 1, 1, 1};
            out->max_cost = formula_size(*f, unit);
            out->neg_type = 1;
            return 0;
        }
        if (planted) return LTL_E_LIMIT;
    }
    return LTL_E_LIMIT;
}

extern "C" int32_t ltl_gen_config(int32_t config, uint64_t seed, int32_t index, ltl_gen_instance *out) {
    memset(out, 0, sizeof(*out));
    Rng g(seed ^ (0x243f6a8885a308d3ull * (uint64_t)(config * 1000003 + index + 1)));
    Draw d;
    std::string planted;
    int max_cost = 10, hamming = 0, neg_type = 1;
    switch (config) {
    case 1:
        d = {10, 10, 3, 4, 8, std::vector<double>(3, 0.5)};
        planted = "F((a)&(X(b)))";
        max_cost = 10;
        break;
    case 2: {
        d = {500, 500, 5, 10, 32, {}};
        for (int a = 0; a < 5; ++a) d.rate.push_back(0.3 + 0.4 * g.uniform());
        planted = "(G((~(a))|(F(b))))&(~((c)U(d)))";
        max_cost = 14;
        break;
    }
    case 3:
        d = {1000, 1000, 8, 64, 64, std::vector<double>(8, 0.5)};
        planted = "((a)U((b)&(X(c))))|(G(d))";
        max_cost = 12;
        hamming = 100;
        break;
    case 4: {
        int W = g.range(20, 400);
        double share = 0.3 + 0.4 * g.uniform();
        int P = std::min(W - 1, std::max(1, (int)std::lround(share * W)));
        int A = g.range(2, 6);
        d = {P, W - P, A, 5, 40, std::vector<double>(A, 0.5)};
        break;
    }
    case 5:
        d = {1000, 1000, 6, 16, 64, std::vector<double>(6, 0.5)};
        break;
    default:
        return LTL_E_ARG;
    }
    const AtomNames names = letters(d.A);
    std::vector<uint8_t> lens;
    std::vector<uint64_t> atoms;
    const long budget = 200L * (d.P + d.N) + 20000;
    if (config == 4 || config == 5) {
        int lo = config == 4 ? 5 : 11, hi = config == 4 ? 9 : 13;
        for (int attempt = 0;; ++attempt) {
            if (attempt > 100000) return LTL_E_LIMIT;
            int size = g.range(lo, hi);
            auto f = random_formula(g, size, d.A);
            if (!fill_classes(g, d, *f, lens, atoms, budget)) continue;
            planted = print_formula(*f, names);
            max_cost = size;
            break;
        }
    } else {
        auto f = parse_formula(planted, names);
        if (!fill_classes(g, d, *f, lens, atoms, budget)) return LTL_E_LIMIT;
    }
    if (config == 3) {
        // Swap 50 positive and 50 negative labels: 5% of 2000 flipped, the
        // class sizes unchanged.
        const int W = d.P + d.N, k = 50;
        std::vector<int> pi(d.P), ni(d.N);
        for (int i = 0; i < d.P; ++i) pi[i] = i;
        for (int i = 0; i < d.N; ++i) ni[i] = d.P + i;
        for (int i = 0; i < k; ++i) std::swap(pi[i], pi[i + (int)(g.next() % (uint64_t)(d.P - i))]);
        for (int i = 0; i < k; ++i) std::swap(ni[i], ni[i + (int)(g.next() % (uint64_t)(d.N - i))]);
        std::vector<int> order;  // new positives: kept positives + flipped negatives
        std::vector<char> flip(W, 0);
        for (int i = 0; i < k; ++i) { flip[pi[i]] = 1; flip[ni[i]] = 1; }
        for (int t = 0; t < d.P; ++t) if (!flip[t]) order.push_back(t);
        for (int i = 0; i < k; ++i) order.push_back(ni[i]);
        for (int t = d.P; t < W; ++t) if (!flip[t]) order.push_back(t);
        for (int i = 0; i < k; ++i) order.push_back(pi[i]);
        std::vector<uint8_t> l2(W);
        std::vector<uint64_t> a2(atoms.size());
        for (int t = 0; t < W; ++t) {
            l2[t] = lens[order[t]];
            for (int a = 0; a < d.A; ++a) a2[(size_t)a * W + t] = atoms[(size_t)a * W + order[t]];
        }
        lens.swap(l2); atoms.swap(a2);
    }
    emit(out, d, lens, atoms, planted);
    out->max_cost = max_cost;
    out->hamming_k = hamming;
    out->neg_type = neg_type;
    return 0;
}

extern "C" void ltl_gen_free(ltl_gen_instance *inst) {
    if (!inst) return;
    free(inst->lens); free(inst->atoms);
    inst->lens = nullptr; inst->atoms = nullptr;
}

// ---------------------------------------------------------------------------
// Trace files (code/ltli6463.cu:1500-1573): the alphabet is taken from the
// last line (letters a-z, columns in file order); the atoms are the sorted
// letters; a '1' in column j puts letter j into the position's token.
// ---------------------------------------------------------------------------
// The reference's trace format (code/ltli6463.cu:1500-1573): positive
// lines, "---", negative lines, "---", operators, "---", alphabet (the last
// line).  Positions are separated by ';', the bits of one position by ','.
// ext = false: the reference's alphabet rule, its letters a-z (as a set, so
// sorted) name the atoms, column j of a position is the j-th letter of the
// line in file order (code/ltli6463.cu:1507-1524); names gets the letters.
// ext = true (more than 26 propositions): if the alphabet line is anything
// but single letters a-z, its comma-separated tokens name the columns in
// file order (at most 64, each a valid atom name, distinct) and names gets
// them as a comma-separated list (the form ltl_problem.names accepts).
static int32_t read_traces(const char *path, int32_t max_traces, int32_t K, bool ext, int32_t *num_pos,
                           int32_t *num_neg, int32_t *num_atoms, int32_t *max_len, uint8_t *lens, uint64_t *atoms,
                           char *names, size_t names_cap) {
    if (!path || max_traces < 0 || K < 1 || K > 4 || !names || names_cap < 1) return LTL_E_ARG;
    const int cap = std::min(64 * K, 255);
    std::ifstream in(path, std::ios::binary);
    if (!in) return LTL_E_ARG;
    std::stringstream ss;
    ss << in.rdbuf();
    std::string text = ss.str();
    std::vector<std::string> lines;
    {
        std::string cur;
        for (char c : text) {
            if (c == '\n') { lines.push_back(cur); cur.clear(); }
            else if (c != '\r') cur += c;
        }
        if (!cur.empty()) lines.push_back(cur);
    }
    // Robust against what editors add: trailing blank lines (the reference
    // reads the alphabet from the last line, code/ltli6463.cu:1511-1526),
    // trailing blanks on a line, CR-LF line ends.
    for (auto &l : lines)
        while (!l.empty() && (l.back() == ' ' || l.back() == '\t')) l.pop_back();
    while (!lines.empty() && lines.back().empty()) lines.pop_back();
    if (lines.empty()) return LTL_E_ARG;
    // column j -> atom col_of[j]
    std::vector<int> col_of;
    std::vector<std::string> tokens;
    {
        std::string cur;
        for (char c : lines.back() + ",") {
            if (c == ',') {
                while (!cur.empty() && cur.front() == ' ') cur.erase(cur.begin());
                while (!cur.empty() && cur.back() == ' ') cur.pop_back();
                tokens.push_back(cur);
                cur.clear();
            } else {
                cur += c;
            }
        }
    }
    bool letters_only = true;
    for (const auto &t : tokens) letters_only = letters_only && t.size() == 1 && t[0] >= 'a' && t[0] <= 'z';
    int A = 0;
    std::string out;
    if (!ext || letters_only) {
        std::string alpha;
        bool inset[26] = {false};
        for (char c : lines.back())
            if (c >= 'a' && c <= 'z') { alpha += c; inset[c - 'a'] = true; }
        int idx[26];
        for (int c = 0; c < 26; ++c)
            if (inset[c]) { out += (char)('a' + c); idx[c] = A++; }
        for (char c : alpha) col_of.push_back(idx[c - 'a']);
    } else {
        for (size_t j = 0; j < tokens.size(); ++j) {
            if (j) out += ',';
            out += tokens[j];
            col_of.push_back((int)j);
        }
        A = (int)tokens.size();
        AtomNames check;
        if (A > 64 || !atom_names(out.c_str(), A, check)) return LTL_E_LIMIT;
    }
    if (out.size() + 1 > names_cap) return LTL_E_LIMIT;
    memcpy(names, out.c_str(), out.size() + 1);
    *num_atoms = A;
    int W = 0, P = -1, section = 0, longest = 0;
    const size_t row = (size_t)max_traces * K;  // atom stride while reading
    std::vector<uint64_t> bits((size_t)std::max(A, 1) * 4);
    for (const auto &line : lines) {
        if (section >= 2) break;
        if (line == "---") { if (section == 0) P = W; ++section; continue; }
        if (W >= max_traces) return LTL_E_LIMIT;
        std::fill(bits.begin(), bits.end(), 0ull);
        int pos = 0, j = 0;
        for (char c : line) {
            if (c == ';') { ++pos; j = 0; }
            else if (c == ',') continue;
            else {
                if (c == '1' && j < (int)col_of.size() && pos < cap)
                    bits[(size_t)col_of[j] * 4 + (pos >> 6)] |= 1ull << (pos & 63);
                ++j;
            }
        }
        if (pos + 1 > cap) return LTL_E_LIMIT;
        lens[W] = (uint8_t)(pos + 1);
        longest = std::max(longest, pos + 1);
        for (int a = 0; a < A; ++a)
            for (int k = 0; k < K; ++k) atoms[(size_t)a * row + (size_t)W * K + k] = bits[(size_t)a * 4 + k];
        ++W;
    }
    if (P < 0) P = W;
    *num_pos = P; *num_neg = W - P;
    if (max_len) *max_len = longest;
    for (int a = 1; a < A; ++a)
        memmove(atoms + (size_t)a * W * K, atoms + (size_t)a * row, (size_t)W * K * 8);
    return 0;
}

extern "C" int32_t ltl_read_trace_file_words(const char *path, int32_t max_traces, int32_t words_per_trace,
                                             int32_t *num_pos, int32_t *num_neg, int32_t *num_atoms,
                                             int32_t *max_len, uint8_t *lens, uint64_t *atoms, char *names) {
    return read_traces(path, max_traces, words_per_trace, false, num_pos, num_neg, num_atoms, max_len, lens, atoms,
                       names, 27);
}

extern "C" int32_t ltl_read_trace_file_names(const char *path, int32_t max_traces, int32_t words_per_trace,
                                             int32_t *num_pos, int32_t *num_neg, int32_t *num_atoms,
                                             int32_t *max_len, uint8_t *lens, uint64_t *atoms, char *names,
                                             int32_t names_cap) {
    return read_traces(path, max_traces, words_per_trace, true, num_pos, num_neg, num_atoms, max_len, lens, atoms,
                       names, names_cap > 0 ? (size_t)names_cap : 0);
}

extern "C" int32_t ltl_read_trace_file(const char *path, int32_t max_traces, int32_t *num_pos,
                                       int32_t *num_neg, int32_t *num_atoms, uint8_t *lens,
                                       uint64_t *atoms, char *names) {
    return ltl_read_trace_file_words(path, max_traces, 1, num_pos, num_neg, num_atoms, nullptr, lens, atoms,
                                     names);
}

extern "C" int32_t ltl_eval_formula(const char *formula, const ltl_problem *pb, uint64_t *out_cs) {
    if (!formula || !pb) return LTL_E_ARG;
    AtomNames names;
    if (!atom_names(pb->names, pb->num_atoms, names)) return LTL_E_ARG;
    auto f = parse_formula(formula, names);
    if (!f) return LTL_E_ARG;
    eval_formula(*f, pb->num_pos + pb->num_neg, pb->words_per_trace, pb->lens, pb->atoms, out_cs);
    return 0;
}

extern "C" int32_t ltl_mismatches(const uint64_t *cs, int32_t num_pos, int32_t num_neg) {
    int32_t m = 0;
    for (int i = 0; i < num_pos; ++i) m += (int32_t)(~cs[i] & 1ull);
    for (int i = num_pos; i < num_pos + num_neg; ++i) m += (int32_t)(cs[i] & 1ull);
    return m;
}
