// dc.cpp -- divide-and-conquer drivers over ltl_learn, in process.
//
// Restates the reference's deterministic and random splitting
// (code/dc.py:188-249 DetSplit, :252-321 RandSplit): a window of at most
// `window` traces is learned directly; larger sets are split, the halves
// learned recursively, and the answers combined with & and |, using formula
// matches on the other traces (LTLmatch, dc.py:185) to decide.  Differences
// by design: sub-problems are gathered on the host and solved by ltl_learn
// on the device (no subprocess, no subInput file); LTLmatch is the
// bit-parallel host evaluator on the full trace set, cached per formula;
// sampling uses a seeded SplitMix64 stream (the reference uses Python's
// random module, reseeded with 0 on every RandSplit call, dc.py:255);
// empty windows, splits that cannot shrink and recursion deeper than 64 end
// with LTL_E_LIMIT instead of recursing forever.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ltl_amd.h"
#include "formula.hpp"

using namespace ltl;

namespace {

using Idx = std::vector<int>;

struct Driver {
    ltl_ctx *ctx;
    const ltl_problem *pb;
    int window;
    int algo;
    uint64_t seed;
    uint64_t rng = 0;
    AtomNames names;   // parsed
    const char *raw = nullptr;  // as the caller gave them (passed to the window calls)
    int W, A, K;  // traces, atoms, words per trace
    int32_t err = 0;
    uint64_t all_ltls = 0;
    int64_t calls = 0;
    std::map<std::string, std::vector<uint8_t>> cache;

    // Position-0 value of `f` on every trace of the full problem.
    const std::vector<uint8_t> *match(const std::string &f) {
        auto it = cache.find(f);
        if (it != cache.end()) return &it->second;
        auto tree = parse_formula(f, names);
        if (!tree) {
            err = LTL_E_ARG;
            return nullptr;
        }
        std::vector<uint64_t> cs((size_t)W * K);
        eval_formula(*tree, W, K, pb->lens, pb->atoms, cs.data());
        std::vector<uint8_t> m(W);
        for (int i = 0; i < W; ++i) m[i] = (uint8_t)(cs[(size_t)i * K] & 1ull);
        return &cache.emplace(f, std::move(m)).first->second;
    }
    bool all_of(const std::string &f, const Idx &s) {
        const auto *m = match(f);
        if (!m) return false;
        for (int i : s)
            if (!(*m)[i]) return false;
        return true;
    }
    bool none_of(const std::string &f, const Idx &s) {
        const auto *m = match(f);
        if (!m) return false;
        for (int i : s)
            if ((*m)[i]) return false;
        return true;
    }
    Idx filter(const std::string &f, const Idx &s, bool want) {
        Idx out;
        const auto *m = match(f);
        if (!m) return out;
        for (int i : s)
            if ((bool)(*m)[i] == want) out.push_back(i);
        return out;
    }
    static Idx minus(const Idx &a, const Idx &b) {
        Idx out;
        for (int i : a)
            if (std::find(b.begin(), b.end(), i) == b.end()) out.push_back(i);
        return out;
    }

    // ltl_learn on the traces pos + neg; "" when none up to max_cost.
    std::string solve(const Idx &pos, const Idx &neg) {
        const int P = (int)pos.size(), N = (int)neg.size(), w = P + N;
        std::vector<uint8_t> lens(w);
        std::vector<uint64_t> atoms((size_t)A * w * K);
        for (int k = 0; k < w; ++k) {
            const int t = k < P ? pos[k] : neg[k - P];
            lens[k] = pb->lens[t];
            for (int a = 0; a < A; ++a)
                for (int j = 0; j < K; ++j)
                    atoms[((size_t)a * w + k) * K + j] = pb->atoms[((size_t)a * W + t) * K + j];
        }
        ltl_problem sub{P, N, A, K, lens.data(), atoms.data(), raw};
        ltl_result r;
        const int32_t st = ltl_learn(ctx, &sub, &r);
        ++calls;
        std::string f;
        if (st == LTL_FOUND) f = r.formula;
        else if (st < 0) err = st;
        all_ltls += r.all_ltls;
        ltl_result_free(&r);
        return f;
    }

    // k of n indices without replacement, in selection order (partial
    // Fisher-Yates over a copy, SplitMix64 draws).
    Idx sample(const Idx &s, int k) {
        Idx a(s);
        for (int i = 0; i < k; ++i) {
            uint64_t z = (rng += 0x9e3779b97f4a7c15ull);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            z ^= z >> 31;
            const int j = i + (int)(z % (uint64_t)(a.size() - i));
            std::swap(a[i], a[j]);
        }
        a.resize(k);
        return a;
    }

    static std::string both(const std::string &a, const std::string &b) { return "(" + a + ")&(" + b + ")"; }

    // code/dc.py:188-249
    std::string det(const Idx &pos, const Idx &neg, int depth) {
        if (err) return "";
        if (depth > 64 || (pos.empty() && neg.empty())) { err = LTL_E_LIMIT; return ""; }
        if ((int)(pos.size() + neg.size()) <= window) {
            std::string f = solve(pos, neg);
            if (err || !f.empty()) return f;
            if (pos.size() + neg.size() <= 1) { err = LTL_E_LIMIT; return ""; }
        }
        const Idx p1(pos.begin(), pos.begin() + pos.size() / 2), p2(pos.begin() + pos.size() / 2, pos.end());
        const Idx n1(neg.begin(), neg.begin() + neg.size() / 2), n2(neg.begin() + neg.size() / 2, neg.end());
        const std::string phi11 = det(p1, n1, depth + 1);
        if (err) return "";
        if (all_of(phi11, p2) && none_of(phi11, n2)) return phi11;
        std::string left;
        if (none_of(phi11, n2)) {
            left = phi11;
        } else {
            const Idx n2a = filter(phi11, n2, true);
            const std::string phi12 = det(p1, n2a, depth + 1);
            if (err) return "";
            left = none_of(phi12, minus(neg, n2a)) ? phi12 : both(phi11, phi12);
            if (all_of(left, p2)) return left;
        }
        const Idx p2m = filter(left, p2, false);
        const std::string phi21 = det(p2m, n1, depth + 1);
        if (err) return "";
        const Idx rest = minus(pos, p2m);
        if (all_of(phi21, rest) && none_of(phi21, n2)) return phi21;
        std::string right;
        if (none_of(phi21, n2)) {
            right = phi21;
        } else {
            const Idx n2b = filter(phi21, n2, true);
            const std::string phi22 = det(p2m, n2b, depth + 1);
            if (err) return "";
            right = none_of(phi22, minus(neg, n2b)) ? phi22 : both(phi21, phi22);
            if (all_of(right, rest)) return right;
        }
        return "(" + left + ")|(" + right + ")";
    }

    // code/dc.py:252-321
    std::string rnd(const Idx &pos, const Idx &neg, int depth) {
        if (err) return "";
        if (depth > 64 || (pos.empty() && neg.empty())) { err = LTL_E_LIMIT; return ""; }
        rng = seed;
        int win = window;
        std::string phi11;
        for (;;) {
            if (win < 1) { err = LTL_E_LIMIT; return ""; }
            Idx p1, n1;
            const int np = (int)pos.size(), nn = (int)neg.size();
            if (np + nn <= win) {
                p1 = pos; n1 = neg;
            } else if (np <= win / 2) {
                p1 = pos; n1 = sample(neg, win - np);
            } else if (nn <= win / 2) {
                n1 = neg; p1 = sample(pos, win - nn);
            } else {
                p1 = sample(pos, win / 2); n1 = sample(neg, win - win / 2);
            }
            phi11 = solve(p1, n1);
            if (err) return "";
            if (!phi11.empty()) break;
            win /= 2;
        }
        const Idx p2 = filter(phi11, pos, false), n2 = filter(phi11, neg, true);
        const Idx p1 = minus(pos, p2), n1 = minus(neg, n2);
        if (p2.empty() && n2.empty()) return phi11;
        std::string left;
        if (n2.empty()) {
            left = phi11;
        } else {
            const std::string phi12 = rnd(p1, n2, depth + 1);
            if (err) return "";
            left = none_of(phi12, n1) ? phi12 : both(phi11, phi12);
            if (all_of(left, p2)) return left;
        }
        const std::string phi21 = rnd(p2, n1, depth + 1);
        if (err) return "";
        if (all_of(phi21, p1) && none_of(phi21, n2)) return phi21;
        std::string right;
        if (none_of(phi21, n2)) {
            right = phi21;
        } else {
            const std::string phi22 = rnd(p2, n2, depth + 1);
            if (err) return "";
            right = none_of(phi22, n1) ? phi22 : both(phi21, phi22);
            if (all_of(right, p1)) return right;
        }
        return "(" + left + ")|(" + right + ")";
    }
};

void to_nodes(const Node &n, std::vector<ltl_node> &out) {
    // breadth first, node 0 = root
    std::vector<const Node *> q{&n};
    out.clear();
    out.push_back({n.op, n.op == LTL_ATOM ? n.atom : -1, -1, -1});
    for (size_t h = 0; h < q.size(); ++h) {
        const Node *x = q[h];
        const Node *kids[2] = {x->a.get(), x->b.get()};
        for (int s = 0; s < 2; ++s) {
            if (!kids[s]) continue;
            const int idx = (int)out.size();
            out.push_back({kids[s]->op, kids[s]->op == LTL_ATOM ? kids[s]->atom : -1, -1, -1});
            if (s == 0) out[h].left = idx; else out[h].right = idx;
            q.push_back(kids[s]);
        }
    }
}

}  // namespace

extern "C" int32_t ltl_learn_dc(ltl_ctx *ctx, const ltl_problem *pb, int32_t window, int32_t algo, uint64_t seed,
                                ltl_result *res) {
    if (!res) return LTL_E_ARG;
    memset(res, 0, sizeof(*res));
    res->found_id = -1;
    if (!ctx || !pb || window < 2 || (algo != LTL_DC_DET && algo != LTL_DC_RAND) || pb->num_atoms < 1 ||
        pb->num_atoms > 26 || pb->num_pos + pb->num_neg < 1)
        return res->status = LTL_E_ARG;
    const auto t0 = std::chrono::steady_clock::now();
    Driver d{ctx, pb, window, algo, seed};
    d.W = pb->num_pos + pb->num_neg;
    d.A = pb->num_atoms;
    d.K = pb->words_per_trace > 1 ? pb->words_per_trace : 1;
    if (!atom_names(pb->names, pb->num_atoms, d.names)) return res->status = LTL_E_ARG;
    d.raw = pb->names;
    Idx pos(pb->num_pos), neg(pb->num_neg);
    for (int i = 0; i < pb->num_pos; ++i) pos[i] = i;
    for (int i = 0; i < pb->num_neg; ++i) neg[i] = pb->num_pos + i;
    const std::string f = algo == LTL_DC_DET ? d.det(pos, neg, 0) : d.rnd(pos, neg, 0);
    res->all_ltls = d.all_ltls;
    res->n_stored = d.calls;
    res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (d.err) return res->status = d.err;
    auto tree = parse_formula(f, d.names);
    if (!tree) return res->status = LTL_E_ARG;
    std::vector<ltl_node> nodes;
    to_nodes(*tree, nodes);
    // costs of the context's options
    ltl_options o;
    ltl_get_options(ctx, &o);
    res->cost = formula_size(*tree, o.cost);
    res->n_nodes = (int32_t)nodes.size();
    res->nodes = (ltl_node *)malloc(nodes.size() * sizeof(ltl_node));
    memcpy(res->nodes, nodes.data(), nodes.size() * sizeof(ltl_node));
    res->formula = strdup(f.c_str());
    return res->status = LTL_FOUND;
}
