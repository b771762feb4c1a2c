// dc.cpp -- divide-and-conquer drivers over the learner, in process.
//
// Restates the reference's deterministic and random splitting
// (code/dc.py:188-249 DetSplit, :252-321 RandSplit): a window of at most
// `window` traces is learned directly; larger sets are split, the halves
// learned recursively, and the answers combined with & and |, using formula
// matches on the other traces (LTLmatch, dc.py:136-186) to decide.
// Differences by design: sub-problems are gathered on the host and learned
// in process (no subprocess, no subInput file), through ltl_learn_batch;
// LTLmatch runs on the device over the full trace set (copied there once per
// call), cached per formula, and the match of a formula the driver builds
// as (a)&(b) or (a)|(b) is the & or | of its parts' cached matches (the
// value at position 0 of a conjunction is the conjunction of the values);
// sampling uses a seeded SplitMix64 stream (the reference uses Python's
// random module, reseeded with 0 on every RandSplit call, dc.py:255); empty
// windows, splits that cannot shrink and recursion deeper than 64 end with
// LTL_E_LIMIT instead of recursing forever.
//
// Several problems (ltl_learn_dc_batch) run in lock-step, one driver thread
// each: a driver that needs a window learned or a formula matched posts the
// request and waits; once every live driver waits, the last to arrive learns
// all posted windows in one ltl_learn_batch call (independent problems: the
// persistent batched launch takes them together) and matches every formula
// they returned in one device launch, then wakes the others.  A driver's
// requests, and so its result, do not depend on the others.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ltl_amd.h"
#include "dc_eval.hpp"
#include "formula.hpp"

using namespace ltl;

namespace {

using Idx = std::vector<int>;
struct Driver;

// Lock-step coordinator of the drivers of one call (see the file comment).
struct Coordinator {
    ltl_ctx *ctx = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    MatchBuf mbuf;
    std::mutex mu;
    std::condition_variable cv;
    int live = 0, waiting = 0;
    struct Req {
        Driver *d = nullptr;
        bool solve = true;             // a window to learn, else a formula to match
        ltl_problem sub{};             // the window (lens and atoms below)
        std::vector<uint8_t> lens;
        std::vector<uint64_t> atoms;
        std::string formula;           // solve: the answer ("" = none); match: the formula
        int32_t status = 0;
        uint64_t all_ltls = 0;
        bool done = false;
    };
    std::vector<Req *> pending;

    void post(Req *r) {
        std::unique_lock<std::mutex> lk(mu);
        pending.push_back(r);
        if (++waiting == live) run_locked();
        else cv.wait(lk, [&] { return r->done; });
    }
    void leave() {
        std::unique_lock<std::mutex> lk(mu);
        --live;
        if (live > 0 && waiting == live) run_locked();
    }
    void run_locked();
    void serve(const std::vector<Req *> &reqs);  // (run_locked's work)
};

struct Driver {
    Coordinator *co;
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
    std::unique_ptr<DevTraces> dev;  // the full trace set on the device
    std::map<std::string, std::vector<uint8_t>> cache;
    std::string result;

    // Position-0 value of `f` on every trace of the full problem.
    const std::vector<uint8_t> *match(const std::string &f) {
        auto it = cache.find(f);
        if (it == cache.end()) {
            Coordinator::Req q;
            q.d = this;
            q.solve = false;
            q.formula = f;
            co->post(&q);
            if (q.status < 0) {
                err = q.status;
                return nullptr;
            }
            it = cache.find(f);
            if (it == cache.end()) {
                err = LTL_E_ARG;
                return nullptr;
            }
        }
        return &it->second;
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

    // The window pos + neg learned (through the coordinator); "" when
    // nothing up to max_cost.
    std::string solve(const Idx &pos, const Idx &neg) {
        const int P = (int)pos.size(), N = (int)neg.size(), w = P + N;
        Coordinator::Req q;
        q.d = this;
        q.lens.resize(w);
        q.atoms.resize((size_t)A * w * K);
        for (int k = 0; k < w; ++k) {
            const int t = k < P ? pos[k] : neg[k - P];
            q.lens[k] = pb->lens[t];
            for (int a = 0; a < A; ++a)
                for (int j = 0; j < K; ++j)
                    q.atoms[((size_t)a * w + k) * K + j] = pb->atoms[((size_t)a * W + t) * K + j];
        }
        q.sub = ltl_problem{P, N, A, K, q.lens.data(), q.atoms.data(), raw};
        co->post(&q);
        ++calls;
        all_ltls += q.all_ltls;
        if (q.status < 0) err = q.status;
        return q.status == LTL_FOUND ? q.formula : std::string();
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

    // (a)&(b) and (a)|(b), with their matches from the parts' (both parts
    // have been matched where the drivers build them)
    std::string combine(const std::string &a, const std::string &b, bool conj) {
        const std::string f = "(" + a + ")" + (conj ? "&" : "|") + "(" + b + ")";
        if (!cache.count(f)) {
            const auto *ma = match(a), *mb = match(b);
            if (!ma || !mb) return f;
            std::vector<uint8_t> m(W);
            for (int i = 0; i < W; ++i) m[i] = conj ? ((*ma)[i] & (*mb)[i]) : ((*ma)[i] | (*mb)[i]);
            cache.emplace(f, std::move(m));
        }
        return f;
    }
    std::string both(const std::string &a, const std::string &b) { return combine(a, b, true); }
    std::string either(const std::string &a, const std::string &b) { return combine(a, b, false); }

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
        return either(left, right);
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
        return either(left, right);
    }
};

void Coordinator::run_locked() {
    std::vector<Req *> reqs;
    reqs.swap(pending);
    waiting -= (int)reqs.size();
    try {
        serve(reqs);
    } catch (...) {  // (host memory): the posted requests fail, the drivers go on
        for (Req *r : reqs)
            if (r->status >= 0) r->status = LTL_E_OOM;
    }
    for (Req *r : reqs) r->done = true;
    cv.notify_all();
}

void Coordinator::serve(const std::vector<Req *> &reqs) {
    hipError_t he = hipSetDevice(device);
    // the posted windows, one batched call
    std::vector<Req *> sv;
    for (Req *r : reqs)
        if (r->solve) sv.push_back(r);
    if (!sv.empty()) {
        std::vector<ltl_problem> pbs;
        for (Req *r : sv) pbs.push_back(r->sub);
        std::vector<ltl_result> rs(sv.size());
        const int32_t st = ltl_learn_batch(ctx, pbs.data(), (int32_t)pbs.size(), rs.data());
        for (size_t k = 0; k < sv.size(); ++k) {
            Req *r = sv[k];
            r->status = st < 0 ? st : rs[k].status;
            if (st >= 0) {
                r->all_ltls = rs[k].all_ltls;
                if (rs[k].status == LTL_FOUND && rs[k].formula) r->formula = rs[k].formula;
                ltl_result_free(&rs[k]);
            }
        }
    }
    // matches of the returned formulas and of the posted ones, one launch
    std::vector<MatchJob> jobs;
    std::vector<std::unique_ptr<Node>> trees;
    std::vector<Req *> owner;
    for (Req *r : reqs) {
        if (r->status < 0 || (r->solve && r->status != LTL_FOUND)) continue;
        if (r->d->cache.count(r->formula)) continue;
        auto tree = parse_formula(r->formula, r->d->names);
        if (!tree) {
            r->status = LTL_E_ARG;
            continue;
        }
        jobs.push_back({r->d->dev.get(), tree.get()});
        trees.push_back(std::move(tree));
        owner.push_back(r);
    }
    std::vector<std::vector<uint8_t>> vals;
    const int32_t e = he != hipSuccess ? LTL_E_HIP : dev_match(jobs, stream, mbuf, vals);
    for (size_t k = 0; k < owner.size(); ++k) {
        if (e) owner[k]->status = e;
        else owner[k]->d->cache.emplace(owner[k]->formula, std::move(vals[k]));
    }
}

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

// (leaves the coordinator exactly once, whatever happens)
void run_driver(Driver &d) {
    try {
        Idx pos(d.pb->num_pos), neg(d.pb->num_neg);
        for (int i = 0; i < d.pb->num_pos; ++i) pos[i] = i;
        for (int i = 0; i < d.pb->num_neg; ++i) neg[i] = d.pb->num_pos + i;
        d.result = d.algo == LTL_DC_DET ? d.det(pos, neg, 0) : d.rnd(pos, neg, 0);
    } catch (...) {
        d.err = LTL_E_OOM;
    }
    d.co->leave();
}

void fill_result(ltl_ctx *ctx, Driver &d, ltl_result *res) {
    res->all_ltls = d.all_ltls;
    res->n_stored = d.calls;
    if (d.err) {
        res->status = d.err;
        return;
    }
    auto tree = parse_formula(d.result, d.names);
    if (!tree) {
        res->status = LTL_E_ARG;
        return;
    }
    std::vector<ltl_node> nodes;
    to_nodes(*tree, nodes);
    // costs of the context's options
    ltl_options o;
    ltl_get_options(ctx, &o);
    res->cost = formula_size(*tree, o.cost);
    res->n_nodes = (int32_t)nodes.size();
    res->nodes = (ltl_node *)malloc(nodes.size() * sizeof(ltl_node));
    memcpy(res->nodes, nodes.data(), nodes.size() * sizeof(ltl_node));
    res->formula = strdup(d.result.c_str());
    res->status = LTL_FOUND;
}

// The drivers of problems [0, n) in lock-step; per-problem results in res.
int32_t run_dc(ltl_ctx *ctx, const ltl_problem *pbs, int32_t n, int32_t window, int32_t algo, uint64_t seed,
               ltl_result *res) {
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n; ++i) {
        memset(&res[i], 0, sizeof(res[i]));
        res[i].found_id = -1;
    }
    if (!ctx || window < 2 || (algo != LTL_DC_DET && algo != LTL_DC_RAND)) {
        for (int i = 0; i < n; ++i) res[i].status = LTL_E_ARG;
        return LTL_E_ARG;
    }
    ltl_options o;
    ltl_get_options(ctx, &o);
    Coordinator co;
    co.ctx = ctx;
    co.device = o.device;
    if (hipSetDevice(o.device) != hipSuccess || hipStreamCreate(&co.stream) != hipSuccess) {
        for (int i = 0; i < n; ++i) res[i].status = LTL_E_HIP;
        return LTL_E_HIP;
    }
    std::vector<std::unique_ptr<Driver>> ds;
    for (int i = 0; i < n; ++i) {
        const ltl_problem *pb = &pbs[i];
        auto d = std::make_unique<Driver>(Driver{&co, pb, window, algo, seed});
        if (pb->num_atoms < 1 || pb->num_atoms > 64 || pb->num_pos < 0 || pb->num_neg < 0 ||
            pb->num_pos + pb->num_neg < 1 || !atom_names(pb->names, pb->num_atoms, d->names)) {
            res[i].status = LTL_E_ARG;
            continue;
        }
        d->W = pb->num_pos + pb->num_neg;
        d->A = pb->num_atoms;
        d->K = pb->words_per_trace > 1 ? pb->words_per_trace : 1;
        d->raw = pb->names;
        d->dev = std::make_unique<DevTraces>();
        if (const int32_t e = dev_traces_upload(pb, co.stream, *d->dev)) {
            res[i].status = e;
            continue;
        }
        ds.push_back(std::move(d));
    }
    co.live = (int)ds.size();
    std::vector<std::thread> th;
    for (size_t k = 1; k < ds.size(); ++k) {
        try {
            th.emplace_back(run_driver, std::ref(*ds[k]));
        } catch (...) {  // no thread: this driver ends with an error
            ds[k]->err = LTL_E_LIMIT;
            co.leave();
        }
    }
    if (!ds.empty()) run_driver(*ds[0]);
    for (auto &t : th) t.join();
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (auto &d : ds) {
        ltl_result *r = &res[d->pb - pbs];
        fill_result(ctx, *d, r);
        r->seconds = sec;
    }
    ds.clear();  // (device copies freed before the stream)
    if (co.mbuf.d) hipFree(co.mbuf.d);
    co.mbuf.d = nullptr;
    hipStreamDestroy(co.stream);
    int32_t ok = 0;
    for (int i = 0; i < n; ++i) ok += res[i].status == LTL_FOUND;
    return ok;
}

}  // namespace

extern "C" int32_t ltl_learn_dc(ltl_ctx *ctx, const ltl_problem *pb, int32_t window, int32_t algo, uint64_t seed,
                                ltl_result *res) {
    if (!res) return LTL_E_ARG;
    if (!pb) {
        memset(res, 0, sizeof(*res));
        res->found_id = -1;
        return res->status = LTL_E_ARG;
    }
    run_dc(ctx, pb, 1, window, algo, seed, res);
    return res->status;
}

extern "C" int32_t ltl_learn_dc_batch(ltl_ctx *ctx, const ltl_problem *pbs, int32_t n, int32_t window, int32_t algo,
                                      uint64_t seed, ltl_result *res) {
    if (n < 0 || (n > 0 && (!pbs || !res))) return LTL_E_ARG;
    // (groups of at most 256 problems: one driver thread each)
    int32_t solved = 0;
    for (int32_t i = 0; i < n; i += 256) {
        const int32_t r = run_dc(ctx, pbs + i, std::min<int32_t>(256, n - i), window, algo, seed, res + i);
        if (r < 0) return r;
        solved += r;
    }
    return solved;
}

extern "C" int32_t ltl_match(ltl_ctx *ctx, const ltl_problem *pb, const char *const *formulas, int32_t n,
                             uint8_t *out) {
    if (!ctx || !pb || n < 0 || (n > 0 && (!formulas || !out))) return LTL_E_ARG;
    if (pb->num_atoms < 1 || pb->num_atoms > 64 || pb->num_pos < 0 || pb->num_neg < 0) return LTL_E_ARG;
    if (n == 0) return 0;
    AtomNames names;
    if (!atom_names(pb->names, pb->num_atoms, names)) return LTL_E_ARG;
    std::vector<std::unique_ptr<Node>> trees;
    for (int i = 0; i < n; ++i) {
        if (!formulas[i]) return LTL_E_ARG;
        trees.push_back(parse_formula(formulas[i], names));
        if (!trees.back()) return LTL_E_ARG;
    }
    ltl_options o;
    ltl_get_options(ctx, &o);
    hipStream_t s = nullptr;
    if (hipSetDevice(o.device) != hipSuccess || hipStreamCreate(&s) != hipSuccess) return LTL_E_HIP;
    int32_t e = 0;
    {
        DevTraces dt;
        e = dev_traces_upload(pb, s, dt);
        std::vector<std::vector<uint8_t>> vals;
        if (!e) {
            std::vector<MatchJob> jobs;
            for (auto &t : trees) jobs.push_back({&dt, t.get()});
            MatchBuf buf;
            e = dev_match(jobs, s, buf, vals);
        }
        const int W = pb->num_pos + pb->num_neg;
        if (!e)
            for (int i = 0; i < n; ++i) memcpy(out + (size_t)i * W, vals[i].data(), (size_t)W);
    }
    hipStreamDestroy(s);
    return e;
}
