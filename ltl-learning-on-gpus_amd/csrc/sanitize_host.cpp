// sanitize_host.cpp -- drives the host-only code of the library (trace file
// readers, formula parser and evaluator, instance generator: host_util.cpp)
// over well-formed and damaged inputs, for a build with AddressSanitizer and
// UndefinedBehaviorSanitizer (make -C ltl-learning-on-gpus_amd/csrc
// sanitize).  Arguments: trace files to read as given and mutated; a
// sanitizer report, or an inconsistent result, ends the run with a nonzero
// status.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include <unistd.h>

#include "../../include/ltl_amd.h"
#include "../../include/ltl_gen.h"

namespace {

uint64_t rng_state = 0x243f6a8885a308d3ull;
uint64_t next() {  // SplitMix64
    uint64_t z = (rng_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
int below(int n) { return (int)(next() % (uint64_t)n); }

int failures = 0;
void fail(const std::string &what) {
    fprintf(stderr, "FAIL: %s\n", what.c_str());
    ++failures;
}

std::string slurp(const char *path) {
    std::ifstream f(path, std::ios::binary);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// Read `text` (written to a temporary file) with the name-list reader; on
// success, evaluate formulas over the traces read.
void read_and_eval(const std::string &text, const char *tmp) {
    {
        std::ofstream f(tmp, std::ios::binary);
        f << text;
    }
    int lines = 1;
    for (char ch : text) lines += ch == '\n';
    const int K = 4;
    std::vector<uint8_t> lens(lines);
    std::vector<uint64_t> atoms((size_t)64 * lines * K);
    std::vector<char> names(64 * 40);
    int32_t P = 0, N = 0, A = 0, longest = 0;
    const int32_t rc = ltl_read_trace_file_names(tmp, lines, K, &P, &N, &A, &longest, lens.data(), atoms.data(),
                                                 names.data(), (int32_t)names.size());
    if (rc != 0) return;
    if (P < 0 || N < 0 || P + N > lines || A < 1 || A > 64 || longest > 255) {
        fail("reader returned sizes out of range");
        return;
    }
    const int W = P + N;
    if (W == 0) return;
    ltl_problem pb{P, N, A, K, lens.data(), atoms.data(), names.data()};
    std::vector<uint64_t> cs((size_t)W * K);
    // the first name (letters: the first character; lists: up to the comma)
    std::string a0(names.data());
    if (a0.find(',') != std::string::npos) a0 = a0.substr(0, a0.find(','));
    else a0 = a0.substr(0, 1);
    const std::string fs[] = {a0, "X(" + a0 + ")", "F(" + a0 + ")", "G(" + a0 + ")", "~(" + a0 + ")",
                              "(" + a0 + ")U(X(" + a0 + "))", "(F(" + a0 + "))&(G(~(" + a0 + ")))"};
    for (const auto &f : fs) {
        if (ltl_eval_formula(f.c_str(), &pb, cs.data()) != 0) fail("formula over the file's first name: " + f);
        const int32_t m = ltl_mismatches(cs.data(), P, N);
        if (m < 0 || m > W) fail("mismatches out of range");
    }
}

std::string mutate(const std::string &s) {
    std::string t = s;
    const int kind = below(8);
    if (t.empty()) return "---\n";
    const size_t at = (size_t)below((int)t.size());
    switch (kind) {
    case 0: t.resize(at); break;                                   // truncated
    case 1: t[at] = "01,;-\n\rabzX_"[below(13)]; break;           // one character
    case 2: t.insert(at, std::string(1 + below(300), "01,;"[below(4)])); break;  // long run
    case 3: t.erase(at, (size_t)below(40)); break;                  // a gap
    case 4: t.insert(at, "\n---\n"); break;                         // extra separator
    case 5: t.insert(at, std::string(1 + below(3), '\0')); break;   // NUL bytes
    case 6: {                                                       // many names
        std::string nm;
        for (int i = 0; i < 60 + below(20); ++i) nm += (i ? "," : "") + std::string("p") + std::to_string(i);
        t += "\n---\n" + nm;
        break;
    }
    default: t += t.substr(at); break;                              // duplicated tail
    }
    return t;
}

// Random strings over the formula grammar's characters, valid formulas and
// their mutations, through the parser and evaluator.
void formulas() {
    const int W = 40, A = 3, K = 2;
    std::vector<uint8_t> lens(W);
    std::vector<uint64_t> atoms((size_t)A * W * K);
    for (int t = 0; t < W; ++t) lens[t] = (uint8_t)(1 + below(128));
    for (auto &w : atoms) w = next();
    for (int t = 0; t < W; ++t)  // bits past the length stay clear
        for (int a = 0; a < A; ++a)
            for (int k = 0; k < K; ++k) {
                const int lo = 64 * k, len = lens[t];
                uint64_t m = len <= lo ? 0 : len - lo >= 64 ? ~0ull : ((1ull << (len - lo)) - 1);
                atoms[((size_t)a * W + t) * K + k] &= m;
            }
    ltl_problem pb{20, 20, A, K, lens.data(), atoms.data(), "abc"};
    std::vector<uint64_t> cs((size_t)W * K);
    const char alpha[] = "abcd()~&|XFGU, ";
    for (int it = 0; it < 20000; ++it) {
        std::string f;
        if (it % 2) {
            const int n = below(40);
            for (int i = 0; i < n; ++i) f += alpha[below((int)sizeof(alpha) - 1)];
        } else {
            // a valid formula, built from the inside out, then perhaps damaged
            f = std::string(1, "abc"[below(3)]);
            for (int d = below(30); d > 0; --d) {
                const int op = below(7);
                const std::string g = std::string(1, "abc"[below(3)]);
                if (op < 4) f = std::string(1, "~XFG"[op]) + "(" + f + ")";
                else f = "(" + f + ")" + "&|U"[op - 4] + "(" + g + ")";
            }
            if (it % 4 == 0 && !f.empty()) f.erase((size_t)below((int)f.size()), 1);
        }
        const int32_t rc = ltl_eval_formula(f.c_str(), &pb, cs.data());
        if (rc != 0 && rc != LTL_E_ARG) fail("parser status " + std::to_string(rc));
    }
}

void generators() {
    for (int config = 1; config <= 5; ++config)
        for (int index = 0; index < 2; ++index) {
            ltl_gen_instance g;
            if (ltl_gen_config(config, 1 + index, index, &g) != 0) {
                fail("ltl_gen_config");
                continue;
            }
            const int W = g.num_pos + g.num_neg;
            ltl_problem pb{g.num_pos, g.num_neg, g.num_atoms, 1, g.lens, g.atoms, g.names};
            std::vector<uint64_t> cs(W);
            if (ltl_eval_formula(g.planted, &pb, cs.data()) != 0) fail("planted formula does not parse");
            else if (config != 3 && ltl_mismatches(cs.data(), g.num_pos, g.num_neg) != 0)
                fail("planted formula does not separate its traces");
            ltl_gen_free(&g);
        }
    for (int it = 0; it < 30; ++it) {
        ltl_gen_instance g;
        const int lo = 1 + below(64), hi = lo + below(65 - lo);
        if (ltl_gen_custom(next(), 1 + below(40), 1 + below(40), 1 + below(6), lo, hi, 0.2 + 0.6 * (below(100) / 100.0),
                           nullptr, 1 + below(9), &g) == 0)
            ltl_gen_free(&g);
    }
}

}  // namespace

int main(int argc, char **argv) {
    const std::string tmp_path = std::string(getenv("TMPDIR") ? getenv("TMPDIR") : "/tmp") + "/ltl_sanitize_host." +
                                 std::to_string((long)getpid());
    const char *tmp = tmp_path.c_str();
    int files = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string text = slurp(argv[i]);
        read_and_eval(text, tmp);
        for (int m = 0; m < 20; ++m) read_and_eval(mutate(text), tmp);
        ++files;
    }
    formulas();
    generators();
    std::remove(tmp);
    printf("sanitize_host: %d files (20 mutations each), 20000 formulas, generators: %d failures\n", files,
           failures);
    return failures ? 1 : 0;
}
