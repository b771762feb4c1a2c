// ltli_main.cpp -- command-line front end over the C ABI.
//
// Same arguments and stdout as the reference binary (code/ltli6463.cu:
// 1575-1681), so that drivers which parse its output (code/dc.py:23-43, 192-194)
// work unchanged:
//
//   ltli [--reference-keys] [--launch-log] <input_file> <c1> .. <c8> <maxCost> <RlxUnqChkTyp> <NegType>
//
// stdout: the traces, "Cost Function: ...", "Number of Traces: N", the
// run statistics, then a last line  LTL: "<formula>"  (or "not_found"), as
// the reference built with MEASUREMENT_MODE prints them.  --launch-log (or
// LTLI_LAUNCH_LOG=1) first prints a line per block of candidates in the
// format of the reference's default build (code/ltli6463.cu:1096, 1225):
//   Cost N  | (op) | AllLTLs: a | StoredLTLs: s | ToBeChecked: n
// with " | test" appended for the blocks of a check pass (tested, not stored)
// and " | ahead" for blocks tested before their level.
// Differences: errors exit with a nonzero status (the reference always exits
// 0), traces may number more than 64 and hold up to 255 positions, an
// alphabet line of comma-separated names instead of letters names up to 64
// propositions, and deduplication is exact for any total trace length
// unless --reference-keys is given (or LTLI_REFERENCE_KEYS=1 is set in the
// environment): then RlxUnqChkTyp selects the reference's relaxed uniqueness
// key (code/ltli6463.cu:57-141) above 126 trace positions, as the reference
// always does.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/ltl_amd.h"

static int usage(const char *argv0) {
    printf("Arguments should be in the form of\n");
    printf("-----------------------------------------------------------------\n");
    printf("%s [--reference-keys] [--launch-log] <input_file_address> <c1> <c2> <c3> <c4> <c5> <c6> <c7> <c8> <maxCost> "
           "<RlxUnqChkTyp> <NegType>\n",
           argv0);
    printf("-----------------------------------------------------------------\n");
    printf("\nFor example\n");
    printf("-----------------------------------------------------------------\n");
    printf("%s ./input.txt 1 1 1 1 1 1 1 1 500 3 2\n", argv0);
    printf("-----------------------------------------------------------------\n");
    return 2;
}

// Column letters of the alphabet line, in file order (the token letters are
// printed in this order, code/ltli6463.cu:1531-1543, 1645-1652).
static std::string file_alphabet(const char *path) {
    std::ifstream f(path);
    std::string line, last;
    while (std::getline(f, line))
        if (!line.empty() && line != "\r") last = line;
    std::string a;
    for (char c : last)
        if (c >= 'a' && c <= 'z') a += c;
    return a;
}

static void print_launch(const ltl_launch_record *r, void *) {
    static const char ops[] = "A~&|XFGU";
    printf("Cost %-2d | (%c) | AllLTLs: %-11lu | StoredLTLs: %-10ld | ToBeChecked: %-10lu %s\n", r->cost,
           r->op >= 0 && r->op < 8 ? ops[r->op] : '?', (unsigned long)r->all_ltls, (long)r->stored,
           (unsigned long)r->to_check, r->pass ? "" : r->launched ? "| test" : "| ahead");
}

static bool env_flag(const char *name) { return getenv(name) && std::atoi(getenv(name)) != 0; }

int main(int argc, char **argv) {
    bool ref_keys = env_flag("LTLI_REFERENCE_KEYS");
    bool launch_log = env_flag("LTLI_LAUNCH_LOG");
    const char *prog = argv[0];
    while (argc > 13) {
        const std::string a = argv[1];
        if (a == "--reference-keys") ref_keys = true;
        else if (a == "--launch-log") launch_log = true;
        else break;
        argv[1] = argv[0];
        ++argv;
        --argc;
    }
    if (argc != 13) return usage(prog);
    bool bad = false;
    for (int i = 2; i < argc - 2; ++i) {
        const int v = std::atoi(argv[i]);
        if (v <= 0 || v > SHRT_MAX) {
            printf("Argument number %d, \"%s\", should be a positive short integer.\n", i, argv[i]);
            bad = true;
        }
    }
    const int rlx = std::atoi(argv[11]), neg = std::atoi(argv[12]);
    if (rlx < 1 || rlx > 3) {
        printf("Argument number 11, RlxUnqChkTyp = \"%s\", should be 1, 2, or 3.\n", argv[11]);
        bad = true;
    }
    if (neg < 1 || neg > 2) {
        printf("Argument number 12, NegType = \"%s\", should be 1, or 2.\n", argv[12]);
        bad = true;
    }
    if (bad) return 2;

    // Traces of up to 255 positions (the reference stops at 63,
    // code/ltli6463.cu:1078): read in four words per trace, then repacked to
    // the fewest words that hold the longest trace (one word up to 64).
    int max_traces = 1;
    {
        std::ifstream f(argv[1], std::ios::binary);
        for (std::string line; std::getline(f, line);) ++max_traces;
    }
    std::vector<uint8_t> lens(max_traces);
    std::vector<uint64_t> atoms((size_t)64 * max_traces * 4);
    std::vector<char> names(64 * 40);
    int32_t P = 0, N = 0, A = 0, longest = 0;
    int32_t rc = ltl_read_trace_file_names(argv[1], max_traces, 4, &P, &N, &A, &longest, lens.data(), atoms.data(),
                                           names.data(), (int32_t)names.size());
    if (rc == LTL_E_ARG) {
        printf("Failed to open the input file.\n");
        return 1;
    }
    if (rc == LTL_E_LIMIT) {
        printf("In this version, The input can have:\n");
        printf("1) Max(len(trace)) = 255.\n");
        printf("2) At most 64 propositions, named by letters or by comma-separated names.\n");
        return 1;
    }
    const int W = P + N;
    const int K = (std::max(longest, 1) + 63) / 64;
    for (size_t x = 0; x < (size_t)A * W; ++x)
        for (int k = 0; k < K; ++k) atoms[x * K + k] = atoms[x * 4 + k];

    ltl_options opt;
    ltl_default_options(&opt);
    for (int i = 0; i < 8; ++i) opt.cost[i] = (uint16_t)std::atoi(argv[i + 2]);
    opt.max_cost = (uint16_t)std::atoi(argv[10]);
    opt.neg_type = (uint8_t)neg;
    if (ref_keys) opt.dedup_mode = (uint8_t)rlx;
    // Device pool: 16 GiB unless LTLI_HBM_BUDGET_GB says otherwise (0 = 75% of
    // the free memory).  Reserving most of a 288 GB device costs seconds per
    // process, against milliseconds of learning for the reference's inputs.
    const char *bud = getenv("LTLI_HBM_BUDGET_GB");
    opt.hbm_budget_bytes = (uint64_t)((bud ? std::atof(bud) : 16.0) * (double)(1ull << 30));
    int32_t st = 0;
    ltl_ctx *ctx = ltl_create(&opt, &st);
    if (!ctx) {
        printf("No usable GPU (error %d).\n", st);
        return 1;
    }
    if (launch_log) ltl_set_launch_log(ctx, print_launch, nullptr);
    ltl_problem pb{P, N, A, K, lens.data(), atoms.data(), names.data()};
    ltl_result res;
    const auto t0 = std::chrono::high_resolution_clock::now();
    rc = ltl_learn(ctx, &pb, &res);
    const auto t1 = std::chrono::high_resolution_clock::now();

    // (name, atom) in the order a position's token lists them: the letters
    // of the alphabet line in file order, or the names in column order
    std::vector<std::pair<std::string, int>> show;
    const std::string nm(names.data());
    if (nm.find(',') == std::string::npos) {
        for (char c : file_alphabet(argv[1])) show.push_back({std::string(1, c), (int)nm.find(c)});
    } else {
        std::string cur;
        for (char c : nm + ",") {
            if (c == ',') { show.push_back({cur, (int)show.size()}); cur.clear(); }
            else cur += c;
        }
    }
    auto print_traces = [&](const char *title, int from, int to) {
        printf("\n%s: \n", title);
        for (int t = from; t < to; ++t) {
            printf("\t");
            for (int k = 0; k < lens[t]; ++k) {
                std::string s;
                for (const auto &na : show) {
                    const int a = na.second;
                    if (a >= 0 && ((atoms[((size_t)a * W + t) * K + k / 64] >> (k % 64)) & 1)) {
                        s += na.first;
                        s += ", ";
                    }
                }
                printf("{%s}\t", s.substr(0, s.size() >= 2 ? s.size() - 2 : 0).c_str());
            }
            printf("\n");
        }
    };
    print_traces("Positive", 0, P);
    print_traces("Negative", P, W);
    printf("\nCost Function: p:%u, ~:%u, &:%u, |:%u, X:%u, F:%u, G:%u, U:%u", opt.cost[0], opt.cost[1],
           opt.cost[2], opt.cost[3], opt.cost[4], opt.cost[5], opt.cost[6], opt.cost[7]);
    printf("\nNumber of Traces: %d", W);
    if (rc >= 0) {
        const double us = (double)std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count();
        printf("\nNumber of All LTLs: %lu", (unsigned long)res.all_ltls);
        printf("\nCost of Final LTL: %d", res.cost);
        printf("\nRunning Time: %f s", us * 0.000001);
        printf("\n\nLTL: \"%s\"\n", res.formula ? res.formula : "not_found");
    } else {
        printf("\n\nLTL: \"see_the_error\"\n");
    }
    ltl_result_free(&res);
    ltl_destroy(ctx);
    return rc < 0 ? 1 : 0;
}
