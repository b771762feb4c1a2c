/*
 * ltl_oracle.h -- CPU oracle for the LTL characteristic-set (CS) enumeration.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  Nothing in the product links against it.
 *
 * It restates, in plain C, the algorithm of the reference learner
 * (code/ltli6463.cu of MojtabaValizadeh/ltl-learning-on-gpus):
 *   - CS encoding: one uint64 per trace, bit k = "holds at position k"
 *     (code/ltli6463.cu:1100-1133); traces longer than 64 positions take K
 *     words each (an extension: the reference stops at 63 positions,
 *     code/ltli6463.cu:1078);
 *   - the seven operators (code/ltli6463.cu:203, 295, 385, 468, 551-556,
 *     638-645, 739-771);
 *   - the target check (code/ltli6463.cu:240-243), extended with a Hamming
 *     tolerance (misclassified traces <= k);
 *   - the cost-level loop, split loops, operator order and NegType
 *     (code/ltli6463.cu:1199-1447);
 *   - the store-overflow "onTheFly"/"lastRound" semantics
 *     (code/ltli6463.cu:1021-1027, 1196-1206, 1442-1443);
 *   - the reference's relaxed uniqueness keys, modes 1/2/3 and the exact
 *     126-bit packing (code/ltli6463.cu:57-141);
 *   - formula printing (code/ltli6463.cu:877-928);
 *   - trace-file reading (code/ltli6463.cu:1500-1573).
 *
 * Canonical order.  Candidates are numbered by a global 64-bit id equal to
 * the running candidate count ("allLTLs", code/ltli6463.cu:1233 etc.): atoms
 * get ids 0..A-1, then each (operator, split) block in the reference's order,
 * binary blocks left-major / right-minor, and for U each pair emits lUr then
 * rUl (code/ltli6463.cu:756-771).  This is the order the reference launches
 * when a whole block fits one launch.  Dedup keeps the first occurrence in
 * this order; the returned formula is the first accepted candidate.
 */
#ifndef LTL_ORACLE_H
#define LTL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_ATOM = 0, ORC_NOT = 1, ORC_AND = 2, ORC_OR = 3,
       ORC_NEXT = 4, ORC_FIN = 5, ORC_GLOB = 6, ORC_UNTIL = 7 };

typedef struct {
    int32_t num_pos, num_neg, num_atoms;
    const uint8_t  *lens;   /* [W], 1..64 * K */
    const uint64_t *atoms;  /* [num_atoms * W * K], atom-major, then trace,
                               then word (word k = positions 64k..64k+63) */
    const char     *names;  /* num_atoms letters (may be NULL: 'a'+i) */
    int32_t words_per_trace; /* K: 0 or 1 = one word per trace, up to 4 */
    int32_t pad_;
} orc_problem;

typedef struct {
    int32_t cost[8];     /* p, ~, &, |, X, F, G, U */
    int32_t max_cost;
    int32_t neg_type;    /* 1: negate any formula, 2: atoms only */
    int32_t hamming_k;   /* accept iff misclassified <= k (0 = exact) */
    int32_t dedup_mode;  /* 0: exact (full CS); 1,2,3: reference keys */
    int32_t keep_store;  /* copy the store into the result */
    int64_t capacity;    /* max stored CSs incl. atoms; 0 = unlimited */
    int32_t threads;     /* >1: OpenMP candidate-parallel blocks */
    int32_t pad_;
} orc_options;

typedef struct {
    int32_t status;      /* 1 found, 0 not found, <0 error */
    int32_t cost;        /* cost of the result, or last level processed */
    uint64_t all_ltls;   /* candidates in all processed blocks */
    int64_t found_id;    /* canonical id of the returned candidate, -1 */
    int32_t root_op, root_left, root_right;
    int32_t on_the_fly;
    int64_t n_stored;
    int32_t n_levels;    /* level_start/level_end hold max_cost+2 entries */
    int32_t pad_;
    int64_t *level_start, *level_end;
    uint64_t *words;     /* [n_stored * W] when keep_store */
    uint8_t  *op;
    int32_t  *left, *right;
    char     *formula;
    double    seconds;   /* wall time of the enumeration */
} orc_result;

/* Operators on one trace word (len = trace length, 1..64). */
uint64_t orc_op_not(uint64_t x, int len);
uint64_t orc_op_next(uint64_t x);
uint64_t orc_op_fin(uint64_t x);
uint64_t orc_op_glob(uint64_t x, int len);
uint64_t orc_op_until(uint64_t l, uint64_t r);

/* Operator op on one trace of K words (word k = positions 64k..64k+63,
 * length len <= 64K); y is the right operand of & | U, else unused. */
void orc_trace_op(int32_t op, const uint64_t *x, const uint64_t *y, int32_t K, int32_t len, uint64_t *out);

/* Misclassified-trace count of a CS (bit 0 vs label; positives first). */
int32_t orc_mismatches(const uint64_t *cs, int32_t num_pos, int32_t num_neg);

/* Reference relaxed-uniqueness key (code/ltli6463.cu:57-141). */
void orc_relaxed_key(const uint64_t *cs, const uint8_t *lens, int32_t W,
                     int32_t mode, int32_t len_sum, uint64_t *h, uint64_t *l);

int32_t orc_enumerate(const orc_problem *pb, const orc_options *opt, orc_result *res);
void orc_result_free(orc_result *res);

/* Textbook per-position evaluation of a formula in the reference grammar.
 * Returns 0 and fills out_cs[W * K], or a negative parse error. */
int32_t orc_eval_formula(const char *formula, const orc_problem *pb, uint64_t *out_cs);
/* Symbol count (= cost under unit costs), or negative on parse error. */
int32_t orc_formula_cost(const char *formula, const int32_t *cost);

/* Syntactic enumeration without dedup, per-position semantics.  Returns the
 * minimal accepted cost, -1 if none <= max_cost, -2 if the formula count
 * exceeds max_formulas. */
int32_t orc_bruteforce_min_cost(const orc_problem *pb, const orc_options *opt,
                                int64_t max_formulas);

/* Parse a trace file in the reference format.  Fills W, lens[] (capacity
 * max_traces), atoms[] (capacity 26*max_traces, atom-major over the sorted
 * alphabet), names[27].  Returns 0 or a negative error. */
int32_t orc_read_trace_file(const char *path, int32_t max_traces,
                            int32_t *num_pos, int32_t *num_neg, int32_t *num_atoms,
                            uint8_t *lens, uint64_t *atoms, char *names);

#ifdef __cplusplus
}
#endif
#endif
