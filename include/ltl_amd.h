/*
 * ltl_amd.h -- C ABI of the MI355X LTL learner (drop-in boundary).
 *
 * The reference learner (MojtabaValizadeh/ltl-learning-on-gpus) exposes its
 * hot path through one in-process C++ function,
 *
 *   std::string LTLI(const unsigned short *costFun, unsigned short maxCost,
 *                    unsigned RlxUnqChkTyp, unsigned NegType,
 *                    std::set<char> alphabet, int &LTLcost,
 *                    std::uint64_t &allLTLs, pos, neg)
 *                                                  (code/ltli6463.cu:1043-1052)
 *
 * and through the CLI that wraps it (code/ltli6463.cu:1575-1681).  This
 * header replaces that function with plain-C entry points: plain pointers and
 * sizes, no C++ or torch types.  All entry points are synchronous on the
 * calling host thread.  One context per host thread; a context is not
 * re-entrant.  Nothing is printed by the library.
 */
#ifndef LTL_AMD_H
#define LTL_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference returns strings ("not_found",
 * "see_the_error", code/ltli6463.cu:1082, 1461) and always exits 0; here
 * every entry point returns one of these. */
#define LTL_FOUND        1   /* a formula was found */
#define LTL_NOT_FOUND    0   /* none up to max_cost ("not_found") */
#define LTL_E_ARG       (-1) /* invalid argument */
#define LTL_E_LIMIT     (-2) /* instance beyond limits ("see_the_error") */
#define LTL_E_OOM       (-3) /* device memory exhausted */
#define LTL_E_HIP       (-4) /* HIP runtime error */
#define LTL_E_COLLISION (-5) /* dedup fingerprint collision detected */
#define LTL_E_NODEVICE  (-6) /* no GPU visible */

/* Operator codes, in the cost-vector order of the reference
 * (costFun = p, ~, &, |, X, F, G, U; code/ltli6463.cu:1142-1150). */
enum { LTL_ATOM = 0, LTL_NOT = 1, LTL_AND = 2, LTL_OR = 3,
       LTL_NEXT = 4, LTL_FIN = 5, LTL_GLOB = 6, LTL_UNTIL = 7 };

/* Dedup modes.  EXACT compares full characteristic sets; REF1..REF3 are the
 * reference's relaxed uniqueness keys (RlxUnqChkTyp, code/ltli6463.cu:57-141). */
enum { LTL_DEDUP_EXACT = 0, LTL_DEDUP_REF1 = 1, LTL_DEDUP_REF2 = 2, LTL_DEDUP_REF3 = 3 };

/* Candidate order within the last cost level. */
enum { LTL_ORDER_PLAIN = 0, LTL_ORDER_LEARNED = 1 };

/* A learning instance.  Replaces the reference's pos/neg token vectors and
 * alphabet set (code/ltli6463.cu:1043-1052): the atom characteristic sets
 * are given already encoded, one uint64 per trace, bit k set iff the atom
 * holds at position k (the encoding of code/ltli6463.cu:1100-1133).
 * Traces 0..num_pos-1 are positive, the rest negative.
 * Traces longer than 64 positions (the reference stops at 63,
 * code/ltli6463.cu:1078) take words_per_trace = K words each, 2 <= K <= 4:
 * word k holds positions 64k..64k+63, atoms[(a*W + i)*K + k], lengths up to
 * 64K (and 255).  K = 0 or 1 is the one-word layout.  Relaxed dedup modes
 * and the batched launch take one-word instances only (ltl_learn_batch
 * hands the others to ltl_learn). */
typedef struct {
    int32_t num_pos;
    int32_t num_neg;
    int32_t num_atoms;
    int32_t words_per_trace;
    const uint8_t  *lens;   /* [num_pos+num_neg], each 1..64*K */
    const uint64_t *atoms;  /* [num_atoms * (num_pos+num_neg) * K], atom-major */
    const char     *names;  /* num_atoms single-character names; or, for more
                               than 26 propositions, a comma-separated list of
                               num_atoms names, each a lowercase letter or _
                               followed by lowercase letters, digits or _
                               ("p0,p1,...,p29"); or NULL (a, b, c, ...).  The
                               returned formula spells atoms this way. */
} ltl_problem;

/* Options.  cost/max_cost/neg_type/dedup_mode carry the reference's
 * costFun, maxCost, NegType and RlxUnqChkTyp (code/ltli6463.cu:1581-1620). */
typedef struct {
    uint16_t cost[8];          /* p, ~, &, |, X, F, G, U; each >= 1 */
    uint16_t max_cost;
    uint8_t  neg_type;         /* 1: negate any formula, 2: atoms only */
    uint8_t  dedup_mode;       /* LTL_DEDUP_* */
    int32_t  hamming_k;        /* accept iff misclassified traces <= k */
    int32_t  order;            /* LTL_ORDER_* */
    int32_t  hash_slots_log2;  /* dedup table size; 0 = sized from budget */
    int32_t  precheck;         /* 1 = test a level on per-trace summaries
                                  before generating it (default) */
    int32_t  group_lanes;      /* lanes per characteristic set; 0 = tuned */
    int64_t  capacity;         /* max stored sets; 0 = sized from budget */
    uint64_t hbm_budget_bytes; /* 0 = 75% of free device memory at the first call */
    int32_t  device;           /* HIP device ordinal */
    int32_t  profile;          /* 1 = time every kernel with HIP events, 2 = the generation
                                  and emit kernels only (fewer dispatch barriers), 3 = the
                                  kinds in profile_kinds only; each timed launch costs a few
                                  microseconds of dispatch barrier */
    uint32_t profile_kinds;    /* profile 3: bit k = kernel kind k of ltl_get_kernel_stats */
    int32_t  reserved[5];
} ltl_options;

/* One node of the returned formula, rebuilt on the host from parent indices
 * (cf. the device BFS and host recursion of code/ltli6463.cu:29-55, 877-928).
 * Node 0 is the root; left/right index into the node array (-1 if none). */
typedef struct {
    int32_t op;      /* LTL_* operator code */
    int32_t atom;    /* atom index for LTL_ATOM, else -1 */
    int32_t left;
    int32_t right;
} ltl_node;

typedef struct {
    int32_t  status;      /* LTL_FOUND / LTL_NOT_FOUND / LTL_E_* */
    int32_t  cost;        /* the reference's LTLcost out-parameter */
    uint64_t all_ltls;    /* the reference's allLTLs out-parameter */
    int64_t  found_id;    /* canonical candidate id of the result, -1 */
    int64_t  n_stored;    /* characteristic sets stored (incl. atoms) */
    int32_t  on_the_fly;  /* store overflowed into check-only mode */
    int32_t  n_nodes;
    ltl_node *nodes;      /* owned by the library */
    char     *formula;    /* reference grammar, or "not_found"; owned */
    double   seconds;     /* wall time of the call */
    double   kernel_seconds; /* summed device time of the enumeration kernels */
    uint64_t hbm_bytes;   /* algorithmic bytes of the enumeration kernels */
    uint64_t candidates;  /* candidates generated (full sets computed) */
    uint64_t checked;     /* candidates tested on summaries only */
    int64_t  words_deferred; /* sets of the last generated level stored with
                              * parents only, neither words nor summaries
                              * (plain order: no later pass reads them); the
                              * next ltl_store_read or ltl_store_summaries
                              * recomputes them */
    int32_t  deferred_succ;  /* bit j: the last level's unary block j (~, X,
                              * F, G over the deferred level) was tested in
                              * that level's generation */
    int32_t  retries;     /* repeats of the call with fresh fingerprint keys
                             after a detected fingerprint collision (the
                             result does not depend on the keys) */
    int64_t  words_unwritten; /* sets of the level below the last generated
                               * one stored with summaries and parents but
                               * without words (plain order, large levels:
                               * no later pass reads them in bulk); the next
                               * ltl_store_read recomputes them */
} ltl_result;

typedef struct ltl_ctx ltl_ctx;

/* Fill *opt with the defaults: unit costs, max_cost 10, neg_type 1, exact
 * dedup, plain order, precheck on. */
void ltl_default_options(ltl_options *opt);

/* Create a context on opt->device.  Device memory is reserved lazily and
 * reused across calls (the reference allocates and frees about 5 GB per
 * call, code/ltli6463.cu:1153-1187, 1486-1493).  Returns NULL on error;
 * *status (if not NULL) receives the LTL_E_* code. */
ltl_ctx *ltl_create(const ltl_options *opt, int32_t *status);

/* Replace the options of an existing context. */
int32_t ltl_set_options(ltl_ctx *ctx, const ltl_options *opt);

/* Learn the minimal-cost formula separating the instance.  Replaces
 * LTLI() (code/ltli6463.cu:1043).  Returns res->status. */
int32_t ltl_learn(ltl_ctx *ctx, const ltl_problem *pb, ltl_result *res);

/* Allocation-free learning call.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same as ltl_learn
 * on a context whose hbm_budget_bytes is LTL_NOALLOC_SCRATCH_BYTES (same
 * store capacity, table and chunking, so the same result), except that the
 * call's device memory is a static scratch region of that size inside the
 * library, one per GPU, held until the process ends: the call allocates no
 * device or pinned memory and creates no events.  Plain order only, no log,
 * no profiling (LTL_E_ARG otherwise).  An instance whose buffers do not fit
 * the region returns LTL_E_LIMIT.  Every call on a GPU shares that GPU's
 * region and leaves its contents there (ltl_store_read of the context reads
 * them until the next call): calls that overlap on one GPU, from any
 * contexts, race. */
#define LTL_NOALLOC_SCRATCH_BYTES (128ull << 20)
int32_t ltl_learn_noalloc(ltl_ctx *ctx, const ltl_problem *pb, ltl_result *res);

/* Learn n independent instances in one persistent launch (no reference
 * counterpart; the reference runs one process per instance through
 * code/dc.py:23-43).  res[i] receives instance i.  Returns the number of
 * instances solved, or a negative LTL_E_* code.  Instances handed to the
 * whole-device path run up to six at once (plain order, no log): the
 * context keeps up to five helper contexts for them, each with an eighth
 * of its HBM budget (a sixteenth of the free memory, at most 8 GiB, when the
 * budget is 0), released by ltl_destroy. */
int32_t ltl_learn_batch(ltl_ctx *ctx, const ltl_problem *pbs, int32_t n, ltl_result *res);

/* Parameters of the batched launch (no reference counterpart).
 *   tile         candidates per workgroup step = threads per workgroup
 *                (256, 512 or 1024)
 *   word_block   loads in flight per lane in the word loops (4 or 8)
 *   workgroups   persistent workgroups (0 = 2 per CU, 1 per CU for tile 1024)
 *   defer_limit  candidates of one cost level above which an instance is
 *                handed to the whole-device path (ltl_learn) */
typedef struct {
    int32_t tile;
    int32_t word_block;
    int32_t workgroups;
    int32_t reserved;
    int64_t defer_limit;
} ltl_batch_params;

typedef struct {
    ltl_batch_params params;
    double seconds;      /* wall time of ltl_learn_batch with these params */
    int32_t deferred;    /* instances handed to ltl_learn */
    int32_t reserved;
} ltl_batch_trial;

typedef struct {
    double seconds;         /* wall time of the last ltl_learn_batch */
    double kernel_seconds;  /* the persistent launch alone */
    int32_t n;              /* instances */
    int32_t on_device;      /* finished inside the persistent launch */
    int32_t deferred;       /* handed to ltl_learn */
    int32_t host;           /* answered on the host (an atom separates) */
} ltl_batch_stats;

int32_t ltl_set_batch_params(ltl_ctx *ctx, const ltl_batch_params *p);
int32_t ltl_get_batch_params(const ltl_ctx *ctx, ltl_batch_params *p);
int32_t ltl_batch_last_stats(const ltl_ctx *ctx, ltl_batch_stats *s);

/* Project autotuner: time ltl_learn_batch on these instances for a grid of
 * (tile, word_block, defer_limit), keep the fastest in the context, and
 * report up to max_trials trials.  Returns the number of trials. */
int32_t ltl_autotune_batch(ltl_ctx *ctx, const ltl_problem *pbs, int32_t n, ltl_batch_trial *trials,
                           int32_t max_trials);

/* Free what a result owns (formula, nodes). */
void ltl_result_free(ltl_result *res);

/* Copy the context's current options. */
int32_t ltl_get_options(const ltl_ctx *ctx, ltl_options *opt);

/* Release the context and its device memory. */
void ltl_destroy(ltl_ctx *ctx);

/* ---- Introspection (no reference counterpart; the reference keeps a
 * commented-out dump of its cache, code/ltli6463.cu:1463-1484) ---------- */

/* Per-kernel statistics of the last call (profile = 1). */
typedef struct {
    char     name[24];
    uint64_t launches;
    double   seconds;      /* summed device time (HIP events) */
    uint64_t alg_bytes;    /* algorithmic bytes: operand reads per candidate,
                              table slot, temporaries, survivor writes */
    uint64_t min_bytes;    /* compulsory bytes: each operand set read once */
    uint64_t items;        /* candidates (gen) or survivors (emit) */
} ltl_kernel_stat;

/* Copy up to max entries; returns the number of kernel kinds. */
int32_t ltl_get_kernel_stats(const ltl_ctx *ctx, ltl_kernel_stat *out, int32_t max);

/* Read stored sets [first, first+count) of the last call, unpacked to one
 * uint64 per trace (cs[k*W + i]), or K per trace (cs[(k*W + i)*K + j]) for
 * K words per trace; op/left/right may be NULL. */
int32_t ltl_store_read(ltl_ctx *ctx, int64_t first, int64_t count, uint64_t *cs,
                       uint8_t *op, int32_t *left, int32_t *right);

/* The per-trace summaries of stored sets [first, first+count) of the last
 * call, in trace order: vector v (0: position 0 holds, 1: position 1 holds,
 * 2: some position holds, 3: every position holds) of set k in
 * sums[(k*4 + v) * ceil(W/64) + i/64], bit i%64 for trace i.  Introspection
 * for tests (no reference counterpart). */
int32_t ltl_store_summaries(ltl_ctx *ctx, int64_t first, int64_t count, uint64_t *sums);

/* Level boundaries of the last call: start[c], end[c] for c < max. */
int32_t ltl_levels(const ltl_ctx *ctx, int64_t *start, int64_t *end, int32_t max);

/* ---- Per-launch log -------------------------------------------------------
 * The reference prints a line per kernel launch (code/ltli6463.cu:1096,
 * 1225, 1259, 1293, 1325, 1357, 1389, 1423):
 *   Cost N | (op) | AllLTLs: a | StoredLTLs: s | ToBeChecked: n
 * With a callback set, ltl_learn and ltl_learn_noalloc call it on the calling
 * thread once for the atoms and once per block of candidates (one operator
 * over one operand level or pair of levels: the reference launches such a
 * block in chunks of its temporary capacity), in enumeration order, before
 * the block's kernels are queued.  ltl_learn_batch logs nothing. */
typedef struct {
    int32_t  cost;       /* the level's cost */
    int32_t  op;         /* LTL_* operator; LTL_ATOM for the atoms */
    int32_t  pass;       /* 1: generation (tested, deduplicated, stored);
                            0: a check pass (tested only) */
    int32_t  launched;   /* 0: registered without a launch (tested ahead of
                            the level, or in the generation of the level below) */
    uint64_t all_ltls;   /* canonical id of the block's first candidate (AllLTLs) */
    int64_t  stored;     /* sets stored when the level started (StoredLTLs) */
    uint64_t to_check;   /* candidates of the block (ToBeChecked) */
} ltl_launch_record;
typedef void (*ltl_launch_fn)(const ltl_launch_record *rec, void *user);
/* Set (fn not NULL) or clear the per-launch callback.  Returns 0 or
 * LTL_E_ARG. */
int32_t ltl_set_launch_log(ltl_ctx *ctx, ltl_launch_fn fn, void *user);

/* ---- Divide and conquer --------------------------------------------------
 * Replaces the reference's Python drivers DetSplit / RandSplit
 * (code/dc.py:188-249, :252-321) that call the CLI once per window: the
 * traces are split recursively until a window of at most `window` traces is
 * learned by ltl_learn (in process, the context's options), and the answers
 * are combined with & and |.  RandSplit samples windows with a seeded
 * SplitMix64 stream, reseeded on every call as dc.py:255 reseeds Python's
 * random module.  res->formula separates all traces; res->cost is its cost
 * under the context's costs (not minimal in general); res->all_ltls sums
 * the sub-calls' counts; res->n_stored is the number of ltl_learn calls. */
enum { LTL_DC_DET = 0, LTL_DC_RAND = 1 };
int32_t ltl_learn_dc(ltl_ctx *ctx, const ltl_problem *pb, int32_t window, int32_t algo, uint64_t seed,
                     ltl_result *res);
/* ltl_learn_dc over n problems at once (the reference runs dc.py once per
 * benchmark file, runDC, code/dc.py:373-387): one driver per problem, in lock-step;
 * whenever every driver waits for a window, the windows of all of them are
 * learned by one ltl_learn_batch call, and the formulas returned are matched
 * on the device in one launch (groups of up to 256 problems, one host thread
 * each).  res[i] is what ltl_learn_dc returns for pbs[i] (res[i].seconds:
 * its group's time).  Returns the number of problems with a formula, or
 * LTL_E_ARG. */
int32_t ltl_learn_dc_batch(ltl_ctx *ctx, const ltl_problem *pbs, int32_t n, int32_t window, int32_t algo,
                           uint64_t seed, ltl_result *res);
/* The reference's LTLmatch (code/dc.py:136-186) on the device: out[i*W + t]
 * = 1 iff formulas[i] (reference output grammar) holds at position 0 of
 * trace t.  Returns 0, LTL_E_ARG (a parse error) or LTL_E_LIMIT (a formula
 * of more than 2^16 nodes in an unbalanced shape, or more than 4 words per
 * trace). */
int32_t ltl_match(ltl_ctx *ctx, const ltl_problem *pb, const char *const *formulas, int32_t n, uint8_t *out);

/* ---- Learned operand ordering ------------------------------------------
 * No counterpart in the reference (its order is fixed by the loops of
 * code/ltli6463.cu:1209-1440).  A scorer is a BF16 MLP (24 -> 128 -> 128 -> 1,
 * FP32 accumulation) run on the matrix cores.  With order = LTL_ORDER_LEARNED
 * the check pass of every cost level tests the unary blocks, then the
 * (64-set x 64-set) tile pairs of the binary blocks in descending score, in
 * rounds; it stops at the first round that accepts a candidate.  Every lower
 * level has been tested completely, so the cost is the same as in plain
 * order; the formula is a separating formula of that cost, not necessarily
 * the one with the smallest canonical id. */
typedef struct ltl_scorer ltl_scorer;

#define LTL_SCORER_FEATURES 24
/* Parameter layout (floats): W1[128][32] (columns 24..31 unused), b1[128],
 * W2[128][128], b2[128], w3[128], b3. */
#define LTL_SCORER_PARAMS (128 * 32 + 128 + 128 * 128 + 128 + 128 + 1)

/* Create a scorer on a device with seeded initial weights; NULL on error. */
ltl_scorer *ltl_scorer_create(int32_t device, uint64_t seed, int32_t *status);
void ltl_scorer_destroy(ltl_scorer *sc);
int32_t ltl_scorer_get_params(ltl_scorer *sc, float *params);
/* Replace the parameters; resets the optimizer state. */
int32_t ltl_scorer_set_params(ltl_scorer *sc, const float *params);
/* y[i] = score of feature row x[i*24 .. i*24+23]. */
int32_t ltl_scorer_forward(ltl_scorer *sc, const float *x, int64_t n, float *y);
/* Gradient of the mean squared error mean_i (y_i - t_i)^2 with respect to
 * the parameters (layout above); *loss = that error.  The BF16 products
 * are those of the forward pass. */
int32_t ltl_scorer_grad(ltl_scorer *sc, const float *x, const float *t, int64_t n, float *grad, float *loss);
/* Train on log rows (24 features, then the label; rows[i*25 + 24]) with
 * Adam (beta 0.9 / 0.999), mini-batches of `batch` rows in a seeded
 * shuffled order, `epochs` passes.  *loss = mean batch loss of the last pass. */
int32_t ltl_scorer_train(ltl_scorer *sc, const float *rows, int64_t n, int32_t epochs, int32_t batch, float lr,
                         uint64_t seed, float *loss);
/* Attach a scorer to a context (same device); NULL detaches. */
int32_t ltl_set_scorer(ltl_ctx *ctx, ltl_scorer *sc);

/* Enumeration log: with max_rows > 0, every following ltl_learn call writes
 * up to per_block rows per binary block of each level's check pass (tile
 * pairs at an even stride over the block): the 24 features of the tile pair
 * and, as the label, 1 - (fewest misclassified traces of any candidate in
 * it) / W.  max_rows = 0 turns logging off. */
int32_t ltl_set_log(ltl_ctx *ctx, int64_t max_rows, int32_t per_block);
/* Copy up to max_rows rows (25 floats each) of the last call; returns the
 * number of rows the call logged. */
int64_t ltl_get_log(ltl_ctx *ctx, float *rows, int64_t max_rows);

/* ---- Host helpers around the boundary (no GPU needed) ------------------ */

/* Read a trace file in the reference format (code/ltli6463.cu:1500-1573):
 * positive lines, "---", negative lines, "---", operators, "---", alphabet.
 * Fills lens[max_traces], atoms[26*max_traces] (atom-major, stride W after
 * the call), names[27].  Returns 0 or LTL_E_ARG / LTL_E_LIMIT. */
int32_t ltl_read_trace_file(const char *path, int32_t max_traces,
                            int32_t *num_pos, int32_t *num_neg, int32_t *num_atoms,
                            uint8_t *lens, uint64_t *atoms, char *names);

/* The same for traces of up to 64*words_per_trace positions (at most 255;
 * the reference stops at 63, code/ltli6463.cu:1078): atoms[26*max_traces*K]
 * in the K-word layout of ltl_problem (stride W*K per atom after the call).
 * *max_len receives the longest trace, so that a caller can repack to one
 * word when it is at most 64. */
int32_t ltl_read_trace_file_words(const char *path, int32_t max_traces, int32_t words_per_trace,
                                  int32_t *num_pos, int32_t *num_neg, int32_t *num_atoms,
                                  int32_t *max_len, uint8_t *lens, uint64_t *atoms, char *names);

/* More than 26 propositions (an extension of the reference format, whose
 * alphabet is the letters a-z, code/ltli6463.cu:1521): when the alphabet line
 * is anything but single letters a-z, its comma-separated tokens (at most 64,
 * each a lowercase letter or _ followed by lowercase letters, digits or _,
 * distinct) name the columns in file order, and names[names_cap] receives
 * them as a comma-separated list for ltl_problem.names.  Files with a
 * letters-only alphabet line read as with ltl_read_trace_file_words.  atoms
 * holds [64 * max_traces * K].  Returns 0 or LTL_E_ARG / LTL_E_LIMIT. */
int32_t ltl_read_trace_file_names(const char *path, int32_t max_traces, int32_t words_per_trace,
                                  int32_t *num_pos, int32_t *num_neg, int32_t *num_atoms,
                                  int32_t *max_len, uint8_t *lens, uint64_t *atoms, char *names,
                                  int32_t names_cap);

/* Evaluate a formula in the reference output grammar on every trace
 * (bit-parallel; the product counterpart of code/dc.py:136-186).  Fills
 * out_cs[W * K] (K = words per trace); returns 0 or LTL_E_ARG on a parse
 * error. */
int32_t ltl_eval_formula(const char *formula, const ltl_problem *pb, uint64_t *out_cs);

/* Misclassified traces of a characteristic set, one word per trace (for K
 * words per trace pass every K-th word; 0 = separates). */
int32_t ltl_mismatches(const uint64_t *cs, int32_t num_pos, int32_t num_neg);

/* Library version string. */
const char *ltl_version(void);

#ifdef __cplusplus
}
#endif
#endif
