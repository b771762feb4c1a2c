/*
 * ltl_gen.h -- seeded synthetic workload generator.
 *
 * No counterpart in the reference (which ships trace files only, under
 * benchmarks/).  Produces the five benchmark workloads of this project:
 *   1  10+10 traces, 3 atoms, lengths U{4..8}, bits Bernoulli(0.5), planted
 *      F(a & X b); unit costs, max cost 10.
 *   2  500+500 traces, 5 atoms, lengths U{10..32}, per-atom rate U[0.3,0.7],
 *      planted G(~a | F b) & ~(c U d); max cost 14.
 *   3  1000+1000 traces, 8 atoms, length 64, Bernoulli(0.5), planted
 *      (a U (b & X c)) | G d, then 50 positive and 50 negative labels
 *      swapped (5% flipped); Hamming tolerance 100; max cost 12.
 *   4  one of 256 independent instances: 20-400 traces, positive share
 *      U[0.3,0.7], 2-6 atoms, lengths U{5..40}, random planted formula of
 *      size 5-9; max cost = planted size.
 *   5  1000+1000 traces, 6 atoms, lengths U{16..64}, random planted
 *      formula of size 11-13; max cost = planted size.
 * Labels come from the planted formula; traces are drawn until both classes
 * are full (positives first in the output).  Random numbers: SplitMix64
 * seeding a xoshiro256** stream (Vigna, public domain).
 */
#ifndef LTL_GEN_H
#define LTL_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t num_pos, num_neg, num_atoms;
    int32_t max_cost;       /* suggested max cost */
    int32_t hamming_k;      /* suggested Hamming tolerance */
    int32_t neg_type;       /* suggested NegType */
    uint8_t  *lens;         /* [W], malloc'd */
    uint64_t *atoms;        /* [num_atoms * W], malloc'd */
    char names[32];
    char planted[512];      /* planted formula, reference grammar */
} ltl_gen_instance;

/* Generate instance `index` of workload `config` (1..5) from `seed`.
 * Returns 0, or LTL_E_ARG for an unknown config. */
int32_t ltl_gen_config(int32_t config, uint64_t seed, int32_t index, ltl_gen_instance *out);

/* Generate a custom instance: explicit sizes and planted formula (NULL for
 * a random formula of size planted_size). */
int32_t ltl_gen_custom(uint64_t seed, int32_t num_pos, int32_t num_neg, int32_t num_atoms,
                       int32_t len_min, int32_t len_max, double bit_rate,
                       const char *planted, int32_t planted_size, ltl_gen_instance *out);

void ltl_gen_free(ltl_gen_instance *inst);

#ifdef __cplusplus
}
#endif
#endif
