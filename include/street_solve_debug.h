/*
 * street_solve_debug.h -- test entry points of libstreet_solve.so.
 *
 * No reference counterpart: these expose internals (the device hand
 * evaluator, raw CFR state after k iterations, the plan's byte accounting)
 * so parity tests can check them against the CPU oracle. Not part of the
 * solver ABI; callers of the solver never need them.
 */
#ifndef STREET_SOLVE_DEBUG_H
#define STREET_SOLVE_DEBUG_H

#include <stdint.h>
#include "street_solve.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Ranks of n 7-card hands (cards[7*i .. 7*i+6]) with the LDS-table evaluator;
 * on_device = 0 runs the same code on the host. */
int ssdbg_eval7(const int *cards, int n, uint32_t *out, int on_device);

/* Device ranks of all 1326 combos on a 5-card board (0 = blocked), combo k of
 * cards a < b at k = 51a - a(a-1)/2 + b - a - 1. */
int ssdbg_combo_ranks(const int *board5, uint32_t *out1326);

/* Device census over all C(52,7) hands: counts per category 0..9, rank sum. */
int ssdbg_eval_census(unsigned long long counts[10], unsigned long long *rank_sum);

/* Raw state after `iterations` CFR iterations (form 0 plain, 1 batched kernel,
 * 2 plain with the BF16 precision showdown of 2-player rivers),
 * in the reference layout [num_nodes][8][max_hands]; root_ev[max_hands]. */
int ssdbg_solve_state(const SSProblem *problem, int iterations, int form, float *regrets, float *strat_sum,
                      float *root_ev);

/* Algorithmic bytes of one CFR iteration and the device footprint. */
int ssdbg_plan_bytes(const SSProblem *problem, int64_t *algo_bytes_per_iter, int64_t *footprint);

/* Device copy of `bytes` bytes (multiple of 16) between two fresh buffers,
 * `repeats` times, by a plain streaming kernel (k_copy): the known-size
 * dispatch that calibrates HBM byte counters. Returns the mean ms per copy in
 * *ms when ms is not NULL. */
int ssdbg_copy(int64_t bytes, int repeats, float *ms);

/* Launch shapes of the phased form, for tests of their equivalence (every
 * shape gives the same results bit for bit); -1 restores the automatic
 * choice. fwd_group_min: forward-sweep levels with at least this many tasks
 * run three tasks per workgroup (0: every level). term_small: 0 never runs a
 * pass's terminal classes as one launch, 1 when they fit (the default).
 * sd_parts: workgroups per two-opponent showdown (1-4) when the classes run
 * as separate launches. Applies to task lists built afterwards; not
 * thread-safe against concurrent solves. */
int ssdbg_set_phased_shape(int fwd_group_min, int term_small, int sd_parts);

/* Python-compatible draws: for t < count and k < num_lists, out[t*num_lists+k]
 * = the index random.Random(seed).choice picks from a list of lens[k] items,
 * drawn in that order (the opponent sampling of ss_leaf_equity). Host only. */
int ssdbg_py_choice(int seed, const int *lens, int num_lists, int count, int *out);

#ifdef __cplusplus
}
#endif
#endif
