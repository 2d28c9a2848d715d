/*
 * street_solve_leaf.h -- depth-limit leaf values computed on the GPU.
 *
 * Replaces the equity and bias loops of python/leaf_values.py, which the
 * reference runs in pure Python on the host before every turn or flop solve:
 *   ss_leaf_equity  <- _accumulate_nplayer_equity (leaf_values.py:573-691)
 *                      over the runout loops of compute_turn_leaf_values
 *                      (:761-793) and compute_flop_leaf_equity (:486-523)
 *   ss_leaf_values  <- the per-leaf bias expansion (:795-816 and :525-567)
 *                      with biased_leaf_value (:91-122)
 * The runout lists (numpy RandomState draws, leaf_values.py:757-763 and
 * :477-492) and the leaf infos (extract_leaf_info_from_tree, :189-262) stay
 * with the caller; poker-solver_amd/leaf_values.py builds both exactly as the
 * reference does and calls these two functions.
 *
 * Cards are 0..51 (rank * 4 + suit). Return 0 on success, a negative
 * SS_ERR_* code otherwise (message in ss_last_error()). Both calls are
 * synchronous and take host buffers; calls from several threads are
 * serialised. The device buffers they use, and a pinned host buffer that
 * large results are read back through, are kept per device between calls
 * (grown to the largest call so far) until the process ends.
 */
#ifndef STREET_SOLVE_LEAF_H
#define STREET_SOLVE_LEAF_H

#include "street_solve.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SS_LEAF_MAX_HANDS 1326
#define SS_LEAF_OPP_COMBO_LIMIT 2000 /* leaf_values.py:629 OPP_COMBO_LIMIT */
#define SS_LEAF_SAMPLE_SEED 42       /* leaf_values.py:635 random.Random(42) */

/* Per-hand showdown equity averaged over runouts.
 *   hands       [num_players][max_hands][2] hole cards (rows >= num_hands[p]
 *               are ignored)
 *   runouts     [num_runouts][5] five-card boards
 *   raw_equity  [num_players][max_hands] out: for hand h of player p, the
 *               mean over the runouts where h and at least one opponent tuple
 *               are unblocked of (wins + sum 1/ties) / tuples, over all
 *               opponent tuples (one unblocked hand per other player, no
 *               shared cards) when there are <= 2000 of them, else over 2000
 *               tuples drawn as random.Random(42).choice per opponent; 0 for
 *               hands never counted. Range weights play no part (as in the
 *               reference). */
int ss_leaf_equity(int num_players, const int *hands, const int *num_hands, int max_hands,
                   const int *runouts, int num_runouts, double *raw_equity);

/* Leaf values from equity.
 *   leaf_idx, leaf_pot  [num_leaf_infos] original leaf index and pot (chips)
 *   leaf_values  [num_leaf_infos * 4^num_players][num_players][max_hands]
 *                out: row leaf_idx * 4^NP + combo holds, for h < num_hands[p],
 *                biased_leaf_value(raw_equity[p][h], pot, s_p, s_opp) with
 *                s_i = (combo >> 2i) & 3 and s_opp the other player's profile
 *                (two players) or round-half-even of the mean of the others'
 *                (3-6 players); rows no leaf info names, and hands beyond
 *                num_hands[p], are 0. A later info with the same leaf_idx
 *                overwrites an earlier one; rows past the array are dropped. */
int ss_leaf_values(int num_players, const double *raw_equity, const int *num_hands, int max_hands,
                   const int *leaf_idx, const int *leaf_pot, int num_leaf_infos, float *leaf_values);

/* Leaf values evaluated from equity straight into the device leaf buffer
 * of subgame `index` of a batch session (ss_batch_create with any leaf
 * values, e.g. zeros; include/street_solve.h), without the round trip of
 * the [rows][NP][max_hands] array through the host: the next ss_batch_run
 * solves with them. raw_equity is ss_leaf_equity's [NP][max_hands] output
 * (max_hands = the subgame's largest range); leaf_idx / leaf_pot as for
 * ss_leaf_values.
 *   tree_order = 0: tree leaf l takes profile row l, as the reference passes
 *     the producer's array to the solver (python/street_solver_gpu.py:
 *     299-311); num_leaf_infos * 4^NP must equal the tree's leaves. The
 *     solve equals one given ss_leaf_values' array as its leaf values.
 *   tree_order = 1 (opt-in): tree leaf l takes row L0 * 4^NP + sum_p
 *     s_p 4^p, where L0 is the original leaf its continuation chain
 *     expands (numbered in node order), s_p the continuation strategy
 *     player p picks along the chain (street_solve.cu:400-455: the first
 *     active player's choice outermost) and 0 for players not in the
 *     chain: the profile the leaf actually stands for, whatever the
 *     number of active players. leaf_idx then names original leaves.
 * No counterpart in the reference: the leaf values there go through host
 * memory, in the producer's digit order (leaf_values.py:552, 801). */
int ss_batch_leaf_values(SSBatch *batch, int index, const double *raw_equity, const int *leaf_idx,
                         const int *leaf_pot, int num_leaf_infos, int tree_order);

/* Allocation-free variant of ss_leaf_equity (same arguments, same results
 * bit for bit). No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code between HIP graph captures, tight
 * serving loops, hosts that forbid hipMalloc). Every device buffer of the
 * call (uploaded hand and runout lists, per-(runout, player) strengths,
 * tuple records, per-runout equity) lives in one fixed array of
 * SS_LEAF_STATIC_SCRATCH_BYTES compiled into the library, one copy per GPU,
 * held until the process ends; the call makes no hipMalloc, hipFree,
 * hipHostMalloc or other allocating call and reads the result back with a
 * plain hipMemcpy. A call that would need more scratch (the need grows with
 * runouts x players x max_hands and the 2000-tuple records) fails with
 * SS_ERR_ARG and says how much it needs; a 3-player flop over 144 runouts at
 * 1176 hands needs about 35 MB. All calls on one GPU share the array:
 * they are serialised by a per-device lock here, and any other code that
 * used the same array concurrently would race with them. The array keeps
 * the last call's contents; every element a call reads was written earlier
 * in that call. */
#define SS_LEAF_STATIC_SCRATCH_BYTES (64u << 20)

int ss_leaf_equity_static(int num_players, const int *hands, const int *num_hands, int max_hands,
                          const int *runouts, int num_runouts, double *raw_equity);

#ifdef __cplusplus
}
#endif

#endif
