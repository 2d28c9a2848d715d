/*
 * street_solve_range.h -- Bayesian range narrowing on the GPU.
 *
 * Replaces the per-hand loop of RangeNarrower.update
 * (python/range_narrowing.py:76-122): after an observed action, every hand's
 * weight is multiplied by P(action | hand) from the solver's average strategy
 * (street_solver_gpu.py:550-553), floored, and rescaled so the largest weight
 * is 1. poker-solver_amd/range_narrowing.py keeps the class around it.
 *
 * Return 0 on success, a negative SS_ERR_* code otherwise (message in
 * ss_last_error()). Synchronous, host buffers.
 */
#ifndef STREET_SOLVE_RANGE_H
#define STREET_SOLVE_RANGE_H

#include "street_solve.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SS_RANGE_MAX_HANDS 1326
#define SS_RANGE_WEIGHT_FLOOR 0.005 /* range_narrowing.py:46 WEIGHT_FLOOR */

/* weights[h] (in/out, float64): w' = w * p, with p = p_action[h] when
 * has_p == NULL or has_p[h] != 0, else p_missing; w' = weight_floor when
 * w' < weight_floor and w > 0; then every w' is divided by max w' when that
 * maximum is > 0 (else left as is). num_hands 0..1326. */
int ss_range_update(int num_hands, double *weights, const double *p_action, const unsigned char *has_p,
                    double p_missing, double weight_floor);

/* The same update chained to a solve held on the device: P(action | hand)
 * is read from the strategy sums of subgame `index` of a batch session
 * (ss_batch_create / ss_batch_run, include/street_solve.h) at decision node
 * `node`, normalised as the solver's output is (street_solve.cu:1597-1610),
 * cut at p_threshold (0.001 reproduces StreetSolverGPU.get_avg_strategy,
 * python/street_solver_gpu.py:484) and 0 for action = -1 (a label the node
 * lacks), without reading the strategy back to the host. solver_hand[h]
 * maps hand h of the narrowed range to the acting player's hand index in
 * the solve, or -1 for a hand the solve does not hold; those take
 * 1 / max(number of distinct probabilities over the solve's hands, 2), as
 * range_narrowing.py:83-84. The result equals ss_range_update on the
 * probabilities the host path builds (RangeNarrower.update with
 * get_avg_strategy_probs, street_solver_gpu.py:550-553) bit for bit.
 * Requires a batch that has been run and a solve whose hand list for the
 * acting player has no repeated combo. Synchronous; weights is a host buffer. */
int ss_batch_range_update(SSBatch *batch, int index, int node, int action, int num_hands, const int *solver_hand,
                          double *weights, double weight_floor, double p_threshold);

#ifdef __cplusplus
}
#endif

#endif
