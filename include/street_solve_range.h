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

#ifdef __cplusplus
}
#endif

#endif
