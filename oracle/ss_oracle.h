/*
 * ss_oracle.h -- CPU oracle for the single-street N-player Linear-CFR solver.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: it restates the
 * reference semantics (src/cuda/street_solve.cu, src/hand_eval.h) on the CPU
 * so tests can compare the HIP library against it. Nothing in the product
 * links, loads or calls it.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   or_eval7           exhaustive category counts over all C(52,7) hands
 *                      (textbook values) + golden vectors from the compiled
 *                      reference hand_eval.h
 *   or_build_tree      golden tree fixtures from the compiled reference
 *                      host builder (street_solve.cu:42-460, 1192-1296)
 *   or_solve (brute)   closed forms on 1-vs-1 hand games, invariants, and the
 *                      reference odometer itself (this IS the reference formula)
 *   or_solve (fast)    brute mode on random small games (k <= 2 opponents)
 *   or_exploitability  exhaustive pure-strategy best response on tiny games
 */
#ifndef SS_ORACLE_H
#define SS_ORACLE_H

#include <stdint.h>
#include "../include/street_solve.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_HANDS 1326

typedef struct {
    const SSNode *nodes;
    const int *children;
    int num_nodes;
    int num_players;
    int num_hands[SS_MAX_PLAYERS];
    const int *hands[SS_MAX_PLAYERS];      /* [nh][2] */
    const float *weights[SS_MAX_PLAYERS];  /* [nh] */
    int board[5];
    int num_board;
    const float *leaf_values;              /* [num_leaves][np][max_h] or NULL */
    int num_leaves;
    const int *frozen_action;              /* [num_nodes] or NULL */
    int is_river;
} OrProblem;

typedef struct {
    float *avg;            /* [num_nodes][8][max_h]: host-normalised average strategy */
    float *root_strategy;  /* [8][max_h]: current strategy at node 0 after the last pass */
    float *root_ev;        /* [max_h]: cfv at node 0 of the last traverser */
    double *regrets;       /* optional [num_nodes][8][max_h] */
    double *strat_sum;     /* optional [num_nodes][8][max_h] */
} OrOutput;

/* terminal evaluation modes */
#define OR_TERM_BRUTE 0    /* reference odometer, O(M^k) per hand */
#define OR_TERM_FAST  1    /* card-removal closed forms (k <= 2), O(M*52) per node */

uint32_t or_eval5(int c0, int c1, int c2, int c3, int c4);
uint32_t or_eval7(const int c[7]);
/* counts[cat] over all C(52,7) hands (cat 1..9) and a sum of all ranks mod 2^64 */
void or_eval7_census(uint64_t counts[10], uint64_t *rank_sum);

int or_build_tree(const int *board, int num_board, int starting_pot, int effective_stack,
                  const float *bet_sizes, int num_bet_sizes, int num_players,
                  const int *acting_order, int is_river, int use_cont_strats,
                  int max_raises, SSTreeData *out);
void or_free_tree(SSTreeData *t);

/* Linear CFR exactly as ss_solve_gpu; fp64 != 0 runs every quantity in double. */
int or_solve(const OrProblem *p, int iterations, int fp64, int term_mode, OrOutput *out);

/* FP64 best response against the average strategy avg ([N][8][max_h]).
 * Returns exploitability in % of pot_ref (mean per-player BR gain per unit of
 * joint deal mass); gains[p] receives each player's gain in chips. */
double or_exploitability(const OrProblem *p, const float *avg, int term_mode,
                         double pot_ref, double *gains);

/* Terminal values of one node for one traverser given explicit reach arrays
 * reach[p][h] (fp64); fills cfv[h] for h < nh[trav]. For kernel-level tests. */
int or_terminal_values(const OrProblem *p, int node, int trav, const double *const *reach,
                       int term_mode, double *cfv);

#ifdef __cplusplus
}
#endif
#endif
