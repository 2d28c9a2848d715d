/*
 * street_solve.h -- C-ABI of the MI355X single-street N-player Linear-CFR
 * subgame solver (libstreet_solve.so).
 *
 * Two layers:
 *
 *  1. The legacy entry points ss_build_tree / ss_solve_gpu / ss_free_tree /
 *     ss_free_output, with struct layouts byte-identical to the reference
 *     (src/cuda/street_solve.cuh:52-143, :172-194) so that the ctypes mirror
 *     in python/street_solver_gpu.py:66-124 binds this library unchanged.
 *     Limits are those of the reference ABI: <= 200 hands per player,
 *     <= 3 raises per street.
 *
 *  2. Extended entry points (ss_*_ex, batched and served forms) with no
 *     reference counterpart: up to 1326 hands per player, a configurable raise
 *     cap, an FP64 exploitability report, many subgames per launch and a
 *     continuous-batching request loop.
 *
 * Every entry point returns 0 on success and a negative code on failure;
 * ss_last_error() returns a message for the calling thread's last failure.
 */
#ifndef STREET_SOLVE_H
#define STREET_SOLVE_H

#include <stdint.h>

/* ---- limits of the legacy ABI (street_solve.cuh:35-42) ------------------ */
#define SS_MAX_PLAYERS     6
#define SS_MAX_HANDS       200
#define SS_MAX_ACTIONS     8
#define SS_MAX_TREE_NODES  4096
#define SS_MAX_BOARD       5
#define SS_MAX_LEAVES      64
#define SS_NUM_CONT_STRATS 4
#define SS_MAX_RAISES      3

/* limits of the extended ABI */
#define SS_EX_MAX_HANDS    1326   /* all two-card combos */
#define SS_EX_MAX_RAISES   8

/* node types (street_solve.cuh:45-48) */
#define SS_NODE_DECISION 0
#define SS_NODE_FOLD     1
#define SS_NODE_SHOWDOWN 2
#define SS_NODE_LEAF     4

/* error codes */
#define SS_OK              0
#define SS_ERR_GENERIC    -1   /* the only failure code of the reference */
#define SS_ERR_ARG        -2
#define SS_ERR_HIP        -3
#define SS_ERR_NOMEM      -4
#define SS_ERR_UNSUPPORTED -5
#define SS_ERR_BUSY       -6
#define SS_ERR_NOTFOUND   -7

/* ---- flat tree node: street_solve.cuh:52-65 (104 bytes) ----------------- */
typedef struct {
    int type;
    int player;
    int num_children;
    int first_child;
    int pot;
    int bets[SS_MAX_PLAYERS];
    int board_cards[5];
    int num_board;
    int leaf_idx;
    int fold_player;
    int num_players;
    int active_players[SS_MAX_PLAYERS];
} SSNode;

/* ---- solver input: street_solve.cuh:69-122 (14600 bytes) --------------- */
typedef struct {
    SSNode *nodes;
    int *children;
    int num_nodes;
    int num_children_total;
    int num_players;
    int hands[SS_MAX_PLAYERS][SS_MAX_HANDS][2];
    float weights[SS_MAX_PLAYERS][SS_MAX_HANDS];
    int num_hands[SS_MAX_PLAYERS];
    int board[SS_MAX_BOARD];
    int num_board;
    float *leaf_values;          /* [num_leaves][num_players][max_hands], caller-owned */
    int num_leaves;
    int *level_order;
    int *node_depth;
    int max_depth;
    int *decision_node_indices;
    int num_decision_nodes;
    int *showdown_node_indices;
    int num_showdown_nodes;
    int *leaf_node_indices;
    int num_leaf_nodes;
    int *fold_node_indices;
    int num_fold_nodes;
    int starting_pot;
    int effective_stack;
    int is_river;
    int *frozen_action;          /* [num_nodes], -1 = free; caller-owned */
} SSTreeData;

/* ---- solver output: street_solve.cuh:126-143 (56 bytes) ----------------- */
typedef struct {
    float *root_strategy;        /* [root_num_actions][num_hands[root_player]] */
    int root_num_actions;
    int root_player;
    float *root_ev;              /* [num_players][max_hands]; row 0 filled */
    float *avg_strategies;       /* [num_avg_nodes][SS_MAX_ACTIONS][max_hands] */
    int *avg_strategy_node_ids;
    int num_avg_nodes;
    int num_players;
    int max_hands;
} SSOutput;

/* ---- extended ABI types -------------------------------------------------- */

/* Per-subgame ranges for the extended entry points (any size <= 1326). */
typedef struct {
    const SSTreeData *tree;              /* built by ss_build_tree[_ex]; its
                                            hands/weights/num_hands are ignored */
    int num_hands[SS_MAX_PLAYERS];
    const int *hands[SS_MAX_PLAYERS];    /* [num_hands[p]][2] card ids 0..51 */
    const float *weights[SS_MAX_PLAYERS];/* [num_hands[p]] root reach */
    const float *leaf_values;            /* [num_leaves][num_players][max_hands] or NULL */
    const int *frozen_action;            /* [num_nodes] or NULL */
} SSProblem;

#define SS_PREC_FP32    0   /* plain form: FP32 regrets, FP32 terminals */
#define SS_PREC_BF16X2  1   /* precision form: 2-player showdown as a BF16
                               hi+lo MFMA contraction, FP32 accumulation */

typedef struct {
    int max_iterations;
    int precision;               /* SS_PREC_* */
    double target_exploitability;/* % of starting pot; <= 0 runs max_iterations */
    int check_every;             /* iterations between FP64 exploitability checks */
    int flags;                   /* SS_FLAG_* */
} SSSolveConfig;

#define SS_FLAG_NO_EXPL     1    /* skip the final exploitability evaluation */
#define SS_FLAG_ADAPTIVE_CHECKS 2 /* first check at check_every, then each subgame's next
                                     check where its exploitability would reach the target
                                     if it fell as fast as t^-2 (rounded down to the
                                     iteration step, at least one step and at most
                                     4 * check_every later); without it checks are every
                                     check_every iterations */
#define SS_FLAG_SPARSE_CHECKS 4   /* checks stay on multiples of check_every, but a subgame
                                     skips those before the iteration where its exploitability
                                     would reach the target at t^-2 decay (at most 3 in a row);
                                     with SS_FLAG_ADAPTIVE_CHECKS this one wins */

typedef struct {
    int iterations;              /* iterations actually run */
    int converged;               /* 1 if target_exploitability was reached */
    double exploitability;       /* % of starting pot (FP64 best response); -1 if not evaluated */
    double nashconv[SS_MAX_PLAYERS]; /* per-player BR gain, chips per unit deal mass */
    double elapsed_ms;           /* wall time of the call (served form: arrival to retirement) */
    double kernel_ms;            /* solve loop on the device: CFR iterations and exploitability
                                    checks, setup and output transfers excluded */
    int64_t algo_bytes;          /* algorithmic HBM bytes of the CFR iterations */
} SSReport;

#ifdef __cplusplus
extern "C" {
#endif

/* ---- legacy ABI ----------------------------------------------------------- */

/* Replaces street_solve.cu:1192-1296 (decl. street_solve.cuh:172-179). */
int ss_build_tree(const int *board, int num_board,
                  int starting_pot, int effective_stack,
                  const float *bet_sizes, int num_bet_sizes,
                  int num_players, const int *acting_order,
                  int is_river, int use_cont_strats,
                  SSTreeData *tree_out);

/* Replaces street_solve.cu:1298-1639 (decl. street_solve.cuh:187-191). */
int ss_solve_gpu(SSTreeData *tree_data, int max_iterations, SSOutput *output);

/* Replaces street_solve.cu:1641-1652. Unlike the reference (:1650) it does NOT
 * free tree_data->leaf_values or frozen_action: both are caller-owned. */
void ss_free_tree(SSTreeData *tree_data);

/* Replaces street_solve.cu:1654-1660. */
void ss_free_output(SSOutput *output);

/* ---- extended ABI (no reference counterpart) ------------------------------ */

/* ss_build_tree with a configurable raise cap (reference: SS_MAX_RAISES = 3). */
int ss_build_tree_ex(const int *board, int num_board,
                     int starting_pot, int effective_stack,
                     const float *bet_sizes, int num_bet_sizes,
                     int num_players, const int *acting_order,
                     int is_river, int use_cont_strats, int max_raises,
                     SSTreeData *tree_out);

/* One subgame, plain form (level-synchronous kernels, one subgame per launch). */
int ss_solve_gpu_ex(const SSProblem *problem, const SSSolveConfig *config,
                    SSOutput *output, SSReport *report);

/* Many subgames advanced together by one persistent kernel per iteration, the
 * iteration loop captured in a hipGraph. outputs/reports have n entries. */
int ss_solve_batch_ex(const SSProblem *problems, int n, const SSSolveConfig *config,
                      SSOutput *outputs, SSReport *reports);

/* Batched form as a session: ss_batch_create uploads every subgame and captures
 * the iteration graph; each ss_batch_run resets regrets and strategy sums and
 * solves all subgames again (inputs stay resident in HBM between runs). */
typedef struct SSBatch SSBatch;
typedef struct {
    double run_ms;               /* wall time of the run (device work + host checks) */
    double cfr_ms;               /* HIP-event time of the CFR graph replays */
    double check_ms;             /* HIP-event time of the FP64 exploitability checks */
    int64_t algo_bytes;          /* algorithmic bytes of all CFR iterations run */
    int launches;                /* persistent-kernel launches (one per iteration) */
    int max_iterations;          /* iterations of the slowest subgame */
    int mode;                    /* strategy the autotuner chose: 0 persistent kernel with one
                                    workgroup per subgame, 1 phased launches over all subgames */
    /* launch shape the autotuner chose for the phased form (-1 / 0: the automatic rule):
       forward levels with >= this many tasks run three per workgroup; 0 = terminal classes
       in separate launches; workgroups per two-opponent showdown; sweep workgroup cap */
    int shape_fwd_group_min, shape_term_small, shape_sd_parts, shape_sweep_grid;
    double tune_ms;              /* time ss_batch_create spent tuning the shape (0 when cached) */
} SSBatchStats;

SSBatch *ss_batch_create(const SSProblem *problems, int n, const SSSolveConfig *config);
int ss_batch_run(SSBatch *batch, SSReport *reports, SSBatchStats *stats);
int ss_batch_output(SSBatch *batch, int index, SSOutput *output);
void ss_batch_destroy(SSBatch *batch);

/* Served form: a continuous-batching request loop on one GPU. */
typedef struct SSServer SSServer;
typedef struct {
    int64_t hbm_budget_bytes;    /* admission limit on resident subgame state */
    int max_slots;               /* max resident subgames */
    int iters_per_step;          /* CFR iterations per graph launch */
    int check_every;             /* iterations between exploitability checks */
} SSServerConfig;

SSServer *ss_server_create(const SSServerConfig *config);
/* Copies everything it needs from problem; returns a ticket >= 1 or < 0. */
int64_t ss_server_submit(SSServer *server, const SSProblem *problem,
                         const SSSolveConfig *config);
/* 1 = finished (output/report filled), 0 = pending, < 0 = error. */
int ss_server_poll(SSServer *server, int64_t ticket, SSOutput *output, SSReport *report);
int ss_server_cancel(SSServer *server, int64_t ticket);
/* Blocks until every admitted or queued request finished. */
int ss_server_drain(SSServer *server);
void ss_server_destroy(SSServer *server);

/* Bytes of device state one problem occupies (admission accounting). */
int64_t ss_problem_footprint(const SSProblem *problem);

/* Message for the calling thread's last failure ("" if none). */
const char *ss_last_error(void);

/* Library version string. */
const char *ss_version(void);

#ifdef __cplusplus
}
#endif

#endif /* STREET_SOLVE_H */
