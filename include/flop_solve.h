/*
 * flop_solve.h -- C ABI of the 2-player multi-street (flop through river)
 * solve with chance nodes.
 *
 * Drop-in for the reference's src/cuda/flop_solve.cuh: the same constants,
 * struct layouts and entry points, so the reference's ctypes bindings
 * (bench/bench_gpu_flop.py:8-28, precompute/run_all.py:37-113) load this
 * library unchanged. Behaviour follows src/cuda/flop_solve.cu; the
 * implementation is this library's own (poker-solver_amd/csrc/ss_flop.hip).
 */
#ifndef SS_FLOP_SOLVE_H
#define SS_FLOP_SOLVE_H

#include <stdint.h>

/* limits, as flop_solve.cuh:32-35 */
#define FS_MAX_HANDS 200      /* hands per player */
#define FS_MAX_ACTIONS 6      /* actions per decision node (row stride of the strategy outputs) */
#define FS_MAX_TREE_NODES 100000
#define FS_MAX_BOARD 5

/* node types, as flop_solve.cuh:38-41 */
#define FS_NODE_DECISION 0
#define FS_NODE_FOLD 1
#define FS_NODE_SHOWDOWN 2
#define FS_NODE_CHANCE 3      /* deals the next board card: one child per card not on the board */

/* one tree node, layout of flop_solve.cuh:45-54 */
typedef struct {
    int type;
    int player;          /* acting player at a decision node, the winner at a fold, else -1 */
    int num_children;    /* actions, or dealt cards at a chance node */
    int first_child;     /* into FSTreeData.children */
    int pot;
    int bets[2];         /* this street's bets */
    int board_cards[5];
    int num_board;
} FSNode;

/* the tree and the ranges, layout of flop_solve.cuh:58-82 */
typedef struct {
    FSNode *nodes;
    int *children;
    int num_nodes;
    int num_children_total;
    int hands[2][FS_MAX_HANDS][2];   /* filled by the caller */
    float weights[2][FS_MAX_HANDS];  /* filled by the caller */
    int num_hands[2];                /* filled by the caller */
    int *level_order;                /* node indices in breadth-first order */
    int *node_depth;
    int max_depth;
    int *decision_node_indices;
    int num_decision_nodes;
    int *showdown_node_indices;
    int num_showdown_nodes;
} FSTreeData;

/* results, layout of flop_solve.cuh:86-113 */
typedef struct {
    float *root_strategy;            /* [root actions][root player's hands]: the current strategy after the last pass */
    int root_num_actions;
    int root_player;
    float *root_ev;                  /* left NULL (the reference fills nothing here) */
    float *all_avg_strategies;       /* extract_all: [num_nodes][FS_MAX_ACTIONS][max_hands] average strategies */
    float *all_cfv;                  /* extract_all: [num_nodes][max_hands] counterfactual values of the last pass */
    int max_hands;
    int *turn_root_indices;          /* extract_all: decision children of the flop-end chance nodes */
    int *turn_root_cards;            /* and the turn card each was dealt */
    int num_turn_roots;
} FSOutput;

#ifdef __cplusplus
extern "C" {
#endif

/* Builds the flop-through-river tree on the host: the flop betting with the
 * given sizes, then every street end expanded into one child per undealt card
 * with a single 0.75-pot size on the turn and the river. Replaces
 * fs_build_tree (flop_solve.cuh:128-135, flop_solve.cu:894-972). Returns 0. */
int fs_build_tree(const int *flop_board, int starting_pot, int effective_stack, const float *bet_sizes,
                  int num_bet_sizes, FSTreeData *tree_out);

/* Linear CFR on the GPU for max_iterations iterations; fills root_strategy,
 * root_num_actions and root_player. Replaces fs_solve_gpu
 * (flop_solve.cuh:146-150, flop_solve.cu:974-1270). Returns 0, or a negative
 * street_solve.h error code (ss_last_error() describes it). */
int fs_solve_gpu(FSTreeData *tree_data, int max_iterations, FSOutput *output);

/* As fs_solve_gpu, and also the average strategies and counterfactual values
 * of every node and the turn roots. Replaces fs_solve_gpu_extract_all
 * (flop_solve.cuh:174-178, flop_solve.cu:1282-1572). */
int fs_solve_gpu_extract_all(FSTreeData *tree_data, int max_iterations, FSOutput *output);

/* Frees what fs_build_tree allocated (flop_solve.cu:1272-1280). */
void fs_free_tree(FSTreeData *tree_data);

/* Frees what the solves allocated (flop_solve.cu:1574-1582). */
void fs_free_output(FSOutput *output);

#ifdef __cplusplus
}
#endif

#endif /* SS_FLOP_SOLVE_H */
