/*
 * fs_oracle.c -- CPU oracle of the 2-player flop-through-river solve with
 * chance nodes (include/flop_solve.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker that the HIP library's fs_solve_gpu /
 * fs_solve_gpu_extract_all are compared against. Nothing in the product links,
 * loads or calls it.
 *
 * Restates src/cuda/flop_solve.cu on the CPU:
 *   - showdown ranks, 0 for a hand that shares a card with the node's board
 *     (fs_precompute_strengths :366-399; the evaluator fs_eval5/fs_eval7
 *     :327-359 is token for token the street solver's ss_eval5/ss_eval7,
 *     restated as or_eval7 in ss_oracle.c);
 *   - one iteration = traverser 0's pass, then traverser 1's
 *     (graph body :1150-1206), each pass:
 *       regret matching at every decision node (fs_batch_regret_match :657-690),
 *       the opponent's reach from its input weights down the tree
 *       (fs_batch_propagate_reach :693-735: the opponent's own nodes scale by
 *       its current strategy, every other node copies),
 *       fold and showdown values (fs_batch_fold_value :737-779,
 *       fs_batch_showdown_value :781-833),
 *       values up the tree with the traverser's regret and strategy-sum
 *       updates and chance averaging (fs_batch_propagate_cfv :835-891);
 *   - the average-strategy normalisation of fs_solve_gpu_extract_all
 *     (:1503-1530).
 * The passes visit the tree depth first instead of level by level: every
 * update touches only its own node's rows, so the order does not change any
 * value. In single precision the products that the device code contracts
 * into fused multiply-adds are written as fmaf here.
 *
 * Pins (tests/test_flop_oracle.py): closed-form equilibrium of a half-street
 * game (bet / bluff / call frequencies), hand-computed values after one
 * iteration, chance averaging and card blocking on hand-built trees.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/flop_solve.h"

uint32_t or_eval7(const int c[7]);

typedef struct {
    float *root_strategy;   /* [root actions][root player's hands]: current strategy after the last pass */
    float *avg;             /* [N][FS_MAX_ACTIONS][max_h]: normalised average strategy (extract_all layout) */
    float *cfv;             /* [N][max_h]: values of traverser 1's (the last) pass */
    double *regrets;        /* optional [N][FS_MAX_ACTIONS][max_h] */
    double *strategy_sum;   /* optional [N][FS_MAX_ACTIONS][max_h] */
} FsOrOutput;

#define FS_NA FS_MAX_ACTIONS

/* ---------------------------------------------------------------------- */
/* state shared by both precisions                                         */

typedef struct {
    const FSTreeData *td;
    int N, max_h, nh[2], starting_pot;
    int *sd_index;          /* [N] rank row of a showdown node (one per distinct board), -1 otherwise */
    uint32_t *rank;         /* [distinct showdown boards][2][max_h] */
    uint64_t *mask[2];      /* [nh[p]] card set of each hand, bit c for card c */
    int *abase, A;          /* first action row of each decision node (-1 elsewhere); rows in all */
} FsCommon;

static int fs_hand_on_board(const FSNode *n, int c0, int c1)
{
    for (int b = 0; b < n->num_board; ++b)
        if (c0 == n->board_cards[b] || c1 == n->board_cards[b]) return 1;
    return 0;
}

static uint64_t fs_board_key(const FSNode *n)
{
    uint64_t k = (uint64_t)n->num_board;
    for (int b = 0; b < n->num_board; ++b) k = k << 6 | (uint64_t)(n->board_cards[b] & 63);
    return k;
}

typedef struct {
    uint64_t key;
    int node;
} FsKeyNode;

static int fs_cmp_key(const void *a, const void *b)
{
    const FsKeyNode *x = (const FsKeyNode *)a, *y = (const FsKeyNode *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->node - y->node;
}

/* Ranks of both players' hands for every showdown board, computed once per
 * distinct board (the showdowns of one river betting tree share it). */
static int common_init(FsCommon *c, const FSTreeData *td)
{
    memset(c, 0, sizeof(*c));
    c->td = td;
    c->N = td->num_nodes;
    c->nh[0] = td->num_hands[0];
    c->nh[1] = td->num_hands[1];
    c->max_h = c->nh[0] > c->nh[1] ? c->nh[0] : c->nh[1];
    c->starting_pot = td->nodes[0].pot;
    c->sd_index = (int *)malloc(sizeof(int) * (size_t)c->N);
    c->abase = (int *)malloc(sizeof(int) * (size_t)c->N);
    if (!c->sd_index || !c->abase) return -1;
    for (int n = 0; n < c->N; ++n) {
        c->abase[n] = td->nodes[n].type == FS_NODE_DECISION ? c->A : -1;
        if (td->nodes[n].type == FS_NODE_DECISION) c->A += td->nodes[n].num_children;
    }
    for (int p = 0; p < 2; ++p) {
        c->mask[p] = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(c->nh[p] ? c->nh[p] : 1));
        if (!c->mask[p]) return -1;
        for (int h = 0; h < c->nh[p]; ++h)
            c->mask[p][h] = 1ull << (td->hands[p][h][0] & 63) | 1ull << (td->hands[p][h][1] & 63);
    }
    FsKeyNode *keys = (FsKeyNode *)malloc(sizeof(FsKeyNode) * (size_t)(c->N ? c->N : 1));
    if (!c->sd_index || !keys) {
        free(keys);
        return -1;
    }
    int nsd = 0;
    for (int n = 0; n < c->N; ++n) {
        c->sd_index[n] = -1;
        if (td->nodes[n].type == FS_NODE_SHOWDOWN) {
            keys[nsd].key = fs_board_key(&td->nodes[n]);
            keys[nsd++].node = n;
        }
    }
    qsort(keys, (size_t)nsd, sizeof(FsKeyNode), fs_cmp_key);
    int nboards = 0;
    for (int i = 0; i < nsd; ++i) {
        if (i == 0 || keys[i].key != keys[i - 1].key) ++nboards;
        c->sd_index[keys[i].node] = nboards - 1;
    }
    c->rank = (uint32_t *)calloc((size_t)(nboards ? nboards : 1) * 2 * c->max_h, sizeof(uint32_t));
    if (!c->rank) {
        free(keys);
        return -1;
    }
    for (int i = 0; i < nsd; ++i) {
        if (i > 0 && keys[i].key == keys[i - 1].key) continue;
        const FSNode *nd = &td->nodes[keys[i].node];
        uint32_t *row = c->rank + (size_t)c->sd_index[keys[i].node] * 2 * c->max_h;
        for (int p = 0; p < 2; ++p)
            for (int h = 0; h < c->nh[p]; ++h) {
                const int c0 = td->hands[p][h][0], c1 = td->hands[p][h][1];
                uint32_t r = 0;
                if (!fs_hand_on_board(nd, c0, c1)) {
                    int cards[7] = {nd->board_cards[0], nd->board_cards[1], nd->board_cards[2], nd->board_cards[3],
                                    nd->board_cards[4], c0, c1};
                    r = or_eval7(cards);
                }
                row[(size_t)p * c->max_h + h] = r;
            }
    }
    free(keys);
    return 0;
}

static void common_free(FsCommon *c)
{
    free(c->sd_index);
    free(c->rank);
    free(c->mask[0]);
    free(c->mask[1]);
    free(c->abase);
}

/* ---------------------------------------------------------------------- */
/* the solve, once in float (as the device) and once in double             */

/* fmaf is correctly rounded either way; the second single-precision copy
 * lets the compiler emit the x86 FMA instruction instead of a library call
 * (picked at run time when the CPU has it; the build disables implicit
 * contraction, so only the written fmaf calls fuse). */
#define FS_REAL float
#define FS_SFX _f32
#define FS_FMA(a, b, c) fmaf((a), (b), (c))
#define FS_ATTR
#include "fs_oracle_impl.inc"
#undef FS_SFX
#undef FS_ATTR

#if defined(__x86_64__)
#define FS_SFX _f32fma
#define FS_ATTR __attribute__((target("fma")))
#include "fs_oracle_impl.inc"
#undef FS_SFX
#undef FS_ATTR
#endif
#undef FS_REAL
#undef FS_FMA

#define FS_REAL double
#define FS_SFX _f64
#define FS_FMA(a, b, c) ((a) * (b) + (c))
#define FS_ATTR
#include "fs_oracle_impl.inc"
#undef FS_REAL
#undef FS_SFX
#undef FS_FMA
#undef FS_ATTR

/*
 * Linear CFR for `iterations` iterations on td (hands and weights filled).
 * fp64 != 0 runs every quantity in double. Output arrays are optional (NULL
 * skips them) and sized as FsOrOutput states with max_h = max(num_hands).
 * Returns 0, or -1 on an allocation failure.
 */
int fs_or_solve(const FSTreeData *td, int iterations, int fp64, FsOrOutput *out)
{
    FsCommon c;
    if (common_init(&c, td)) {
        common_free(&c);
        return -1;
    }
    int rc;
    if (fp64)
        rc = fs_solve_f64(&c, iterations, out);
#if defined(__x86_64__)
    else if (__builtin_cpu_supports("fma"))
        rc = fs_solve_f32fma(&c, iterations, out);
#endif
    else
        rc = fs_solve_f32(&c, iterations, out);
    common_free(&c);
    return rc;
}

/* Showdown rank of both players' hands at node n (0 when blocked by its board),
 * as the solve uses them: out[2][max_h]. */
int fs_or_ranks(const FSTreeData *td, int n, uint32_t *out)
{
    FsCommon c;
    if (common_init(&c, td)) {
        common_free(&c);
        return -1;
    }
    if (n < 0 || n >= c.N || c.sd_index[n] < 0) {
        common_free(&c);
        return -1;
    }
    memcpy(out, c.rank + (size_t)c.sd_index[n] * 2 * c.max_h, sizeof(uint32_t) * 2 * c.max_h);
    common_free(&c);
    return 0;
}
