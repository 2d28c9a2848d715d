/*
 * ss_oracle.c -- CPU oracle (test infrastructure only; see ss_oracle.h).
 *
 * Restates, on the CPU and in plain C:
 *   - the 5/7-card evaluator of src/hand_eval.h:40-156 (= the device copy
 *     ss_eval5/ss_eval7 at src/cuda/street_solve.cu:467-510);
 *   - the single-street tree builder of src/cuda/street_solve.cu:42-460 and
 *     its BFS / classification pass :1192-1296;
 *   - the Linear-CFR loop of ss_solve_gpu, street_solve.cu:1298-1612, with
 *     every kernel :575-1185 (cfr_impl.inc);
 *   - an FP64 best response / exploitability, which the reference lacks on
 *     this path (definition in cfr_impl.inc).
 */
#include "ss_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================
 * Evaluator -- hand_eval.h:40-156. Encoding cat<<20 | rank nibbles.
 * ====================================================================== */

uint32_t or_eval5(int c0, int c1, int c2, int c3, int c4)
{
    int cards[5] = {c0, c1, c2, c3, c4};
    int rk[5], cnt[13];
    memset(cnt, 0, sizeof(cnt));
    int flush = 1;
    for (int i = 0; i < 5; i++) {
        rk[i] = cards[i] >> 2;
        cnt[rk[i]]++;
        if ((cards[i] & 3) != (cards[0] & 3)) flush = 0;
    }
    /* descending ranks */
    for (int i = 1; i < 5; i++)
        for (int j = i; j > 0 && rk[j - 1] < rk[j]; j--) {
            int t = rk[j]; rk[j] = rk[j - 1]; rk[j - 1] = t;
        }
    int distinct = rk[0] != rk[1] && rk[1] != rk[2] && rk[2] != rk[3] && rk[3] != rk[4];
    int straight = 0, high = rk[0];
    if (distinct && rk[0] - rk[4] == 4) straight = 1;
    if (rk[0] == 12 && rk[1] == 3 && rk[2] == 2 && rk[3] == 1 && rk[4] == 0) {
        straight = 1; high = 3;                   /* wheel: five-high */
    }
    uint32_t five = ((uint32_t)rk[0] << 16) | (rk[1] << 12) | (rk[2] << 8) | (rk[3] << 4) | rk[4];
    if (straight && flush) return (9u << 20) | ((uint32_t)high << 16);
    if (flush) return (6u << 20) | five;
    if (straight) return (5u << 20) | ((uint32_t)high << 16);

    int quad = -1, trip = -1, pa = -1, pb = -1;
    for (int r = 12; r >= 0; r--) {
        if (cnt[r] == 4) quad = r;
        else if (cnt[r] == 3) trip = r;
        else if (cnt[r] == 2) { if (pa < 0) pa = r; else pb = r; }
    }
    /* remaining singles, high to low, excluding the given ranks */
    int kick[5], nk = 0;
    for (int r = 12; r >= 0; r--)
        if (cnt[r] > 0 && r != quad && r != trip && r != pa && r != pb) kick[nk++] = r;
    if (quad >= 0) return (8u << 20) | ((uint32_t)quad << 16) | ((uint32_t)kick[0] << 12);
    if (trip >= 0 && pa >= 0) return (7u << 20) | ((uint32_t)trip << 16) | ((uint32_t)pa << 12);
    if (trip >= 0) return (4u << 20) | ((uint32_t)trip << 16) | (kick[0] << 12) | (kick[1] << 8);
    if (pa >= 0 && pb >= 0) return (3u << 20) | ((uint32_t)pa << 16) | (pb << 12) | (kick[0] << 8);
    if (pa >= 0) return (2u << 20) | ((uint32_t)pa << 16) | (kick[0] << 12) | (kick[1] << 8) | (kick[2] << 4);
    return (1u << 20) | five;
}

uint32_t or_eval7(const int c[7])
{
    /* all 21 five-card subsets: skip the pair (i, j) */
    uint32_t best = 0;
    for (int i = 0; i < 7; i++)
        for (int j = i + 1; j < 7; j++) {
            int s[5], k = 0;
            for (int t = 0; t < 7; t++)
                if (t != i && t != j) s[k++] = c[t];
            uint32_t v = or_eval5(s[0], s[1], s[2], s[3], s[4]);
            if (v > best) best = v;
        }
    return best;
}

void or_eval7_census(uint64_t counts[10], uint64_t *rank_sum)
{
    uint64_t cnt[10] = {0}, sum = 0;
#pragma omp parallel for schedule(dynamic) reduction(+:sum)
    for (int a = 0; a < 46; a++) {
        uint64_t lc[10] = {0};
        int c[7];
        c[0] = a;
        for (c[1] = a + 1; c[1] < 47; c[1]++)
        for (c[2] = c[1] + 1; c[2] < 48; c[2]++)
        for (c[3] = c[2] + 1; c[3] < 49; c[3]++)
        for (c[4] = c[3] + 1; c[4] < 50; c[4]++)
        for (c[5] = c[4] + 1; c[5] < 51; c[5]++)
        for (c[6] = c[5] + 1; c[6] < 52; c[6]++) {
            uint32_t v = or_eval7(c);
            lc[v >> 20]++;
            sum += v;
        }
#pragma omp critical
        for (int k = 0; k < 10; k++) cnt[k] += lc[k];
    }
    for (int k = 0; k < 10; k++) counts[k] = cnt[k];
    *rank_sum = sum;
}

/* ======================================================================
 * Tree builder -- street_solve.cu:42-460 (rules), :1192-1296 (layout).
 * Node order = allocation order of a depth-first build; a decision node's
 * child-id block is appended after all its subtrees are built.
 * ====================================================================== */

typedef struct {
    SSNode *nodes; int nn, ncap;
    int *ch; int nch, chcap;
    int leaves;
} OrTB;

typedef struct {
    int is_river, np, max_raises;
    const int *order;
    const float *sizes; int nsizes;
    const int *board; int nb;
} OrBuildCtx;

static int tb_node(OrTB *tb)
{
    if (tb->nn == tb->ncap) {
        tb->ncap = tb->ncap ? 2 * tb->ncap : 512;
        tb->nodes = (SSNode *)realloc(tb->nodes, (size_t)tb->ncap * sizeof(SSNode));
    }
    SSNode *n = &tb->nodes[tb->nn];
    memset(n, 0, sizeof(*n));
    n->player = -1; n->leaf_idx = -1; n->fold_player = -1;
    return tb->nn++;
}

static int tb_children(OrTB *tb, int k)
{
    while (tb->nch + k > tb->chcap) {
        tb->chcap = tb->chcap ? 2 * tb->chcap : 2048;
        tb->ch = (int *)realloc(tb->ch, (size_t)tb->chcap * sizeof(int));
    }
    int s = tb->nch;
    tb->nch += k;
    return s;
}

static void fill_state(SSNode *n, const OrBuildCtx *cx, int pot, const int *bets, const int *active)
{
    n->pot = pot;
    for (int i = 0; i < cx->np; i++) n->bets[i] = bets[i];
    for (int i = 0; i < cx->nb && i < 5; i++) n->board_cards[i] = cx->board[i];
    n->num_board = cx->nb;
    n->num_players = 0;
    for (int i = 0; i < cx->np; i++) {
        n->active_players[i] = active[i];
        if (active[i]) n->num_players++;
    }
}

static int next_in_order(const OrBuildCtx *cx, const int *active, int cur)
{
    for (int i = 1; i <= cx->np; i++) {
        int idx = (cur + i) % cx->np;
        int p = cx->order[idx];
        if (p < cx->np && active[p]) return idx;
    }
    return -1;
}

static int build_rec(OrTB *tb, const OrBuildCtx *cx, int oi, int pot, int stack,
                     const int *bets, const int *acted, const int *active, int raises)
{
    int np = cx->np, nact = 0, mx = 0;
    for (int i = 0; i < np; i++)
        if (active[i]) { nact++; if (bets[i] > mx) mx = bets[i]; }

    if (nact <= 1) {                                  /* :155-166 */
        int id = tb_node(tb);
        SSNode *n = &tb->nodes[id];
        n->type = SS_NODE_FOLD;
        for (int i = 0; i < np; i++) if (active[i]) { n->player = i; break; }
        fill_state(n, cx, pot, bets, active);
        return id;
    }
    int complete = 1;                                 /* :126-135 */
    for (int i = 0; i < np; i++)
        if (active[i] && (!acted[i] || bets[i] != mx)) { complete = 0; break; }
    if (complete) {                                   /* :169-182 */
        int id = tb_node(tb);
        SSNode *n = &tb->nodes[id];
        if (cx->is_river) n->type = SS_NODE_SHOWDOWN;
        else { n->type = SS_NODE_LEAF; n->leaf_idx = tb->leaves++; }
        fill_state(n, cx, pot, bets, active);
        return id;
    }
    int actor = cx->order[oi];
    if (!active[actor]) {                             /* :188-207 */
        int nx = next_in_order(cx, active, oi);
        if (nx < 0) {
            int id = tb_node(tb);
            tb->nodes[id].type = SS_NODE_SHOWDOWN;
            fill_state(&tb->nodes[id], cx, pot, bets, active);
            return id;
        }
        return build_rec(tb, cx, nx, pot, stack, bets, acted, active, raises);
    }

    int id = tb_node(tb);
    tb->nodes[id].type = SS_NODE_DECISION;
    tb->nodes[id].player = actor;
    fill_state(&tb->nodes[id], cx, pot, bets, active);

    int to_call = mx - bets[actor];
    if (to_call < 0) to_call = 0;
    int nx = next_in_order(cx, active, oi);
    if (nx < 0) nx = oi;

    int kids[16], nk = 0;
    int nb2[SS_MAX_PLAYERS], na2[SS_MAX_PLAYERS], nact2[SS_MAX_PLAYERS];

    if (to_call > 0) {                                /* fold, :229-241 */
        memcpy(nact2, active, sizeof(int) * np);
        nact2[actor] = 0;
        kids[nk++] = build_rec(tb, cx, nx, pot, stack, bets, acted, nact2, raises);
    }
    {                                                 /* check / call, :244-260 */
        memcpy(nb2, bets, sizeof(int) * np);
        memcpy(na2, acted, sizeof(int) * np);
        nb2[actor] = mx;
        na2[actor] = 1;
        kids[nk++] = build_rec(tb, cx, nx, pot + to_call, stack, nb2, na2, active, raises);
    }
    if (raises < cx->max_raises && nk < 14) {         /* bet / raise, :263-322 */
        int allin = 0;
        for (int i = 0; i < cx->nsizes && nk < 14; i++) {
            int amt;
            /* float arithmetic with C truncation, as :267-270 */
            if (to_call == 0) amt = (int)(cx->sizes[i] * (float)pot);
            else amt = to_call + (int)(cx->sizes[i] * (float)(pot + to_call));
            if (amt >= stack) amt = stack;
            if (amt <= to_call) continue;
            if (amt >= stack) {
                if (allin) continue;
                allin = 1;
            }
            memcpy(nb2, bets, sizeof(int) * np);
            memcpy(na2, acted, sizeof(int) * np);
            nb2[actor] = mx + amt;                    /* over-counts by to_call, :284 */
            na2[actor] = 1;
            for (int q = 0; q < np; q++)
                if (q != actor && active[q]) na2[q] = 0;
            kids[nk++] = build_rec(tb, cx, nx, pot + amt, stack, nb2, na2, active, raises + 1);
        }
        if (!allin && stack > to_call && nk < 14) {
            memcpy(nb2, bets, sizeof(int) * np);
            memcpy(na2, acted, sizeof(int) * np);
            nb2[actor] = mx + stack;
            na2[actor] = 1;
            for (int q = 0; q < np; q++)
                if (q != actor && active[q]) na2[q] = 0;
            kids[nk++] = build_rec(tb, cx, nx, pot + stack, stack, nb2, na2, active, raises + 1);
        }
    }
    int base = tb_children(tb, nk);
    for (int i = 0; i < nk; i++) tb->ch[base + i] = kids[i];
    tb->nodes[id].first_child = base;
    tb->nodes[id].num_children = nk;
    return id;
}

/* street_solve.cu:336-460: each pre-expansion leaf becomes a chain of
 * decisions, one level per active player (lowest index first), 4 children
 * each; leaves are renumbered from 0 in creation order. */
static void expand_leaves(OrTB *tb, int np)
{
    int orig = tb->nn, nl = 0;
    int *lid = (int *)malloc(sizeof(int) * (orig ? orig : 1));
    for (int i = 0; i < orig; i++)
        if (tb->nodes[i].type == SS_NODE_LEAF) lid[nl++] = i;
    tb->leaves = 0;
    for (int k = 0; k < nl; k++) {
        SSNode tmpl = tb->nodes[lid[k]];            /* state copied to the chain */
        int act[SS_MAX_PLAYERS], na = 0;
        for (int p = 0; p < np; p++) if (tmpl.active_players[p]) act[na++] = p;
        if (na == 0) continue;
        tb->nodes[lid[k]].type = SS_NODE_DECISION;
        tb->nodes[lid[k]].player = act[0];
        tb->nodes[lid[k]].leaf_idx = -1;
        int *cur = (int *)malloc(sizeof(int));
        cur[0] = lid[k];
        int ncur = 1;
        for (int lev = 0; lev < na; lev++) {
            int last = (lev == na - 1);
            int *nxt = (int *)malloc(sizeof(int) * ncur * SS_NUM_CONT_STRATS);
            int nn = 0;
            for (int pi = 0; pi < ncur; pi++) {
                int par = cur[pi];
                if (lev > 0) {
                    tb->nodes[par].type = SS_NODE_DECISION;
                    tb->nodes[par].player = act[lev];
                }
                int kid[SS_NUM_CONT_STRATS];
                for (int s = 0; s < SS_NUM_CONT_STRATS; s++) {
                    int c = tb_node(tb);
                    SSNode *cn = &tb->nodes[c];
                    if (last) { cn->type = SS_NODE_LEAF; cn->leaf_idx = tb->leaves++; }
                    cn->pot = tmpl.pot;
                    memcpy(cn->bets, tmpl.bets, sizeof(cn->bets));
                    for (int b = 0; b < tmpl.num_board && b < 5; b++) cn->board_cards[b] = tmpl.board_cards[b];
                    cn->num_board = tmpl.num_board;
                    cn->num_players = 0;
                    for (int p = 0; p < np; p++) {
                        cn->active_players[p] = tmpl.active_players[p];
                        if (tmpl.active_players[p]) cn->num_players++;
                    }
                    kid[s] = c;
                    if (!last) nxt[nn++] = c;
                }
                int base = tb_children(tb, SS_NUM_CONT_STRATS);
                for (int s = 0; s < SS_NUM_CONT_STRATS; s++) tb->ch[base + s] = kid[s];
                tb->nodes[par].first_child = base;
                tb->nodes[par].num_children = SS_NUM_CONT_STRATS;
            }
            free(cur);
            cur = nxt;
            ncur = nn;
        }
        free(cur);
    }
    free(lid);
}

int or_build_tree(const int *board, int num_board, int starting_pot, int effective_stack,
                  const float *bet_sizes, int num_bet_sizes, int num_players,
                  const int *acting_order, int is_river, int use_cont_strats,
                  int max_raises, SSTreeData *out)
{
    memset(out, 0, sizeof(*out));
    if (num_players < 1 || num_players > SS_MAX_PLAYERS) return -1;
    OrTB tb;
    memset(&tb, 0, sizeof(tb));
    OrBuildCtx cx = {is_river, num_players, max_raises, acting_order, bet_sizes,
                     num_bet_sizes, board, num_board};
    int bets[SS_MAX_PLAYERS] = {0}, acted[SS_MAX_PLAYERS] = {0}, active[SS_MAX_PLAYERS] = {0};
    for (int i = 0; i < num_players; i++) active[i] = 1;
    build_rec(&tb, &cx, 0, starting_pot, effective_stack, bets, acted, active, 0);
    if (use_cont_strats && !is_river && tb.leaves > 0) expand_leaves(&tb, num_players);

    int N = tb.nn;
    out->nodes = tb.nodes;
    out->children = tb.ch;
    out->num_nodes = N;
    out->num_children_total = tb.nch;
    out->num_leaves = tb.leaves;
    out->is_river = is_river;
    out->starting_pot = starting_pot;
    out->effective_stack = effective_stack;
    out->num_players = num_players;
    for (int i = 0; i < num_board && i < SS_MAX_BOARD; i++) out->board[i] = board[i];
    out->num_board = num_board;

    out->level_order = (int *)malloc(sizeof(int) * N);
    out->node_depth = (int *)calloc(N, sizeof(int));
    int head = 0, tail = 0;
    out->level_order[tail++] = 0;                    /* BFS, :1238-1259 */
    while (head < tail) {
        int n = out->level_order[head++];
        int d = out->node_depth[n];
        if (d > out->max_depth) out->max_depth = d;
        for (int a = 0; a < tb.nodes[n].num_children; a++) {
            int c = tb.ch[tb.nodes[n].first_child + a];
            out->node_depth[c] = d + 1;
            out->level_order[tail++] = c;
        }
    }
    int cnt[5] = {0};
    for (int i = 0; i < N; i++) cnt[tb.nodes[i].type]++;
    out->num_decision_nodes = cnt[SS_NODE_DECISION];
    out->num_fold_nodes = cnt[SS_NODE_FOLD];
    out->num_showdown_nodes = cnt[SS_NODE_SHOWDOWN];
    out->num_leaf_nodes = cnt[SS_NODE_LEAF];
    out->decision_node_indices = (int *)malloc(sizeof(int) * (cnt[0] + 1));
    out->fold_node_indices = (int *)malloc(sizeof(int) * (cnt[1] + 1));
    out->showdown_node_indices = (int *)malloc(sizeof(int) * (cnt[2] + 1));
    out->leaf_node_indices = (int *)malloc(sizeof(int) * (cnt[4] + 1));
    int di = 0, fi = 0, si = 0, li = 0;
    for (int i = 0; i < N; i++) {
        switch (tb.nodes[i].type) {
        case SS_NODE_DECISION: out->decision_node_indices[di++] = i; break;
        case SS_NODE_FOLD: out->fold_node_indices[fi++] = i; break;
        case SS_NODE_SHOWDOWN: out->showdown_node_indices[si++] = i; break;
        case SS_NODE_LEAF: out->leaf_node_indices[li++] = i; break;
        }
    }
    return 0;
}

void or_free_tree(SSTreeData *t)
{
    free(t->nodes); free(t->children); free(t->level_order); free(t->node_depth);
    free(t->decision_node_indices); free(t->fold_node_indices);
    free(t->showdown_node_indices); free(t->leaf_node_indices);
    memset(t, 0, sizeof(*t));
}

/* ======================================================================
 * Shared problem-level tables for the CFR restatement.
 * ====================================================================== */

typedef struct {
    const OrProblem *p;
    int N, NP, maxh;
    int *bfs;            /* BFS order of all nodes */
    int nbfs;
    uint8_t board_mask[52];
    uint32_t *str;       /* [NP][maxh] hand strength at showdown (0 = board-blocked) */
    uint32_t combo_str[52][52];
    /* strength-sorted combo lists for the closed forms */
    int ng;              /* number of live combos */
    int (*glist)[2];     /* [ng] cards, ascending strength */
    uint32_t *gstr;      /* [ng] */
    int clen[52];
    int clist[52][52];   /* per card: partner cards ascending strength */
    uint32_t cstr[52][52];/* per card: strengths of clist entries */
} OrTables;

static int card_blocked(const OrTables *t, int c) { return c < 0 || c > 51 || t->board_mask[c]; }

static void tables_init(OrTables *t, const OrProblem *p)
{
    memset(t, 0, sizeof(*t));
    t->p = p;
    t->N = p->num_nodes;
    t->NP = p->num_players;
    for (int q = 0; q < t->NP; q++) if (p->num_hands[q] > t->maxh) t->maxh = p->num_hands[q];
    for (int b = 0; b < p->num_board; b++) t->board_mask[p->board[b]] = 1;
    t->bfs = (int *)malloc(sizeof(int) * t->N);
    t->bfs[0] = 0;
    int head = 0, tail = 1;
    while (head < tail) {
        int n = t->bfs[head++];
        const SSNode *nd = &p->nodes[n];
        for (int a = 0; a < nd->num_children; a++) t->bfs[tail++] = p->children[nd->first_child + a];
    }
    t->nbfs = tail;
    t->str = (uint32_t *)calloc((size_t)t->NP * (t->maxh ? t->maxh : 1), sizeof(uint32_t));
    if (p->is_river && p->num_board == 5) {
        for (int a = 0; a < 52; a++)
            for (int b = 0; b < 52; b++) {
                if (a == b || t->board_mask[a] || t->board_mask[b]) continue;
                int c7[7] = {p->board[0], p->board[1], p->board[2], p->board[3], p->board[4], a, b};
                t->combo_str[a][b] = or_eval7(c7);
            }
        for (int q = 0; q < t->NP; q++)
            for (int h = 0; h < p->num_hands[q]; h++) {
                int a = p->hands[q][2 * h], b = p->hands[q][2 * h + 1];
                t->str[q * t->maxh + h] = (card_blocked(t, a) || card_blocked(t, b) || a == b)
                                              ? 0 : t->combo_str[a][b];
            }
        /* global list */
        t->glist = (int (*)[2])malloc(sizeof(int) * 2 * 1326);
        t->gstr = (uint32_t *)malloc(sizeof(uint32_t) * 1326);
        for (int a = 0; a < 52; a++)
            for (int b = a + 1; b < 52; b++)
                if (t->combo_str[a][b]) {
                    int i = t->ng++;
                    t->glist[i][0] = a; t->glist[i][1] = b; t->gstr[i] = t->combo_str[a][b];
                }
        for (int i = 1; i < t->ng; i++)          /* insertion sort by strength */
            for (int j = i; j > 0 && t->gstr[j - 1] > t->gstr[j]; j--) {
                uint32_t s = t->gstr[j]; t->gstr[j] = t->gstr[j - 1]; t->gstr[j - 1] = s;
                int c0 = t->glist[j][0], c1 = t->glist[j][1];
                t->glist[j][0] = t->glist[j - 1][0]; t->glist[j][1] = t->glist[j - 1][1];
                t->glist[j - 1][0] = c0; t->glist[j - 1][1] = c1;
            }
        for (int c = 0; c < 52; c++) {
            int n = 0;
            for (int d = 0; d < 52; d++) if (t->combo_str[c][d]) t->clist[c][n++] = d;
            for (int i = 1; i < n; i++)
                for (int j = i; j > 0 && t->combo_str[c][t->clist[c][j - 1]] > t->combo_str[c][t->clist[c][j]]; j--) {
                    int s = t->clist[c][j]; t->clist[c][j] = t->clist[c][j - 1]; t->clist[c][j - 1] = s;
                }
            t->clen[c] = n;
            for (int i = 0; i < n; i++) t->cstr[c][i] = t->combo_str[c][t->clist[c][i]];
        }
    }
}

static void tables_free(OrTables *t)
{
    free(t->bfs); free(t->str); free(t->glist); free(t->gstr);
}

/* count of entries with strength < s (lt) or <= s (le) in a sorted array */
static int lower_count(const uint32_t *v, int n, uint32_t s, int le)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (le ? v[mid] <= s : v[mid] < s) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* ---- CFR in float (reference-faithful rounding) and double ---- */
#define REAL float
#define SFX(x) x##_f32
#include "cfr_impl.inc"
#undef REAL
#undef SFX

#define REAL double
#define SFX(x) x##_f64
#define OR_FP64_INSTANCE 1
#include "cfr_impl.inc"
#undef REAL
#undef SFX

int or_solve(const OrProblem *p, int iterations, int fp64, int term_mode, OrOutput *out)
{
    if (!p || p->num_players < 1 || p->num_players > SS_MAX_PLAYERS || p->num_nodes < 1) return -1;
    return fp64 ? solve_f64(p, iterations, term_mode, out) : solve_f32(p, iterations, term_mode, out);
}
