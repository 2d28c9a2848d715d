/* oracle_driver.c -- drives the CPU oracle (oracle/ss_oracle.c) under
 * AddressSanitizer and UndefinedBehaviorSanitizer (tests/sanitize/Makefile).
 *
 * Test infrastructure. Builds trees for 2-5 players (river, turn with
 * continuation leaves), solves small random games in FP32 and FP64 with both
 * terminal modes, evaluates exploitability and single terminal nodes, and
 * checks that strategies are distributions. Prints "SANITIZE_OK <checks>"
 * and exits 0 when every check held. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ss_oracle.h"

static int g_checks = 0;
#define CHECK(c)                                                                      \
    do {                                                                              \
        ++g_checks;                                                                   \
        if (!(c)) {                                                                   \
            fprintf(stderr, "check failed: %s (%s:%d)\n", #c, __FILE__, __LINE__);    \
            exit(3);                                                                  \
        }                                                                             \
    } while (0)

static unsigned g_seed = 20261204u;
static unsigned next_u(void) {
    g_seed = g_seed * 1664525u + 1013904223u;
    return g_seed >> 8;
}

static void game(int np, int nb, int hp, int cont) {
    int board[5] = {3, 17, 29, 40, 51};
    float sizes[2] = {0.5f, 1.0f};
    int order[6] = {0, 1, 2, 3, 4, 5};
    SSTreeData t;
    CHECK(or_build_tree(board, nb, 1000, 3000, sizes, 2, np, order, nb == 5, cont, 2, &t) == 0);
    CHECK(t.num_nodes > 1);
    int hands[6][40][2];
    float w[6][40];
    int used[52] = {0};
    for (int k = 0; k < nb; ++k) used[board[k]] = 1;
    OrProblem p;
    memset(&p, 0, sizeof(p));
    p.nodes = t.nodes;
    p.children = t.children;
    p.num_nodes = t.num_nodes;
    p.num_players = np;
    for (int q = 0; q < np; ++q) {
        int n = 0;
        while (n < hp) {
            int a = (int)(next_u() % 52), b = (int)(next_u() % 52);
            if (a == b || used[a] || used[b]) continue;
            int dup = 0;
            for (int i = 0; i < n; ++i)
                dup |= (hands[q][i][0] == a && hands[q][i][1] == b) || (hands[q][i][0] == b && hands[q][i][1] == a);
            if (dup) continue;
            hands[q][n][0] = a < b ? a : b;
            hands[q][n][1] = a < b ? b : a;
            w[q][n] = (float)(next_u() % 1000) / 1000.0f + 0.01f;
            ++n;
        }
        p.num_hands[q] = hp;
        p.hands[q] = &hands[q][0][0];
        p.weights[q] = w[q];
    }
    for (int k = 0; k < nb; ++k) p.board[k] = board[k];
    p.num_board = nb;
    p.is_river = nb == 5;
    float *leaf = NULL;
    if (t.num_leaves > 0) {
        size_t n = (size_t)t.num_leaves * np * hp;
        leaf = (float *)malloc(sizeof(float) * n);
        for (size_t i = 0; i < n; ++i) leaf[i] = (float)((int)(next_u() % 200) - 100);
        p.leaf_values = leaf;
        p.num_leaves = t.num_leaves;
    }
    size_t cells = (size_t)t.num_nodes * 8 * hp;
    float *avg = (float *)calloc(cells, sizeof(float));
    float *rs = (float *)calloc(8 * hp, sizeof(float));
    float *ev = (float *)calloc(hp, sizeof(float));
    double *rg = (double *)calloc(cells, sizeof(double));
    double *ss = (double *)calloc(cells, sizeof(double));
    for (int fp64 = 0; fp64 < 2; ++fp64)
        for (int mode = 0; mode < 2; ++mode) {
            if (mode == OR_TERM_FAST && np > 3) continue;   /* closed forms cover <= 2 opponents */
            OrOutput o = {avg, rs, ev, rg, ss};
            CHECK(or_solve(&p, 4, fp64, mode, &o) == 0);
            for (int n = 0; n < t.num_nodes; ++n) {
                if (t.nodes[n].type != 0) continue;
                int q = t.nodes[n].player;
                for (int h = 0; h < hp; ++h) {
                    double s = 0;
                    for (int a = 0; a < t.nodes[n].num_children; ++a) s += avg[((size_t)n * 8 + a) * hp + h];
                    CHECK(fabs(s - 1.0) < 1e-4);
                }
                (void)q;
            }
            double gains[6];
            double e = or_exploitability(&p, avg, mode, 1000.0, gains);
            CHECK(e >= -1e-9 && e == e);
        }
    double *reach[6];
    double rbuf[6][40];
    for (int q = 0; q < np; ++q) {
        for (int h = 0; h < hp; ++h) rbuf[q][h] = 0.5;
        reach[q] = rbuf[q];
    }
    double cfv[40];
    for (int n = 0; n < t.num_nodes; ++n)
        if (t.nodes[n].type != 0)
            CHECK(or_terminal_values(&p, n, n % np, (const double *const *)reach, OR_TERM_BRUTE, cfv) == 0);
    free(avg);
    free(rs);
    free(ev);
    free(rg);
    free(ss);
    free(leaf);
    or_free_tree(&t);
}

int main(void) {
    game(2, 5, 12, 0);
    game(3, 5, 7, 0);
    game(4, 5, 4, 0);
    game(5, 5, 3, 0);
    game(2, 4, 9, 1);
    game(3, 4, 5, 1);
    int c7[7] = {0, 4, 8, 12, 16, 20, 51};
    CHECK(or_eval7(c7) != 0);
    printf("SANITIZE_OK %d checks\n", g_checks);
    return 0;
}
