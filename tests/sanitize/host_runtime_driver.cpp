// host_runtime_driver.cpp -- drives the host code of libstreet_solve under
// AddressSanitizer and UndefinedBehaviorSanitizer (tests/sanitize/Makefile).
//
// Test infrastructure. Runs on a machine without a GPU: every call that
// reaches the device returns SS_ERR_HIP, so the parts exercised are the tree
// builder (ss_build_tree / ss_build_tree_ex / ss_free_tree, with structural
// checks of every tree), the planner behind ss_problem_footprint and
// ssdbg_plan_bytes, the argument checks of every entry point, the host plans
// of the leaf calls (runout and tuple lists, Python-compatible draws) and the
// error and cleanup paths of the solve, batch and server entry points.
// Prints "SANITIZE_OK <checks>" and exits 0 when every check held; any
// sanitizer report aborts the process with a non-zero status.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <random>
#include <vector>

#include "street_solve.h"
#include "street_solve_debug.h"
#include "street_solve_leaf.h"
#include "street_solve_range.h"

static int g_checks = 0;

#define CHECK(cond)                                                                 \
    do {                                                                            \
        ++g_checks;                                                                 \
        if (!(cond)) {                                                              \
            std::fprintf(stderr, "check failed: %s (%s:%d)\n", #cond, __FILE__, __LINE__); \
            std::exit(3);                                                           \
        }                                                                           \
    } while (0)

// Every child id in range, every non-root node the child of exactly one
// node, level_order a permutation in non-decreasing depth, the per-type index
// lists consistent with the node types, leaf indices 0..num_leaves-1 once each.
static void check_tree(const SSTreeData &t, int np) {
    const int N = t.num_nodes;
    CHECK(N >= 1 && t.nodes && t.children && t.level_order && t.node_depth);
    std::vector<int> parent_count(N, 0), seen(N, 0), leaf_seen(t.num_leaves > 0 ? t.num_leaves : 1, 0);
    int ndec = 0, nsd = 0, nleaf = 0, nfold = 0;
    for (int n = 0; n < N; ++n) {
        const SSNode &nd = t.nodes[n];
        if (nd.type == SS_NODE_DECISION) {
            ++ndec;
            CHECK(nd.num_children >= 1 && nd.num_children <= SS_MAX_ACTIONS);
            CHECK(nd.player >= 0 && nd.player < np);
            CHECK(nd.first_child >= 0 && nd.first_child + nd.num_children <= t.num_children_total);
            for (int a = 0; a < nd.num_children; ++a) {
                int c = t.children[nd.first_child + a];
                CHECK(c > 0 && c < N);
                ++parent_count[c];
                CHECK(t.node_depth[c] == t.node_depth[n] + 1);
            }
        } else {
            CHECK(nd.num_children == 0);
            if (nd.type == SS_NODE_SHOWDOWN) ++nsd;
            if (nd.type == SS_NODE_FOLD) {
                ++nfold;
                CHECK(nd.player >= 0 && nd.player < np);
            }
            if (nd.type == SS_NODE_LEAF) {
                ++nleaf;
                CHECK(nd.leaf_idx >= 0 && nd.leaf_idx < t.num_leaves);
                ++leaf_seen[nd.leaf_idx];
            }
        }
    }
    CHECK(parent_count[0] == 0);
    for (int n = 1; n < N; ++n) CHECK(parent_count[n] == 1);
    int prev = 0;
    for (int i = 0; i < N; ++i) {
        int n = t.level_order[i];
        CHECK(n >= 0 && n < N && !seen[n]);
        seen[n] = 1;
        CHECK(t.node_depth[n] >= prev && t.node_depth[n] <= t.max_depth);
        prev = t.node_depth[n];
    }
    CHECK(ndec == t.num_decision_nodes && nsd == t.num_showdown_nodes);
    CHECK(nleaf == t.num_leaf_nodes && nfold == t.num_fold_nodes);
    for (int i = 0; i < t.num_decision_nodes; ++i) CHECK(t.nodes[t.decision_node_indices[i]].type == SS_NODE_DECISION);
    for (int i = 0; i < t.num_showdown_nodes; ++i) CHECK(t.nodes[t.showdown_node_indices[i]].type == SS_NODE_SHOWDOWN);
    for (int i = 0; i < t.num_fold_nodes; ++i) CHECK(t.nodes[t.fold_node_indices[i]].type == SS_NODE_FOLD);
    for (int i = 0; i < t.num_leaf_nodes; ++i) CHECK(t.nodes[t.leaf_node_indices[i]].type == SS_NODE_LEAF);
    if (t.num_leaves > 0 && nleaf == t.num_leaves)
        for (int l = 0; l < t.num_leaves; ++l) CHECK(leaf_seen[l] == 1);
}

// Random unblocked, distinct hands for every player: int32 [nh][2] each.
struct Ranges {
    std::vector<std::vector<int>> hands;
    std::vector<std::vector<float>> weights;
};

static Ranges random_ranges(std::mt19937 &rng, const int *board, int nb, int np, int nh) {
    std::vector<std::pair<int, int>> combos;
    for (int a = 0; a < 52; ++a)
        for (int b = a + 1; b < 52; ++b) {
            bool blocked = false;
            for (int k = 0; k < nb; ++k) blocked |= board[k] == a || board[k] == b;
            if (!blocked) combos.emplace_back(a, b);
        }
    Ranges r;
    std::uniform_real_distribution<float> u(0.f, 1.f);
    for (int p = 0; p < np; ++p) {
        std::shuffle(combos.begin(), combos.end(), rng);
        std::vector<int> h;
        std::vector<float> w;
        for (int i = 0; i < nh && i < (int)combos.size(); ++i) {
            h.push_back(combos[i].first);
            h.push_back(combos[i].second);
            w.push_back(u(rng));
        }
        r.hands.push_back(h);
        r.weights.push_back(w);
    }
    return r;
}

static SSProblem problem_of(const SSTreeData &t, const Ranges &r, const std::vector<float> &leaf) {
    SSProblem p;
    std::memset(&p, 0, sizeof(p));
    p.tree = &t;
    for (size_t q = 0; q < r.hands.size(); ++q) {
        p.num_hands[q] = (int)r.weights[q].size();
        p.hands[q] = r.hands[q].data();
        p.weights[q] = r.weights[q].data();
    }
    p.leaf_values = leaf.empty() ? nullptr : leaf.data();
    return p;
}

static void trees_and_plans(std::mt19937 &rng) {
    const int river[5] = {3, 17, 29, 40, 51}, turn[4] = {3, 17, 29, 40}, flop[3] = {3, 17, 29};
    const float sizes_a[2] = {0.5f, 1.0f}, sizes_b[4] = {0.25f, 0.5f, 1.0f, 1.5f}, sizes_c[1] = {0.33f};
    struct Cfg {
        const int *board;
        int nb, np, cont, raises, nsizes, pot, stack;
        const float *sizes;
    };
    std::vector<Cfg> cfgs;
    for (int np = 2; np <= 6; ++np) {
        cfgs.push_back({river, 5, np, 0, 2, 2, 10000, 40000, sizes_a});
        cfgs.push_back({river, 5, np, 0, np <= 4 ? 3 : 1, 1, 1500, 900, sizes_c});
        if (np <= 4) cfgs.push_back({river, 5, np, 0, 2, 4, 5000, 100000, sizes_b});
        if (np <= 3) {
            cfgs.push_back({turn, 4, np, 1, 2, 2, 10000, 40000, sizes_a});
            cfgs.push_back({turn, 4, np, 0, 1, 1, 10000, 5000, sizes_c});
        }
        if (np == 2) cfgs.push_back({flop, 3, np, 1, 2, 2, 10000, 40000, sizes_a});
    }
    cfgs.push_back({river, 5, 3, 0, 0, 0, 10000, 40000, nullptr});   // no sizes: all-in only
    for (const Cfg &c : cfgs) {
        int order[SS_MAX_PLAYERS];
        for (int i = 0; i < c.np; ++i) order[i] = (i * 5 + 1) % c.np;   // a rotation, valid for every np here
        bool perm = true;
        std::vector<int> cnt(c.np, 0);
        for (int i = 0; i < c.np; ++i) ++cnt[order[i]];
        for (int i = 0; i < c.np; ++i) perm &= cnt[i] == 1;
        if (!perm)
            for (int i = 0; i < c.np; ++i) order[i] = i;
        SSTreeData t;
        int rc = ss_build_tree_ex(c.board, c.nb, c.pot, c.stack, c.sizes, c.nsizes, c.np, order, c.nb == 5,
                                  c.cont, c.raises, &t);
        CHECK(rc == 0);
        check_tree(t, c.np);
        // three or more opponents run the enumerating terminals, which the
        // planner admits up to 2^24 opponent tuples per hand: small ranges there
        const int nh = c.np <= 3 ? 40 + (int)(rng() % 160) : (c.np == 4 ? 60 : (c.np == 5 ? 25 : 12));
        Ranges r = random_ranges(rng, c.board, c.nb, c.np, nh);
        std::vector<float> leaf;
        if (t.num_leaves > 0) leaf.assign((size_t)t.num_leaves * c.np * nh, 0.25f);
        SSProblem p = problem_of(t, r, leaf);
        int64_t fp = ss_problem_footprint(&p);
        if (fp <= 0) std::fprintf(stderr, "np %d nh %d: %s\n", c.np, nh, ss_last_error());
        CHECK(fp > 0);
        if (c.np >= 5) {   // full ranges there exceed the tuple guard: refused, not attempted
            Ranges big = random_ranges(rng, c.board, c.nb, c.np, 1000);
            SSProblem pb = problem_of(t, big, leaf);
            CHECK(ss_problem_footprint(&pb) == SS_ERR_UNSUPPORTED);
        }
        int64_t algo = 0, fp2 = 0;
        CHECK(ssdbg_plan_bytes(&p, &algo, &fp2) == 0);
        CHECK(algo > 0 && fp2 == fp);
        // the legacy entry point on the same tree: hands in the fixed slots
        if (nh <= SS_MAX_HANDS) {
            for (int q = 0; q < c.np; ++q) {
                t.num_hands[q] = nh;
                for (int h = 0; h < nh; ++h) {
                    t.hands[q][h][0] = r.hands[q][2 * h];
                    t.hands[q][h][1] = r.hands[q][2 * h + 1];
                    t.weights[q][h] = r.weights[q][h];
                }
            }
            t.leaf_values = leaf.empty() ? nullptr : leaf.data();
            SSOutput out;
            std::memset(&out, 0, sizeof(out));
            int src = ss_solve_gpu(&t, 2, &out);
            CHECK(src != 0 || out.avg_strategies != nullptr);   // no GPU here: an error, never a crash
            if (src == 0) ss_free_output(&out);
        }
        ss_free_tree(&t);
        CHECK(t.nodes == nullptr && t.num_nodes == 0);
    }
}

static void argument_errors(std::mt19937 &rng) {
    const int river[5] = {3, 17, 29, 40, 51};
    const float sizes[2] = {0.5f, 1.0f};
    const int order[6] = {0, 1, 2, 3, 4, 5};
    SSTreeData t;
    CHECK(ss_build_tree_ex(river, 5, 10000, 40000, sizes, 2, 1, order, 1, 0, 2, &t) < 0);
    CHECK(ss_build_tree_ex(river, 5, 10000, 40000, sizes, 2, 7, order, 1, 0, 2, &t) < 0);
    CHECK(ss_build_tree_ex(nullptr, 5, 10000, 40000, sizes, 2, 2, order, 1, 0, 2, &t) < 0);
    CHECK(ss_build_tree_ex(river, 5, 10000, 40000, sizes, 2, 2, nullptr, 1, 0, 2, &t) < 0);
    CHECK(ss_build_tree_ex(river, 5, 10000, 40000, sizes, 2, 2, order, 1, 0, 99, &t) < 0);
    CHECK(ss_build_tree_ex(river, 5, 10000, 40000, sizes, 2, 2, order, 1, 0, 2, nullptr) < 0);
    CHECK(std::strlen(ss_last_error()) > 0);

    CHECK(ss_build_tree_ex(river, 5, 10000, 40000, sizes, 2, 2, order, 1, 0, 2, &t) == 0);
    Ranges r = random_ranges(rng, river, 5, 2, 30);
    SSProblem p = problem_of(t, r, {});
    CHECK(ss_problem_footprint(nullptr) < 0);
    // a hand holding a board card is valid input (a dead hand, as in the
    // reference); a hand with a repeated card or a card out of range is not
    std::vector<int> bad = r.hands[0];
    bad[0] = river[2];
    p.hands[0] = bad.data();
    CHECK(ss_problem_footprint(&p) > 0);
    bad = r.hands[0];
    bad[1] = bad[0];
    p.hands[0] = bad.data();
    CHECK(ss_problem_footprint(&p) < 0);
    bad = r.hands[0];
    bad[2] = 52;
    p.hands[0] = bad.data();
    CHECK(ss_problem_footprint(&p) < 0);
    p.hands[0] = r.hands[0].data();
    p.num_hands[0] = 1327;
    CHECK(ss_problem_footprint(&p) < 0);
    p.num_hands[0] = 30;
    p.weights[1] = nullptr;
    CHECK(ss_problem_footprint(&p) < 0);
    p.weights[1] = r.weights[1].data();
    CHECK(ss_problem_footprint(&p) > 0);

    // device-bound entry points: without a GPU an error, never a crash or leak
    SSSolveConfig cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    cfg.max_iterations = 3;
    SSOutput out;
    SSReport rep;
    std::memset(&out, 0, sizeof(out));
    int rc = ss_solve_gpu_ex(&p, &cfg, &out, &rep);
    if (rc == 0) ss_free_output(&out);
    SSOutput outs[1];
    SSReport reps[1];
    std::memset(outs, 0, sizeof(outs));
    rc = ss_solve_batch_ex(&p, 1, &cfg, outs, reps);
    if (rc == 0) ss_free_output(&outs[0]);
    SSBatch *b = ss_batch_create(&p, 1, &cfg);
    if (b) ss_batch_destroy(b);
    cfg.precision = 7;
    CHECK(ss_solve_gpu_ex(&p, &cfg, &out, &rep) < 0);
    SSServerConfig sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.max_slots = 4;
    sc.iters_per_step = 5;
    sc.check_every = 10;
    SSServer *s = ss_server_create(&sc);
    if (s) ss_server_destroy(s);
    ss_free_tree(&t);
    SSOutput empty;
    std::memset(&empty, 0, sizeof(empty));
    ss_free_output(&empty);
    ss_free_output(nullptr);
    ss_free_tree(nullptr);
}

static void leaf_and_range_host_paths(std::mt19937 &rng) {
    // ss_leaf_equity / ss_leaf_equity_static: the host plan (unblocked lists,
    // tuple sources, Python-compatible draws over 2000 tuples) runs in full
    // before the first device call
    const int np = 3, mh = 60, nrun = 6;
    std::vector<int> hands((size_t)np * mh * 2), nh(np, mh), runs;
    std::vector<int> deck(52);
    for (int i = 0; i < 52; ++i) deck[i] = i;
    std::shuffle(deck.begin(), deck.end(), rng);
    const int board4[4] = {deck[0], deck[1], deck[2], deck[3]};
    for (int p = 0; p < np; ++p)
        for (int h = 0; h < mh; ++h) {
            int a = deck[4 + (p * 7 + h) % 48], c = deck[4 + (p * 7 + h + 1 + h % 5) % 48];
            if (a == c) c = deck[4 + (p * 7 + h + 2) % 48];
            hands[(p * mh + h) * 2] = a;
            hands[(p * mh + h) * 2 + 1] = c;
        }
    for (int r = 0; r < nrun; ++r) {
        for (int k = 0; k < 4; ++k) runs.push_back(board4[k]);
        runs.push_back(deck[4 + 40 + r]);
    }
    std::vector<double> raw((size_t)np * mh, -1.0);
    int rc = ss_leaf_equity(np, hands.data(), nh.data(), mh, runs.data(), nrun, raw.data());
    CHECK(rc == 0 || rc == SS_ERR_HIP);
    rc = ss_leaf_equity_static(np, hands.data(), nh.data(), mh, runs.data(), nrun, raw.data());
    CHECK(rc == 0 || rc == SS_ERR_HIP);
    CHECK(ss_leaf_equity(1, hands.data(), nh.data(), mh, runs.data(), nrun, raw.data()) == SS_ERR_ARG);
    CHECK(ss_leaf_equity(np, nullptr, nh.data(), mh, runs.data(), nrun, raw.data()) == SS_ERR_ARG);
    std::vector<int> bad_runs = runs;
    bad_runs[1] = bad_runs[0];
    CHECK(ss_leaf_equity_static(np, hands.data(), nh.data(), mh, bad_runs.data(), nrun, raw.data()) == SS_ERR_ARG);
    // zero runouts: answered on the host, all zeros
    CHECK(ss_leaf_equity_static(np, hands.data(), nh.data(), mh, runs.data(), 0, raw.data()) == 0);
    for (double v : raw) CHECK(v == 0.0);
    std::vector<int> idx = {0, 2}, pot = {1000, 1500};
    std::vector<float> vals((size_t)2 * 64 * np * mh);
    rc = ss_leaf_values(np, raw.data(), nh.data(), mh, idx.data(), pot.data(), 2, vals.data());
    CHECK(rc == 0 || rc == SS_ERR_HIP);
    idx[1] = -1;
    CHECK(ss_leaf_values(np, raw.data(), nh.data(), mh, idx.data(), pot.data(), 2, vals.data()) == SS_ERR_ARG);
    std::vector<int> lens = {7, 3, 11};
    std::vector<int> draws(3 * 50);
    CHECK(ssdbg_py_choice(42, lens.data(), 3, 50, draws.data()) == 0);
    for (int t = 0; t < 50; ++t)
        for (int k = 0; k < 3; ++k) CHECK(draws[t * 3 + k] >= 0 && draws[t * 3 + k] < lens[k]);
    // range update argument checks
    std::vector<double> w(10, 1.0), pa(10, 0.5);
    std::vector<unsigned char> has(10, 1);
    CHECK(ss_range_update(-1, w.data(), pa.data(), has.data(), 0.5, 0.005) < 0);
    CHECK(ss_range_update(10, nullptr, pa.data(), has.data(), 0.5, 0.005) < 0);
    // host evaluator against itself on the device path's host twin
    std::vector<int> cards7 = {0, 4, 8, 12, 16, 20, 51};
    uint32_t r7 = 0;
    CHECK(ssdbg_eval7(cards7.data(), 1, &r7, 0) == 0 && r7 != 0);
}

int main() {
    std::mt19937 rng(20261204);
    trees_and_plans(rng);
    argument_errors(rng);
    leaf_and_range_host_paths(rng);
    std::printf("SANITIZE_OK %d checks; version %s\n", g_checks, ss_version());
    return 0;
}
