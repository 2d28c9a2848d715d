// ss_tree.cpp -- single-street betting-tree builder behind ss_build_tree.
//
// Produces the reference layout bit-for-bit (pinned by tests/golden/trees.json.gz,
// written from the reference's own builder). The betting rules and their quirks
// follow src/cuda/street_solve.cu:
//   fold child only when facing a bet (:229-241); check/call always (:244-260);
//   bets while raises < cap and < 14 children (:263); sizes truncated from float
//   products (:267-270), capped at the stack, dropped when <= to_call, one all-in
//   (:272-278); the raiser's bet becomes max_bet + amount (:284) and the stack is
//   never decremented (:290); explicit all-in when no size reached it (:302-321);
//   fold terminal once one player remains, its winner in `player` (:155-166);
//   round complete -> showdown (river) or depth-limit leaf (:169-182).
// Node ids are allocation order of a depth-first build; a decision node's block
// of child ids is appended to `children` after all of its subtrees (:325-328).
// Continuation expansion (:336-460) and BFS/classification (:1238-1287) follow.
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ss_internal.h"

namespace ss {

namespace {

struct Builder {
    int is_river, np, max_raises;
    const int *order;
    const float *sizes;
    int nsizes;
    const int *board;
    int nb;

    std::vector<SSNode> nodes;
    std::vector<int> kids;
    int leaves = 0;

    int new_node(int type, int pot, const int *bets, const int *active) {
        SSNode n;
        std::memset(&n, 0, sizeof(n));
        n.type = type;
        n.player = -1;
        n.leaf_idx = -1;
        n.fold_player = -1;
        n.pot = pot;
        for (int i = 0; i < np; ++i) n.bets[i] = bets[i];
        for (int i = 0; i < nb && i < 5; ++i) n.board_cards[i] = board[i];
        n.num_board = nb;
        for (int i = 0; i < np; ++i) {
            n.active_players[i] = active[i];
            n.num_players += active[i] ? 1 : 0;
        }
        nodes.push_back(n);
        return (int)nodes.size() - 1;
    }

    // Position in acting order of the next active player after `cur`.
    int next_seat(const int *active, int cur) const {
        for (int step = 1; step <= np; ++step) {
            int idx = (cur + step) % np;
            int p = order[idx];
            if (p < np && active[p]) return idx;
        }
        return -1;
    }

    struct State {
        int bets[SS_MAX_PLAYERS];
        int acted[SS_MAX_PLAYERS];
        int active[SS_MAX_PLAYERS];
    };

    int build(int seat, int pot, int stack, const State &st, int raises) {
        int live = 0, top = 0;
        for (int i = 0; i < np; ++i)
            if (st.active[i]) {
                ++live;
                if (st.bets[i] > top) top = st.bets[i];
            }
        if (live <= 1) {
            int id = new_node(SS_NODE_FOLD, pot, st.bets, st.active);
            for (int i = 0; i < np; ++i)
                if (st.active[i]) { nodes[id].player = i; break; }
            return id;
        }
        bool done = true;
        for (int i = 0; i < np && done; ++i)
            if (st.active[i] && (!st.acted[i] || st.bets[i] != top)) done = false;
        if (done) {
            if (is_river) return new_node(SS_NODE_SHOWDOWN, pot, st.bets, st.active);
            int id = new_node(SS_NODE_LEAF, pot, st.bets, st.active);
            nodes[id].leaf_idx = leaves++;
            return id;
        }
        int actor = order[seat];
        if (!st.active[actor]) {
            int nx = next_seat(st.active, seat);
            if (nx < 0) return new_node(SS_NODE_SHOWDOWN, pot, st.bets, st.active);
            return build(nx, pot, stack, st, raises);
        }
        int id = new_node(SS_NODE_DECISION, pot, st.bets, st.active);
        nodes[id].player = actor;

        int to_call = top - st.bets[actor];
        if (to_call < 0) to_call = 0;
        int nx = next_seat(st.active, seat);
        if (nx < 0) nx = seat;

        int child[16], nc = 0;
        if (to_call > 0) {
            State f = st;
            f.active[actor] = 0;
            child[nc++] = build(nx, pot, stack, f, raises);
        }
        {
            State c = st;
            c.bets[actor] = top;
            c.acted[actor] = 1;
            child[nc++] = build(nx, pot + to_call, stack, c, raises);
        }
        auto raise_to = [&](int amount) {
            State r = st;
            r.bets[actor] = top + amount;
            r.acted[actor] = 1;
            for (int q = 0; q < np; ++q)
                if (q != actor && st.active[q]) r.acted[q] = 0;
            child[nc++] = build(nx, pot + amount, stack, r, raises + 1);
        };
        if (raises < max_raises && nc < 14) {
            bool shoved = false;
            for (int i = 0; i < nsizes && nc < 14; ++i) {
                int amount = to_call == 0 ? (int)(sizes[i] * (float)pot)
                                          : to_call + (int)(sizes[i] * (float)(pot + to_call));
                if (amount >= stack) amount = stack;
                if (amount <= to_call) continue;
                if (amount >= stack) {
                    if (shoved) continue;
                    shoved = true;
                }
                raise_to(amount);
            }
            if (!shoved && stack > to_call && nc < 14) raise_to(stack);
        }
        int base = (int)kids.size();
        for (int i = 0; i < nc; ++i) kids.push_back(child[i]);
        nodes[id].first_child = base;
        nodes[id].num_children = nc;
        return id;
    }

    // Each pre-expansion leaf becomes a decision chain, one level per active
    // player in seat-index order, SS_NUM_CONT_STRATS children per node.
    void expand_leaves() {
        std::vector<int> old;
        for (int i = 0; i < (int)nodes.size(); ++i)
            if (nodes[i].type == SS_NODE_LEAF) old.push_back(i);
        leaves = 0;
        for (int root : old) {
            SSNode tmpl = nodes[root];
            int who[SS_MAX_PLAYERS], nwho = 0;
            for (int p = 0; p < np; ++p)
                if (tmpl.active_players[p]) who[nwho++] = p;
            if (nwho == 0) continue;
            nodes[root].type = SS_NODE_DECISION;
            nodes[root].player = who[0];
            nodes[root].leaf_idx = -1;
            std::vector<int> frontier(1, root), next;
            for (int lev = 0; lev < nwho; ++lev) {
                bool last = lev == nwho - 1;
                next.clear();
                for (int par : frontier) {
                    if (lev > 0) {
                        nodes[par].type = SS_NODE_DECISION;
                        nodes[par].player = who[lev];
                    }
                    int c4[SS_NUM_CONT_STRATS];
                    for (int s = 0; s < SS_NUM_CONT_STRATS; ++s) {
                        // intermediate nodes stay type 0 with player -1 until
                        // the next level assigns them (as the reference does)
                        int c = new_node(last ? SS_NODE_LEAF : SS_NODE_DECISION, tmpl.pot, tmpl.bets,
                                         tmpl.active_players);
                        std::memcpy(nodes[c].bets, tmpl.bets, sizeof(tmpl.bets));
                        if (last) nodes[c].leaf_idx = leaves++;
                        c4[s] = c;
                        if (!last) next.push_back(c);
                    }
                    int base = (int)kids.size();
                    for (int s = 0; s < SS_NUM_CONT_STRATS; ++s) kids.push_back(c4[s]);
                    nodes[par].first_child = base;
                    nodes[par].num_children = SS_NUM_CONT_STRATS;
                }
                frontier.swap(next);
            }
        }
    }
};

template <typename T>
T *dup_array(const std::vector<T> &v, size_t extra = 0) {
    T *p = (T *)std::malloc(sizeof(T) * (v.size() + extra) + 1);
    if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

}  // namespace

int build_tree(const int *board, int num_board, int starting_pot, int effective_stack,
               const float *bet_sizes, int num_bet_sizes, int num_players, const int *acting_order,
               int is_river, int use_cont_strats, int max_raises, SSTreeData *out) {
    if (!out) return set_error(SS_ERR_ARG, "tree_out is NULL");
    std::memset(out, 0, sizeof(*out));
    if (num_players < 2 || num_players > SS_MAX_PLAYERS)
        return set_error(SS_ERR_ARG, "num_players must be 2..6");
    if (num_board < 0 || num_board > SS_MAX_BOARD || (num_board > 0 && !board))
        return set_error(SS_ERR_ARG, "bad board");
    if (!acting_order) return set_error(SS_ERR_ARG, "acting_order is NULL");
    if (num_bet_sizes < 0 || (num_bet_sizes > 0 && !bet_sizes)) return set_error(SS_ERR_ARG, "bad bet sizes");
    if (max_raises < 0 || max_raises > SS_EX_MAX_RAISES) return set_error(SS_ERR_ARG, "max_raises out of range");
    for (int i = 0; i < num_players; ++i)
        if (acting_order[i] < 0) return set_error(SS_ERR_ARG, "negative seat in acting_order");

    Builder b{is_river, num_players, max_raises, acting_order, bet_sizes, num_bet_sizes, board, num_board};
    b.nodes.reserve(512);
    Builder::State st;
    std::memset(&st, 0, sizeof(st));
    for (int i = 0; i < num_players; ++i) st.active[i] = 1;
    b.build(0, starting_pot, effective_stack, st, 0);
    if (use_cont_strats && !is_river && b.leaves > 0) b.expand_leaves();

    const int N = (int)b.nodes.size();
    out->nodes = dup_array(b.nodes);
    out->children = dup_array(b.kids);
    out->num_nodes = N;
    out->num_children_total = (int)b.kids.size();
    out->num_leaves = b.leaves;
    out->is_river = is_river;
    out->starting_pot = starting_pot;
    out->effective_stack = effective_stack;
    out->num_players = num_players;
    for (int i = 0; i < num_board; ++i) out->board[i] = board[i];
    out->num_board = num_board;

    std::vector<int> bfs, depth(N, 0);
    bfs.reserve(N);
    bfs.push_back(0);
    int maxd = 0;
    for (size_t h = 0; h < bfs.size(); ++h) {
        const SSNode &nd = b.nodes[bfs[h]];
        int d = depth[bfs[h]];
        if (d > maxd) maxd = d;
        for (int a = 0; a < nd.num_children; ++a) {
            int c = b.kids[nd.first_child + a];
            depth[c] = d + 1;
            bfs.push_back(c);
        }
    }
    out->level_order = dup_array(bfs);
    out->node_depth = dup_array(depth);
    out->max_depth = maxd;

    std::vector<int> byType[5];
    for (int i = 0; i < N; ++i) byType[b.nodes[i].type].push_back(i);
    out->decision_node_indices = dup_array(byType[SS_NODE_DECISION], 1);
    out->num_decision_nodes = (int)byType[SS_NODE_DECISION].size();
    out->showdown_node_indices = dup_array(byType[SS_NODE_SHOWDOWN], 1);
    out->num_showdown_nodes = (int)byType[SS_NODE_SHOWDOWN].size();
    out->leaf_node_indices = dup_array(byType[SS_NODE_LEAF], 1);
    out->num_leaf_nodes = (int)byType[SS_NODE_LEAF].size();
    out->fold_node_indices = dup_array(byType[SS_NODE_FOLD], 1);
    out->num_fold_nodes = (int)byType[SS_NODE_FOLD].size();
    if (!out->nodes || !out->children || !out->level_order || !out->node_depth)
        return set_error(SS_ERR_NOMEM, "out of host memory building the tree");
    return SS_OK;
}

void free_tree(SSTreeData *td) {
    if (!td) return;
    std::free(td->nodes);
    std::free(td->children);
    std::free(td->level_order);
    std::free(td->node_depth);
    std::free(td->decision_node_indices);
    std::free(td->showdown_node_indices);
    std::free(td->leaf_node_indices);
    std::free(td->fold_node_indices);
    // leaf_values and frozen_action belong to the caller: never freed here
    // (the reference frees leaf_values at street_solve.cu:1650, which is a
    // Python-owned ctypes buffer in python/street_solver_gpu.py:311-313).
    std::memset(td, 0, sizeof(*td));
}

}  // namespace ss
