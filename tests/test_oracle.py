"""The CPU oracle pinned against golden vectors, closed forms and brute force."""

import ctypes
import gzip
import itertools
import json
import os

import numpy as np
import pytest

import oracle_bind as ob
import ss_abi
import workloads as wl

HERE = os.path.dirname(os.path.abspath(__file__))
TREE_KEYS = ["num_nodes", "num_children_total", "num_leaves", "max_depth", "nodes", "children", "level_order",
             "node_depth", "decision", "showdown", "leaf", "fold"]


def golden_trees():
    return json.load(gzip.open(os.path.join(HERE, "golden", "trees.json.gz"), "rt"))["trees"]


# ---------------------------------------------------------------- evaluator
def test_eval7_golden_vectors():
    E = json.load(open(os.path.join(HERE, "golden", "eval_vectors.json")))
    for h, v in zip(E["hands7"], E["values"]):
        assert ob.eval7(h) == v
    for rb in E["river_boards"]:
        bd = rb["board"]
        combos = [(x, y) for x in range(52) for y in range(x + 1, 52) if x not in bd and y not in bd]
        assert [ob.eval7(bd + [x, y]) for x, y in combos] == rb["values"]


@pytest.mark.slow
def test_eval7_census_textbook_and_reference():
    """All C(52,7) hands: category counts are the textbook values and equal the
    reference census; the rank sum equals the reference's."""
    C = json.load(open(os.path.join(HERE, "golden", "eval_census.json")))
    counts = (ctypes.c_uint64 * 10)()
    s = ctypes.c_uint64()
    ob.lib().or_eval7_census(counts, ctypes.byref(s))
    textbook = [0, 23294460, 58627800, 31433400, 6461620, 6180020, 4047644, 3473184, 224848, 41584]
    assert list(counts) == textbook == C["category_counts"]
    assert s.value == C["rank_sum_mod_2_64"]


def test_eval5_examples():
    c = lambda s: ob.lib().or_eval5(*[("23456789TJQKA".index(t[0]) * 4 + "cdhs".index(t[1])) for t in s.split()])
    assert c("As Ks Qs Js Ts") == (9 << 20) | (12 << 16)
    assert c("Ah 2c 3d 4s 5h") == (5 << 20) | (3 << 16)           # wheel: five-high
    assert c("Ah Ad Ac Kh Kd") == (7 << 20) | (12 << 16) | (11 << 12)
    assert c("2h 2d 3c 3h 4d") == (3 << 20) | (1 << 16) | (0 << 12) | (2 << 8)


# ---------------------------------------------------------------- tree builder
@pytest.mark.parametrize("t", golden_trees(), ids=lambda t: t["name"])
def test_oracle_builder_matches_reference(t):
    td = ob.build_tree(t["board"], t["pot"], t["stack"], t["bet_sizes"], t["num_players"], t["acting_order"],
                       t["is_river"], t["use_cont_strats"], 3)
    d = ss_abi.tree_to_dict(td)
    for k in TREE_KEYS:
        assert d[k] == t[k], k
    ob.lib().or_free_tree(ctypes.byref(td))


def test_tree_sizes_at_two_raises():
    """Counts at max 2 raises (configs 1 and 2) from an independent count of
    the builder rules: the 2-player 100 BB river with sizes 0.5 and 1 pot has 51 nodes, 18 decision, 17 showdown, 16 fold; the 3-player 150 BB river with 0.33 and 0.75 has 373 / 156 / 154 / 63."""
    for board, pot, stack, sizes, np_, want in [
        ([50, 23, 4, 30, 5], 10000, 40000, [0.5, 1.0], 2, (51, 18, 17, 16)),
        ([50, 23, 4, 30, 5], 15000, 50000, [0.33, 0.75], 3, (373, 156, 154, 63)),
    ]:
        td = ob.build_tree(board, pot, stack, sizes, np_, max_raises=2)
        assert (td.num_nodes, td.num_decision_nodes, td.num_showdown_nodes, td.num_fold_nodes) == want


def _prune_equal(t3, t2, n3, n2, raises=0):
    """The raise-cap-2 tree equals the cap-3 tree with every bet/raise child
    removed at nodes that already saw 2 raises."""
    a, b = t3.nodes[n3], t2.nodes[n2]
    assert (a.type, a.player, a.pot, list(a.bets), list(a.active_players)) == \
           (b.type, b.player, b.pot, list(b.bets), list(b.active_players))
    if a.type != 0:
        return
    kids3 = [t3.children[a.first_child + i] for i in range(a.num_children)]
    kids2 = [t2.children[b.first_child + i] for i in range(b.num_children)]
    keep = []
    mx = max(a.bets[p] for p in range(6) if a.active_players[p]) if any(a.active_players) else 0
    for c in kids3:
        cn = t3.nodes[c]
        is_raise = cn.bets[a.player] > mx
        if is_raise and raises >= 2:
            continue
        keep.append((c, raises + (1 if is_raise else 0)))
    assert len(keep) == len(kids2)
    for (c3, r), c2 in zip(keep, kids2):
        _prune_equal(t3, t2, c3, c2, r)


@pytest.mark.parametrize("np_", [2, 3])
def test_raise_cap_two_is_pruned_cap_three(np_):
    board = [50, 23, 4, 30, 5]
    t3 = ob.build_tree(board, 10000, 40000, [0.5, 1.0], np_, max_raises=3)
    t2 = ob.build_tree(board, 10000, 40000, [0.5, 1.0], np_, max_raises=2)
    _prune_equal(t3, t2, 0, 0)


# ---------------------------------------------------------------- CFR restatement
def _problem(seed, npl, hp, nb=5, leaves=False):
    sp = wl.small_spot(seed, num_players=npl, hands_per_player=hp, board_cards=nb, use_leaves=leaves)
    td = ob.build_tree(sp.board, sp.pot, sp.stack, sp.bet_sizes, npl, max_raises=2,
                       use_cont=0 if nb == 5 else 1)
    lv = sp.leaf_values(td.num_leaves)
    return ob.Problem(td, sp.board, sp.hands, sp.weights, leaf_values=lv)


@pytest.mark.parametrize("seed,npl,hp,nb,leaves", [(0, 2, 10, 5, False), (1, 3, 6, 5, False),
                                                   (2, 2, 8, 4, True), (3, 3, 5, 4, True)])
def test_fast_terminals_equal_brute_force(seed, npl, hp, nb, leaves):
    """Closed-form card removal == the reference odometer, node by node, on
    random reach (with exact zeros), every traverser."""
    prob = _problem(seed, npl, hp, nb, leaves)
    r = np.random.default_rng(seed)
    td = prob.tree
    for n in range(td.num_nodes):
        if td.nodes[n].type == 0:
            continue
        for t in range(npl):
            reach = [r.uniform(0, 1, hp) * (r.uniform(0, 1, hp) < 0.8) for _ in range(npl)]
            a = ob.terminal_values(prob, n, t, reach, ob.TERM_BRUTE)
            b = ob.terminal_values(prob, n, t, reach, ob.TERM_FAST)
            assert np.allclose(a, b, rtol=1e-10, atol=1e-9), (n, t)


def _cards(txt):
    return ["23456789TJQKA".index(t[0]) * 4 + "cdhs".index(t[1]) for t in txt.split()]


def test_one_hand_each_closed_form():
    """1 hand vs 1 hand on the river (the payoff check of
    bench/prove_correctness.c:24-67): the hand that always loses folds to every
    bet (call loses the call, fold loses nothing more) and the hand that
    always wins never folds."""
    board = _cards("Ac Qc 2c 7d 4h")
    nuts = np.array([_cards("Jc Tc")], np.int32)   # club flush
    weak = np.array([_cards("3d 4d")], np.int32)   # pair of fours
    td = ob.build_tree(board, 1000, 5000, [0.75], 2, max_raises=2)
    prob = ob.Problem(td, board, [nuts, weak], [np.ones(1, np.float32)] * 2)
    res = ob.solve(prob, 300, fp64=True, mode=ob.TERM_BRUTE)
    # reach of each player under the average strategy (info sets the opponent
    # never reaches carry zero counterfactual value and stay uniform)
    reach = np.ones((td.num_nodes, 2))
    order = [0]
    for n in order:
        nd = td.nodes[n]
        for i in range(nd.num_children):
            c = td.children[nd.first_child + i]
            reach[c] = reach[n]
            if nd.type == 0:
                reach[c, nd.player] *= res["avg"][n, i, 0]
            order.append(c)
    checked = 0
    for n in range(td.num_nodes):
        nd = td.nodes[n]
        if nd.type != 0 or reach[n, 1 - nd.player] < 1e-3:
            continue
        kids = [td.children[nd.first_child + i] for i in range(nd.num_children)]
        for i, c in enumerate(kids):
            if td.nodes[c].type == 1:   # fold child
                f = res["avg"][n, i, 0]
                assert (f > 0.99) if nd.player == 1 else (f < 0.01), (n, nd.player, f)
                checked += 1
    assert checked >= 2


def test_strategies_are_distributions():
    prob = _problem(4, 3, 6)
    res = ob.solve(prob, 20, fp64=True, mode=ob.TERM_FAST)
    td = prob.tree
    for n in range(td.num_nodes):
        nd = td.nodes[n]
        if nd.type != 0:
            continue
        s = res["avg"][n, :nd.num_children, :len(prob.hands[nd.player])].sum(0)
        assert np.allclose(s, 1.0, atol=1e-6)


def _exhaustive_gain(prob, avg, p):
    """Best-response gain of player p found by enumerating every pure strategy
    of p (one action per decision node, chosen per hand), scored exactly with
    the reference's brute-force terminal values; minus p's value under avg."""
    td = prob.tree
    N = td.num_nodes
    order = [0]
    i = 0
    while i < len(order):
        nd = td.nodes[order[i]]
        order += [td.children[nd.first_child + k] for k in range(nd.num_children)]
        i += 1
    reach = [None] * prob.num_players
    for q in range(prob.num_players):
        if q == p:
            continue
        rr = np.zeros((N, len(prob.hands[q])))
        rr[0] = prob.weights[q]
        for n in order:
            nd = td.nodes[n]
            if nd.type != 0:
                continue
            for k in range(nd.num_children):
                c = td.children[nd.first_child + k]
                rr[c] = rr[n] * (avg[n, k, :len(prob.hands[q])] if nd.player == q else 1.0)
        reach[q] = rr
    term = {n: ob.terminal_values(prob, n, p, [reach[q][n] if q != p else None for q in range(prob.num_players)],
                                  ob.TERM_BRUTE)
            for n in range(N) if td.nodes[n].type != 0}
    infos = [n for n in range(N) if td.nodes[n].type == 0 and td.nodes[n].player == p]
    gain = 0.0
    for h in range(len(prob.hands[p])):
        def val(n, pick):
            nd = td.nodes[n]
            if nd.type != 0:
                return term[n][h]
            kids = [td.children[nd.first_child + k] for k in range(nd.num_children)]
            if nd.player == p:
                if pick is None:
                    return sum(avg[n, k, h] * val(c, None) for k, c in enumerate(kids))
                return val(kids[pick[n]], pick)
            return sum(val(c, pick) for c in kids)
        best = max(val(0, dict(zip(infos, combo)))
                   for combo in itertools.product(*[range(td.nodes[n].num_children) for n in infos]))
        gain += prob.weights[p][h] * (best - val(0, None))
    return gain


def test_exploitability_matches_exhaustive_pure_best_response():
    board = _cards("Ac Qc 2c 7d 4h")
    hands = [np.array([_cards("Jc Tc"), _cards("3d 4d")], np.int32),
             np.array([_cards("As Ks"), _cards("5s 6s")], np.int32)]
    w = [np.array([0.6, 0.4], np.float32), np.array([0.5, 0.5], np.float32)]
    td = ob.build_tree(board, 1000, 2000, [1.0], 2, max_raises=1)
    prob = ob.Problem(td, board, hands, w)
    res = ob.solve(prob, 7, fp64=True, mode=ob.TERM_BRUTE)
    e, gains = ob.exploitability(prob, res["avg"], mode=ob.TERM_BRUTE)
    for p in range(2):
        assert abs(gains[p] - _exhaustive_gain(prob, res["avg"], p)) < 1e-6 * max(1.0, abs(gains[p]))
    assert e > 0


def test_exploitability_decreases():
    prob = _problem(9, 2, 12)
    es = []
    for it in (5, 50, 400):
        res = ob.solve(prob, it, fp64=True, mode=ob.TERM_FAST)
        es.append(ob.exploitability(prob, res["avg"])[0])
    assert es[0] > es[1] > es[2] >= 0
