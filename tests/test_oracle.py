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


# ---------------------------------------------------------------- hand-derived pins of the CFR loop
# A 9-node two-player river (pot 100, stack 50, one size 0.5 pot, max 2 raises):
#   0 P0: check -> 1, bet 50 -> 6      1 P1: check -> 2 (showdown, pot 100), bet 50 -> 3
#   3 P0: fold -> 4 (P1 wins pot 150), call -> 5 (showdown, pot 200)
#   6 P1: fold -> 7 (P0 wins pot 150), call -> 8 (showdown, pot 200)
# P0 holds 6h6s (a seven-high straight), P1 holds AcAd (the wheel): P0 wins
# every showdown. Raw weights 2.0 (P0) and 0.5 (P1), deliberately not 1 and
# not normalised. Every number below was worked out by hand from
# street_solve.cu:575-614 (regret matching), :650-683 (reach), :704-1005
# (terminal values), :1155-1175 (backup: R <- (R + dR) * t/(t+1),
# SS += t * sigma), :1471-1537 (passes P0 then P1, each re-matching every
# node first) and :1563-1578 (root outputs).
TINY_BOARD = "2c 3d 4h 5s 7c"


def _tiny_2p():
    board = _cards(TINY_BOARD)
    td = ob.build_tree(board, 100, 50, [0.5], 2, max_raises=2)
    assert td.num_nodes == 9
    hands = [np.array([_cards("6h 6s")], np.int32), np.array([_cards("Ac Ad")], np.int32)]
    assert ob.eval7(board + _cards("6h 6s")) > ob.eval7(board + _cards("Ac Ad"))
    return ob.Problem(td, board, hands, [np.array([2.0], np.float32), np.array([0.5], np.float32)])


# (node, action) -> literal after iteration 1 and after iteration 2
TINY_REGRETS = {
    1: {(0, 0): -4.6875, (0, 1): 4.6875, (3, 0): -9.375, (3, 1): 9.375,
        (1, 0): 0.0, (1, 1): 0.0, (6, 0): 25.0, (6, 1): -25.0},
    2: {(0, 0): 125.0 / 24.0, (0, 1): 3.125, (3, 0): -31.25, (3, 1): 6.25,
        (1, 0): 125.0 / 6.0, (1, 1): -125.0 / 6.0, (6, 0): 50.0 / 3.0, (6, 1): -125.0 / 3.0},
}
TINY_STRAT_SUMS = {
    1: {(0, 0): 0.5, (0, 1): 0.5, (3, 0): 0.5, (3, 1): 0.5, (1, 0): 0.5, (1, 1): 0.5, (6, 0): 0.5, (6, 1): 0.5},
    2: {(0, 0): 0.5, (0, 1): 2.5, (3, 0): 0.5, (3, 1): 2.5, (1, 0): 1.5, (1, 1): 1.5, (6, 0): 2.5, (6, 1): 0.5},
}
TINY_ROOT = {1: ((0.0, 1.0), -50.0), 2: ((0.625, 0.375), -31.25)}


@pytest.mark.parametrize("k", [1, 2])
def test_cfr_state_hand_derived(k):
    """Regrets and strategy sums after k iterations equal the hand-derived
    literals. Mistakes this catches: the discount applied before the regret
    add (R <- R*t/(t+1) + dR gives R[3] = (-18.75, 18.75) after one
    iteration); a reach-weighted strategy sum (P0's weight 2 doubles SS[0]);
    a strategy sum without the t factor (SS[0] = (0.5, 1.5) after two);
    the traverser's own reach in its values, or normalised weights (every
    regret scales)."""
    prob = _tiny_2p()
    r = ob.solve(prob, k, fp64=True, mode=ob.TERM_BRUTE, want_state=True)
    for (n, a), want in TINY_REGRETS[k].items():
        assert r["regrets"][n, a, 0] == pytest.approx(want, abs=1e-12), (k, n, a)
    for (n, a), want in TINY_STRAT_SUMS[k].items():
        assert r["strat_sum"][n, a, 0] == pytest.approx(want, abs=1e-12), (k, n, a)
    # the fast (closed-form) terminals give the same numbers
    f = ob.solve(prob, k, fp64=True, mode=ob.TERM_FAST, want_state=True)
    assert np.array_equal(f["regrets"], r["regrets"]) and np.array_equal(f["strat_sum"], r["strat_sum"])


@pytest.mark.parametrize("k", [1, 2])
def test_root_outputs_hand_derived(k):
    """root_strategy is the root actor's strategy from the LAST traverser
    pass's regret matching ((0.625, 0.375) after two iterations; the first
    pass's would be (0, 1)); root_ev is the node-0 value of the LAST traverser
    (player NP-1 = P1: -50 after one iteration; P0's would be +53.125)."""
    prob = _tiny_2p()
    r = ob.solve(prob, k, fp64=True, mode=ob.TERM_BRUTE)
    (s0, s1), ev = TINY_ROOT[k]
    assert r["root_strategy"][0, 0] == pytest.approx(s0, abs=1e-6)
    assert r["root_strategy"][1, 0] == pytest.approx(s1, abs=1e-6)
    assert r["root_ev"][0] == pytest.approx(ev, abs=1e-5)


def test_average_strategy_hand_derived():
    """Host normalisation (street_solve.cu:1597-1610) of the strategy sums
    after two iterations: (1/6, 5/6) at P0's nodes 0 and 3, uniform at node 1,
    (5/6, 1/6) at node 6."""
    r = ob.solve(_tiny_2p(), 2, fp64=True, mode=ob.TERM_BRUTE)
    want = {0: (1 / 6, 5 / 6), 1: (0.5, 0.5), 3: (1 / 6, 5 / 6), 6: (5 / 6, 1 / 6)}
    for n, (a0, a1) in want.items():
        assert r["avg"][n, 0, 0] == pytest.approx(a0, abs=1e-6)
        assert r["avg"][n, 1, 0] == pytest.approx(a1, abs=1e-6)


# ---------------------------------------------------------------- hand-derived terminal values
def _three_player_tree(stack=50):
    # 25 nodes for stack 50 (pot 100, size 0.5): node 10 = showdown after
    # bet-call-call (pot 250, bets 50 each); node 6 = fold won by P2 with
    # P0 and P1 folded (pot 150, bets (0, 0, 50)).
    return ob.build_tree(_cards("Kc Kd 7h 7s 2c"), 100, stack, [0.5], 3, max_raises=2)


def _tv(td, board, hands, reach, node, trav, mode):
    prob = ob.Problem(td, board, [np.array(h, np.int32).reshape(-1, 2) for h in hands],
                      [np.ones(len(h), np.float32) for h in hands])
    return ob.terminal_values(prob, node, trav, [np.array(r, np.float64) for r in reach], mode)


@pytest.mark.parametrize("mode", [ob.TERM_BRUTE, ob.TERM_FAST])
def test_showdown_three_way_and_two_way_ties(mode):
    """Board Kc Kd 7h 7s 2c. P0 holds Ac3h (kings and sevens, ace kicker); P1
    holds As4h (ties, reach 0.3) or QhJh (loses, 0.7); P2 holds Ah5c (ties,
    0.2) or 3d4d (loses, 0.5). At node 10 (pot 250, everyone in for 50),
    P0's value is, by tuple:
      3-way tie   (250/3 - 50) * 0.06 = 2
      2-way ties  (250/2 - 50) * (0.15 + 0.14) = 21.75
      outright    (250 - 50) * 0.35 = 70
    total 93.75 (street_solve.cu:972-990). Catches pot/2 for every tie
    (96.25) and ties paid without subtracting the bet."""
    td = _three_player_tree()
    nd = td.nodes[10]
    assert nd.type == 2 and nd.pot == 250 and list(nd.bets)[:3] == [50, 50, 50]
    board = _cards("Kc Kd 7h 7s 2c")
    h0, a1, b1, a2, b2 = (_cards(x) for x in ("Ac 3h", "As 4h", "Qh Jh", "Ah 5c", "3d 4d"))
    s = lambda h: ob.eval7(board + h)
    assert s(h0) == s(a1) == s(a2) and s(b1) < s(h0) and s(b2) < s(h0)
    v = _tv(td, board, [[h0], [a1, b1], [a2, b2]], [[1.0], [0.3, 0.7], [0.2, 0.5]], 10, 0, mode)
    assert v[0] == pytest.approx(93.75, abs=1e-9)


def test_two_player_tie_pays_half_pot_minus_bet():
    """Two players at node 8 of the tiny river (pot 200, bets 50 each): a tie
    pays 200/2 - 50 = 50, a win 150, so opponent reach 0.4 tying and 0.6
    losing gives 0.4 * 50 + 0.6 * 150 = 110 (street_solve.cu:886-903)."""
    board = _cards(TINY_BOARD)
    td = ob.build_tree(board, 100, 50, [0.5], 2, max_raises=2)
    h0, tie, lose = _cards("6h 6s"), _cards("6c 6d"), _cards("Ac Ad")
    assert ob.eval7(board + h0) == ob.eval7(board + tie) > ob.eval7(board + lose)
    for mode in (ob.TERM_BRUTE, ob.TERM_FAST):
        v = _tv(td, board, [[h0], [tie, lose]], [[1.0], [0.4, 0.6]], 8, 0, mode)
        assert v[0] == pytest.approx(110.0, abs=1e-9)


@pytest.mark.parametrize("mode", [ob.TERM_BRUTE, ob.TERM_FAST])
def test_folded_traverser_at_showdown_gets_zero(mode):
    """A player who folded earlier gets 0 at a later showdown even though
    chips it put in are in the pot (street_solve.cu:857): the value is not
    -bet. The node is found in a deeper 3-player tree (stack 1000) as a
    showdown where an inactive player has a positive bet."""
    td = _three_player_tree(stack=1000)
    board = _cards("Kc Kd 7h 7s 2c")
    found = None
    for n in range(td.num_nodes):
        nd = td.nodes[n]
        if nd.type == 2:
            for p in range(3):
                if not nd.active_players[p] and nd.bets[p] > 0:
                    found = (n, p)
                    break
        if found:
            break
    assert found is not None
    n, p = found
    hands = [[_cards("Ac 3h")], [_cards("Qh Jh")], [_cards("3d 4d")]]
    v = _tv(td, board, hands, [[1.0], [1.0], [1.0]], n, p, mode)
    assert v[0] == 0.0 and td.nodes[n].bets[p] > 0


@pytest.mark.parametrize("mode", [ob.TERM_BRUTE, ob.TERM_FAST])
def test_fold_counts_folded_players_reach(mode):
    """Node 6 (pot 150, P2 wins with P0 and P1 folded): traverser P2's value is
    (150 - 50) times the joint reach of ALL non-traversers, folded ones
    included, over card-disjoint pairs (street_solve.cu:735-743, 764-825).
    P0 holds x (0.4) or y (0.6); P1 holds u (0.5) or v (0.25), v sharing a
    card with x: mass = 0.4*0.5 + 0.6*(0.5 + 0.25) = 0.65, value 65.
    Catches 'active opponents only' (100) and a missing opponent-opponent
    conflict check (75)."""
    td = _three_player_tree()
    nd = td.nodes[6]
    assert nd.type == 1 and nd.player == 2 and nd.pot == 150 and nd.bets[2] == 50
    assert not nd.active_players[0] and not nd.active_players[1]
    board = _cards("Kc Kd 7h 7s 2c")
    x, y, u, v = _cards("Ah 9d"), _cards("Qs Jd"), _cards("Tc 8c"), _cards("9d 6s")
    h2 = _cards("4c 5c")
    val = _tv(td, board, [[x, y], [u, v], [h2]], [[0.4, 0.6], [0.5, 0.25], [1.0]], 6, 2, mode)
    assert val[0] == pytest.approx(65.0, abs=1e-9)


def test_leaf_value_with_no_active_opponent_has_mass_one():
    """A leaf where no opponent of the traverser is active takes mass 1
    (street_solve.cu:1056-1058): value = the traverser's leaf value. The
    builder never makes such a leaf, so a two-player turn leaf is edited to
    leave only the traverser active."""
    board = _cards("2c 3d 4h 5s")
    td = ob.build_tree(board, 100, 50, [0.5], 2, use_cont=0, max_raises=2)
    leaf = next(n for n in range(td.num_nodes) if td.nodes[n].type == 4)
    td.nodes[leaf].active_players[1] = 0
    hands = [np.array([_cards("Ac Ad")], np.int32), np.array([_cards("Kc Kd")], np.int32)]
    lv = np.zeros((td.num_leaves, 2, 1), np.float32)
    lv[td.nodes[leaf].leaf_idx, 0, 0] = 37.5
    prob = ob.Problem(td, board, hands, [np.ones(1, np.float32)] * 2, leaf_values=lv)
    for mode in (ob.TERM_BRUTE, ob.TERM_FAST):
        v = ob.terminal_values(prob, leaf, 0, [np.array([1.0]), np.array([0.25])], mode)
        assert v[0] == pytest.approx(37.5, abs=1e-9)
    td.nodes[leaf].active_players[1] = 1
    for mode in (ob.TERM_BRUTE, ob.TERM_FAST):
        v = ob.terminal_values(prob, leaf, 0, [np.array([1.0]), np.array([0.25])], mode)
        assert v[0] == pytest.approx(37.5 * 0.25, abs=1e-9)
