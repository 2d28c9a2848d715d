"""CPU tests of the leaf-value producers (python/leaf_values.py counterpart).

The oracle (oracle/leaf_oracle.py) is pinned by the reference's own outputs
(tests/golden/leaf_values.npz and leaf_infos.json, written by
make_leaf_golden.py from the reference module) and by hand-derived answers:
a two-outer river, boards every hand ties on (two and three ways), opponent
tuples that always collide, hand-computed bias values and the round-half-even
opponent profile. The library's host-side pieces (the Python-compatible
draws, argument checks, runout lists, the leaf-info walk over a tree the
library built) are checked here without a GPU; the GPU parity tests are in
test_gpu_leaf.py.
"""

import ctypes
import gzip
import json
import os
import random
import sys

import numpy as np
import pytest

import ss_abi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import leaf_oracle as lo  # noqa: E402
import leaf_values as lv  # noqa: E402

GOLD = np.load(os.path.join(REPO, "tests", "golden", "leaf_values.npz"))
NAMES = [str(n) for n in GOLD["names"]]
RANKS, SUITS = "23456789TJQKA", "cdhs"


def card(s):
    return RANKS.index(s[0]) * 4 + SUITS.index(s[1])


def cards(s):
    return [card(c) for c in s.split()]


def golden_case(name):
    g = lambda k: GOLD[f"{name}__{k}"]
    nh = [int(x) for x in g("nh")]
    hc = g("hands")
    hands = [[(int(hc[p, h, 0]), int(hc[p, h, 1])) for h in range(nh[p])] for p in range(len(nh))]
    return (str(g("kind")), [int(c) for c in g("board")], hands, nh, int(g("max_hands")),
            [tuple(int(v) for v in r) for r in g("infos")], g("out"))


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_reference_leaf_values(name):
    kind, board, hands, nh, mh, infos, ref = golden_case(name)
    f = lo.turn_leaf_values if kind == "turn" else lo.flop_leaf_equity
    out = f(board, hands, nh, mh, infos)
    assert out.shape == ref.shape
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


def _trees():
    with gzip.open(os.path.join(REPO, "tests", "golden", "trees.json.gz"), "rt") as f:
        return {t["name"]: t for t in json.load(f)["trees"]}


def _extract_cases():
    with open(os.path.join(REPO, "tests", "golden", "leaf_infos.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _extract_cases(), ids=lambda c: f"{c['tree']}-np{c['num_players']}")
def test_oracle_extract_matches_reference(case):
    got = lo.extract_leaf_infos(_trees()[case["tree"]], case["num_players"])
    assert [[i, p, list(b)] for i, p, b in got] == case["leaf_infos"]


def test_oracle_known_answers():
    # Ah As vs Kd Kc on 2c 7d 9h Js: no straight or flush can come, KK wins on
    # the two remaining kings, the 4 rivers holding a hole card are skipped.
    raw = lo.runout_equity([[(card("Ah"), card("As"))], [(card("Kd"), card("Kc"))]], [1, 1],
                           lo.river_runouts(cards("2c 7d 9h Js"), [1, 1]), 1)
    assert raw[0, 0] == 42 / 44 and raw[1, 0] == 2 / 44
    # Identical off-suit holdings under four spades of Broadway: every river ties.
    b = cards("Ts Js Qs Ks")
    raw = lo.runout_equity([[(card("2c"), card("3c"))], [(card("2d"), card("3d"))]], [1, 1],
                           lo.river_runouts(b, [1, 1]), 1)
    assert raw[0, 0] == 0.5 and raw[1, 0] == 0.5
    raw = lo.runout_equity([[(card("2c"), card("3c"))], [(card("2d"), card("3d"))], [(card("2h"), card("3h"))]],
                           [1, 1, 1], lo.river_runouts(b, [1, 1, 1]), 1)
    assert np.allclose(raw, 1.0 / 3, rtol=0, atol=1e-14)  # mean of 44 rounded thirds
    # Opponents that always share a card: no tuple is ever counted.
    raw = lo.runout_equity([[(card("Ah"), card("As"))], [(card("Kd"), card("Kc"))], [(card("Kd"), card("Qc"))]],
                           [1, 1, 1], lo.river_runouts(cards("2c 7d 9h Js"), [1, 1, 1]), 1)
    assert np.all(raw == 0.0)


def test_bias_model_known_values():
    for f in (lo.biased_leaf_value, lv.biased_leaf_value):
        assert f(1.0, 100, 0, 1) == pytest.approx(40.25, abs=1e-12)   # 0.35*50 + 0.65*(0.5*70)
        assert f(0.25, 200, 3, 3) == pytest.approx(-98.0, abs=1e-12)  # -0.25 * 200 * 1.96
        assert f(0.5, 1000, 2, 0) == 0.0
        assert f(0.0, 300, 1, 2) == pytest.approx(-105.0, abs=1e-12)  # -0.5 * 300 * 0.7
    # Opponent profile for 3+ players: round half to even of the others' mean.
    assert lo.opp_proxy([0, 0, 1], 0) == 0     # 0.5 -> 0
    assert lo.opp_proxy([0, 1, 2], 0) == 2     # 1.5 -> 2
    assert lo.opp_proxy([0, 3, 2], 0) == 2     # 2.5 -> 2
    assert lo.opp_proxy([0, 1, 2], 1) == 1
    assert lo.opp_proxy([3, 3, 3, 3], 0) == 3


def test_bias_strategy_and_categories():
    cats = lv.classify_actions(5, True)
    assert cats == {0: "fold", 1: "call", 2: "raise", 3: "raise", 4: "raise"}
    assert lv.classify_actions(3, False) == {0: "call", 1: "raise", 2: "raise"}
    s = np.array([[0.5, 0.25, 0.25], [0.0, 0.0, 0.0]])
    b = lv.bias_strategy(s, 3, lv.classify_actions(3, False))
    assert np.allclose(b[0], [0.5 / 3.0, 1.25 / 3.0, 1.25 / 3.0])
    assert np.allclose(b[1], [1 / 3.0] * 3)
    assert np.array_equal(lv.bias_strategy(s, 0, cats), s)


@pytest.mark.parametrize("lens", [[1128], [1081, 1050], [7, 1, 300, 13, 2], [2] * 5])
def test_library_draws_equal_python_random(lib, lens):
    n, cnt = len(lens), 2000
    out = (ctypes.c_int * (n * cnt))()
    assert lib.ssdbg_py_choice(42, (ctypes.c_int * n)(*lens), n, cnt, out) == 0
    rng = random.Random(42)
    assert list(out) == [rng.choice(range(lens[k])) for _ in range(cnt) for k in range(n)]


def test_runout_lists_equal_oracle():
    for nh in ([10, 10], [51, 3]):
        assert lv.turn_runouts(cards("Qs As 2d 7h"), nh) == lo.river_runouts(cards("Qs As 2d 7h"), nh)
        assert lv.flop_runouts(cards("Qs As 2d"), nh) == lo.flop_runouts(cards("Qs As 2d"), nh)
    assert len(lv.turn_runouts(cards("Qs As 2d 7h"), [51, 3])) == 12
    assert len(lv.flop_runouts(cards("Qs As 2d"), [10, 10])) == 49 * 48


def test_leaf_arguments_rejected_without_gpu(lib):
    hands = (ctypes.c_int * 4)(0, 1, 2, 3)
    run = (ctypes.c_int * 5)(10, 11, 12, 13, 14)
    raw = (ctypes.c_double * 2)()
    for npl, nh, mh, rr in [(7, (1, 1), 1, run), (2, (1, 1), 0, run), (2, (2, 1), 1, run),
                            (2, (1, 1), 1, (ctypes.c_int * 5)(10, 11, 12, 13, 13))]:
        nha = (ctypes.c_int * 6)(*nh)
        assert lib.ss_leaf_equity(npl, hands, nha, mh, rr, 1, raw) == -2
    # (runout, player) pairs are a grid dimension: more than 65535 is refused up front
    assert lib.ss_leaf_equity(2, hands, (ctypes.c_int * 2)(1, 1), 1, run, 32768, raw) == -2
    assert "65535" in ss_abi.last_error(lib)
    bad = (ctypes.c_int * 4)(0, 0, 2, 3)
    assert lib.ss_leaf_equity(2, bad, (ctypes.c_int * 2)(1, 1), 1, run, 1, raw) == -2
    assert "cards" in ss_abi.last_error(lib)
    out = (ctypes.c_float * 64)()
    assert lib.ss_leaf_values(2, raw, (ctypes.c_int * 2)(1, 1), 1, (ctypes.c_int * 1)(-1),
                              (ctypes.c_int * 1)(100), 1, out) == -2
    with pytest.raises(ValueError):
        lv.compute_turn_leaf_values(cards("Qs As 2d 7h"), [], 10, 100)
    with pytest.raises(ValueError):
        lv.compute_turn_leaf_values(cards("Qs As 2d 7h"), [], 10, 100, player_hands=[[(0, 1, 1.0)]] * 7)


def test_extract_on_library_tree_matches_reference(lib):
    trees = _trees()
    for case in _extract_cases():
        t = trees[case["tree"]]
        td = ss_abi.SSTreeData()
        b, s, o = t["board"], t["bet_sizes"], t["acting_order"]
        rc = lib.ss_build_tree((ctypes.c_int * len(b))(*b), len(b), t["pot"], t["stack"],
                               (ctypes.c_float * len(s))(*s), len(s), t["num_players"],
                               (ctypes.c_int * len(o))(*o), t["is_river"], t["use_cont_strats"], ctypes.byref(td))
        assert rc == 0
        try:
            got = lv.extract_leaf_info_from_tree(td, num_players=case["num_players"])
            assert [[li.leaf_idx, li.pot, list(li.bets)] for li in got] == case["leaf_infos"]
        finally:
            lib.ss_free_tree(ctypes.byref(td))
