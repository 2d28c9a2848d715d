"""GPU parity of the flop-through-river solve (include/flop_solve.h) against
the flop-solve oracle (oracle/fs_oracle.c), through the C ABI.

The device kernels perform the oracle's single-precision operations in the
same order, with the same fused multiply-adds, so the outputs are compared
bit for bit: a one-ulp difference could otherwise flip a near-zero regret
and diverge from there. Ranks come from different evaluators (only their
order is used, and both are pinned to the same hand ranking).
"""

import ctypes

import numpy as np
import pytest

import flop_solve
import oracle_bind as ob
import ss_abi
from test_flop_oracle import chance_tree, half_street_game

pytestmark = pytest.mark.gpu

QsAs2d = [43, 51, 1]


def ranges(board, n0, n1, seed):
    combos = [(c0, c1) for c0 in range(52) for c1 in range(c0 + 1, 52) if not {c0, c1} & set(board)]
    rng = np.random.default_rng(seed)
    h0 = [combos[i] for i in rng.choice(len(combos), n0, replace=False)]
    h1 = [combos[i] for i in rng.choice(len(combos), n1, replace=False)]
    return h0, h1, rng.uniform(0.1, 1.0, n0).astype(np.float32), rng.uniform(0.1, 1.0, n1).astype(np.float32)


def gpu_vs_oracle(tree, iters):
    print(f"flop parity: {tree.td.num_nodes} nodes, hands {tree.td.num_hands[0]}/{tree.td.num_hands[1]}, "
          f"{iters} iterations", flush=True)
    got = flop_solve.solve(tree, iters, extract_all=True)
    want = ob.fs_solve(tree.td, iters)
    mh = max(tree.td.num_hands[0], tree.td.num_hands[1])
    assert got["max_hands"] == mh
    for k, g, w in (("root_strategy", got["root_strategy"], want["root_strategy"]),
                    ("avg", got["all_avg_strategies"], want["avg"]), ("cfv", got["all_cfv"], want["cfv"])):
        if not np.array_equal(g, w):
            d = np.abs(g - w)
            i = np.unravel_index(d.argmax(), d.shape)
            raise AssertionError(f"{k}: {np.count_nonzero(d)} of {d.size} differ, max {d.max()} at {i}: "
                                 f"gpu {g[i]} oracle {w[i]}")
    return got, want


def test_flop_hand_built_tree_values():
    t = chance_tree()
    got, _ = gpu_vs_oracle(t, 1)
    cfv = got["all_cfv"]
    assert cfv[:3, 0].tolist() == [-10.0, 0.0, -10.0]
    assert cfv[9].tolist() == [-20.0, 0.0]
    assert cfv[10:15, 0].tolist() == [-100.0, 100.0, 0.0, 0.0, -100.0]
    assert got["root_strategy"][:, 0].tolist() == [0.0, 1.0]


def test_flop_half_street_equilibrium():
    t = half_street_game(100, 50)
    got, _ = gpu_vs_oracle(t, 3000)
    avg = got["all_avg_strategies"]
    assert avg[0, 1, 0] > 0.995
    assert abs(avg[0, 1, 1] - 1 / 3) < 0.01 and abs(avg[2, 1, 0] - 2 / 3) < 0.01


@pytest.mark.parametrize("iters", [1, 4])
def test_flop_reference_tree_parity(iters):
    """The reference benchmark's tree (Qs As 2d, pot 650, stack 9750, 75%)."""
    tree = flop_solve.FlopTree(QsAs2d, 650, 9750, [0.75])
    tree.set_hands(*ranges(QsAs2d, 12, 12, seed=iters))
    gpu_vs_oracle(tree, iters)


def test_flop_unequal_ranges_short_stack():
    """Short stacks (all-ins on the flop, 60 chips behind) and hand counts on
    either side of a 64-hand row."""
    board = [45, 21, 0]   # Kh 7d 2c
    tree = flop_solve.FlopTree(board, 100, 60, [0.5, 1.0])
    tree.set_hands(*ranges(board, 5, 70, seed=11))
    gpu_vs_oracle(tree, 2)
    tree.set_hands(*ranges(board, 70, 5, seed=12))
    gpu_vs_oracle(tree, 2)


@pytest.mark.parametrize("n0,n1", [(200, 48), (48, 200)])
def test_flop_max_hands(n0, n1):
    """200 hands (the ABI's limit) for one player on the smallest tree (a stack
    of 0 on the flop: check-check to the turn)."""
    board = [51, 47, 43]   # As Ks Qs
    tree = flop_solve.FlopTree(board, 100, 0, [1.0])
    tree.set_hands(*ranges(board, n0, n1, seed=5))
    gpu_vs_oracle(tree, 2)


def test_flop_turn_roots_and_plain_solve():
    tree = flop_solve.FlopTree(QsAs2d, 650, 9750, [0.75])
    tree.set_hands(*ranges(QsAs2d, 6, 9, seed=2))
    a = tree.arrays()
    nodes, kids = a["nodes"], a["children"]
    want_roots, want_cards = [], []
    for n in np.flatnonzero((nodes[:, 0] == flop_solve.FS_NODE_CHANCE) & (nodes[:, 12] == 3)):
        for k in kids[nodes[n, 3]:nodes[n, 3] + nodes[n, 2]]:
            if nodes[k, 0] == flop_solve.FS_NODE_DECISION:
                want_roots.append(k)
                want_cards.append(nodes[k, 10])
    full = flop_solve.solve(tree, 3, extract_all=True)
    assert full["turn_root_indices"].tolist() == want_roots
    assert full["turn_root_cards"].tolist() == want_cards
    assert len(want_roots) % 49 == 0 and len(want_roots) > 0
    plain = flop_solve.solve(tree, 3)
    assert np.array_equal(plain["root_strategy"], full["root_strategy"])
    assert plain["root_player"] == 0 and plain["root_num_actions"] == nodes[0, 2]


def test_flop_graph_and_stream_launches_agree(monkeypatch):
    tree = flop_solve.FlopTree([45, 21, 0], 100, 400, [0.5, 1.0])
    tree.set_hands(*ranges([45, 21, 0], 10, 10, seed=9))
    a = flop_solve.solve(tree, 3, extract_all=True)
    monkeypatch.setenv("SS_NO_GRAPH", "1")
    b = flop_solve.solve(tree, 3, extract_all=True)
    for k in ("root_strategy", "all_avg_strategies", "all_cfv"):
        assert np.array_equal(a[k], b[k]), k


def test_flop_zero_iterations():
    t = half_street_game()
    got = flop_solve.solve(t, 0, extract_all=True)
    assert not got["root_strategy"].any()             # the strategy buffer starts at zero
    assert np.all(got["all_avg_strategies"][0, :2, :2] == 0.5) and not got["all_cfv"].any()


def test_flop_bad_inputs_are_rejected():
    lib = flop_solve.bind(ss_abi.load_library())
    out = flop_solve.FSOutput()

    def rejected(td, iters=1):
        rc = lib.fs_solve_gpu(ctypes.byref(td), iters, ctypes.byref(out))
        assert rc != 0 and ss_abi.last_error(), rc

    t = half_street_game()
    t.td.num_hands[1] = 0
    rejected(t.td)
    t = half_street_game()
    t.td.num_hands[0] = 201
    rejected(t.td)
    rejected(half_street_game().td, iters=-1)
    t = half_street_game()
    t.nodes[2].num_children = 7       # more actions than FS_MAX_ACTIONS, past the child list
    rejected(t.td)
    t = half_street_game()
    t.children[1] = 1                 # node 1 listed twice: not a tree
    rejected(t.td)
    t = half_street_game()
    t.nodes[1].num_board = 4          # a showdown without a full board
    rejected(t.td)
    t = half_street_game()
    t.td.hands[0][0][1] = t.td.hands[0][0][0]   # a hand with one card twice
    rejected(t.td)
    assert lib.fs_solve_gpu(None, 1, ctypes.byref(out)) != 0
    # a valid call still works after the rejections
    assert flop_solve.solve(half_street_game(), 2)["root_num_actions"] == 2
