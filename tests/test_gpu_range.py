"""GPU range narrowing (ss_range_update) against the oracle, bit for bit:
full 1326-combo and board-filtered ranges with drawn probabilities, hands
missing from the probabilities, products under the floor, zero weights, an
all-zero range, an empty range, and a chain of updates as between streets."""

import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import range_oracle as ro  # noqa: E402
import range_narrowing as rn  # noqa: E402

pytestmark = pytest.mark.gpu


def run(hands, probs):
    n = rn.RangeNarrower()
    n.set_initial_range("v", hands)
    n.update("v", "bet", probs)
    return n._ranges["v"]


@pytest.mark.parametrize("seed,board,missing", [(1, None, 0.0), (2, [3, 17, 22, 40, 51], 0.1), (3, None, 0.5)])
def test_update_equals_oracle(lib, seed, board, missing):
    rng = np.random.default_rng(seed)
    hands = [(a, b, float(w)) for (a, b, _), w in zip(rn.generate_all_hands(board),
                                                       rng.uniform(0, 1, 1326))]
    hands[5] = (hands[5][0], hands[5][1], 0.0)
    probs = {}
    for a, b, _ in hands:
        if rng.uniform() >= missing:
            probs[(a, b)] = float(rng.choice([0.0, 0.0005, rng.uniform(0, 1)]))
    print(f"range update seed={seed} hands={len(hands)} probs={len(probs)}", flush=True)
    got = run(hands, probs)
    assert got == ro.range_update(hands, probs)


def test_update_edges_and_chain(lib):
    assert run([], {}) == []
    assert run([(0, 1, 0.0), (2, 3, 0.0)], {(0, 1): 0.5}) == ro.range_update([(0, 1, 0.0), (2, 3, 0.0)], {(0, 1): 0.5})
    rng = np.random.default_rng(9)
    n = rn.RangeNarrower()
    n.set_uniform_range("v")
    ref = list(n._ranges["v"])
    for street in range(3):
        probs = {(a, b): float(rng.uniform()) for a, b, _ in ref}
        n.update("v", f"a{street}", probs)
        ref = ro.range_update(ref, probs)
        assert n._ranges["v"] == ref
    assert n.get_action_log("v") == ["a0", "a1", "a2"]


@pytest.mark.parametrize("npl,seed", [(2, 31), (3, 32)])
def test_update_from_resident_solution_equals_host_path(lib, npl, seed):
    """RangeNarrower.update_from_solution reads P(action | hand) on the GPU from
    a resident solve (StreetSolverGPU(..., resident=True)) and equals the host
    path RangeNarrower.update(..., solver.get_avg_strategy_probs(...))
    (python/street_solver_gpu.py:550-553, range_narrowing.py:76-122) bit for
    bit: for every action label of the player's first decision node, for a
    label the node lacks (every probability 0), with the narrowed range
    holding all 1326 combos (board-blocked ones and combos outside the solve's
    range take the missing-hand probability) and a hand stored high card
    first (not found by key, as in the reference), and for a chain of two
    updates."""
    from street_solver_gpu import StreetSolverGPU, _build_labels_at_node
    import workloads as wl
    rng = np.random.default_rng(seed)
    board = [int(c) for c in rng.choice(52, 5, replace=False)]
    combos = wl.all_combos(board)
    ranges = []
    for _ in range(npl):
        pick = np.sort(rng.choice(len(combos), 300, replace=False))
        ranges.append([(int(combos[k][0]), int(combos[k][1]), float(rng.uniform(0.05, 1.0))) for k in pick])
    s = StreetSolverGPU(board=board, player_ranges=ranges, pot_bb=20.0, stack_bb=80.0, bet_sizes=[0.5, 1.0],
                        max_raises=2, resident=True)
    s.solve(iterations=25)
    start = [(a, b, float(w)) for (a, b, _), w in zip(rn.generate_all_hands(None), rng.uniform(0, 1, 1326))]
    a0, b0, w0 = start[10]
    start[10] = (b0, a0, w0)                       # stored high card first
    start[11] = (start[11][0], start[11][1], 0.0)  # a zero weight stays zero
    for player in range(npl):
        node = next(int(n) for n in s._sol.node_ids if s._view.actor[int(n)] == player)
        labels = _build_labels_at_node(s._tree, node)
        assert len(set(labels)) == len(labels)
        for label in labels + ["Raise 999%"]:
            print(f"resident narrowing np={npl} player={player} label={label}", flush=True)
            host, dev = rn.RangeNarrower(), rn.RangeNarrower()
            host.set_initial_range("v", start)
            dev.set_initial_range("v", start)
            for _ in range(2):
                host.update("v", label, s.get_avg_strategy_probs(player, label))
                dev.update_from_solution("v", label, s, player)
                assert host._ranges["v"] == dev._ranges["v"], (player, label)
            assert host.get_action_log("v") == dev.get_action_log("v") == [label, label]
    s.free()
