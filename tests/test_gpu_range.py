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
