"""CPU tests of range narrowing (python/range_narrowing.py counterpart).

The oracle (oracle/range_oracle.py) is pinned by hand-derived updates: the
product, the floor, the max-rescale, the missing-hand default of
1 / max(#distinct probabilities, 2) and the all-zero range. The mirror's host
pieces (combo generation, blueprint parsing, fold removal, copies) are
checked without a GPU; the GPU update is compared with the oracle in
test_gpu_range.py."""

import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import range_oracle as ro  # noqa: E402
import range_narrowing as rn  # noqa: E402


def test_oracle_known_updates():
    r = [(0, 1, 1.0), (2, 3, 0.5), (4, 5, 0.2), (6, 7, 0.0)]
    # p = 0.5, 0.1, missing -> 1/3 (distinct values {0.5, 0.1, 0.0}), 0.0
    probs = {(0, 1): 0.5, (2, 3): 0.1, (6, 7): 0.0}
    out = ro.range_update(r, probs)
    # products 0.5, 0.05, 0.2/3, 0 (weight 0 gets no floor); max 0.5
    assert out[0] == (0, 1, 1.0)
    assert out[1][2] == pytest.approx(0.1, abs=1e-15)
    assert out[2][2] == pytest.approx((0.2 / 3) / 0.5, abs=1e-15)
    assert out[3][2] == 0.0
    # the floor: a tiny product is lifted to 0.005 before the rescale
    out = ro.range_update([(0, 1, 1.0), (2, 3, 1.0)], {(0, 1): 0.8, (2, 3): 0.001})
    assert out[1][2] == pytest.approx(0.005 / 0.8, abs=1e-15)
    # every hand at zero: no rescale
    assert ro.range_update([(0, 1, 0.0)], {(0, 1): 0.7}) == [(0, 1, 0.0)]
    # one distinct value -> missing hands use 1/2
    out = ro.range_update([(0, 1, 1.0), (2, 3, 1.0)], {(0, 1): 0.25})
    assert out == [(0, 1, 0.5), (2, 3, 1.0)]


def test_mirror_host_pieces():
    assert len(rn.generate_all_hands()) == 1326
    assert len(rn.generate_all_hands([0, 1, 2, 3, 4])) == 1081
    bp = {"AhKs": {"actions": [{"action": "Check", "frequency": 0.25}, {"action": "Bet 75%", "frequency": 0.75}]},
          "2c2d": {"actions": [{"action": "Check", "frequency": 1.0}]}}
    probs = rn.make_blueprint_probs(bp, "bet")
    assert probs == {(47, 50): 0.75, (0, 1): 0.0}  # Ah = 50, Ks = 47; "bet" matches "Bet 75%"
    n = rn.RangeNarrower()
    n.set_initial_range("villain", [(0, 1, 1.0), (2, 3, 0.5)])
    n.remove_folded_hands("villain", {(0, 1): 0.96})
    assert n.get_weighted_hands("villain") == [(2, 3, 0.5)]
    c = n.copy()
    c.set_uniform_range("villain")
    assert n.get_hand_count("villain") == 1 and c.get_hand_count("villain") == 1326
    n.update("nobody", "bet", {})  # unknown player: no-op, as the reference
    assert n.get_action_log("nobody") == []
