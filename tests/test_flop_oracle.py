"""Pins of the flop-solve oracle (oracle/fs_oracle.c) on the CPU.

Each check follows from the game's rules, not from the oracle's code:
  - the half-street clairvoyance game's closed-form equilibrium (bluff
    B / (P + B), call P / (P + B), value bets always);
  - every value, regret and strategy sum after one iteration of a hand-built
    tree with chance nodes, computed by hand below from the hand ranks:
    chance averaging, card blocking at showdowns, the fold payoffs
    (starting pot / 2 plus the street bet) and the fold's unchecked traverser
    board conflict (flop_solve.cu:737-779), regret matching, the discount
    t / (t + 1) and the t-weighted strategy sums.
Single and double precision are not compared on large trees: regret matching
is discontinuous where a regret is near zero, so the two can pick different
strategies at such nodes after the first pass.
"""

import numpy as np
import pytest

import flop_solve
import oracle_bind as ob
from flop_solve import FS_NODE_CHANCE as CH, FS_NODE_DECISION as DE, FS_NODE_FOLD as FO, FS_NODE_SHOWDOWN as SD

# cards: rank * 4 + suit, suits c d h s
C2c, C3c, C4d, C7d, C9h, CTd, CJc, CJd, CJh, CKh, CKd, CKs, CQc, CAc = 0, 4, 9, 21, 30, 33, 36, 37, 38, 46, 45, 47, 40, 48
TURN = [C2c, C7d, C9h, CJc]
RIVER = TURN + [CKs]


def half_street_game(P=100, B=50):
    """P0 holds the nuts (Kh Kd, trips) or air (3c 4d, king high), P1 a bluff
    catcher (Jh Td, a pair of jacks) on 2c 7d 9h Jc Ks. P0 checks (showdown)
    or bets B; P1 folds or calls."""
    rows = [
        (DE, 0, [1, 2], P, (0, 0), RIVER),
        (SD, -1, [], P, (0, 0), RIVER),
        (DE, 1, [3, 4], P + B, (B, 0), RIVER),
        (FO, 0, [], P + B, (B, 0), RIVER),
        (SD, -1, [], P + 2 * B, (B, B), RIVER),
    ]
    return ob.FsCustomTree(rows, [(CKh, CKd), (C3c, C4d)], [(CJh, CTd)])


@pytest.mark.parametrize("fp64", [True, False])
def test_half_street_game_reaches_closed_form_equilibrium(fp64):
    P, B = 100, 50
    t = half_street_game(P, B)
    res = ob.fs_solve(t.td, 3000, fp64=fp64)
    avg = res["avg"]
    assert avg[0, 1, 0] > 0.995                                  # the nuts always bet
    assert abs(avg[0, 1, 1] - B / (P + B)) < 0.01                 # bluff B / (P + B) = 1/3
    assert abs(avg[2, 1, 0] - P / (P + B)) < 0.01                 # call P / (P + B) = 2/3
    assert np.allclose(avg[0, :2].sum(0), 1) and np.allclose(avg[2, :2, :1].sum(0), 1)


def chance_tree():
    """Root: P0 checks (a chance node deals five rivers to pot-100 showdowns)
    or bets 50; P1 folds (the fold node's board holds P0's Kd) or calls (a
    chance node deals the same rivers to pot-200 showdowns)."""
    rivers = [CKs, CJd, CTd, CKd, CAc]
    rows = [(DE, 0, [1, 2], 100, (0, 0), TURN),
            (CH, -1, [3, 4, 5, 6, 7], 100, (0, 0), TURN),
            (DE, 1, [8, 9], 150, (50, 0), TURN)]
    rows += [(SD, -1, [], 100, (0, 0), TURN + [r]) for r in rivers]
    rows += [(FO, 0, [], 150, (50, 0), TURN + [CKd]),
             (CH, -1, [10, 11, 12, 13, 14], 200, (50, 50), TURN)]
    rows += [(SD, -1, [], 200, (50, 50), TURN + [r]) for r in rivers]
    # P0: Kh Kd. P1: Jh Td (weight 0.75) and Kh Qc (0.5, shares Kh with P0)
    return ob.FsCustomTree(rows, [(CKh, CKd)], [(CJh, CTd), (CKh, CQc)], [1.0], [0.75, 0.5])


@pytest.mark.parametrize("fp64", [True, False])
def test_one_iteration_by_hand(fp64):
    t = chance_tree()
    res = ob.fs_solve(t.td, 1, fp64=fp64, want_state=True)
    # Traverser 0 (strategies uniform). P1's Kh Qc shares a card with P0 and
    # never counts. Rivers Ks / Jd / Td / Kd / Ac: P0 trips (wins), P1 trips
    # (loses), P1's Td on the board (0), P0's Kd on the board (0), pair of
    # kings against jacks (wins): per unit of P1 reach +1, -1, 0, 0, +1.
    #   check chance: 0.75 * 50 * (1 - 1 + 1) / 5 = 7.5
    #   fold: P1 reaches it with 0.75 / 2 = 0.375, P0 wins 100 / 2 + 0 = 50:
    #         18.75 (P0's own Kd on the fold node's board is not checked)
    #   call chance: 0.375 * 100 * 1 / 5 = 7.5; P1's node: 18.75 + 7.5 = 26.25
    #   root: 0.5 * 7.5 + 0.5 * 26.25 = 16.875;
    #   regrets (7.5 - 16.875) / 2, (26.25 - 16.875) / 2
    R, S = res["regrets"], res["strategy_sum"]
    assert R[0, :2, 0].tolist() == [-4.6875, 4.6875]
    assert S[0, :2, 0].tolist() == [0.5, 0.5]
    # Traverser 1: P0 now checks with 0 and bets with 1. Jh Td against Kh Kd
    # behind the call: -1, +1, 0 (Td blocked), 0 (P0 blocked), -1 at 100:
    # -20. The fold: P0's Kd is on the fold node's board, so P0 drops out: 0.
    # P1's node: 0.5 * 0 + 0.5 * -20 = -10; regrets (0 + 10) / 2, (-20 + 10) / 2.
    assert R[2, :2, 0].tolist() == [5.0, -5.0]
    assert R[2, :2, 1].tolist() == [0.0, 0.0]
    assert S[2, :2, :2].tolist() == [[0.5, 0.5], [0.5, 0.5]]
    cfv = res["cfv"]
    assert cfv[:, 0].tolist()[:3] == [-10.0, 0.0, -10.0]
    assert cfv[8].tolist() == [0.0, 0.0] and cfv[9].tolist() == [-20.0, 0.0]
    assert cfv[10:15, 0].tolist() == [-100.0, 100.0, 0.0, 0.0, -100.0]
    assert res["root_strategy"][:, 0].tolist() == [0.0, 1.0]
    assert res["avg"][0, :2, 0].tolist() == [0.5, 0.5]
