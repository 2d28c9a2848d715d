"""The session layer end to end on the GPU: a two-player hand from flop to
river through SessionStore.handle (create, deal, action, solve, get, delete)
with every solve on the shared served loop. Checked against independent
paths: narrowing equals RangeNarrower.update with the probabilities of a
plain-form solve of the same street (StreetSolverGPU.get_avg_strategy_probs;
the served result equals the plain one bit for bit), actions taken on the
street are frozen in the next solve (one-hot at the frozen node), and the
freeze resets when the street changes."""

import numpy as np
import pytest

import range_narrowing as rn
import session as ss
from street_solver_gpu import StreetSolverGPU

pytestmark = pytest.mark.gpu

RANGES = {"BTN": {"AA": 1.0, "KK": 1.0, "QQ": 0.9, "JJ": 0.8, "AKs": 1.0, "AQs": 0.7, "KQs": 0.6, "T9s": 0.5,
                  "AKo": 0.8, "76s": 0.4, "55": 0.5},
          "BB": {"TT": 1.0, "99": 1.0, "88": 0.8, "AJs": 0.9, "KJs": 0.8, "QJs": 0.7, "JTs": 0.9, "A5s": 0.6,
                 "AQo": 0.7, "KQo": 0.6, "65s": 0.5, "22": 0.4}}


def _plain_probs(s_before, store, sess, position, label):
    """P(label | hand) from an independent plain-form solve of the river the
    session last solved (same tree, ranges and iterations)."""
    players = sess.players()
    board = [ss._cards.card_to_int(c) for c in sess.board]
    blocked = set(board)
    ranges = [[h for h in s_before.get_weighted_hands(p) if h[0] not in blocked and h[1] not in blocked]
              for p in players]
    stack = min(sess.stacks_bb[p] for p in players)
    sol = StreetSolverGPU(board=board, player_ranges=ranges, pot_bb=sess.pot_bb, stack_bb=stack,
                          bet_sizes=sess.bet_sizes, max_raises=sess.max_raises, extended=True)
    sol.solve(iterations=sess.iterations)
    probs = sol.get_avg_strategy_probs(players.index(position), label)
    sol.free()
    return probs


def test_session_flop_to_river():
    store = ss.SessionStore()
    try:
        code, r = store.handle("POST", "/api/session", {"positions": ["BB", "BTN"], "ranges": RANGES,
                                                         "pot_bb": 6.5, "stacks_bb": 97.5, "bet_sizes": [0.5, 1.0],
                                                         "iterations": 60})
        assert code == 201, r
        sid = r["session_id"]
        sess = store.sessions[sid]
        print("session: deal river board at once", flush=True)
        code, r = store.handle("POST", f"/api/session/{sid}/deal", {"cards": ["Qs", "8d", "2c", "7h", "3s"]})
        assert code == 200, r
        sol = r["solve"]
        assert r["state"]["street"] == "river" and sol["iterations"] == 60 and sol["exploitability_pct_pot"] >= 0
        assert sol["to_act"]["position"] == "BB" and set(sol["strategies"]) == {"BB", "BTN"}
        for strat in sol["strategies"]["BB"].values():
            assert abs(sum(strat.values()) - 1.0) < 2e-3
        labels = sorted({k for st in sol["strategies"]["BB"].values() for k in st})
        assert "Check" in labels
        # BB checks: narrowing with the solve's strategy at BB's first node
        before = sess.narrower.copy()
        want = _plain_probs(before, store, sess, "BB", "Check")
        code, r = store.handle("POST", f"/api/session/{sid}/action", {"position": "BB", "action": "Check"})
        assert code == 200 and r["narrowed"], r
        ref = before.copy()
        ref.update("BB", "Check", want)
        assert sess.narrower._ranges["BB"] == ref._ranges["BB"]
        # re-solve with BB's check frozen: BB's root strategy is all check
        code, r = store.handle("POST", f"/api/session/{sid}/solve", {})
        assert code == 200, r
        assert r["solve"]["to_act"]["position"] == "BTN"
        root = r["solve"]["strategies"]["BB"]
        assert all(st == {"Check": 1.0} for st in root.values()), list(root.items())[:3]
        bet = next(k for k in sorted({k for st in r["solve"]["to_act"]["strategy"].values() for k in st})
                   if k.startswith("Bet"))
        code, r = store.handle("POST", f"/api/session/{sid}/action", {"position": "BTN", "action": bet})
        assert code == 200 and r["state"]["pot_bb"] > 6.5
        code, r = store.handle("POST", f"/api/session/{sid}/action", {"position": "BB", "action": "Fold"})
        assert code == 200 and r["state"]["active"] == ["BTN"]
        assert store.handle("POST", f"/api/session/{sid}/deal", {"cards": ["Ah"]})[0] == 400   # board full
        code, r = store.handle("GET", f"/api/session/{sid}")
        assert code == 200 and [h[2] for h in r["state"]["history"]] == ["Check", bet, "Fold"]
        assert store.handle("DELETE", f"/api/session/{sid}")[0] == 200
    finally:
        store.close()


def test_session_turn_street_with_leaf_values_and_street_reset():
    """Flop -> turn: the turn solve carries the leaf producers' values (two
    players: the producer's rows match the tree's expanded leaves), and the
    actions of the flop are not frozen on the turn."""
    store = ss.SessionStore()
    try:
        code, r = store.handle("POST", "/api/session", {"positions": ["BB", "BTN"], "ranges": RANGES,
                                                         "pot_bb": 10.0, "stacks_bb": 60.0, "bet_sizes": [0.75],
                                                         "iterations": 40})
        sid = r["session_id"]
        print("session: deal turn", flush=True)
        code, r = store.handle("POST", f"/api/session/{sid}/deal", {"cards": ["Qs", "8d", "2c", "7h"]})
        assert code == 200, r
        assert r["state"]["street"] == "turn" and r["solve"]["leaf_values"] == "leaf_values.py"
        code, r = store.handle("POST", f"/api/session/{sid}/action", {"position": "BB", "action": "Check"})
        code, r = store.handle("POST", f"/api/session/{sid}/action", {"position": "BTN", "action": "Check"})
        assert r["state"]["actions_this_street"] == [["BB", "Check"], ["BTN", "Check"]]
        print("session: deal river", flush=True)
        code, r = store.handle("POST", f"/api/session/{sid}/deal", {"cards": ["3s"]})
        assert code == 200 and r["state"]["street"] == "river" and r["state"]["actions_this_street"] == []
        root = r["solve"]["strategies"]["BB"]
        assert any(len(st) > 1 or "Check" not in st for st in root.values())   # nothing frozen on the river
    finally:
        store.close()
