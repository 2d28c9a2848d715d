"""Bayesian range narrowing between streets, the per-hand update on the GPU.

Drop-in for the reference's python/range_narrowing.py: the same class and
functions. RangeNarrower.update (range_narrowing.py:76-122) runs its
per-hand product, floor and max-rescale in libstreet_solve.so
(ss_range_update, include/street_solve_range.h); the bookkeeping around it
stays here. P(action | hand) comes from the solver's average strategy at the
player's first decision node (street_solver_gpu.StreetSolverGPU.
get_avg_strategy_probs, as street_solver_gpu.py:550-553).

  RangeNarrower            range_narrowing.py:36-173
  generate_all_hands       range_narrowing.py:176-194
  make_blueprint_probs     range_narrowing.py:197-225
"""

import ctypes

import numpy as np

import cards as _cards
import ss_abi


class RangeNarrower:
    """Per-player weighted ranges [(card0, card1, weight)], narrowed by
    P(hand | action) ∝ P(action | hand) · P(hand) after each observed action."""

    WEIGHT_FLOOR = 0.005

    def __init__(self):
        self._ranges = {}
        self._action_log = {}

    def set_initial_range(self, player, hands):
        self._ranges[player] = list(hands)
        self._action_log[player] = []

    def set_uniform_range(self, player):
        """All 1326 combos at weight 1."""
        self._ranges[player] = [(c0, c1, 1.0) for c0 in range(52) for c1 in range(c0 + 1, 52)]
        self._action_log[player] = []

    def update(self, player, action, blueprint_probs):
        """w' = w · P(action | hand) (hands missing from blueprint_probs use
        1 / max(number of distinct probabilities, 2)); a hand with w > 0 keeps
        at least WEIGHT_FLOOR; then all weights are divided by the largest."""
        if player not in self._ranges:
            return
        hands = self._ranges[player]
        n_approx = max(len(set(blueprint_probs.values())), 2)
        n = len(hands)
        if n:
            w = np.array([h[2] for h in hands], dtype=np.float64)
            p = np.zeros(n, dtype=np.float64)
            has = np.zeros(n, dtype=np.uint8)
            for i, (c0, c1, _) in enumerate(hands):
                v = blueprint_probs.get((c0, c1))
                if v is not None:
                    p[i] = v
                    has[i] = 1
            lib = ss_abi.load_library()
            rc = lib.ss_range_update(n, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     has.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
                                     1.0 / n_approx, self.WEIGHT_FLOOR)
            if rc != 0:
                raise RuntimeError(f"ss_range_update failed ({rc}): {ss_abi.last_error(lib)}")
            self._ranges[player] = [(c0, c1, float(x)) for (c0, c1, _), x in zip(hands, w)]
        self._action_log.setdefault(player, []).append(action)

    def update_from_solution(self, player, action, solver, solver_player, action_label=None):
        """update(player, action, solver.get_avg_strategy_probs(solver_player,
        label)) with P(label | hand) read on the GPU from the solve's resident
        strategy sums (StreetSolverGPU(..., resident=True)) instead of a host
        dictionary: ss_batch_range_update normalises the strategy sums at the
        solver player's first decision node as the solver's output does, cuts
        them at 0.001, and applies this update; the weights equal the host
        path's bit for bit. label defaults to action."""
        return solver.narrow_range(self, player, solver_player, action if action_label is None else action_label)

    def _device_update(self, player, action, batch, node, action_index, solver_keys, threshold):
        if player not in self._ranges:
            return
        hands = self._ranges[player]
        n = len(hands)
        if n:
            w = np.array([h[2] for h in hands], dtype=np.float64)
            # the host path looks each hand up by its (c0, c1) key as stored
            sh = np.array([solver_keys.get((c0, c1), -1) for c0, c1, _ in hands], dtype=np.int32)
            lib = ss_abi.load_library()
            rc = lib.ss_batch_range_update(batch, 0, int(node), int(action_index), n,
                                           sh.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                           w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           self.WEIGHT_FLOOR, float(threshold))
            if rc != 0:
                raise RuntimeError(f"ss_batch_range_update failed ({rc}): {ss_abi.last_error(lib)}")
            self._ranges[player] = [(c0, c1, float(x)) for (c0, c1, _), x in zip(hands, w)]
        self._action_log.setdefault(player, []).append(action)

    def remove_folded_hands(self, player, blueprint_probs):
        """Drop hands whose P(fold | hand) is at least 0.95."""
        if player not in self._ranges:
            return
        self._ranges[player] = [(c0, c1, w) for c0, c1, w in self._ranges[player]
                                if blueprint_probs.get((c0, c1), 0) < 0.95]

    def get_weighted_hands(self, player):
        """The range sorted by weight, largest first."""
        return sorted(self._ranges.get(player, []), key=lambda x: -x[2])

    def get_hand_count(self, player):
        return sum(1 for _, _, w in self._ranges.get(player, []) if w > self.WEIGHT_FLOOR)

    def get_action_log(self, player):
        return self._action_log.get(player, [])

    def copy(self):
        new = RangeNarrower()
        new._ranges = {p: list(h) for p, h in self._ranges.items()}
        new._action_log = {p: list(l) for p, l in self._action_log.items()}
        return new


def generate_all_hands(board_cards=None):
    """All combos c0 < c1 avoiding the board, weight 1."""
    blocked = set(board_cards or [])
    return [(c0, c1, 1.0) for c0 in range(52) if c0 not in blocked
            for c1 in range(c0 + 1, 52) if c1 not in blocked]


def make_blueprint_probs(hand_strategies, action):
    """{(card0, card1): frequency of the first action whose name contains
    `action` (case-insensitive), else 0} from {"AhKs": {"actions": [{"action":
    name, "frequency": f}, ...]}}."""
    probs = {}
    for hand_str, strat in hand_strategies.items():
        c0, c1 = _cards.card_to_int(hand_str[:2]), _cards.card_to_int(hand_str[2:])
        p = 0.0
        for info in strat.get("actions", []):
            if action.lower() in info.get("action", "").lower():
                p = info.get("frequency", 0.0)
                break
        probs[(min(c0, c1), max(c0, c1))] = p
    return probs
