"""Depth-limit leaf values for turn and flop subgames, computed on the GPU.

Drop-in for the producer functions of the reference's python/leaf_values.py:
the same names, arguments, return shapes and errors, with the equity loops
and the 4^NP bias expansion running in libstreet_solve.so
(include/street_solve_leaf.h). Only the small host steps stay here: the
runout lists (numpy RandomState draws, as the reference makes them), the
leaf-info walk over the built tree, and the scalar bias helpers.

  biased_leaf_value            leaf_values.py:91-122
  classify_actions             leaf_values.py:125-149
  bias_strategy                leaf_values.py:152-182
  extract_leaf_info_from_tree  leaf_values.py:189-262
  compute_flop_leaf_equity     leaf_values.py:408-570
  compute_turn_leaf_values     leaf_values.py:694-816

The output array is laid out as the reference's: [len(leaf_infos) * 4^NP]
[NP][max_hands] float32, row leaf_idx * 4^NP + sum_i s_i * 4^i. That profile
order (player 0 least significant) is the reference producer's; the tree
builder's continuation chains put the first active player's choice most
significant, a mismatch of the reference that this module keeps so its
output equals the reference producer's.

compute_flop_leaf_values (leaf_values.py:265-405) needs a blueprint store of
per-action EVs and is not provided.
"""

import ctypes
import functools
from collections import namedtuple
from typing import List, Optional, Tuple

import numpy as np

import ss_abi

BIAS_FOLD_PROB = [0.0, 0.35, 0.0, 0.0]
BIAS_POT_FACTOR = [1.0, 0.7, 1.0, 1.4]
MAX_RUNOUT_SAMPLES = 12

LeafInfo = namedtuple("LeafInfo", ["leaf_idx", "pot", "bets"])


def biased_leaf_value(equity, leaf_pot, s_self, s_opp):
    """Value of one hand at one leaf for bias profiles (s_self, s_opp): with
    the opponent's fold probability it wins half the leaf pot, otherwise it
    takes (equity - 0.5) of the pot scaled by both profiles' pot factors."""
    f = BIAS_FOLD_PROB[s_opp]
    showdown_pot = leaf_pot * (BIAS_POT_FACTOR[s_self] * BIAS_POT_FACTOR[s_opp])
    return f * (0.5 * leaf_pot) + (1.0 - f) * ((equity - 0.5) * showdown_pot)


def classify_actions(num_actions: int, first_action_is_fold: bool):
    """Action index -> 'fold' | 'call' | 'raise' for a street-root action list
    ([Fold, Call, raises...] when facing a bet, else [Check, bets...])."""
    if first_action_is_fold:
        cats = {0: "fold", 1: "call"}
        cats.update({a: "raise" for a in range(2, num_actions)})
    else:
        cats = {0: "call"}
        cats.update({a: "raise" for a in range(1, num_actions)})
    return cats


def bias_strategy(strategy: np.ndarray, bias_type: int, action_categories: dict,
                  multiplier: float = 5.0) -> np.ndarray:
    """Multiply the probabilities of the biased category by `multiplier` and
    renormalise each row (uniform where a row sums to <= 1e-10); bias 0
    returns a copy."""
    if bias_type == 0:
        return strategy.copy()
    target = {1: "fold", 2: "call", 3: "raise"}[bias_type]
    out = strategy.copy()
    for a, cat in action_categories.items():
        if a < out.shape[1] and cat == target:
            out[:, a] *= multiplier
    rows = out.sum(axis=1, keepdims=True)
    return np.where(rows > 1e-10, out / np.maximum(rows, 1e-10), 1.0 / out.shape[1])


_DECISION, _LEAF = 0, 4


def extract_leaf_info_from_tree(tree_data, num_players=2) -> List[LeafInfo]:
    """One LeafInfo per player-0 decision node that heads a num_players-deep
    chain of 4-way decision nodes ending in leaf terminals: the original leaf
    index (its first terminal's leaf_idx // 4^num_players), the node's pot and
    its first num_players bets; sorted by leaf index. tree_data is an
    ss_abi.SSTreeData filled by ss_build_tree (or anything with the same
    nodes / children / num_nodes fields)."""
    nodes, children = tree_data.nodes, tree_data.children
    cont = 4 ** num_players

    def chain(i, depth):
        n = nodes[i]
        if n.type != _DECISION or n.num_children != 4:
            return False
        kids = [children[n.first_child + c] for c in range(4)]
        if depth == num_players - 1:
            return all(nodes[k].type == _LEAF for k in kids)
        return all(chain(k, depth + 1) for k in kids)

    out = []
    for i in range(tree_data.num_nodes):
        n = nodes[i]
        if n.type != _DECISION or n.player != 0 or n.num_children != 4 or not chain(i, 0):
            continue
        cur = i
        for _ in range(num_players):
            cur = children[nodes[cur].first_child]
        out.append(LeafInfo(leaf_idx=nodes[cur].leaf_idx // cont, pot=n.pot,
                            bets=tuple(n.bets[p] for p in range(num_players))))
    out.sort(key=lambda x: x.leaf_idx)
    return out


def _players(player_hands, oop_hands, ip_hands, fname):
    if player_hands is None:
        if oop_hands is None or ip_hands is None:
            raise ValueError(f"{fname} requires either player_hands or both oop_hands and ip_hands")
        player_hands = [oop_hands, ip_hands]
    n = len(player_hands)
    if n < 2 or n > 6:
        raise ValueError(f"{fname} supports 2-6 players, got {n}")
    return player_hands


@functools.lru_cache(maxsize=4096)
def _drawn(cards, seed):
    # The draw depends only on the ordered card list and the seed.
    return tuple(int(c) for c in sorted(np.random.RandomState(seed).choice(list(cards), MAX_RUNOUT_SAMPLES,
                                                                            replace=False)))


def _sample(cards, seed, large):
    if len(cards) > MAX_RUNOUT_SAMPLES and large:
        return list(_drawn(tuple(cards), seed))
    return list(cards)


def _equity(player_hands, runouts, max_hands, lib):
    """ss_leaf_equity: float64 [NP][max_hands] runout-averaged equity."""
    npl = len(player_hands)
    nh = [len(h) for h in player_hands]
    if max(nh) > max_hands:
        raise IndexError(f"a player has {max(nh)} hands, more than max_hands={max_hands}")
    hands = np.zeros((npl, max_hands, 2), dtype=np.int32)
    for p, hs in enumerate(player_hands):
        if hs:
            try:
                rows = np.asarray(hs, dtype=np.float64)[:, :2]
            except (ValueError, TypeError):
                rows = [(hand[0], hand[1]) for hand in hs]
            hands[p, :len(hs)] = rows
    nh_a = np.array(nh, dtype=np.int32)
    run = np.ascontiguousarray(np.array(runouts, dtype=np.int32).reshape(-1, 5))
    raw = np.zeros((npl, max_hands), dtype=np.float64)
    ip = ctypes.POINTER(ctypes.c_int)
    rc = lib.ss_leaf_equity(npl, hands.ctypes.data_as(ip), nh_a.ctypes.data_as(ip), max_hands,
                            run.ctypes.data_as(ip), len(run), raw.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if rc != 0:
        raise RuntimeError(f"ss_leaf_equity failed ({rc}): {ss_abi.last_error(lib)}")
    return raw, nh_a


def _leaf_values(player_hands, runouts, leaf_infos, max_hands, lib=None):
    lib = lib or ss_abi.load_library()
    npl = len(player_hands)
    ip = ctypes.POINTER(ctypes.c_int)
    raw, nh_a = _equity(player_hands, runouts, max_hands, lib)
    idx = np.array([li.leaf_idx for li in leaf_infos], dtype=np.int32)
    pot = np.array([li.pot for li in leaf_infos], dtype=np.int32)
    # Every element is written by the library (rows no leaf info names and
    # hands past a player's count as 0).
    out = np.empty((len(leaf_infos) * 4 ** npl, npl, max_hands), dtype=np.float32)
    rc = lib.ss_leaf_values(npl, raw.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nh_a.ctypes.data_as(ip),
                            max_hands, idx.ctypes.data_as(ip), pot.ctypes.data_as(ip), len(leaf_infos),
                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    if rc != 0:
        raise RuntimeError(f"ss_leaf_values failed ({rc}): {ss_abi.last_error(lib)}")
    return out


def turn_runouts(board_4, nh):
    """River runouts of a turn board: all 48, or 12 drawn by
    RandomState(42) when any range has more than 50 hands."""
    rivers = [c for c in range(52) if c not in set(board_4)]
    return [list(board_4) + [r] for r in _sample(rivers, 42, any(n > 50 for n in nh))]


def flop_runouts(flop_board, nh):
    """(turn, river) runouts of a flop: all, or 12 turns drawn by
    RandomState(42) and 12 rivers per turn by RandomState(43 + turn) when any
    range has more than 50 hands."""
    dead = set(flop_board)
    large = any(n > 50 for n in nh)
    out = []
    for t in _sample([c for c in range(52) if c not in dead], 42, large):
        rivers = [c for c in range(52) if c not in dead and c != t]
        out.extend(list(flop_board) + [t, r] for r in _sample(rivers, 43 + t, large))
    return out


def compute_turn_leaf_values(
    board_4: List[int],
    leaf_infos: List[LeafInfo],
    max_hands: int,
    starting_pot: int,
    player_hands: Optional[List[List[Tuple[int, int, float]]]] = None,
    oop_hands: Optional[List[Tuple[int, int, float]]] = None,
    ip_hands: Optional[List[Tuple[int, int, float]]] = None,
) -> np.ndarray:
    """Leaf values of a turn subgame (2-6 players): per-hand showdown equity
    over the river runouts against the other players' hands, then
    biased_leaf_value at each leaf's pot for all 4^NP bias profiles.
    Returns float32 [len(leaf_infos) * 4^NP, NP, max_hands]."""
    player_hands = _players(player_hands, oop_hands, ip_hands, "compute_turn_leaf_values")
    nh = [len(h) for h in player_hands]
    return _leaf_values(player_hands, turn_runouts(board_4, nh), leaf_infos, max_hands)


def compute_flop_leaf_equity(
    flop_board: List[int],
    leaf_infos: List[LeafInfo],
    max_hands: int,
    starting_pot: int,
    player_hands: Optional[List[List[Tuple[int, int, float]]]] = None,
    oop_hands: Optional[List[Tuple[int, int, float]]] = None,
    ip_hands: Optional[List[Tuple[int, int, float]]] = None,
) -> np.ndarray:
    """Leaf values of a flop subgame (2-6 players) from equity over the
    (turn, river) runouts; same layout as compute_turn_leaf_values."""
    player_hands = _players(player_hands, oop_hands, ip_hands, "compute_flop_leaf_equity")
    nh = [len(h) for h in player_hands]
    return _leaf_values(player_hands, flop_runouts(flop_board, nh), leaf_infos, max_hands)


def tree_leaf_infos(tree_data, num_players):
    """One LeafInfo per original depth-limit leaf of a continuation-expanded
    tree, numbered in node order (the leaf_idx the tree-order layout of
    ss_batch_leaf_values uses): each expanded leaf climbs its chain (one level
    per active player, street_solve.cu:400-455) to the head node, whose pot
    and first num_players bets the info carries."""
    nodes, children = tree_data.nodes, tree_data.children
    parent = {}
    for i in range(tree_data.num_nodes):
        n = nodes[i]
        if n.type == _DECISION:
            for a in range(n.num_children):
                parent[children[n.first_child + a]] = i
    heads = set()
    for i in range(tree_data.num_nodes):
        n = nodes[i]
        if n.type != _LEAF:
            continue
        cur = i
        for _ in range(sum(1 for q in range(num_players) if n.active_players[q])):
            cur = parent[cur]
        heads.add(cur)
    return [LeafInfo(leaf_idx=k, pot=nodes[h].pot, bets=tuple(nodes[h].bets[q] for q in range(num_players)))
            for k, h in enumerate(sorted(heads))]


def load_leaf_values_into_batch(batch, index, board, leaf_infos, player_hands, tree_order=False, lib=None):
    """The leaf values of subgame `index` of an ss_api.Batch computed on the
    GPU and written straight into its device leaf buffer
    (ss_batch_leaf_values): equity over the turn's river runouts (4-card
    board) or the flop's (turn, river) runouts, drawn as
    compute_turn_leaf_values / compute_flop_leaf_equity draw them, then the
    bias profiles. tree_order=False gives the solve the same leaf values as
    passing compute_*'s array (the reference's layout); True uses the
    tree-order profiles (tree_leaf_infos numbering)."""
    lib = lib or ss_abi.load_library()
    max_hands = max(len(h) for h in player_hands)
    nh = [len(h) for h in player_hands]
    runouts = turn_runouts(board, nh) if len(board) == 4 else flop_runouts(board, nh)
    raw, _ = _equity(player_hands, runouts, max_hands, lib)
    idx = np.array([li.leaf_idx for li in leaf_infos], dtype=np.int32)
    pot = np.array([li.pot for li in leaf_infos], dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    rc = lib.ss_batch_leaf_values(batch.h, int(index), raw.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  idx.ctypes.data_as(ip), pot.ctypes.data_as(ip), len(leaf_infos), int(tree_order))
    if rc != 0:
        raise RuntimeError(f"ss_batch_leaf_values failed ({rc}): {ss_abi.last_error(lib)}")

