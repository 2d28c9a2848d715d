"""CPU restatement of the reference's depth-limit leaf-value producers
(python/leaf_values.py). TEST INFRASTRUCTURE ONLY: tests/, the smoke check
and bench.py's CPU-baseline leg use it as the checker; the library never
imports it.

Functions and the reference lines they restate:
  biased_leaf_value     leaf_values.py:84-122 (bias tables and formula)
  river_runouts         leaf_values.py:748-765 (turn leaves: river cards,
                        12 drawn by numpy RandomState(42) when any range has
                        more than 50 hands)
  flop_runouts          leaf_values.py:474-497 (flop leaves: 12 turns by
                        RandomState(42), 12 rivers per turn by
                        RandomState(43 + turn))
  runout_equity         leaf_values.py:499-523 / :767-793 with
                        _accumulate_nplayer_equity :573-691
  values_from_equity    leaf_values.py:525-567 / :795-816
  turn_leaf_values      compute_turn_leaf_values :694-816
  flop_leaf_equity      compute_flop_leaf_equity :408-570
  extract_leaf_infos    extract_leaf_info_from_tree :189-262 (over the
                        plain-list tree snapshot of ss_abi.tree_to_dict)

Hand ranks come from the oracle's own C evaluator (liboracle.so or_eval7);
only the order of ranks matters here, and it is the poker order the
reference's _eval7_py tuple ranks also give.

The per-hand sum of 1 and 1/ties over opponent tuples is a sequential float64
sum in tuple order (np.add.accumulate), as the reference's loop adds them.
"""

import ctypes
import itertools
import os
import random

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

BIAS_FOLD_PROB = [0.0, 0.35, 0.0, 0.0]
BIAS_POT_FACTOR = [1.0, 0.7, 1.0, 1.4]
OPP_COMBO_LIMIT = 2000
MAX_SAMPLES = 12

_LIB = None


def _eval_lib():
    global _LIB
    if _LIB is None:
        p = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(p):
            raise RuntimeError(f"{p} missing: run `make -C oracle`")
        _LIB = ctypes.CDLL(p)
        _LIB.or_eval7.restype = ctypes.c_uint32
        _LIB.or_eval7.argtypes = [ctypes.POINTER(ctypes.c_int)]
    return _LIB


def eval7(cards):
    return int(_eval_lib().or_eval7((ctypes.c_int * 7)(*cards)))


def biased_leaf_value(equity, leaf_pot, s_self, s_opp):
    """leaf_values.py:91-122: fold equity from a fold-biased opponent plus a
    pot-scaled showdown share."""
    f = BIAS_FOLD_PROB[s_opp]
    showdown_pot = leaf_pot * (BIAS_POT_FACTOR[s_self] * BIAS_POT_FACTOR[s_opp])
    return f * (0.5 * leaf_pot) + (1.0 - f) * ((equity - 0.5) * showdown_pot)


def _any_large(nh):
    return any(n > 50 for n in nh)


def river_runouts(board4, nh):
    dead = set(board4)
    rivers = [c for c in range(52) if c not in dead]
    if len(rivers) > MAX_SAMPLES and _any_large(nh):
        rivers = sorted(np.random.RandomState(42).choice(rivers, MAX_SAMPLES, replace=False))
    return [list(board4) + [int(r)] for r in rivers]


def flop_runouts(flop, nh):
    dead = set(flop)
    turns = [c for c in range(52) if c not in dead]
    large = _any_large(nh)
    if len(turns) > MAX_SAMPLES and large:
        turns = sorted(np.random.RandomState(42).choice(turns, MAX_SAMPLES, replace=False))
    out = []
    for t in turns:
        rivers = [c for c in range(52) if c not in dead and c != t]
        if len(rivers) > MAX_SAMPLES and large:
            rivers = sorted(np.random.RandomState(43 + t).choice(rivers, MAX_SAMPLES, replace=False))
        out.extend(list(flop) + [int(t), int(r)] for r in rivers)
    return out


def _mask(c0, c1):
    return (1 << c0) | (1 << c1)


def runout_equity(hands, nh, runouts, max_hands):
    """hands[p][h] = (c0, c1); returns raw equity [np][max_hands] float64."""
    npl = len(nh)
    raw = np.zeros((npl, max_hands))
    cnt = np.zeros((npl, max_hands))
    for board in runouts:
        bset = set(board)
        valid, strength = [], []
        for p in range(npl):
            v = [h for h in range(nh[p]) if hands[p][h][0] not in bset and hands[p][h][1] not in bset]
            valid.append(v)
            strength.append({h: eval7(list(board) + [hands[p][h][0], hands[p][h][1]]) for h in v})
        for p in range(npl):
            if not valid[p]:
                continue
            opps = [q for q in range(npl) if q != p]
            lists = [valid[q] for q in opps]
            total = 1
            for l in lists:
                total *= len(l)
            if total == 0:
                continue
            if total <= OPP_COMBO_LIMIT:
                tuples = list(itertools.product(*lists))
            else:
                rng = random.Random(42)
                tuples = [tuple(rng.choice(l) for l in lists) for _ in range(OPP_COMBO_LIMIT)]
            C = len(tuples)
            cmask = np.zeros(C, dtype=np.uint64)
            ok = np.ones(C, dtype=bool)
            best = np.zeros(C, dtype=np.int64)
            mult = np.zeros(C, dtype=np.int64)
            for t, tup in enumerate(tuples):
                m = 0
                seen = 0
                ranks = []
                for q, h in zip(opps, tup):
                    hm = _mask(*hands[q][h][:2])
                    if seen & hm:
                        ok[t] = False
                    seen |= hm
                    ranks.append(strength[q][h])
                    m |= hm
                cmask[t] = m
                best[t] = max(ranks)
                mult[t] = ranks.count(best[t])
            hv = valid[p]
            hs = np.array([strength[p][h] for h in hv], dtype=np.int64)
            hm = np.array([_mask(*hands[p][h][:2]) for h in hv], dtype=np.uint64)
            use = ok[None, :] & ((hm[:, None] & cmask[None, :]) == 0)
            add = np.where(hs[:, None] > best[None, :], 1.0,
                           np.where(hs[:, None] == best[None, :], 1.0 / (mult[None, :] + 1), 0.0))
            add = np.where(use, add, 0.0)
            acc = np.add.accumulate(add, axis=1)[:, -1]
            nv = use.sum(axis=1)
            for i, h in enumerate(hv):
                if nv[i] > 0:
                    raw[p, h] += acc[i] / nv[i]
                    cnt[p, h] += 1.0
    good = cnt > 0
    raw[good] /= cnt[good]
    return raw


def opp_proxy(biases, p):
    npl = len(biases)
    if npl > 2:
        other = sum(biases[q] for q in range(npl) if q != p)
        return max(0, min(3, int(round(other / (npl - 1)))))
    return biases[1 - p]


def values_from_equity(raw, nh, max_hands, leaf_infos):
    """leaf_infos: [(leaf_idx, pot)]; returns [len * 4^np][np][max_hands] float32."""
    npl = len(nh)
    cont = 4 ** npl
    total = len(leaf_infos) * cont
    out = np.zeros((total, npl, max_hands), dtype=np.float32)
    for idx, pot in leaf_infos:
        for combo in range(cont):
            flat = idx * cont + combo
            if flat >= total:
                continue
            biases = [(combo // (4 ** i)) % 4 for i in range(npl)]
            for p in range(npl):
                so = opp_proxy(biases, p)
                for h in range(min(nh[p], max_hands)):
                    out[flat, p, h] = biased_leaf_value(raw[p, h], pot, biases[p], so)
    return out


def turn_leaf_values(board4, hands, nh, max_hands, leaf_infos):
    raw = runout_equity(hands, nh, river_runouts(board4, nh), max_hands)
    return values_from_equity(raw, nh, max_hands, leaf_infos)


def flop_leaf_equity(flop, hands, nh, max_hands, leaf_infos):
    raw = runout_equity(hands, nh, flop_runouts(flop, nh), max_hands)
    return values_from_equity(raw, nh, max_hands, leaf_infos)


# Node row of the tree snapshot: type, player, num_children, first_child, pot,
# bets[6], board[5], num_board, leaf_idx, ...
_T, _PL, _NC, _FC, _POT, _BETS, _LEAF = 0, 1, 2, 3, 4, 5, 17
_DECISION, _LEAF_TYPE = 0, 4


def extract_leaf_infos(tree, num_players):
    """leaf_values.py:189-262: every player-0 decision node heading a
    num_players-deep chain of 4-way decision nodes over leaf terminals gives
    (orig leaf index = first terminal's leaf_idx // 4^num_players, pot,
    bets[0..num_players)), sorted by leaf index."""
    nodes, children = tree["nodes"], tree["children"]
    cont = 4 ** num_players

    def chain(i, depth):
        n = nodes[i]
        if n[_T] != _DECISION or n[_NC] != 4:
            return False
        kids = [children[n[_FC] + c] for c in range(4)]
        if depth == num_players - 1:
            return all(nodes[k][_T] == _LEAF_TYPE for k in kids)
        return all(chain(k, depth + 1) for k in kids)

    out = []
    for i in range(tree["num_nodes"]):
        n = nodes[i]
        if n[_T] != _DECISION or n[_PL] != 0 or n[_NC] != 4:
            continue
        if chain(i, 0):
            cur = i
            for _ in range(num_players):
                cur = children[nodes[cur][_FC]]
            out.append((nodes[cur][_LEAF] // cont, n[_POT], tuple(n[_BETS + p] for p in range(num_players))))
    out.sort(key=lambda x: x[0])
    return out
