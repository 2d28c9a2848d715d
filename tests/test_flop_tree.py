"""Flop-through-river tree builder and C ABI of include/flop_solve.h on the CPU.

fs_build_tree is host code: these tests run it without a GPU and compare it
with the reference's own builder (src/cuda/flop_solve.cu:1-323, :894-972)
through the committed per-array digests in golden/flop_trees.json and, where
the reference is mounted, against a CPU build of its source directly.
"""

import ctypes
import json
import os
import random
import re

import numpy as np
import pytest

import flop_solve
from flop_solve import FS_NODE_CHANCE, FS_NODE_DECISION, FS_NODE_FOLD, FS_NODE_SHOWDOWN

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
REF_BUILDER = os.path.join(REPO, "oracle", "_ref", "fs_host_ref.so")
REF_SRC = "/root/reference/src/cuda"

SPOTS = json.load(open(os.path.join(HERE, "golden", "flop_trees.json")))["spots"]
ARRAYS = ["nodes", "children", "level_order", "node_depth", "decision", "showdown"]


def digest(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


def test_struct_layouts_match_reference_header():
    # flop_solve.cuh on x86-64 (checked against the header compiled with g++)
    assert ctypes.sizeof(flop_solve.FSNode) == 52
    assert ctypes.sizeof(flop_solve.FSTreeData) == 4888
    assert ctypes.sizeof(flop_solve.FSOutput) == 72
    t = {f: getattr(flop_solve.FSTreeData, f).offset for f, _ in flop_solve.FSTreeData._fields_}
    assert (t["hands"], t["weights"], t["num_hands"], t["level_order"]) == (24, 3224, 4824, 4832)
    assert (t["max_depth"], t["decision_node_indices"], t["num_showdown_nodes"]) == (4848, 4856, 4880)
    o = {f: getattr(flop_solve.FSOutput, f).offset for f, _ in flop_solve.FSOutput._fields_}
    assert (o["all_avg_strategies"], o["max_hands"], o["num_turn_roots"]) == (24, 40, 64)


@pytest.mark.parametrize("spot", SPOTS, ids=lambda s: s["name"])
def test_builder_matches_reference_digests(lib, spot):
    tree = flop_solve.FlopTree(spot["flop"], spot["starting_pot"], spot["effective_stack"], spot["bet_sizes"])
    arr = tree.arrays()
    assert tree.td.num_nodes == spot["num_nodes"]
    assert tree.td.num_children_total == spot["num_children_total"]
    assert tree.td.num_decision_nodes == spot["num_decision"]
    assert tree.td.num_showdown_nodes == spot["num_showdown"]
    assert arr["max_depth"] == spot["max_depth"]
    assert arr["nodes"][:64].tolist() == spot["head_nodes"]
    for k in ARRAYS:
        assert digest(arr[k]) == spot["sha256"][k], k


def test_tree_structure_invariants(lib):
    # Qs As 2d, the reference benchmark's spot (bench/bench_gpu_flop.py:52-58)
    tree = flop_solve.FlopTree([43, 51, 1], 650, 9750, [0.75])
    a = tree.arrays()
    nodes, kids = a["nodes"], a["children"]
    typ, nch, first, nb = nodes[:, 0], nodes[:, 2], nodes[:, 3], nodes[:, 12]
    boards = nodes[:, 7:12]
    # every node but the root is exactly one node's child
    assert np.array_equal(np.sort(kids), np.arange(1, len(nodes)))
    ch = np.flatnonzero(typ == FS_NODE_CHANCE)
    assert len(ch) > 0 and (nb[ch] < 5).all()
    for n in ch[:: max(1, len(ch) // 200)]:
        dealt = [boards[k][nb[n]] for k in kids[first[n]:first[n] + nch[n]]]
        on_board = set(boards[n][:nb[n]].tolist())
        assert dealt == [c for c in range(52) if c not in on_board]
        for k in kids[first[n]:first[n] + nch[n]]:
            assert nb[k] == nb[n] + 1 and boards[k][:nb[n]].tolist() == boards[n][:nb[n]].tolist()
            assert nodes[k][4] == nodes[n][4]          # the pot carries over
            assert nodes[k][5] == nodes[k][6] == 0     # the street's bets start level at 0
    sd = np.flatnonzero(typ == FS_NODE_SHOWDOWN)
    assert (nb[sd] == 5).all() and np.array_equal(sd, a["showdown"])
    dec = np.flatnonzero(typ == FS_NODE_DECISION)
    assert np.array_equal(dec, a["decision"]) and (nch[dec] >= 1).all() and (nch[dec] <= 6).all()
    fold = typ == FS_NODE_FOLD
    assert (nch[fold] == 0).all() and (nch[typ == FS_NODE_SHOWDOWN] == 0).all()
    # 49 turn cards under each flop-end chance node, 48 rivers under each turn-end one
    assert set(nch[ch[nb[ch] == 3]].tolist()) == {49} and set(nch[ch[nb[ch] == 4]].tolist()) == {48}
    # BFS order and depths
    assert a["level_order"][0] == 0 and np.all(np.diff(a["node_depth"][a["level_order"]]) >= 0)
    assert a["max_depth"] == a["node_depth"].max()


@pytest.mark.skipif(not os.path.exists(REF_BUILDER), reason="reference builder not compiled here")
def test_builder_matches_compiled_reference_on_random_spots(lib):
    ref = flop_solve.bind(ctypes.CDLL(REF_BUILDER))
    rng = random.Random(7)
    for _ in range(6):
        flop = rng.sample(range(52), 3)
        pot = rng.choice([50, 100, 333, 1000])
        stack = rng.choice([0, 40, 150, 400, 2500, 12000])
        sizes = sorted(rng.sample([0.25, 0.33, 0.5, 0.66, 0.75, 1.0, 1.5, 2.0], rng.randint(0, 3)))
        ours = flop_solve.FlopTree(flop, pot, stack, sizes).arrays()
        theirs = flop_solve.FlopTree(flop, pot, stack, sizes, lib=ref).arrays()
        for k in ARRAYS + ["max_depth"]:
            assert np.array_equal(ours[k], theirs[k]), (flop, pot, stack, sizes, k)


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_SRC, "flop_solve.cu")), reason="reference not mounted")
def test_flop_evaluator_is_the_street_evaluator():
    """fs_eval5/fs_eval7 (flop_solve.cu:327-359) and ss_eval5/ss_eval7
    (street_solve.cu:467-510) are the same function, token for token, so one
    pinned restatement (the oracle's or_eval7) serves both."""
    def tokens(path, lo, hi, prefix):
        src = "".join(open(path).readlines()[lo - 1:hi])
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src).replace(prefix, "X_")
        return re.findall(r"[A-Za-z_][A-Za-z0-9_]*|\d+u?|\S", src)
    fs = tokens(os.path.join(REF_SRC, "flop_solve.cu"), 327, 359, "fs_")
    ss = tokens(os.path.join(REF_SRC, "street_solve.cu"), 467, 510, "ss_")
    assert fs == ss


def test_free_tree_clears_and_bad_arguments(lib):
    lib = flop_solve.bind(lib)
    tree = flop_solve.FlopTree([0, 5, 10], 100, 100, [1.0])
    assert tree.td.num_nodes > 0
    tree.free()
    assert not tree.td.nodes and tree.td.num_nodes == 0
    td = flop_solve.FSTreeData()
    assert lib.fs_build_tree(None, 100, 100, None, 0, ctypes.byref(td)) != 0
    sizes = (ctypes.c_float * 1)(0.5)
    assert lib.fs_build_tree((ctypes.c_int * 3)(0, 5, 10), 100, 100, sizes, -1, ctypes.byref(td)) != 0
    # no solve call here: the solves need a GPU


def test_set_hands_limits(lib):
    tree = flop_solve.FlopTree([0, 5, 10], 100, 100, [1.0])
    with pytest.raises(ValueError):
        tree.set_hands([(1, 2)] * 201, [(3, 4)])
    tree.set_hands([(12, 13), (14, 15)], [(16, 17)], [0.5, 2.0])
    assert list(tree.td.num_hands) == [2, 1]
    assert tree.td.weights[0][1] == 2.0 and tree.td.hands[1][0][1] == 17
