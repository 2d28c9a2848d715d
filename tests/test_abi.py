"""The C-ABI boundary on the CPU: exports, struct layouts, builder parity,
host evaluator, argument checking, and the Python drop-in wrapper's host logic.
No GPU is touched here (tree building and the host evaluator need none)."""

import ctypes
import glob
import gzip
import json
import os
import re

import numpy as np
import pytest

import cards
import ss_abi
import ss_api

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
TREE_KEYS = ["num_nodes", "num_children_total", "num_leaves", "max_depth", "nodes", "children", "level_order",
             "node_depth", "decision", "showdown", "leaf", "fold"]


def declared_symbols():
    names = set()
    for h in sorted(glob.glob(os.path.join(REPO, "include", "*.h"))):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b((?:ss|ssdbg|fs)_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol(lib):
    syms = declared_symbols()
    assert "ss_build_tree" in syms and "ss_solve_gpu" in syms and "ss_batch_run" in syms
    assert "ss_leaf_equity" in syms and "ss_leaf_values" in syms
    for s in syms:
        assert hasattr(lib, s), s


def test_ctypes_mirror_covers_public_header():
    pub = [s for s in declared_symbols() if s.startswith("ss_")]
    assert sorted(pub) == sorted(ss_abi.SIGNATURES)


def test_struct_layouts_match_reference_abi():
    # street_solve.cuh on x86-64: SSNode 104 B, SSTreeData 14600 B, SSOutput 56 B
    assert ctypes.sizeof(ss_abi.SSNode) == 104
    assert ctypes.sizeof(ss_abi.SSTreeData) == 14600
    assert ctypes.sizeof(ss_abi.SSOutput) == 56
    offs = {f: getattr(ss_abi.SSTreeData, f).offset for f, _ in ss_abi.SSTreeData._fields_}
    assert offs["num_players_field"] == 24 and offs["hands"] == 28 and offs["weights"] == 9628
    assert offs["num_hands"] == 14428 and offs["board"] == 14452 and offs["leaf_values"] == 14480
    assert offs["level_order"] == 14496 and offs["decision_node_indices"] == 14520
    assert offs["starting_pot"] == 14580 and offs["is_river"] == 14588 and offs["frozen_action"] == 14592
    o = {f: getattr(ss_abi.SSOutput, f).offset for f, _ in ss_abi.SSOutput._fields_}
    assert o["root_ev"] == 16 and o["avg_strategies"] == 24 and o["num_avg_nodes"] == 40 and o["max_hands"] == 48


def golden_trees():
    return json.load(gzip.open(os.path.join(HERE, "golden", "trees.json.gz"), "rt"))["trees"]


@pytest.mark.parametrize("t", golden_trees(), ids=lambda t: t["name"])
def test_product_builder_matches_reference(lib, t):
    td = ss_abi.SSTreeData()
    b, s, o = t["board"], t["bet_sizes"], t["acting_order"]
    rc = lib.ss_build_tree((ctypes.c_int * len(b))(*b), len(b), t["pot"], t["stack"], (ctypes.c_float * len(s))(*s),
                           len(s), t["num_players"], (ctypes.c_int * len(o))(*o), t["is_river"],
                           t["use_cont_strats"], ctypes.byref(td))
    assert rc == 0
    d = ss_abi.tree_to_dict(td)
    for k in TREE_KEYS:
        assert d[k] == t[k], k
    lib.ss_free_tree(ctypes.byref(td))
    assert not td.nodes   # struct zeroed


def test_builder_ex_raise_cap_equals_oracle():
    import oracle_bind as ob
    for np_, sizes in [(2, [0.5, 1.0]), (3, [0.33, 0.75]), (4, [1.0])]:
        for cap in (1, 2, 4):
            t = ss_api.Tree([50, 23, 4, 30, 5], 15000, 50000, sizes, np_, max_raises=cap)
            o = ob.build_tree([50, 23, 4, 30, 5], 15000, 50000, sizes, np_, max_raises=cap)
            a, b = ss_abi.tree_to_dict(t.td), ss_abi.tree_to_dict(o)
            for k in TREE_KEYS:
                assert a[k] == b[k], (np_, cap, k)


def test_host_evaluator_matches_reference_vectors(lib):
    E = json.load(open(os.path.join(HERE, "golden", "eval_vectors.json")))
    h = np.array(E["hands7"], dtype=np.int32).ravel()
    n = len(E["values"])
    out = np.zeros(n, np.uint32)
    lib.ssdbg_eval7.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int]
    assert lib.ssdbg_eval7(h.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n,
                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 0) == 0
    assert (out == np.array(E["values"], dtype=np.uint32)).all()


def test_bad_arguments_are_rejected(lib):
    td = ss_abi.SSTreeData()
    board = (ctypes.c_int * 5)(0, 1, 2, 3, 4)
    sizes = (ctypes.c_float * 1)(0.5)
    order = (ctypes.c_int * 7)(0, 1, 2, 3, 4, 5, 6)
    assert lib.ss_build_tree(board, 5, 100, 400, sizes, 1, 7, order, 1, 0, ctypes.byref(td)) < 0
    assert "num_players" in ss_abi.last_error(lib)
    assert lib.ss_build_tree(board, 5, 100, 400, sizes, 1, 2, None, 1, 0, ctypes.byref(td)) < 0
    assert lib.ss_build_tree_ex(board, 5, 100, 400, sizes, 1, 2, order, 1, 0, 99, ctypes.byref(td)) < 0
    # legacy solve: more hands than the legacy ABI allows (checked before any device call)
    assert lib.ss_build_tree(board, 5, 100, 400, sizes, 1, 2, order, 1, 0, ctypes.byref(td)) == 0
    td.num_hands[0] = 201
    out = ss_abi.SSOutput()
    assert lib.ss_solve_gpu(ctypes.byref(td), 10, ctypes.byref(out)) < 0
    assert "legacy ABI" in ss_abi.last_error(lib)
    lib.ss_free_tree(ctypes.byref(td))
    assert lib.ss_solve_gpu(None, 10, ctypes.byref(out)) < 0


def test_free_tree_keeps_caller_buffers(lib):
    """ss_free_tree must not free leaf_values / frozen_action (caller-owned)."""
    td = ss_abi.SSTreeData()
    board = (ctypes.c_int * 4)(0, 9, 20, 33)
    sizes = (ctypes.c_float * 1)(0.5)
    order = (ctypes.c_int * 2)(0, 1)
    assert lib.ss_build_tree(board, 4, 1000, 5000, sizes, 1, 2, order, 0, 1, ctypes.byref(td)) == 0
    leaf = (ctypes.c_float * 16)(*range(16))
    td.leaf_values = ctypes.cast(leaf, ctypes.POINTER(ctypes.c_float))
    lib.ss_free_tree(ctypes.byref(td))
    assert list(leaf)[:3] == [0.0, 1.0, 2.0]


def test_footprint_and_algo_bytes(lib):
    import workloads as wl
    sp = wl.river_2p()
    p = ss_api.Problem.from_spot(sp)
    assert p.tree.num_nodes == 51
    a, f = ss_api.plan_bytes(p)
    # B_iter: 4 * M * per-pass rows (regret match, reach, terminals, backup), M = 1081
    assert a % (4 * 1081) == 0 and 3.0e6 < a < 5.0e6
    assert f > 4 * 1088 * (2 * 2 * 51)
    assert lib.ss_problem_footprint(ctypes.byref(p.struct())) == f


# ---------------------------------------------------------------- wrapper host logic
def test_range_parser_matches_reference_contract():
    assert cards.card_to_int("Ah") == 50 and cards.int_to_card(50) == "Ah"
    with pytest.raises(ValueError):
        cards.card_to_int("Zz")
    r = cards.parse_range_string("AKs,QQ:0.5,AhKh,7c2d")
    assert len(r) == 4 + 6 + 1
    kd, ad = cards.card_to_int("Kd"), cards.card_to_int("Ad")
    assert (min(kd, ad), max(kd, ad), 1.0) in r
    assert all(a < b for a, b, _ in r)
    assert sum(1 for _, _, w in r if w == 0.5) == 6
    assert len(cards.parse_range_string("AKo")) == 12
    assert cards.parse_range_string("") == []


def test_labels_and_trimming_without_gpu(lib):
    from street_solver_gpu import StreetSolverGPU, _build_labels_at_node
    s = StreetSolverGPU(board=["Qs", "As", "2d", "7h", "4c"], oop_range="AhKh,QhQc,JhTh,6h5h,3c3d",
                        ip_range="AcKc,8c8d,Tc9c,KhJh,2h2c", pot_bb=10.0, stack_bb=50.0, bet_sizes=[0.33, 0.75])
    labels = _build_labels_at_node(s._tree, 0)
    assert labels == ["Check", "Bet 33%", "Bet 75%", "Bet 500%"]
    node = s._tree.nodes[s._tree.children[s._tree.nodes[0].first_child + 1]]
    assert node.player == 1
    kids = _build_labels_at_node(s._tree, s._tree.children[s._tree.nodes[0].first_child + 1])
    assert kids[0] == "Fold" and kids[1] == "Call" and kids[2].startswith("Raise")
    s.set_frozen_actions(["Bet 33%"], hero_player=0)
    assert s._frozen_arr[0] == 1 and (s._frozen_arr[1:] == -1).all()
    s.free()
    # more than 200 combos are trimmed to the 200 heaviest in legacy mode
    import workloads as wl
    combos = wl.all_combos([0, 5, 10, 15, 20])
    rng = np.random.default_rng(0)
    w = rng.uniform(0, 1, len(combos))
    rg = [(int(a), int(b), float(x)) for (a, b), x in zip(combos, w)]
    s2 = StreetSolverGPU(board=[0, 5, 10, 15, 20], player_ranges=[rg, rg], pot_bb=10, stack_bb=40)
    assert len(s2.player_hands[0]) == 200
    assert min(x for _, _, x in s2.player_hands[0]) >= sorted(w)[-200] - 1e-9
    s2.free()
