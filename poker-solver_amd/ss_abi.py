"""ctypes mirror of include/street_solve.h and the loader of libstreet_solve.so.

The legacy structs keep the field order and sizes of the reference ctypes
mirror (python/street_solver_gpu.py:66-124), which itself mirrors
src/cuda/street_solve.cuh:52-143: SSNode 104 B, SSTreeData 14600 B,
SSOutput 56 B on x86-64.
"""

import ctypes
import os

SS_MAX_PLAYERS = 6
SS_MAX_HANDS = 200
SS_MAX_ACTIONS = 8
SS_MAX_BOARD = 5
SS_EX_MAX_HANDS = 1326
SS_MAX_RAISES = 3

SS_NODE_DECISION = 0
SS_NODE_FOLD = 1
SS_NODE_SHOWDOWN = 2
SS_NODE_LEAF = 4

SS_PREC_FP32 = 0
SS_PREC_BF16X2 = 1

SS_FLAG_NO_EXPL = 1
SS_FLAG_ADAPTIVE_CHECKS = 2
SS_FLAG_SPARSE_CHECKS = 4

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "build", "street_solve.so")


class SSNode(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int),
        ("player", ctypes.c_int),
        ("num_children", ctypes.c_int),
        ("first_child", ctypes.c_int),
        ("pot", ctypes.c_int),
        ("bets", ctypes.c_int * SS_MAX_PLAYERS),
        ("board_cards", ctypes.c_int * 5),
        ("num_board", ctypes.c_int),
        ("leaf_idx", ctypes.c_int),
        ("fold_player", ctypes.c_int),
        ("num_players", ctypes.c_int),
        ("active_players", ctypes.c_int * SS_MAX_PLAYERS),
    ]


class SSTreeData(ctypes.Structure):
    _fields_ = [
        ("nodes", ctypes.POINTER(SSNode)),
        ("children", ctypes.POINTER(ctypes.c_int)),
        ("num_nodes", ctypes.c_int),
        ("num_children_total", ctypes.c_int),
        ("num_players_field", ctypes.c_int),
        ("hands", ((ctypes.c_int * 2) * SS_MAX_HANDS) * SS_MAX_PLAYERS),
        ("weights", (ctypes.c_float * SS_MAX_HANDS) * SS_MAX_PLAYERS),
        ("num_hands", ctypes.c_int * SS_MAX_PLAYERS),
        ("board", ctypes.c_int * SS_MAX_BOARD),
        ("num_board", ctypes.c_int),
        ("leaf_values", ctypes.POINTER(ctypes.c_float)),
        ("num_leaves", ctypes.c_int),
        ("level_order", ctypes.POINTER(ctypes.c_int)),
        ("node_depth", ctypes.POINTER(ctypes.c_int)),
        ("max_depth", ctypes.c_int),
        ("decision_node_indices", ctypes.POINTER(ctypes.c_int)),
        ("num_decision_nodes", ctypes.c_int),
        ("showdown_node_indices", ctypes.POINTER(ctypes.c_int)),
        ("num_showdown_nodes", ctypes.c_int),
        ("leaf_node_indices", ctypes.POINTER(ctypes.c_int)),
        ("num_leaf_nodes", ctypes.c_int),
        ("fold_node_indices", ctypes.POINTER(ctypes.c_int)),
        ("num_fold_nodes", ctypes.c_int),
        ("starting_pot", ctypes.c_int),
        ("effective_stack", ctypes.c_int),
        ("is_river", ctypes.c_int),
        ("frozen_action", ctypes.POINTER(ctypes.c_int)),
    ]


class SSOutput(ctypes.Structure):
    _fields_ = [
        ("root_strategy", ctypes.POINTER(ctypes.c_float)),
        ("root_num_actions", ctypes.c_int),
        ("root_player", ctypes.c_int),
        ("root_ev", ctypes.POINTER(ctypes.c_float)),
        ("avg_strategies", ctypes.POINTER(ctypes.c_float)),
        ("avg_strategy_node_ids", ctypes.POINTER(ctypes.c_int)),
        ("num_avg_nodes", ctypes.c_int),
        ("num_players_out", ctypes.c_int),
        ("max_hands", ctypes.c_int),
    ]


class SSProblem(ctypes.Structure):
    _fields_ = [
        ("tree", ctypes.POINTER(SSTreeData)),
        ("num_hands", ctypes.c_int * SS_MAX_PLAYERS),
        ("hands", ctypes.POINTER(ctypes.c_int) * SS_MAX_PLAYERS),
        ("weights", ctypes.POINTER(ctypes.c_float) * SS_MAX_PLAYERS),
        ("leaf_values", ctypes.POINTER(ctypes.c_float)),
        ("frozen_action", ctypes.POINTER(ctypes.c_int)),
    ]


class SSSolveConfig(ctypes.Structure):
    _fields_ = [
        ("max_iterations", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("target_exploitability", ctypes.c_double),
        ("check_every", ctypes.c_int),
        ("flags", ctypes.c_int),
    ]


class SSReport(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("exploitability", ctypes.c_double),
        ("nashconv", ctypes.c_double * SS_MAX_PLAYERS),
        ("elapsed_ms", ctypes.c_double),
        ("kernel_ms", ctypes.c_double),
        ("algo_bytes", ctypes.c_int64),
    ]


class SSBatchStats(ctypes.Structure):
    _fields_ = [
        ("run_ms", ctypes.c_double),
        ("cfr_ms", ctypes.c_double),
        ("check_ms", ctypes.c_double),
        ("algo_bytes", ctypes.c_int64),
        ("launches", ctypes.c_int),
        ("max_iterations", ctypes.c_int),
        ("mode", ctypes.c_int),
        ("shape_fwd_group_min", ctypes.c_int),
        ("shape_term_small", ctypes.c_int),
        ("shape_sd_parts", ctypes.c_int),
        ("shape_sweep_grid", ctypes.c_int),
        ("tune_ms", ctypes.c_double),
    ]


class SSServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_int64),
        ("max_slots", ctypes.c_int),
        ("iters_per_step", ctypes.c_int),
        ("check_every", ctypes.c_int),
    ]


# Every symbol include/street_solve.h, street_solve_leaf.h and
# street_solve_range.h declare, with its ctypes signature.
_P = ctypes.POINTER
SIGNATURES = {
    "ss_build_tree": (ctypes.c_int, [_P(ctypes.c_int), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     _P(ctypes.c_float), ctypes.c_int, ctypes.c_int,
                                     _P(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                     _P(SSTreeData)]),
    "ss_solve_gpu": (ctypes.c_int, [_P(SSTreeData), ctypes.c_int, _P(SSOutput)]),
    "ss_free_tree": (None, [_P(SSTreeData)]),
    "ss_free_output": (None, [_P(SSOutput)]),
    "ss_build_tree_ex": (ctypes.c_int, [_P(ctypes.c_int), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        _P(ctypes.c_float), ctypes.c_int, ctypes.c_int,
                                        _P(ctypes.c_int), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        _P(SSTreeData)]),
    "ss_solve_gpu_ex": (ctypes.c_int, [_P(SSProblem), _P(SSSolveConfig), _P(SSOutput), _P(SSReport)]),
    "ss_solve_batch_ex": (ctypes.c_int, [_P(SSProblem), ctypes.c_int, _P(SSSolveConfig),
                                         _P(SSOutput), _P(SSReport)]),
    "ss_batch_create": (ctypes.c_void_p, [_P(SSProblem), ctypes.c_int, _P(SSSolveConfig)]),
    "ss_batch_run": (ctypes.c_int, [ctypes.c_void_p, _P(SSReport), _P(SSBatchStats)]),
    "ss_batch_output": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(SSOutput)]),
    "ss_batch_destroy": (None, [ctypes.c_void_p]),
    "ss_server_create": (ctypes.c_void_p, [_P(SSServerConfig)]),
    "ss_server_submit": (ctypes.c_int64, [ctypes.c_void_p, _P(SSProblem), _P(SSSolveConfig)]),
    "ss_server_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, _P(SSOutput), _P(SSReport)]),
    "ss_server_cancel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "ss_server_drain": (ctypes.c_int, [ctypes.c_void_p]),
    "ss_server_destroy": (None, [ctypes.c_void_p]),
    "ss_problem_footprint": (ctypes.c_int64, [_P(SSProblem)]),
    "ss_leaf_equity": (ctypes.c_int, [ctypes.c_int, _P(ctypes.c_int), _P(ctypes.c_int), ctypes.c_int,
                                      _P(ctypes.c_int), ctypes.c_int, _P(ctypes.c_double)]),
    "ss_leaf_equity_static": (ctypes.c_int, [ctypes.c_int, _P(ctypes.c_int), _P(ctypes.c_int), ctypes.c_int,
                                             _P(ctypes.c_int), ctypes.c_int, _P(ctypes.c_double)]),
    "ss_leaf_values": (ctypes.c_int, [ctypes.c_int, _P(ctypes.c_double), _P(ctypes.c_int), ctypes.c_int,
                                      _P(ctypes.c_int), _P(ctypes.c_int), ctypes.c_int, _P(ctypes.c_float)]),
    "ss_range_update": (ctypes.c_int, [ctypes.c_int, _P(ctypes.c_double), _P(ctypes.c_double),
                                       _P(ctypes.c_ubyte), ctypes.c_double, ctypes.c_double]),
    "ss_batch_leaf_values": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(ctypes.c_double), _P(ctypes.c_int),
                                            _P(ctypes.c_int), ctypes.c_int, ctypes.c_int]),
    "ss_batch_range_update": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             _P(ctypes.c_int), _P(ctypes.c_double), ctypes.c_double,
                                             ctypes.c_double]),
    "ss_last_error": (ctypes.c_char_p, []),
    "ss_version": (ctypes.c_char_p, []),
}

_LIB = None


def load_library(path=None):
    """Load libstreet_solve.so (once) and declare every signature.

    Raises FileNotFoundError when the library has not been built: there is no
    fallback implementation.
    """
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(
            f"street_solve library not found at {p}; build it with "
            f"`python poker-solver_amd/build_lib.py`")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def last_error(lib=None):
    lib = lib or load_library()
    msg = lib.ss_last_error()
    return msg.decode() if msg else ""


def tree_to_dict(td):
    """Plain-Python snapshot of an SSTreeData (for fixtures and comparisons)."""
    n = td.num_nodes
    nodes = []
    for i in range(n):
        nd = td.nodes[i]
        nodes.append([nd.type, nd.player, nd.num_children, nd.first_child, nd.pot]
                     + list(nd.bets) + list(nd.board_cards)
                     + [nd.num_board, nd.leaf_idx, nd.fold_player, nd.num_players]
                     + list(nd.active_players))
    lst = lambda ptr, k: [ptr[i] for i in range(k)]
    return {
        "num_nodes": n,
        "num_children_total": td.num_children_total,
        "num_leaves": td.num_leaves,
        "max_depth": td.max_depth,
        "nodes": nodes,
        "children": lst(td.children, td.num_children_total),
        "level_order": lst(td.level_order, n),
        "node_depth": lst(td.node_depth, n),
        "decision": lst(td.decision_node_indices, td.num_decision_nodes),
        "showdown": lst(td.showdown_node_indices, td.num_showdown_nodes),
        "leaf": lst(td.leaf_node_indices, td.num_leaf_nodes),
        "fold": lst(td.fold_node_indices, td.num_fold_nodes),
    }
