"""ctypes binding of include/flop_solve.h: the 2-player flop-through-river
solve with chance nodes.

The structures follow the C header field for field (the reference's own
bindings, bench/bench_gpu_flop.py:8-28 and precompute/run_all.py:37-113, read
the same layout). `FlopTree` owns one built tree; `solve` runs fs_solve_gpu or
fs_solve_gpu_extract_all and returns numpy arrays.
"""

import ctypes

import numpy as np

import ss_abi

FS_MAX_HANDS = 200
FS_MAX_ACTIONS = 6
FS_NODE_DECISION, FS_NODE_FOLD, FS_NODE_SHOWDOWN, FS_NODE_CHANCE = 0, 1, 2, 3


class FSNode(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("player", ctypes.c_int), ("num_children", ctypes.c_int),
                ("first_child", ctypes.c_int), ("pot", ctypes.c_int), ("bets", ctypes.c_int * 2),
                ("board_cards", ctypes.c_int * 5), ("num_board", ctypes.c_int)]


class FSTreeData(ctypes.Structure):
    _fields_ = [("nodes", ctypes.POINTER(FSNode)), ("children", ctypes.POINTER(ctypes.c_int)),
                ("num_nodes", ctypes.c_int), ("num_children_total", ctypes.c_int),
                ("hands", ((ctypes.c_int * 2) * FS_MAX_HANDS) * 2),
                ("weights", (ctypes.c_float * FS_MAX_HANDS) * 2), ("num_hands", ctypes.c_int * 2),
                ("level_order", ctypes.POINTER(ctypes.c_int)), ("node_depth", ctypes.POINTER(ctypes.c_int)),
                ("max_depth", ctypes.c_int), ("decision_node_indices", ctypes.POINTER(ctypes.c_int)),
                ("num_decision_nodes", ctypes.c_int), ("showdown_node_indices", ctypes.POINTER(ctypes.c_int)),
                ("num_showdown_nodes", ctypes.c_int)]


class FSOutput(ctypes.Structure):
    _fields_ = [("root_strategy", ctypes.POINTER(ctypes.c_float)), ("root_num_actions", ctypes.c_int),
                ("root_player", ctypes.c_int), ("root_ev", ctypes.POINTER(ctypes.c_float)),
                ("all_avg_strategies", ctypes.POINTER(ctypes.c_float)),
                ("all_cfv", ctypes.POINTER(ctypes.c_float)), ("max_hands", ctypes.c_int),
                ("turn_root_indices", ctypes.POINTER(ctypes.c_int)),
                ("turn_root_cards", ctypes.POINTER(ctypes.c_int)), ("num_turn_roots", ctypes.c_int)]


def bind(lib):
    """Declare the fs_* signatures on a loaded library (ours, or any with the same C ABI)."""
    tree_p, out_p = ctypes.POINTER(FSTreeData), ctypes.POINTER(FSOutput)
    lib.fs_build_tree.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_float), ctypes.c_int, tree_p]
    lib.fs_build_tree.restype = ctypes.c_int
    lib.fs_free_tree.argtypes = [tree_p]
    lib.fs_free_tree.restype = None
    if hasattr(lib, "fs_solve_gpu"):
        for name in ("fs_solve_gpu", "fs_solve_gpu_extract_all"):
            getattr(lib, name).argtypes = [tree_p, ctypes.c_int, out_p]
            getattr(lib, name).restype = ctypes.c_int
        lib.fs_free_output.argtypes = [out_p]
        lib.fs_free_output.restype = None
    return lib


def _lib():
    return bind(ss_abi.load_library())


NODE_FIELDS = 13   # ints per FSNode


class FlopTree:
    """One tree built by `fs_build_tree` of `lib` (default: this library)."""

    def __init__(self, flop, starting_pot, effective_stack, bet_sizes, lib=None):
        self.lib = lib if lib is not None else _lib()
        self.td = FSTreeData()
        sizes = (ctypes.c_float * max(1, len(bet_sizes)))(*bet_sizes)
        rc = self.lib.fs_build_tree((ctypes.c_int * 3)(*flop), int(starting_pot), int(effective_stack), sizes,
                                    len(bet_sizes), ctypes.byref(self.td))
        if rc != 0:
            raise RuntimeError(f"fs_build_tree failed: {rc}")

    def set_hands(self, hands0, hands1, weights0=None, weights1=None):
        for p, (hands, w) in enumerate(((hands0, weights0), (hands1, weights1))):
            if len(hands) > FS_MAX_HANDS:
                raise ValueError(f"player {p}: {len(hands)} hands (at most {FS_MAX_HANDS})")
            for h, (c0, c1) in enumerate(hands):
                self.td.hands[p][h][0], self.td.hands[p][h][1] = int(c0), int(c1)
                self.td.weights[p][h] = 1.0 if w is None else float(w[h])
            self.td.num_hands[p] = len(hands)

    @property
    def num_nodes(self):
        return self.td.num_nodes

    def arrays(self):
        """The tree as numpy arrays: nodes [N][13] (the FSNode fields in order), children,
        level_order, node_depth, decision and showdown node indices, and max_depth."""
        n = self.td.num_nodes
        nodes = np.ctypeslib.as_array(ctypes.cast(self.td.nodes, ctypes.POINTER(ctypes.c_int)),
                                      shape=(n * NODE_FIELDS,)).reshape(n, NODE_FIELDS).copy()

        def ints(ptr, count):
            return np.ctypeslib.as_array(ptr, shape=(count,)).copy() if count else np.zeros(0, np.int32)

        return {"nodes": nodes, "children": ints(self.td.children, self.td.num_children_total),
                "level_order": ints(self.td.level_order, n), "node_depth": ints(self.td.node_depth, n),
                "decision": ints(self.td.decision_node_indices, self.td.num_decision_nodes),
                "showdown": ints(self.td.showdown_node_indices, self.td.num_showdown_nodes),
                "max_depth": self.td.max_depth}

    def free(self):
        if self.td.nodes:
            self.lib.fs_free_tree(ctypes.byref(self.td))

    def __del__(self):
        try:
            self.free()
        except Exception:  # noqa: BLE001 - interpreter teardown
            pass


def solve(tree, iterations, extract_all=False):
    """Run the GPU solve on `tree`; returns a dict of numpy arrays."""
    lib = _lib()
    out = FSOutput()
    fn = lib.fs_solve_gpu_extract_all if extract_all else lib.fs_solve_gpu
    rc = fn(ctypes.byref(tree.td), int(iterations), ctypes.byref(out))
    if rc != 0:
        raise RuntimeError(f"flop solve failed ({rc}): {ss_abi.last_error()}")
    try:
        res = {"root_player": out.root_player, "root_num_actions": out.root_num_actions}
        nh = tree.td.num_hands[out.root_player]
        na = out.root_num_actions
        res["root_strategy"] = (np.ctypeslib.as_array(out.root_strategy, shape=(na * nh,)).reshape(na, nh).copy()
                                if na else np.zeros((0, nh), np.float32))
        if extract_all:
            n, mh = tree.td.num_nodes, out.max_hands
            res["max_hands"] = mh
            res["all_avg_strategies"] = np.ctypeslib.as_array(
                out.all_avg_strategies, shape=(n * FS_MAX_ACTIONS * mh,)).reshape(n, FS_MAX_ACTIONS, mh).copy()
            res["all_cfv"] = np.ctypeslib.as_array(out.all_cfv, shape=(n * mh,)).reshape(n, mh).copy()
            k = out.num_turn_roots
            res["turn_root_indices"] = np.ctypeslib.as_array(out.turn_root_indices, shape=(k,)).copy()
            res["turn_root_cards"] = np.ctypeslib.as_array(out.turn_root_cards, shape=(k,)).copy()
        return res
    finally:
        lib.fs_free_output(ctypes.byref(out))
