"""Python host layer over the extended C-ABI of libstreet_solve.so.

Builds trees (ss_build_tree_ex), packs ranges into SSProblem structs and runs
the plain (ss_solve_gpu_ex), batched (ss_solve_batch_ex) and served
(ss_server_*) forms. Every call goes to the HIP library; there is no
CPU fallback, and a missing library raises FileNotFoundError.
"""

import ctypes

import numpy as np

import ss_abi
from ss_abi import SSOutput, SSProblem, SSReport, SSSolveConfig, SSTreeData, SSServerConfig

_P = ctypes.POINTER


class SolveError(RuntimeError):
    pass


def _check(rc, what):
    if rc < 0:
        raise SolveError(f"{what} failed ({rc}): {ss_abi.last_error()}")
    return rc


class Tree:
    """An SSTreeData owned by the library (freed with ss_free_tree)."""

    def __init__(self, board, pot, stack, bet_sizes, num_players, acting_order=None,
                 is_river=None, use_cont=None, max_raises=ss_abi.SS_MAX_RAISES):
        self.lib = ss_abi.load_library()
        board = [int(c) for c in board]
        if acting_order is None:
            acting_order = list(range(num_players))
        if is_river is None:
            is_river = 1 if len(board) == 5 else 0
        if use_cont is None:
            use_cont = 0 if is_river else 1
        self.td = SSTreeData()
        _check(self.lib.ss_build_tree_ex(
            (ctypes.c_int * len(board))(*board), len(board), int(pot), int(stack),
            (ctypes.c_float * max(1, len(bet_sizes)))(*bet_sizes), len(bet_sizes), int(num_players),
            (ctypes.c_int * len(acting_order))(*acting_order), int(is_river), int(use_cont),
            int(max_raises), ctypes.byref(self.td)), "ss_build_tree_ex")
        self.board = board

    @property
    def num_nodes(self):
        return self.td.num_nodes

    @property
    def num_leaves(self):
        return self.td.num_leaves

    def free(self):
        if self.td is not None and self.td.nodes:
            self.lib.ss_free_tree(ctypes.byref(self.td))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Problem:
    """Tree + ranges (+ leaf values, frozen actions) for one subgame."""

    def __init__(self, tree, hands, weights, leaf_values=None, frozen=None):
        self.tree = tree
        self.hands = [np.ascontiguousarray(h, dtype=np.int32).reshape(-1, 2) for h in hands]
        self.weights = [np.ascontiguousarray(w, dtype=np.float32) for w in weights]
        self.leaf_values = None if leaf_values is None else np.ascontiguousarray(leaf_values, dtype=np.float32)
        self.frozen = None if frozen is None else np.ascontiguousarray(frozen, dtype=np.int32)
        self.num_players = len(self.hands)
        self.max_h = max(len(h) for h in self.hands)

    @classmethod
    def from_spot(cls, spot, use_cont=None):
        """Problem for a workloads.Spot (tree built with the spot's raise cap)."""
        t = Tree(spot.board, spot.pot, spot.stack, spot.bet_sizes, spot.num_players,
                 spot.acting_order, use_cont=use_cont, max_raises=spot.max_raises)
        lv = spot.leaf_values(t.num_leaves)
        return cls(t, spot.hands, spot.weights, leaf_values=lv)

    def struct(self):
        s = SSProblem()
        s.tree = ctypes.pointer(self.tree.td)
        for p in range(self.num_players):
            s.num_hands[p] = len(self.hands[p])
            s.hands[p] = self.hands[p].ctypes.data_as(_P(ctypes.c_int))
            s.weights[p] = self.weights[p].ctypes.data_as(_P(ctypes.c_float))
        if self.leaf_values is not None:
            s.leaf_values = self.leaf_values.ctypes.data_as(_P(ctypes.c_float))
        if self.frozen is not None:
            s.frozen_action = self.frozen.ctypes.data_as(_P(ctypes.c_int))
        return s


class Result:
    """Copy of an SSOutput + SSReport as numpy arrays (output freed)."""

    def __init__(self, out, rep, problem):
        lib = ss_abi.load_library()
        mh = out.max_hands
        D = out.num_avg_nodes
        self.num_players = out.num_players_out
        self.max_hands = mh
        self.node_ids = np.array([out.avg_strategy_node_ids[i] for i in range(D)], dtype=np.int32)
        self.avg = np.ctypeslib.as_array(out.avg_strategies, shape=(max(D, 1) * 8 * mh,)).copy() \
            .reshape(max(D, 1), 8, mh)[:D]
        self.root_ev = np.ctypeslib.as_array(out.root_ev, shape=(self.num_players * mh,)).copy()
        rnh = len(problem.hands[out.root_player]) if 0 <= out.root_player < problem.num_players else 0
        n = out.root_num_actions * rnh
        self.root_strategy = (np.ctypeslib.as_array(out.root_strategy, shape=(max(n, 1),)).copy()[:n]
                              .reshape(out.root_num_actions, rnh))
        self.root_player = out.root_player
        self.iterations = rep.iterations if rep is not None else None
        self.converged = bool(rep.converged) if rep is not None else None
        self.exploitability = rep.exploitability if rep is not None else None
        self.nashconv = list(rep.nashconv)[:self.num_players] if rep is not None else None
        self.elapsed_ms = rep.elapsed_ms if rep is not None else None
        self.kernel_ms = rep.kernel_ms if rep is not None else None
        self.algo_bytes = rep.algo_bytes if rep is not None else None
        lib.ss_free_output(ctypes.byref(out))

    def avg_full(self, num_nodes):
        """Average strategy in the node-indexed layout [N][8][max_h]."""
        a = np.zeros((num_nodes, 8, self.max_hands), np.float32)
        a[self.node_ids] = self.avg
        return a


def config(max_iterations, target=0.0, check_every=0, precision=ss_abi.SS_PREC_FP32, flags=0):
    c = SSSolveConfig()
    c.max_iterations = int(max_iterations)
    c.precision = int(precision)
    c.target_exploitability = float(target)
    c.check_every = int(check_every)
    c.flags = int(flags)
    return c


def solve_plain(problem, iterations, target=0.0, check_every=0, flags=0, precision=ss_abi.SS_PREC_FP32):
    lib = ss_abi.load_library()
    out, rep = SSOutput(), SSReport()
    st = problem.struct()
    cfg = config(iterations, target, check_every, precision=precision, flags=flags)
    _check(lib.ss_solve_gpu_ex(ctypes.byref(st), ctypes.byref(cfg), ctypes.byref(out), ctypes.byref(rep)),
           "ss_solve_gpu_ex")
    return Result(out, rep, problem)


def solve_batch(problems, iterations, target=0.0, check_every=0, flags=0):
    lib = ss_abi.load_library()
    n = len(problems)
    structs = (SSProblem * n)(*[p.struct() for p in problems])
    outs = (SSOutput * n)()
    reps = (SSReport * n)()
    cfg = config(iterations, target, check_every, flags=flags)
    _check(lib.ss_solve_batch_ex(structs, n, ctypes.byref(cfg), outs, reps), "ss_solve_batch_ex")
    return [Result(outs[i], reps[i], problems[i]) for i in range(n)]


class Batch:
    """Batched form as a session: upload once (ss_batch_create), then solve all
    subgames from scratch on every run() with inputs resident in HBM."""

    def __init__(self, problems, iterations, target=0.0, check_every=0, flags=0, precision=ss_abi.SS_PREC_FP32):
        self.lib = ss_abi.load_library()
        self.problems = list(problems)
        self.n = len(self.problems)
        self._structs = (SSProblem * self.n)(*[p.struct() for p in self.problems])
        cfg = config(iterations, target, check_every, precision=precision, flags=flags)
        self.h = self.lib.ss_batch_create(self._structs, self.n, ctypes.byref(cfg))
        if not self.h:
            raise SolveError(f"ss_batch_create failed: {ss_abi.last_error()}")

    def run(self):
        reps = (SSReport * self.n)()
        stats = ss_abi.SSBatchStats()
        _check(self.lib.ss_batch_run(self.h, reps, ctypes.byref(stats)), "ss_batch_run")
        return [reps[i] for i in range(self.n)], stats

    def output(self, i, report=None):
        out = SSOutput()
        _check(self.lib.ss_batch_output(self.h, int(i), ctypes.byref(out)), "ss_batch_output")
        return Result(out, report, self.problems[i])

    def close(self):
        if self.h:
            self.lib.ss_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_state(problem, iterations, form=0):
    """Raw regrets / strategy sums [N][8][max_h] and root cfv after k iterations."""
    lib = ss_abi.load_library()
    N, mh = problem.tree.num_nodes, problem.max_h
    R = np.zeros((N, 8, mh), np.float32)
    S = np.zeros((N, 8, mh), np.float32)
    ev = np.zeros(mh, np.float32)
    st = problem.struct()
    fp = _P(ctypes.c_float)
    lib.ssdbg_solve_state.restype = ctypes.c_int
    _check(lib.ssdbg_solve_state(ctypes.byref(st), int(iterations), int(form), R.ctypes.data_as(fp),
                                 S.ctypes.data_as(fp), ev.ctypes.data_as(fp)), "ssdbg_solve_state")
    return R, S, ev


def plan_bytes(problem):
    lib = ss_abi.load_library()
    a, f = ctypes.c_int64(), ctypes.c_int64()
    st = problem.struct()
    _check(lib.ssdbg_plan_bytes(ctypes.byref(st), ctypes.byref(a), ctypes.byref(f)), "ssdbg_plan_bytes")
    return a.value, f.value


class Server:
    """Continuous-batching request loop (ss_server_*)."""

    def __init__(self, hbm_budget_bytes=0, max_slots=256, iters_per_step=10, check_every=50):
        self.lib = ss_abi.load_library()
        c = SSServerConfig()
        c.hbm_budget_bytes = int(hbm_budget_bytes)
        c.max_slots = int(max_slots)
        c.iters_per_step = int(iters_per_step)
        c.check_every = int(check_every)
        self.h = self.lib.ss_server_create(ctypes.byref(c))
        if not self.h:
            raise SolveError(f"ss_server_create failed: {ss_abi.last_error()}")
        self.problems = {}

    def submit(self, problem, iterations, target=0.005 * 100, check_every=0, flags=0,
               precision=ss_abi.SS_PREC_FP32):
        st = problem.struct()
        cfg = config(iterations, target, check_every, precision=precision, flags=flags)
        t = self.lib.ss_server_submit(self.h, ctypes.byref(st), ctypes.byref(cfg))
        _check(t, "ss_server_submit")
        self.problems[t] = problem
        return t

    def poll(self, ticket):
        out, rep = SSOutput(), SSReport()
        rc = self.lib.ss_server_poll(self.h, ticket, ctypes.byref(out), ctypes.byref(rep))
        _check(rc, "ss_server_poll")
        if rc == 0:
            return None
        return Result(out, rep, self.problems.pop(ticket))

    def cancel(self, ticket):
        return _check(self.lib.ss_server_cancel(self.h, ticket), "ss_server_cancel")

    def drain(self):
        return _check(self.lib.ss_server_drain(self.h), "ss_server_drain")

    def close(self):
        if self.h:
            self.lib.ss_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
