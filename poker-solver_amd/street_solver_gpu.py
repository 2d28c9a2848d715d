"""StreetSolverGPU -- drop-in for python/street_solver_gpu.py on MI355X.

The public surface is the reference class's (python/street_solver_gpu.py:
173-596): constructor arguments, attribute names, methods, action labels and
the 0.001 output threshold, because callers (hud_solver.py:834-857) parse
them. The implementation is organised differently:

  * ``_TreeView`` copies the built tree once into numpy arrays (node type,
    actor, pot, bets, child lists); action labels and tree walks read those
    arrays instead of ctypes structs.
  * ``_Solution`` wraps the library's output arrays as numpy views once per
    solve and indexes decision nodes by id and by player.
  * One walker (``_TreeView.follow``) serves both the A3 freeze
    (``set_frozen_actions``) and the strategy lookup (``get_strategy``).
  * Native handles (tree and output) are owned by ``_NativeHandles`` and
    released exactly once, by ``free()`` or by a ``weakref.finalize`` hook.

Input preparation keeps the reference's behaviour: board-blocked combos are
dropped and ranges above the cap are cut to the heaviest combos (stable by
weight, :224-229); pot and stack in BB are converted with SCALE = 100 and
truncated (:263-264); continuation strategies are used only off the river
(:272); leaf values come from the explicit array, else from
leaf_value_fn(tree, hands, max_h, pot), else zeros with a warning off the
river (:296-319). ``ss_free_tree`` here never frees the caller-owned
leaf-value or frozen-action buffers (the reference frees leaf_values at
src/cuda/street_solve.cu:1650, a buffer this object owns).

extended=True lifts the 200-combo cap (up to 1326) and solves through
ss_solve_gpu_ex, which also reports the FP64 exploitability. resident=True
(implies extended) keeps the solve on the GPU as a batch session, so that
range narrowing can read the average strategy there (narrow_range).
"""

import ctypes
import sys
import weakref

import numpy as np

import ss_abi
from cards import card_to_int, int_to_card, parse_range_string  # noqa: F401
from ss_abi import SSOutput, SSProblem, SSReport, SSSolveConfig, SSTreeData

SS_MAX_PLAYERS = ss_abi.SS_MAX_PLAYERS
SS_MAX_HANDS = ss_abi.SS_MAX_HANDS
SS_MAX_ACTIONS = ss_abi.SS_MAX_ACTIONS
SS_MAX_BOARD = ss_abi.SS_MAX_BOARD
SCALE = 100
OUTPUT_THRESHOLD = 0.001        # python/street_solver_gpu.py:418, :484

_SEAT_NAMES = {"oop": 0, "ip": 1}
_NODE_INTS = ctypes.sizeof(ss_abi.SSNode) // ctypes.sizeof(ctypes.c_int)   # 26


def _get_lib():
    return ss_abi.load_library()


class _TreeView:
    """Read-only numpy copy of the node table and child lists of a built tree."""

    def __init__(self, td):
        n = td.num_nodes
        raw = np.ctypeslib.as_array(ctypes.cast(td.nodes, ctypes.POINTER(ctypes.c_int)),
                                    shape=(n * _NODE_INTS,)).reshape(n, _NODE_INTS)
        self.kind = raw[:, 0].copy()
        self.actor = raw[:, 1].copy()
        self.nkids = raw[:, 2].copy()
        self.first = raw[:, 3].copy()
        self.pot = raw[:, 4].copy()
        self.bets = raw[:, 5:5 + SS_MAX_PLAYERS].copy()
        m = td.num_children_total
        self.kids = (np.ctypeslib.as_array(td.children, shape=(m,)).copy() if m > 0
                     else np.zeros(0, np.int32))
        self.num_nodes = n

    def children(self, node):
        f = int(self.first[node])
        return self.kids[f:f + int(self.nkids[node])]

    def labels(self, node):
        """Action labels of a decision node (python/street_solver_gpu.py:128-169):
        a fold child is "Fold"; a showdown or leaf child is "Call" when the
        actor faces a bet, else "Check"; a decision child is named by how far
        the actor's bet grew, in percent of the pot (Python round())."""
        who = int(self.actor[node])
        own = int(self.bets[node, who]) if who >= 0 else 0
        facing = (int(self.bets[node].max()) - own) > 0 if who >= 0 else False
        pot = int(self.pot[node])
        out = []
        for i, c in enumerate(self.children(node)):
            k = int(self.kind[c])
            if k == ss_abi.SS_NODE_FOLD:
                out.append("Fold")
            elif k in (ss_abi.SS_NODE_SHOWDOWN, ss_abi.SS_NODE_LEAF):
                out.append("Call" if facing else "Check")
            elif k == ss_abi.SS_NODE_DECISION:
                grow = (int(self.bets[c, who]) if who >= 0 else 0) - own
                verb = "Raise" if facing else "Bet"
                if grow == 0:
                    out.append("Check")
                elif pot > 0:
                    out.append(f"{verb} {int(round(100 * grow / pot))}%")
                else:
                    out.append(verb)
            else:
                out.append(f"Action {i}")
        return out

    def follow(self, actions, same):
        """Walk from the root along action labels; yields (node, action index)
        for every matched step and returns through StopIteration's value the
        node reached. Stops at a non-decision node or an unmatched label.
        ``same(tree_label, wanted)`` decides a match."""
        node = 0
        for want in actions:
            if node >= self.num_nodes or int(self.kind[node]) != ss_abi.SS_NODE_DECISION:
                break
            names = self.labels(node)
            hit = next((i for i, nm in enumerate(names) if same(nm, want)), None)
            if hit is None:
                break
            yield node, hit
            node = int(self.children(node)[hit])
        return node

    def walk_to(self, actions, same):
        it = self.follow(actions, same)
        while True:
            try:
                next(it)
            except StopIteration as stop:
                return stop.value


def _build_labels_at_node(tree_data, node_idx):
    """Module-level label helper with the reference's signature (:128)."""
    return _TreeView(tree_data).labels(node_idx)


def _exact_label(tree_label, wanted):          # set_frozen_actions (:349)
    return tree_label.lower() == wanted.lower()


def _loose_label(tree_label, wanted):          # get_strategy walk (:457-460)
    return tree_label.lower().strip() == wanted.lower().strip()


def _filter_range(rng, blocked, cap):
    hands = parse_range_string(rng) if isinstance(rng, str) else list(rng)
    hands = [h for h in hands if h[0] not in blocked and h[1] not in blocked]
    if len(hands) > cap:
        hands = sorted(hands, key=lambda h: -h[2])[:cap]    # stable: ties keep order
    return hands


class _NativeHandles:
    """Owns the library-allocated tree and output; ``release`` runs once."""

    def __init__(self, lib):
        self.lib = lib
        self.tree = SSTreeData()
        self.output = SSOutput()
        self.tree_built = False
        self.output_live = False
        self.batch = None          # resident solve (ss_batch_* session), resident=True only

    def drop_output(self):
        if self.output_live:
            self.output_live = False
            self.lib.ss_free_output(ctypes.byref(self.output))
        if self.batch:
            batch, self.batch = self.batch, None
            self.lib.ss_batch_destroy(batch)

    def release(self):
        self.drop_output()
        if self.tree_built:
            self.tree_built = False
            self.lib.ss_free_tree(ctypes.byref(self.tree))


class _Solution:
    """numpy views of one SSOutput plus decision-node indexes."""

    def __init__(self, out, view, hand_counts, max_h):
        self.root_player = out.root_player
        na = out.root_num_actions
        nh = hand_counts[self.root_player] if 0 <= self.root_player < len(hand_counts) else 0
        self.root = (np.ctypeslib.as_array(out.root_strategy, shape=(na * nh,)).reshape(na, nh)
                     if na * nh > 0 else np.zeros((na, 0), np.float32))
        d = out.num_avg_nodes
        self.node_ids = (np.ctypeslib.as_array(out.avg_strategy_node_ids, shape=(d,)).copy()
                         if d > 0 else np.zeros(0, np.int32))
        self.avg = (np.ctypeslib.as_array(out.avg_strategies, shape=(d * SS_MAX_ACTIONS * max_h,))
                    .reshape(d, SS_MAX_ACTIONS, max_h) if d > 0 else np.zeros((0, SS_MAX_ACTIONS, max_h)))
        self.row_of_node = {}
        self.first_row_of_player = {}
        for row, nid in enumerate(self.node_ids.tolist()):
            if int(view.kind[nid]) != ss_abi.SS_NODE_DECISION:
                continue
            self.row_of_node[nid] = row
            self.first_row_of_player.setdefault(int(view.actor[nid]), row)


def _as_dict(labels, probs):
    """{label: p} for p above the threshold, compared in float64 as the
    reference compares Python floats."""
    p64 = np.asarray(probs, dtype=np.float64)
    return {labels[a]: float(p64[a]) for a in range(min(len(labels), len(p64)))
            if p64[a] > OUTPUT_THRESHOLD}


class StreetSolverGPU:
    """N-player single-street Linear-CFR solver on the GPU (2-6 players)."""

    def __init__(self, board, pot_bb, stack_bb, oop_range=None, ip_range=None, player_ranges=None,
                 acting_order=None, bet_sizes=None, leaf_values=None, use_cont_strats=True,
                 leaf_value_fn=None, extended=False, max_raises=None, resident=False):
        self.lib = _get_lib()
        self.resident = bool(resident)
        self.extended = bool(extended) or self.resident
        self._freed = False
        self._solved = False
        self._sol = None
        self.board_cards = [card_to_int(c) if isinstance(c, str) else int(c) for c in board]
        self.num_board = len(self.board_cards)
        self.is_river = self.num_board == 5
        if player_ranges is not None:
            ranges = list(player_ranges)
        elif oop_range is not None and ip_range is not None:
            ranges = [oop_range, ip_range]
        else:
            raise ValueError("Provide oop_range+ip_range or player_ranges")
        cap = ss_abi.SS_EX_MAX_HANDS if self.extended else SS_MAX_HANDS
        blocked = frozenset(self.board_cards)
        self.player_hands = [_filter_range(r, blocked, cap) for r in ranges]
        self.num_players = len(self.player_hands)
        if self.num_players >= 1:
            self.oop_hands = self.player_hands[0]
        if self.num_players >= 2:
            self.ip_hands = self.player_hands[1]
        self.bet_sizes = [0.33, 0.75, 1.5] if bet_sizes is None else list(bet_sizes)
        self.acting_order = list(range(self.num_players)) if acting_order is None else list(acting_order)
        self._index_of = [None] * self.num_players

        self._handles = _NativeHandles(self.lib)
        self._tree = self._handles.tree
        self._output = self._handles.output
        self._build_tree(int(pot_bb * SCALE), int(stack_bb * SCALE), use_cont_strats, max_raises)
        self._finalizer = weakref.finalize(self, self._handles.release)
        self._view = _TreeView(self._tree)
        self._install_ranges()
        self._max_hands = max(len(h) for h in self.player_hands)
        self._leaf_arr = self._leaf_buffer(leaf_values, leaf_value_fn, int(pot_bb * SCALE))
        if self._leaf_arr is not None:
            self._tree.leaf_values = self._leaf_arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        self._frozen_arr = None
        self._report = SSReport()

    # -- construction helpers -------------------------------------------------
    def _build_tree(self, pot, stack, use_cont_strats, max_raises):
        nb = self.num_board
        args = [(ctypes.c_int * nb)(*self.board_cards), nb, pot, stack,
                (ctypes.c_float * max(1, len(self.bet_sizes)))(*self.bet_sizes), len(self.bet_sizes),
                self.num_players, (ctypes.c_int * len(self.acting_order))(*self.acting_order),
                int(self.is_river), int(bool(use_cont_strats) and not self.is_river)]
        if max_raises is None:
            rc = self.lib.ss_build_tree(*args, ctypes.byref(self._tree))
        else:
            rc = self.lib.ss_build_tree_ex(*args, int(max_raises), ctypes.byref(self._tree))
        if rc != 0:
            raise RuntimeError(f"ss_build_tree failed: {ss_abi.last_error()}")
        self._handles.tree_built = True

    def _install_ranges(self):
        """Hands and weights as int32 [nh, 2] / float32 [nh] arrays; in legacy
        mode also written into the struct's fixed 200-combo slots, as the
        reference wrapper fills them after the build (:286-291)."""
        self._hand_arrs, self._weight_arrs = [], []
        for p, hs in enumerate(self.player_hands):
            cards_arr = np.array([h[:2] for h in hs], dtype=np.int32).reshape(-1, 2)
            w_arr = np.array([h[2] for h in hs], dtype=np.float32)
            self._hand_arrs.append(cards_arr)
            self._weight_arrs.append(w_arr)
            if self.extended:
                continue
            n = len(hs)
            self._tree.num_hands[p] = n
            slot_c = np.ctypeslib.as_array(self._tree.hands[p])          # [200, 2] view
            slot_w = np.ctypeslib.as_array(self._tree.weights[p])        # [200] view
            slot_c[:n] = cards_arr
            slot_w[:n] = w_arr

    def _leaf_buffer(self, leaf_values, leaf_value_fn, pot):
        L = self._tree.num_leaves
        if L <= 0:
            return None
        if leaf_values is None and leaf_value_fn is not None:
            leaf_values = leaf_value_fn(self._tree, self.player_hands, self._max_hands, pot)
        want = L * self.num_players * self._max_hands
        if leaf_values is not None:
            flat = np.ascontiguousarray(np.asarray(leaf_values, dtype=np.float32).reshape(-1))
            if flat.size != want:
                raise ValueError(f"leaf_values mismatch: {flat.size} vs {want} "
                                 f"(leaves={L}, players={self.num_players}, max_hands={self._max_hands})")
            return flat
        if not self.is_river:
            print("[WARNING] Zero leaf values — no blueprint data.", file=sys.stderr)
            return np.zeros(want, np.float32)
        return None

    # -- A3 freezing (python/street_solver_gpu.py:327-366) --------------------
    def set_frozen_actions(self, frozen_actions, hero_player):
        frozen = np.full(self._view.num_nodes, -1, dtype=np.int32)
        for node, act in self._view.follow(frozen_actions, _exact_label):
            if int(self._view.actor[node]) == hero_player:
                frozen[node] = act
        self._frozen_arr = frozen
        self._tree.frozen_action = frozen.ctypes.data_as(ctypes.POINTER(ctypes.c_int))

    # -- solving ----------------------------------------------------------------
    def _problem(self):
        prob = SSProblem()
        prob.tree = ctypes.pointer(self._tree)
        for p in range(self.num_players):
            prob.num_hands[p] = len(self._hand_arrs[p])
            prob.hands[p] = self._hand_arrs[p].ctypes.data_as(ctypes.POINTER(ctypes.c_int))
            prob.weights[p] = self._weight_arrs[p].ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        if self._leaf_arr is not None:
            prob.leaf_values = self._leaf_arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        if self._frozen_arr is not None:
            prob.frozen_action = self._frozen_arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        return prob

    def solve(self, iterations=200):
        if self._freed:
            raise RuntimeError("solver already freed")
        self._handles.drop_output()       # a second solve replaces the first output
        self._sol, self._solved = None, False
        if self.resident:
            rc = self._solve_resident(int(iterations))
        elif self.extended:
            cfg = SSSolveConfig()
            cfg.max_iterations = int(iterations)
            rc = self.lib.ss_solve_gpu_ex(ctypes.byref(self._problem()), ctypes.byref(cfg),
                                          ctypes.byref(self._output), ctypes.byref(self._report))
        else:
            rc = self.lib.ss_solve_gpu(ctypes.byref(self._tree), int(iterations), ctypes.byref(self._output))
        if rc != 0:
            raise RuntimeError(f"ss_solve_gpu failed: {ss_abi.last_error()}")
        self._handles.output_live = True
        self._sol = _Solution(self._output, self._view, [len(h) for h in self.player_hands],
                              self._max_hands)
        self._solved = True

    def _solve_resident(self, iterations):
        """Solve as a one-subgame batch session whose state stays on the GPU
        (ss_batch_create / ss_batch_run / ss_batch_output) until the next
        solve or free(), so later steps can read it there."""
        cfg = SSSolveConfig()
        cfg.max_iterations = iterations
        prob = self._problem()
        batch = self.lib.ss_batch_create(ctypes.byref(prob), 1, ctypes.byref(cfg))
        if not batch:
            return -1
        self._handles.batch = batch
        rc = self.lib.ss_batch_run(batch, ctypes.byref(self._report), ctypes.byref(ss_abi.SSBatchStats()))
        if rc == 0:
            rc = self.lib.ss_batch_output(batch, 0, ctypes.byref(self._output))
        return rc

    def narrow_range(self, narrower, narrower_player, player, action_label, threshold=OUTPUT_THRESHOLD):
        """RangeNarrower.update(narrower_player, action_label,
        self.get_avg_strategy_probs(player, action_label)) computed on the GPU
        from this solve's resident strategy sums (resident=True), with the
        same result bit for bit; see RangeNarrower.update_from_solution."""
        sol = self._require_solution()
        if not self._handles.batch:
            raise RuntimeError("narrow_range needs a resident solve (StreetSolverGPU(..., resident=True))")
        pi = self._resolve_player(player)
        row = sol.first_row_of_player.get(pi)
        if row is None:
            raise ValueError(f"player {pi} has no decision node")
        node = int(sol.node_ids[row])
        names = self._view.labels(node)
        hits = [a for a, nm in enumerate(names) if nm == action_label]
        if len(hits) > 1:
            raise ValueError(f"label {action_label!r} names {len(hits)} actions at node {node}")
        action = hits[0] if hits else -1
        # hand h of the narrowed range -> index of the same (c0, c1) key in
        # this solve's range (keys are (low, high) as get_avg_strategy builds them)
        keys = {}
        for i, h in enumerate(self.player_hands[pi]):
            k = (min(h[0], h[1]), max(h[0], h[1]))
            if k in keys:
                raise ValueError(f"player {pi}'s range repeats combo {k}")
            keys[k] = i
        return narrower._device_update(narrower_player, action_label, self._handles.batch, node, action, keys,
                                       threshold)

    @property
    def exploitability(self):
        """FP64 exploitability in % of the starting pot (extended mode only)."""
        return self._report.exploitability if self.extended else None

    # -- strategy access (python/street_solver_gpu.py:375-553) ----------------
    def _require_solution(self):
        if not self._solved:
            raise RuntimeError("Must call solve() first")
        return self._sol

    def _row_strategy(self, row, hand_idx):
        if row is None:
            return {}
        node = int(self._sol.node_ids[row])
        return _as_dict(self._view.labels(node), self._sol.avg[row, :int(self._view.nkids[node]), hand_idx])

    def get_strategy(self, player, hand_str=None, hand_idx=None, frozen_actions=None):
        """Strategy {label: frequency} of one hand at the player's current
        decision: with no action history, the root's last-iteration strategy
        when the player acts first, else the player's first decision node's
        average strategy; with a history, the average strategy at the node the
        history leads to if the player acts there, else the first one."""
        sol = self._require_solution()
        pi = self._resolve_player(player)
        if hand_idx is None:
            hand_idx = self._find_hand(pi, hand_str)
        if not frozen_actions:
            if sol.root_player == pi:
                return _as_dict(self._view.labels(0), sol.root[:, hand_idx])
            return self._row_strategy(sol.first_row_of_player.get(pi), hand_idx)
        node = self._view.walk_to(frozen_actions, _loose_label)
        if int(self._view.kind[node]) == ss_abi.SS_NODE_DECISION and int(self._view.actor[node]) == pi:
            return self._row_strategy(sol.row_of_node.get(node), hand_idx)
        return self._row_strategy(sol.first_row_of_player.get(pi), hand_idx)

    @staticmethod
    def _match_action_label(tree_label, frozen_label):
        return _loose_label(tree_label, frozen_label)

    def get_all_strategies(self, player):
        """{hand index: strategy} for every hand whose strategy is non-empty."""
        self._require_solution()
        pi = self._resolve_player(player)
        every = ((h, self.get_strategy(pi, hand_idx=h)) for h in range(len(self.player_hands[pi])))
        return {h: s for h, s in every if s}

    def get_avg_strategy(self, player):
        """Average strategy at the player's first decision node, keyed by the
        (low card, high card) combo: {(c0, c1): {label: frequency}}."""
        sol = self._require_solution()
        pi = self._resolve_player(player)
        row = sol.first_row_of_player.get(pi)
        if row is None:
            return {}
        return {(min(h[0], h[1]), max(h[0], h[1])): self._row_strategy(row, i)
                for i, h in enumerate(self.player_hands[pi])}

    def get_avg_strategy_probs(self, player, action_label):
        """P(action | hand) for every hand, 0 where the action is below the
        threshold (the input of range narrowing)."""
        return {combo: s.get(action_label, 0.0) for combo, s in self.get_avg_strategy(player).items()}

    def _resolve_player(self, player):
        if isinstance(player, int):
            return player
        return _SEAT_NAMES[player] if player in _SEAT_NAMES else int(player)

    def _find_hand(self, player_idx, hand_str):
        if self._index_of[player_idx] is None:
            index = {}
            for i, h in enumerate(self.player_hands[player_idx]):
                index.setdefault((min(h[0], h[1]), max(h[0], h[1])), i)
            self._index_of[player_idx] = index
        a, b = card_to_int(hand_str[:2]), card_to_int(hand_str[2:])
        i = self._index_of[player_idx].get((min(a, b), max(a, b)))
        if i is None:
            raise ValueError(f"Hand {hand_str} not in player {player_idx}'s range")
        return i

    # -- lifetime ---------------------------------------------------------------
    def free(self):
        """Release the native tree and output (idempotent; also runs when the
        object is garbage-collected)."""
        self._freed = True
        self._sol = None
        self._solved = False
        fin = self.__dict__.get("_finalizer")
        if fin is not None:
            fin()
