"""StreetSolverGPU -- drop-in for python/street_solver_gpu.py on MI355X.

Same constructor, methods, action labels and output thresholds as the
reference class (python/street_solver_gpu.py:173-596), backed by
libstreet_solve.so through the legacy C-ABI (ss_build_tree / ss_solve_gpu /
ss_free_tree / ss_free_output). Behaviour kept from the reference:
  - board-blocked combos are dropped, ranges above 200 combos are trimmed to
    the 200 heaviest (stable by weight) (:224-229);
  - pot / stack in BB are converted with SCALE = 100 and truncated (:263-264);
  - continuation strategies only off the river (:272);
  - leaf values: explicit array, else leaf_value_fn(tree, hands, max_h, pot),
    else zeros with a warning off the river (:296-319);
  - labels "Fold" / "Check" / "Call" / "Bet {pct}%" / "Raise {pct}%" (:128-169)
    and the 0.001 output threshold (:418, :484).
Difference: ss_free_tree here never frees the caller-owned leaf-value or
frozen-action buffers (the reference frees leaf_values at
src/cuda/street_solve.cu:1650, a buffer this class owns).

extended=True lifts the 200-combo cap (up to 1326) and uses ss_solve_gpu_ex.
"""

import ctypes
import sys

import numpy as np

import ss_abi
from cards import card_to_int, int_to_card, parse_range_string  # noqa: F401
from ss_abi import SSOutput, SSProblem, SSReport, SSSolveConfig, SSTreeData

SS_MAX_PLAYERS = ss_abi.SS_MAX_PLAYERS
SS_MAX_HANDS = ss_abi.SS_MAX_HANDS
SS_MAX_ACTIONS = ss_abi.SS_MAX_ACTIONS
SS_MAX_BOARD = ss_abi.SS_MAX_BOARD
SCALE = 100


def _get_lib():
    return ss_abi.load_library()


def _build_labels_at_node(tree_data, node_idx):
    """Action labels of a decision node (python/street_solver_gpu.py:128-169)."""
    node = tree_data.nodes[node_idx]
    acting = node.player
    acting_bet = node.bets[acting] if acting >= 0 else 0
    top = max(node.bets[p] for p in range(SS_MAX_PLAYERS))
    to_call = top - acting_bet if acting >= 0 else 0
    labels = []
    for i in range(node.num_children):
        child = tree_data.nodes[tree_data.children[node.first_child + i]]
        if child.type == ss_abi.SS_NODE_FOLD:
            labels.append("Fold")
        elif child.type in (ss_abi.SS_NODE_SHOWDOWN, ss_abi.SS_NODE_LEAF):
            labels.append("Call" if to_call > 0 else "Check")
        elif child.type == ss_abi.SS_NODE_DECISION:
            amount = (child.bets[acting] if acting >= 0 else 0) - acting_bet
            if amount == 0:
                labels.append("Check")
            elif node.pot > 0:
                pct = int(round(100 * amount / node.pot))
                labels.append(f"Raise {pct}%" if to_call > 0 else f"Bet {pct}%")
            else:
                labels.append("Bet" if to_call == 0 else "Raise")
        else:
            labels.append(f"Action {i}")
    return labels


class StreetSolverGPU:
    """N-player single-street Linear-CFR solver on the GPU (2-6 players)."""

    def __init__(self, board, pot_bb, stack_bb, oop_range=None, ip_range=None, player_ranges=None,
                 acting_order=None, bet_sizes=None, leaf_values=None, use_cont_strats=True,
                 leaf_value_fn=None, extended=False, max_raises=None):
        self.lib = _get_lib()
        self.extended = bool(extended)
        self.board_cards = [card_to_int(c) for c in board] if isinstance(board[0], str) else list(board)
        self.num_board = len(self.board_cards)
        self.is_river = self.num_board == 5
        blocked = set(self.board_cards)
        if player_ranges is not None:
            ranges = list(player_ranges)
        elif oop_range is not None and ip_range is not None:
            ranges = [oop_range, ip_range]
        else:
            raise ValueError("Provide oop_range+ip_range or player_ranges")
        cap = ss_abi.SS_EX_MAX_HANDS if self.extended else SS_MAX_HANDS
        self.num_players = len(ranges)
        self.player_hands = []
        for r in ranges:
            hands = parse_range_string(r) if isinstance(r, str) else list(r)
            hands = [(c0, c1, w) for c0, c1, w in hands if c0 not in blocked and c1 not in blocked]
            if len(hands) > cap:
                hands.sort(key=lambda x: -x[2])
                hands = hands[:cap]
            self.player_hands.append(hands)
        if self.num_players >= 1:
            self.oop_hands = self.player_hands[0]
        if self.num_players >= 2:
            self.ip_hands = self.player_hands[1]
        self.bet_sizes = [0.33, 0.75, 1.5] if bet_sizes is None else list(bet_sizes)
        self.acting_order = list(range(self.num_players)) if acting_order is None else list(acting_order)
        pot, stack = int(pot_bb * SCALE), int(stack_bb * SCALE)

        self._tree = SSTreeData()
        use_cs = 1 if (use_cont_strats and not self.is_river) else 0
        barr = (ctypes.c_int * self.num_board)(*self.board_cards)
        sarr = (ctypes.c_float * max(1, len(self.bet_sizes)))(*self.bet_sizes)
        oarr = (ctypes.c_int * len(self.acting_order))(*self.acting_order)
        if max_raises is None:
            err = self.lib.ss_build_tree(barr, self.num_board, pot, stack, sarr, len(self.bet_sizes),
                                         self.num_players, oarr, 1 if self.is_river else 0, use_cs,
                                         ctypes.byref(self._tree))
        else:
            err = self.lib.ss_build_tree_ex(barr, self.num_board, pot, stack, sarr, len(self.bet_sizes),
                                            self.num_players, oarr, 1 if self.is_river else 0, use_cs,
                                            int(max_raises), ctypes.byref(self._tree))
        if err != 0:
            raise RuntimeError(f"ss_build_tree failed: {ss_abi.last_error()}")

        # hands and weights (legacy struct slots hold the first 200; the
        # extended path passes arrays)
        self._hand_arrs, self._weight_arrs = [], []
        for p in range(self.num_players):
            hs = self.player_hands[p]
            self._hand_arrs.append(np.array([[c0, c1] for c0, c1, _ in hs], dtype=np.int32).reshape(-1, 2))
            self._weight_arrs.append(np.array([w for _, _, w in hs], dtype=np.float32))
            if not self.extended:
                self._tree.num_hands[p] = len(hs)
                for i, (c0, c1, w) in enumerate(hs):
                    self._tree.hands[p][i][0] = c0
                    self._tree.hands[p][i][1] = c1
                    self._tree.weights[p][i] = w
        max_h = max(len(h) for h in self.player_hands)
        self._max_hands = max_h

        if leaf_values is None and leaf_value_fn is not None and self._tree.num_leaves > 0:
            leaf_values = leaf_value_fn(self._tree, self.player_hands, max_h, pot)
        self._leaf_arr = None
        if leaf_values is not None and self._tree.num_leaves > 0:
            expected = self._tree.num_leaves * self.num_players * max_h
            flat = np.ascontiguousarray(np.asarray(leaf_values, dtype=np.float32).ravel())
            if len(flat) != expected:
                raise ValueError(f"leaf_values mismatch: {len(flat)} vs {expected} "
                                 f"(leaves={self._tree.num_leaves}, players={self.num_players}, "
                                 f"max_hands={max_h})")
            self._leaf_arr = flat
        elif not self.is_river and self._tree.num_leaves > 0:
            print("[WARNING] Zero leaf values — no blueprint data.", file=sys.stderr)
            self._leaf_arr = np.zeros(self._tree.num_leaves * self.num_players * max_h, np.float32)
        if self._leaf_arr is not None:
            self._tree.leaf_values = self._leaf_arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

        self._frozen_arr = None
        self._output = SSOutput()
        self._report = SSReport()
        self._solved = False
        self._freed = False

    # -- A3 freezing (python/street_solver_gpu.py:327-366)
    def set_frozen_actions(self, frozen_actions, hero_player):
        n = self._tree.num_nodes
        frozen = np.full(n, -1, dtype=np.int32)
        cur = 0
        for label in frozen_actions:
            node = self._tree.nodes[cur]
            if node.type != ss_abi.SS_NODE_DECISION:
                break
            labels = _build_labels_at_node(self._tree, cur)
            hit = next((i for i, l in enumerate(labels) if l.lower() == label.lower()), -1)
            if hit < 0:
                break
            if node.player == hero_player:
                frozen[cur] = hit
            cur = self._tree.children[node.first_child + hit]
        self._frozen_arr = frozen
        self._tree.frozen_action = frozen.ctypes.data_as(ctypes.POINTER(ctypes.c_int))

    def solve(self, iterations=200):
        if self.extended:
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
            cfg = SSSolveConfig()
            cfg.max_iterations = int(iterations)
            err = self.lib.ss_solve_gpu_ex(ctypes.byref(prob), ctypes.byref(cfg), ctypes.byref(self._output),
                                           ctypes.byref(self._report))
        else:
            err = self.lib.ss_solve_gpu(ctypes.byref(self._tree), int(iterations), ctypes.byref(self._output))
        if err != 0:
            raise RuntimeError(f"ss_solve_gpu failed: {ss_abi.last_error()}")
        self._solved = True

    @property
    def exploitability(self):
        """FP64 exploitability in % of the starting pot (extended mode only)."""
        return self._report.exploitability if self.extended else None

    # -- strategy access (python/street_solver_gpu.py:375-553)
    def get_strategy(self, player, hand_str=None, hand_idx=None, frozen_actions=None):
        if not self._solved:
            raise RuntimeError("Must call solve() first")
        pi = self._resolve_player(player)
        if hand_idx is None:
            hand_idx = self._find_hand(pi, hand_str)
        if not frozen_actions:
            if self._output.root_player == pi:
                nh = len(self.player_hands[pi])
                labels = _build_labels_at_node(self._tree, 0)
                out = {}
                for a in range(min(self._output.root_num_actions, len(labels))):
                    f = self._output.root_strategy[a * nh + hand_idx]
                    if f > 0.001:
                        out[labels[a]] = float(f)
                return out
            return self._get_strategy_at_player_node(pi, hand_idx)
        cur, k = 0, 0
        while k < len(frozen_actions) and cur < self._tree.num_nodes:
            node = self._tree.nodes[cur]
            if node.type != ss_abi.SS_NODE_DECISION:
                break
            labels = _build_labels_at_node(self._tree, cur)
            want = frozen_actions[k]
            hit = next((c for c in range(node.num_children)
                        if c < len(labels) and self._match_action_label(labels[c], want)), -1)
            if hit < 0:
                break
            cur = self._tree.children[node.first_child + hit]
            k += 1
        node = self._tree.nodes[cur]
        if node.type == ss_abi.SS_NODE_DECISION and node.player == pi:
            return self._get_strategy_at_node(cur, pi, hand_idx)
        return self._get_strategy_at_player_node(pi, hand_idx)

    @staticmethod
    def _match_action_label(tree_label, frozen_label):
        return tree_label.lower().strip() == frozen_label.lower().strip()

    def _avg_row(self, di, na, labels, hand_idx):
        mh = self._max_hands
        out = {}
        for a in range(min(na, len(labels))):
            f = self._output.avg_strategies[di * SS_MAX_ACTIONS * mh + a * mh + hand_idx]
            if f > 0.001:
                out[labels[a]] = float(f)
        return out

    def _get_strategy_at_node(self, node_idx, player_idx, hand_idx):
        for di in range(self._output.num_avg_nodes):
            nid = self._output.avg_strategy_node_ids[di]
            if nid != node_idx:
                continue
            node = self._tree.nodes[nid]
            if node.player != player_idx or node.type != ss_abi.SS_NODE_DECISION:
                continue
            return self._avg_row(di, node.num_children, _build_labels_at_node(self._tree, nid), hand_idx)
        return {}

    def _get_strategy_at_player_node(self, player_idx, hand_idx):
        for di in range(self._output.num_avg_nodes):
            nid = self._output.avg_strategy_node_ids[di]
            node = self._tree.nodes[nid]
            if node.player != player_idx or node.type != ss_abi.SS_NODE_DECISION:
                continue
            return self._avg_row(di, node.num_children, _build_labels_at_node(self._tree, nid), hand_idx)
        return {}

    def get_all_strategies(self, player):
        if not self._solved:
            raise RuntimeError("Must call solve() first")
        pi = self._resolve_player(player)
        out = {}
        for h in range(len(self.player_hands[pi])):
            s = self.get_strategy(pi, hand_idx=h)
            if s:
                out[h] = s
        return out

    def get_avg_strategy(self, player):
        if not self._solved:
            raise RuntimeError("Must call solve() first")
        pi = self._resolve_player(player)
        for di in range(self._output.num_avg_nodes):
            nid = self._output.avg_strategy_node_ids[di]
            node = self._tree.nodes[nid]
            if node.type != ss_abi.SS_NODE_DECISION or node.player != pi:
                continue
            labels = _build_labels_at_node(self._tree, nid)
            out = {}
            for h, (c0, c1, _) in enumerate(self.player_hands[pi]):
                out[(min(c0, c1), max(c0, c1))] = self._avg_row(di, node.num_children, labels, h)
            return out
        return {}

    def get_avg_strategy_probs(self, player, action_label):
        return {k: v.get(action_label, 0.0) for k, v in self.get_avg_strategy(player).items()}

    def _resolve_player(self, player):
        if isinstance(player, int):
            return player
        if player == "oop":
            return 0
        if player == "ip":
            return 1
        return int(player)

    def _find_hand(self, player_idx, hand_str):
        c0, c1 = card_to_int(hand_str[:2]), card_to_int(hand_str[2:])
        key = (min(c0, c1), max(c0, c1))
        for i, (h0, h1, _) in enumerate(self.player_hands[player_idx]):
            if (min(h0, h1), max(h0, h1)) == key:
                return i
        raise ValueError(f"Hand {hand_str} not in player {player_idx}'s range")

    def free(self):
        if getattr(self, "_freed", False):
            return
        self._freed = True
        tree = getattr(self, "_tree", None)
        if tree is not None and getattr(tree, "nodes", None):
            try:
                self.lib.ss_free_tree(ctypes.byref(tree))
            except Exception:
                pass
        out = getattr(self, "_output", None)
        if out is not None and getattr(self, "_solved", False):
            try:
                self.lib.ss_free_output(ctypes.byref(out))
            except Exception:
                pass

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
