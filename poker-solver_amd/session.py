"""Hand sessions in front of the served solver: the request / session wire format.

The reference plans a session-based API (CONTEXT_NEXT_SESSION.md:21-61):
server-side hand sessions that carry ranges, pot, stacks and board across
streets; 169 hand classes expanded to 1326 combos on create; Bayesian range
narrowing on every action (RangeNarrower.update with the last solve's
average strategy); A3 strategy freezing of the actions taken on the current
street, reset when the street changes (python/hud_solver.py:133, 339, 544,
862-864); 30-minute expiry and at most 100 sessions. Its routes are

    POST   /api/session               create
    POST   /api/session/{id}/deal     add board cards, solve the new street
    POST   /api/session/{id}/action   a player acts: narrowing, pot / stacks
    POST   /api/session/{id}/solve    re-solve the street with the actions so far frozen
    GET    /api/session/{id}          state
    DELETE /api/session/{id}          end

SessionStore.handle(method, path, body) serves exactly these with JSON-ready
dicts and HTTP status codes; it does no networking (a web framework maps its
routes onto handle()). Every solve is submitted to one shared continuous-
batching server (ss_server_*, include/street_solve.h), so the solves of
concurrent sessions share the GPU. Bodies:

  create  {"positions": ["BTN", "BB"], "ranges": {pos: {"AKs": 1.0, ...} or
           [[c0, c1, w], ...]}, "pot_bb": 6.5, "stacks_bb": 97.5 or {pos: bb},
           "bet_sizes": [0.33, 0.75, 1.5], "iterations": 1000, "max_raises": 2}
          -> {"session_id", "state"}
  deal    {"cards": ["Qs", "As", "2d"]} -> {"state", "solve"}
  action  {"position": "BB", "action": "Bet 75%"} -> {"state", "narrowed": bool}
  solve   {} -> {"state", "solve"}
  "solve" = {"exploitability_pct_pot", "iterations", "to_act", "strategies":
             {position: {"AhKs": {"Check": 0.6, "Bet 75%": 0.4}, ...}}}:
             each active player's average strategy at its first decision node
             of the street's tree (the node the reference narrows with), and
             "to_act" the strategy at the current node for the player to act.
"""

import threading
import time
import uuid

import numpy as np

import cards as _cards
import leaf_values as _lv
import range_narrowing as rn
import ss_api
from street_solver_gpu import OUTPUT_THRESHOLD, SCALE, _TreeView, _loose_label

RANKS = _cards.RANKS
SESSION_TTL_S = 30 * 60
MAX_SESSIONS = 100
STREET_OF_BOARD = {0: "preflop", 3: "flop", 4: "turn", 5: "river"}


# ---------------------------------------------------------------- 169 <-> 1326
def hand_classes():
    """The 169 preflop classes, strongest rank first: "AA", "AKs", "AKo", ..."""
    out = []
    for i in range(12, -1, -1):
        for j in range(i, -1, -1):
            if i == j:
                out.append(RANKS[i] * 2)
            else:
                out += [RANKS[i] + RANKS[j] + "s", RANKS[i] + RANKS[j] + "o"]
    return out


def class_of(c0, c1):
    """Class name of a two-card combo (card = rank * 4 + suit)."""
    r0, r1 = c0 >> 2, c1 >> 2
    hi, lo = max(r0, r1), min(r0, r1)
    if hi == lo:
        return RANKS[hi] * 2
    return RANKS[hi] + RANKS[lo] + ("s" if (c0 & 3) == (c1 & 3) else "o")


def expand_classes(weights):
    """{class: weight} -> [(c0, c1, weight)] over all 1326 combos (c0 < c1, in
    generate_all_hands order); classes not named get weight 0."""
    return [(c0, c1, float(weights.get(class_of(c0, c1), 0.0))) for c0, c1, _ in rn.generate_all_hands()]


def collapse_combos(hands):
    """[(c0, c1, w)] -> {class: mean weight of its listed combos} (the
    169-class view of a narrowed range)."""
    acc = {}
    for c0, c1, w in hands:
        s = acc.setdefault(class_of(c0, c1), [0.0, 0])
        s[0] += w
        s[1] += 1
    return {k: v[0] / v[1] for k, v in acc.items()}


class SessionError(Exception):
    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


class HandSession:
    """One hand: positions in acting order, narrowed ranges, pot, stacks,
    board, and this street's actions."""

    def __init__(self, positions, ranges, pot_bb, stacks_bb, bet_sizes, iterations, max_raises):
        if len(positions) < 2 or len(positions) > 6 or len(set(positions)) != len(positions):
            raise SessionError(400, "positions: 2-6 distinct names")
        self.id = uuid.uuid4().hex
        self.positions = list(positions)
        self.active = {p: True for p in positions}
        self.narrower = rn.RangeNarrower()
        for p in positions:
            r = ranges.get(p)
            if r is None:
                raise SessionError(400, f"no range for {p}")
            if isinstance(r, dict):
                self.narrower.set_initial_range(p, expand_classes(r))
            else:
                self.narrower.set_initial_range(p, [(int(a), int(b), float(w)) for a, b, w in r])
        self.pot_bb = float(pot_bb)
        st = stacks_bb if isinstance(stacks_bb, dict) else {p: stacks_bb for p in positions}
        self.stacks_bb = {p: float(st[p]) for p in positions}
        self.bet_sizes = list(bet_sizes)
        self.iterations = int(iterations)
        self.max_raises = int(max_raises)
        self.board = []
        self.street_bets = {p: 0.0 for p in positions}
        self.actions = []            # this street: (position, label) in order
        self.history = []            # every street: (street, position, label)
        self.last = None             # last solve of this street: dict (see _solve)
        self.touched = time.monotonic()

    @property
    def street(self):
        return STREET_OF_BOARD.get(len(self.board), "invalid")

    def players(self):
        """Active positions in acting order (the solver's player indices)."""
        return [p for p in self.positions if self.active[p]]

    def state(self):
        return {"session_id": self.id, "street": self.street, "board": list(self.board),
                "pot_bb": self.pot_bb, "stacks_bb": dict(self.stacks_bb),
                "active": [p for p in self.positions if self.active[p]],
                "actions_this_street": [list(a) for a in self.actions],
                "history": [list(h) for h in self.history],
                "range_sizes": {p: self.narrower.get_hand_count(p) for p in self.positions}}


class SessionStore:
    """In-memory sessions; one shared ss_api.Server runs every solve."""

    def __init__(self, server=None, ttl_s=SESSION_TTL_S, max_sessions=MAX_SESSIONS, target_pct_pot=0.0,
                 clock=time.monotonic):
        self._server = server
        self._own_server = server is None
        self.ttl_s = ttl_s
        self.max_sessions = max_sessions
        self.target = float(target_pct_pot)
        self.clock = clock
        self.sessions = {}
        self.lock = threading.Lock()

    @property
    def server(self):
        if self._server is None:
            self._server = ss_api.Server(max_slots=256, iters_per_step=10, check_every=50)
        return self._server

    def close(self):
        if self._own_server and self._server is not None:
            self._server.close()
            self._server = None

    # -- routing -------------------------------------------------------------
    def handle(self, method, path, body=None):
        """(status, response dict) for one request of the session API."""
        parts = [p for p in path.strip("/").split("/") if p]
        try:
            if parts[:2] != ["api", "session"]:
                raise SessionError(404, f"no route {path}")
            rest = parts[2:]
            if method == "POST" and not rest:
                return 201, self.create(**(body or {}))
            if not rest:
                raise SessionError(405, f"{method} {path}")
            sid = rest[0]
            if method == "GET" and len(rest) == 1:
                return 200, {"state": self.get(sid).state()}
            if method == "DELETE" and len(rest) == 1:
                self.delete(sid)
                return 200, {"deleted": sid}
            if method == "POST" and len(rest) == 2 and rest[1] == "deal":
                return 200, self.deal(sid, **(body or {}))
            if method == "POST" and len(rest) == 2 and rest[1] == "action":
                return 200, self.action(sid, **(body or {}))
            if method == "POST" and len(rest) == 2 and rest[1] == "solve":
                return 200, self.solve(sid)
            raise SessionError(404, f"no route {method} {path}")
        except SessionError as e:
            return e.status, {"error": str(e)}
        except TypeError as e:        # unknown or missing body fields
            return 400, {"error": str(e)}

    # -- store -----------------------------------------------------------------
    def _purge(self):
        now = self.clock()
        for sid in [s for s, x in self.sessions.items() if now - x.touched > self.ttl_s]:
            del self.sessions[sid]

    def get(self, sid):
        with self.lock:
            self._purge()
            s = self.sessions.get(sid)
            if s is None:
                raise SessionError(404, f"no session {sid}")
            s.touched = self.clock()
            return s

    def delete(self, sid):
        with self.lock:
            if self.sessions.pop(sid, None) is None:
                raise SessionError(404, f"no session {sid}")

    def create(self, positions, ranges, pot_bb, stacks_bb, bet_sizes=(0.33, 0.75, 1.5), iterations=1000,
               max_raises=2):
        s = HandSession(positions, ranges, pot_bb, stacks_bb, bet_sizes, iterations, max_raises)
        s.touched = self.clock()
        with self.lock:
            self._purge()
            if len(self.sessions) >= self.max_sessions:
                raise SessionError(429, f"{self.max_sessions} sessions open")
            self.sessions[s.id] = s
        return {"session_id": s.id, "state": s.state()}

    # -- hand flow -----------------------------------------------------------------
    def deal(self, sid, cards):
        s = self.get(sid)
        try:
            new = [_cards.card_to_int(c) if isinstance(c, str) else int(c) for c in cards]
        except ValueError as e:
            raise SessionError(400, str(e))
        board = [_cards.card_to_int(c) for c in s.board] + new
        if len(set(board)) != len(board) or len(board) not in (3, 4, 5) or not new:
            raise SessionError(400, f"cannot deal {cards} onto {s.board}")
        if len(s.players()) < 2:
            raise SessionError(409, "the hand is over")
        s.board = [_cards.int_to_card(c) for c in board]
        # a new street: its own bets, actions and frozen set (hud_solver.py:543-544)
        s.street_bets = {p: 0.0 for p in s.positions}
        s.actions = []
        s.last = None
        return {"state": s.state(), "solve": self._solve(s)}

    def action(self, sid, position, action):
        s = self.get(sid)
        if s.street in ("preflop", "invalid"):
            raise SessionError(409, "deal a board first")
        if position not in s.active or not s.active[position]:
            raise SessionError(400, f"{position} is not in the hand")
        label = str(action)
        narrowed = False
        # narrowing with the last solve's average strategy at the player's
        # first decision node (hud_solver.py:307-343, street_solver_gpu.py:550-553)
        if s.last is not None and position in s.last["players"]:
            probs = self._probs(s.last, s.last["players"].index(position), label)
            if probs is not None:
                s.narrower.update(position, label, probs)
                narrowed = True
        self._account(s, position, label)
        s.actions.append((position, label))
        s.history.append((s.street, position, label))
        return {"state": s.state(), "narrowed": narrowed}

    def solve(self, sid):
        s = self.get(sid)
        if s.street in ("preflop", "invalid"):
            raise SessionError(409, "deal a board first")
        return {"state": s.state(), "solve": self._solve(s)}

    # -- internals ------------------------------------------------------------------
    @staticmethod
    def _account(s, position, label):
        """Pot and stacks after an action named by its tree label: "Fold" leaves
        the hand, "Check" adds nothing, "Call" matches the street's largest
        bet, "Bet x%" / "Raise x%" grow the player's bet by x% of the pot."""
        word = label.split()[0].lower()
        top = max(s.street_bets.values())
        grow = 0.0
        if word == "fold":
            s.active[position] = False
        elif word == "call":
            grow = top - s.street_bets[position]
        elif word in ("bet", "raise"):
            try:
                pct = float(label.split()[1].rstrip("%"))
            except (IndexError, ValueError):
                raise SessionError(400, f"cannot size {label!r}")
            grow = pct / 100.0 * s.pot_bb
        elif word != "check":
            raise SessionError(400, f"unknown action {label!r}")
        grow = min(grow, s.stacks_bb[position])
        s.street_bets[position] += grow
        s.stacks_bb[position] -= grow
        s.pot_bb += grow

    def _problem(self, s):
        """Tree and ranges of the current street for the active players (in
        acting order), hands blocked by the board dropped, A3 freezing of every
        action taken on this street."""
        players = s.players()
        board = [_cards.card_to_int(c) for c in s.board]
        blocked = set(board)
        hands, weights = [], []
        for p in players:
            hs = [(a, b, w) for a, b, w in s.narrower.get_weighted_hands(p) if a not in blocked and b not in blocked]
            if not hs:
                raise SessionError(409, f"{p}'s range is empty on this board")
            hands.append(np.array([(a, b) for a, b, _ in hs], np.int32))
            weights.append(np.array([w for _, _, w in hs], np.float32))
        stack = int(min(s.stacks_bb[p] for p in players) * SCALE)
        tree = ss_api.Tree(board, int(s.pot_bb * SCALE), stack, s.bet_sizes, len(players),
                           max_raises=s.max_raises)
        view = _TreeView(tree.td)
        frozen = np.full(tree.num_nodes, -1, np.int32)
        labels = [a for _, a in s.actions]
        node = 0
        for who, want in s.actions:
            if node >= view.num_nodes or int(view.kind[node]) != 0:
                break
            names = view.labels(node)
            hit = next((i for i, nm in enumerate(names) if _loose_label(nm, want)), None)
            if hit is None:
                break
            if players[int(view.actor[node])] == who:
                frozen[node] = hit
            node = int(view.children(node)[hit])
        leaf, leaf_src = None, None
        if tree.num_leaves > 0:
            leaf, leaf_src = self._leaf_values(s, tree, board, hands, weights, len(players))
        prob = ss_api.Problem(tree, hands, weights, leaf_values=leaf, frozen=frozen)
        return prob, view, players, node, labels, leaf_src

    @staticmethod
    def _leaf_values(s, tree, board, hands, weights, npl):
        """Depth-limit leaf values of a turn or flop street, as the reference's
        HUD pipeline passes them (hud_solver.py:820-857): the GPU producers of
        leaf_values.py over the tree's leaf infos; when the producer's rows do
        not match the tree's expanded leaves (some leaves with fewer active
        players), zeros, as the reference wrapper uses when it has no leaf
        values (street_solver_gpu.py:315-319)."""
        mh = max(len(h) for h in hands)
        want = tree.num_leaves * npl * mh
        ph = [[(int(a), int(b), float(w)) for (a, b), w in zip(h, ws)] for h, ws in zip(hands, weights)]
        infos = _lv.extract_leaf_info_from_tree(tree.td, npl)
        pot = int(s.pot_bb * SCALE)
        if len(board) == 4:
            out = _lv.compute_turn_leaf_values(board, infos, mh, pot, player_hands=ph)
        else:
            out = _lv.compute_flop_leaf_equity(board, infos, mh, pot, player_hands=ph)
        if out.size == want:
            return np.ascontiguousarray(out.reshape(tree.num_leaves, npl, mh), np.float32), "leaf_values.py"
        return np.zeros((tree.num_leaves, npl, mh), np.float32), "zero"

    def _solve(self, s):
        prob, view, players, node, labels, leaf_src = self._problem(s)
        ticket = self.server.submit(prob, s.iterations, target=self.target)
        res = None
        while res is None:
            res = self.server.poll(ticket)
            if res is None:
                time.sleep(0.0005)
        rows = {int(n): i for i, n in enumerate(res.node_ids)}
        s.last = {"players": players, "res": res, "view": view, "hands": prob.hands, "rows": rows}
        strategies = {}
        for q, p in enumerate(players):
            first = next((int(n) for n in res.node_ids if int(view.actor[int(n)]) == q), None)
            if first is not None:
                strategies[p] = self._strategy_dicts(s.last, first)
        to_act = None
        if node < view.num_nodes and int(view.kind[node]) == 0:
            to_act = {"position": players[int(view.actor[node])], "node": node,
                      "strategy": self._strategy_dicts(s.last, node)}
        return {"exploitability_pct_pot": res.exploitability, "iterations": res.iterations,
                "leaf_values": leaf_src, "to_act": to_act, "strategies": strategies}

    @staticmethod
    def _strategy_dicts(last, node):
        view, res = last["view"], last["res"]
        q = int(view.actor[node])
        names = view.labels(node)
        row = res.avg[last["rows"][node]]
        out = {}
        for h, (a, b) in enumerate(last["hands"][q]):
            p64 = row[:len(names), h].astype(np.float64)
            out[_cards.int_to_card(int(a)) + _cards.int_to_card(int(b))] = {
                names[k]: float(p64[k]) for k in range(len(names)) if p64[k] > OUTPUT_THRESHOLD}
        return out

    @staticmethod
    def _probs(last, q, label):
        """{(c0, c1): P(label | hand)} at solver player q's first decision node,
        0 where the label is absent or at most 0.001 (the reference's
        get_avg_strategy_probs); None when q has no decision node."""
        view, res = last["view"], last["res"]
        first = next((int(n) for n in res.node_ids if int(view.actor[int(n)]) == q), None)
        if first is None:
            return None
        names = view.labels(first)
        row = res.avg[last["rows"][first]]
        hits = [k for k, nm in enumerate(names) if nm == label]
        out = {}
        for h, (a, b) in enumerate(last["hands"][q]):
            v = 0.0
            for k in hits:       # a dict keeps the last label, as get_avg_strategy's does
                x = float(np.float64(row[k, h]))
                if x > OUTPUT_THRESHOLD:
                    v = x
            out[(min(int(a), int(b)), max(int(a), int(b)))] = v
        return out
