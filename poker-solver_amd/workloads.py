"""Synthetic subgame generators for the five benchmark workloads.

Every generator draws from numpy.random.Generator(PCG64(seed)). Hands are all
two-card combos not on the board, in (c0 < c1) lexicographic order -- the
order of python/range_narrowing.py:176-193 (generate_all_hands). Chips use
SCALE = 100 per big blind (python/street_solver_gpu.py:37, 263-264), so
"pot 100" is 10000 chips.

  river_2p      plain 2-player river: weights U[0,1] normalised, pot 100 BB,
                stacks 400 BB, sizes {0.5, 1.0} + all-in, max 2 raises
  river_3p      plain 3-player river: each range 40% nonzero U[0,1],
                normalised, pot 150 BB, stacks 500 BB, sizes {0.33, 0.75}
  batch_mix     heterogeneous rivers: players U{2..max_players}, 1-4 sizes
                from {0.25, 0.5, 0.75, 1.0, 1.5}, density U[0.1, 1.0], pot
                U[50, 300] BB, stacks U[200, 1000] BB, max 2 raises
  served_mix    Poisson stream: 70% rivers as batch_mix, 30% turns (4-card
                board, 1128 combos, 2-3 players) with continuation leaves whose
                values are N(0, 0.3 pot) per (leaf, player, hand)
"""

import numpy as np

SCALE = 100
BASE_SEED = 20261203


def rng_for(seed):
    return np.random.Generator(np.random.PCG64(seed))


def all_combos(board):
    b = set(int(c) for c in board)
    out = [(c0, c1) for c0 in range(52) if c0 not in b for c1 in range(c0 + 1, 52) if c1 not in b]
    return np.array(out, dtype=np.int32)


class Spot:
    """One subgame: board, tree parameters, ranges, optional leaf values."""

    def __init__(self, board, pot, stack, bet_sizes, num_players, hands, weights,
                 max_raises=2, acting_order=None, leaf_sigma=None, seed=0, kind="river"):
        self.board = [int(c) for c in board]
        self.pot = int(pot)
        self.stack = int(stack)
        self.bet_sizes = [float(x) for x in bet_sizes]
        self.num_players = int(num_players)
        self.hands = hands
        self.weights = weights
        self.max_raises = int(max_raises)
        self.acting_order = list(range(num_players)) if acting_order is None else list(acting_order)
        self.leaf_sigma = leaf_sigma      # None, or std-dev of leaf values in chips
        self.seed = seed
        self.kind = kind

    @property
    def is_river(self):
        return len(self.board) == 5

    def leaf_values(self, num_leaves):
        """N(0, leaf_sigma) per (leaf, player, hand), float32 [L, NP, max_h]."""
        if num_leaves == 0 or self.leaf_sigma is None:
            return None
        mh = max(len(h) for h in self.hands)
        r = rng_for(self.seed + 7919)
        return (r.standard_normal((num_leaves, self.num_players, mh)) * self.leaf_sigma).astype(np.float32)


def _ranges(r, board, num_players, density):
    combos = all_combos(board)
    hands, weights = [], []
    for _ in range(num_players):
        w = r.uniform(0.0, 1.0, len(combos))
        if density < 1.0:
            k = int(round(density * len(combos)))
            keep = np.zeros(len(combos), bool)
            keep[r.choice(len(combos), k, replace=False)] = True
            w = np.where(keep, w, 0.0)
        s = w.sum()
        w = w / s if s > 0 else w
        hands.append(combos.copy())
        weights.append(w.astype(np.float32))
    return hands, weights


def river_2p(seed=BASE_SEED + 1000, board=None):
    r = rng_for(seed)
    if board is None:
        board = r.choice(52, 5, replace=False)
    hands, weights = _ranges(r, board, 2, 1.0)
    return Spot(board, 100 * SCALE, 400 * SCALE, [0.5, 1.0], 2, hands, weights, seed=seed)


def river_3p(seed=BASE_SEED + 2000):
    r = rng_for(seed)
    board = r.choice(52, 5, replace=False)
    hands, weights = _ranges(r, board, 3, 0.4)
    return Spot(board, 150 * SCALE, 500 * SCALE, [0.33, 0.75], 3, hands, weights, seed=seed)


SIZE_POOL = [0.25, 0.5, 0.75, 1.0, 1.5]


def mix_river(seed, max_players=6):
    r = rng_for(seed)
    np_ = int(r.integers(2, max_players + 1))
    k = int(r.integers(1, 5))
    sizes = sorted(r.choice(SIZE_POOL, k, replace=False).tolist())
    density = float(r.uniform(0.1, 1.0))
    pot = int(round(r.uniform(50, 300) * SCALE))
    stack = int(round(r.uniform(200, 1000) * SCALE))
    board = r.choice(52, 5, replace=False)
    hands, weights = _ranges(r, board, np_, density)
    return Spot(board, pot, stack, sizes, np_, hands, weights, seed=seed, kind="river")


def batch_mix(n=256, seed=BASE_SEED + 4000, max_players=6):
    return [mix_river(seed + i, max_players) for i in range(n)]


def mix_turn(seed, max_players=3):
    r = rng_for(seed)
    np_ = int(r.integers(2, max_players + 1))
    k = int(r.integers(1, 3))
    sizes = sorted(r.choice(SIZE_POOL, k, replace=False).tolist())
    density = float(r.uniform(0.1, 1.0))
    pot = int(round(r.uniform(50, 300) * SCALE))
    stack = int(round(r.uniform(200, 1000) * SCALE))
    board = r.choice(52, 4, replace=False)
    hands, weights = _ranges(r, board, np_, density)
    return Spot(board, pot, stack, sizes, np_, hands, weights, leaf_sigma=0.3 * pot,
                seed=seed, kind="turn")


def served_stream(n, rate_per_s, seed=BASE_SEED + 5000, river_max_players=6, turn_max_players=3):
    """(arrival_time_s, Spot) pairs: Poisson arrivals, 70% river / 30% turn."""
    r = rng_for(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += float(r.exponential(1.0 / rate_per_s))
        if r.uniform() < 0.7:
            out.append((t, mix_river(seed + 100000 + i, river_max_players)))
        else:
            out.append((t, mix_turn(seed + 200000 + i, turn_max_players)))
    return out


def small_spot(seed, num_players=2, hands_per_player=6, board_cards=5, pot=1000, stack=3000,
               bet_sizes=(0.5, 1.0), max_raises=2, use_leaves=False):
    """Tiny random subgame (few hands) for brute-force parity checks."""
    r = rng_for(seed)
    board = r.choice(52, board_cards, replace=False)
    combos = all_combos(board)
    hands, weights = [], []
    for _ in range(num_players):
        idx = np.sort(r.choice(len(combos), hands_per_player, replace=False))
        hands.append(combos[idx].copy())
        weights.append(r.uniform(0.05, 1.0, hands_per_player).astype(np.float32))
    return Spot(board, pot, stack, list(bet_sizes), num_players, hands, weights,
                max_raises=max_raises, leaf_sigma=(0.3 * pot if use_leaves else None), seed=seed,
                kind="river" if board_cards == 5 else "turn")


def turn_leaf_workload(lib):
    """Leaf values of one 3-player turn subgame at full ranges: board Kd 7c 2h
    9s, every player holding all 1128 unblocked combos, pot 1500, stack 5000,
    one bet size (0.75 pot) with continuation leaves; the leaf infos come from
    the tree the library builds (28 leaves, 64 bias profiles each). Returns
    (board, leaf_infos, player_hands, max_hands)."""
    import ctypes
    import ss_abi
    import leaf_values as lv
    board = [45, 20, 2, 31]
    td = ss_abi.SSTreeData()
    rc = lib.ss_build_tree((ctypes.c_int * 4)(*board), 4, 1500, 5000, (ctypes.c_float * 1)(0.75), 1, 3,
                           (ctypes.c_int * 3)(0, 1, 2), 0, 1, ctypes.byref(td))
    if rc != 0:
        raise RuntimeError(f"ss_build_tree failed: {ss_abi.last_error(lib)}")
    try:
        infos = lv.extract_leaf_info_from_tree(td, num_players=3)
    finally:
        lib.ss_free_tree(ctypes.byref(td))
    combos = [(int(a), int(b), 1.0) for a, b in all_combos(board)]
    return board, infos, [combos] * 3, len(combos)


def flop_leaf_workload(lib):
    """Leaf values of one 3-player flop subgame at full ranges: flop Kd 7c 2h,
    every player holding all 1176 unblocked combos, pot 1500, stack 5000, one
    bet size (0.75 pot) with continuation leaves. Returns (board, leaf_infos,
    player_hands, max_hands)."""
    import ctypes
    import ss_abi
    import leaf_values as lv
    board = [45, 20, 2]
    td = ss_abi.SSTreeData()
    rc = lib.ss_build_tree((ctypes.c_int * 3)(*board), 3, 1500, 5000, (ctypes.c_float * 1)(0.75), 1, 3,
                           (ctypes.c_int * 3)(0, 1, 2), 0, 1, ctypes.byref(td))
    if rc != 0:
        raise RuntimeError(f"ss_build_tree failed: {ss_abi.last_error(lib)}")
    try:
        infos = lv.extract_leaf_info_from_tree(td, num_players=3)
    finally:
        lib.ss_free_tree(ctypes.byref(td))
    combos = [(int(a), int(b), 1.0) for a, b in all_combos(board)]
    return board, infos, [combos] * 3, len(combos)
