"""GPU parity of the leaf-value producers (include/street_solve_leaf.h).

Every output is compared bit for bit: the float64 equity sums run in the
reference's order, the bias formula is evaluated without contraction, and the
float32 rounding is the reference's assignment into a float32 array. Checked
against the reference's own outputs (tests/golden/leaf_values.npz), the
oracle at full ranges (1128 turn combos per player for 2, 3 and 6 players;
1326 listed combos with the board-blocked ones; a 3-player flop over its 144
sampled runouts), hand-derived answers, the edge cases (an empty range,
no runouts), and 24 seeded random cases (turn 2-6 players, flop 2-3 players,
range sizes on both sides of the sampling thresholds, random leaf infos).
"""

import ctypes
import os
import sys

import numpy as np
import pytest

import ss_abi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import leaf_oracle as lo  # noqa: E402
import leaf_values as lv  # noqa: E402
from test_leaf import GOLD, NAMES, card, cards, golden_case  # noqa: E402

pytestmark = pytest.mark.gpu


def gpu_equity(lib, hands, nh, runouts, mh):
    npl = len(nh)
    h = np.zeros((npl, mh, 2), dtype=np.int32)
    for p in range(npl):
        for i in range(nh[p]):
            h[p, i] = hands[p][i][:2]
    run = np.array(runouts, dtype=np.int32).reshape(-1, 5)
    raw = np.zeros((npl, mh))
    ip = ctypes.POINTER(ctypes.c_int)
    rc = lib.ss_leaf_equity(npl, h.ctypes.data_as(ip), np.array(nh, dtype=np.int32).ctypes.data_as(ip), mh,
                            run.ctypes.data_as(ip), len(run), raw.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, ss_abi.last_error(lib)
    return raw


def turn_hands(rng, board, nh, all_combos=False):
    free = [c for c in range(52) if all_combos or c not in board]
    combos = [(a, b) for i, a in enumerate(free) for b in free[i + 1:]]
    out = []
    for n in nh:
        pick = rng.permutation(len(combos))[:n]
        out.append([combos[k] for k in pick])
    return out


@pytest.mark.parametrize("name", NAMES)
def test_leaf_values_equal_reference(lib, name):
    kind, board, hands, nh, mh, infos, ref = golden_case(name)
    print(f"leaf golden {name}", flush=True)
    lis = [lv.LeafInfo(i, p, ()) for i, p in infos]
    ph = [[(a, b, 1.0) for a, b in hs] for hs in hands]
    f = lv.compute_turn_leaf_values if kind == "turn" else lv.compute_flop_leaf_equity
    out = f(board, lis, mh, 1000, player_hands=ph)
    assert out.shape == ref.shape
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


def test_known_answers(lib):
    b = cards("2c 7d 9h Js")
    raw = gpu_equity(lib, [[(card("Ah"), card("As"))], [(card("Kd"), card("Kc"))]], [1, 1],
                     lo.river_runouts(b, [1, 1]), 1)
    assert raw[0, 0] == 42 / 44 and raw[1, 0] == 2 / 44
    t = cards("Ts Js Qs Ks")
    h3 = [[(card("2c"), card("3c"))], [(card("2d"), card("3d"))], [(card("2h"), card("3h"))]]
    raw = gpu_equity(lib, h3, [1, 1, 1], lo.river_runouts(t, [1, 1, 1]), 1)
    assert np.array_equal(raw, lo.runout_equity(h3, [1, 1, 1], lo.river_runouts(t, [1, 1, 1]), 1))
    assert np.allclose(raw, 1.0 / 3, rtol=0, atol=1e-14)
    hc = [[(card("Ah"), card("As"))], [(card("Kd"), card("Kc"))], [(card("Kd"), card("Qc"))]]
    assert np.all(gpu_equity(lib, hc, [1, 1, 1], lo.river_runouts(b, [1, 1, 1]), 1) == 0.0)


@pytest.mark.parametrize("npl,all_combos", [(2, False), (3, False), (6, False), (2, True)])
def test_full_ranges_equal_oracle(lib, npl, all_combos):
    rng = np.random.default_rng(7000 + npl)
    board = cards("Kd 7c 2h 9s")
    nh = [1326 if all_combos else 1128] * npl
    hands = turn_hands(rng, board, nh, all_combos)
    runs = lv.turn_runouts(board, nh)
    print(f"leaf full ranges np={npl} hands={nh[0]} runouts={len(runs)}", flush=True)
    raw = gpu_equity(lib, hands, nh, runs, nh[0])
    ref = lo.runout_equity(hands, nh, runs, nh[0])
    assert np.array_equal(raw, ref)
    assert raw.min() >= 0.0 and raw.max() <= 1.0
    if npl == 3:
        infos = [(0, 1500)]
        out = lv._leaf_values([[(a, b, 1.0) for a, b in hs] for hs in hands], runs,
                              [lv.LeafInfo(0, 1500, ())], nh[0], lib)
        assert np.array_equal(out, lo.values_from_equity(ref, nh, nh[0], infos))


def test_edge_cases(lib):
    board = cards("Qs As 2d 7h")
    rng = np.random.default_rng(11)
    hands = turn_hands(rng, board, [20, 0, 15])
    nh = [20, 0, 15]
    runs = lo.river_runouts(board, nh)
    raw = gpu_equity(lib, hands, nh, runs, 20)
    assert np.array_equal(raw, lo.runout_equity(hands, nh, runs, 20))
    assert np.all(raw == 0.0)  # an empty opponent range leaves no tuple for anyone
    raw = gpu_equity(lib, turn_hands(rng, board, [5, 5]), [5, 5], [], 5)
    assert np.all(raw == 0.0)
    out = lv._leaf_values([[(a, b, 1.0) for a, b in hs] for hs in turn_hands(rng, board, [3, 2])],
                          lo.river_runouts(board, [3, 2]), [], 4, lib)
    assert out.shape == (0, 2, 4)


def test_flop_sampled_runouts_equal_oracle(lib):
    rng = np.random.default_rng(8100)
    board = cards("Kd 7c 2h")
    nh = [300, 300, 300]
    hands = turn_hands(rng, board, nh)
    runs = lv.flop_runouts(board, nh)
    assert len(runs) == 144
    print(f"leaf flop np=3 hands=300 runouts={len(runs)}", flush=True)
    raw = gpu_equity(lib, hands, nh, runs, 300)
    assert np.array_equal(raw, lo.runout_equity(hands, nh, runs, 300))


def _random_case(seed):
    rng = np.random.default_rng(9000 + seed)
    npl = int(rng.integers(2, 7))
    flop = bool(rng.integers(0, 2)) and npl <= 3
    board = [int(c) for c in rng.choice(52, 3 if flop else 4, replace=False)]
    if npl >= 5 or flop:
        sizes = [int(rng.integers(1, 12)) for _ in range(npl)]
        if not flop and rng.integers(0, 2):
            sizes[0] = 51 + int(rng.integers(0, 20))  # > 50 hands: 12 sampled rivers
    else:
        sizes = [int(rng.integers(0, 90)) for _ in range(npl)]
    hands = turn_hands(rng, board, sizes)
    mh = max(max(sizes), 1) + int(rng.integers(0, 3))
    infos = [(int(rng.integers(0, 4)), int(rng.integers(100, 5000))) for _ in range(int(rng.integers(1, 4)))]
    return flop, board, hands, sizes, mh, infos


@pytest.mark.parametrize("seed", range(24))
def test_random_cases_equal_oracle(lib, seed):
    flop, board, hands, nh, mh, infos = _random_case(seed)
    print(f"leaf random seed={seed} flop={flop} nh={nh} max_hands={mh} infos={infos}", flush=True)
    ph = [[(a, b, 1.0) for a, b in hs] for hs in hands]
    lis = [lv.LeafInfo(i, p, ()) for i, p in infos]
    if flop:
        out = lv.compute_flop_leaf_equity(board, lis, mh, 1000, player_hands=ph)
        ref = lo.flop_leaf_equity(board, hands, nh, mh, infos)
    else:
        out = lv.compute_turn_leaf_values(board, lis, mh, 1000, player_hands=ph)
        ref = lo.turn_leaf_values(board, hands, nh, mh, infos)
    assert out.shape == ref.shape
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


def test_large_output_read_back_equals_oracle(lib):
    """An output above the library's 4 MiB pinned read-back threshold (3
    players x 1128 hands x 8 leaf infos x 64 profiles = 6.9 MB, split over
    the host copy threads) equals the oracle's values on the GPU's own equity
    in every element, bit for bit."""
    rng = np.random.default_rng(8300)
    board = cards("Kd 7c 2h 9s")
    nh = [1128, 1128, 1128]
    hands = turn_hands(rng, board, nh)
    runs = lv.turn_runouts(board, nh)
    infos = [(i, 700 + 150 * i) for i in range(8)]
    print(f"leaf large read-back np=3 hands=1128 infos={len(infos)}", flush=True)
    out = lv._leaf_values([[(a, b, 1.0) for a, b in hs] for hs in hands], runs,
                          [lv.LeafInfo(i, p, ()) for i, p in infos], 1128, lib)
    assert out.nbytes >= 4 << 20
    raw = gpu_equity(lib, hands, nh, runs, 1128)
    ref = lo.values_from_equity(raw, nh, 1128, infos)
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


def _turn_problem(seed, npl, nh, zeros):
    """A turn subgame with continuation leaves: ss_api Problem with the given
    leaf values (or zeros), its hands as (c0, c1, w) lists, the tree."""
    import ss_api
    import workloads as wl
    rng = np.random.default_rng(seed)
    board = [int(c) for c in rng.choice(52, 4, replace=False)]
    combos = wl.all_combos(board)
    ph = []
    for _ in range(npl):
        pick = np.sort(rng.choice(len(combos), nh, replace=False))
        ph.append([(int(combos[k][0]), int(combos[k][1]), float(rng.uniform(0.1, 1.0))) for k in pick])
    tree = ss_api.Tree(board, 1200, 6000, [0.75], npl, use_cont=1, max_raises=1)
    hands = [np.array([(a, b) for a, b, _ in h], np.int32) for h in ph]
    weights = [np.array([w for _, _, w in h], np.float32) for h in ph]
    return board, ph, tree, hands, weights


def _solve_with(tree, hands, weights, leaf, iters=12):
    import ss_api
    prob = ss_api.Problem(tree, hands, weights, leaf_values=leaf)
    b = ss_api.Batch([prob], iters, flags=1)
    return prob, b


def test_leaf_values_into_resident_solve_reference_order(lib):
    """ss_batch_leaf_values (reference order) gives the solve exactly the leaf
    values of compute_turn_leaf_values' host array: the solves are equal bit
    for bit (2-player turn, 200 combos per player, continuation leaves)."""
    board, ph, tree, hands, weights = _turn_problem(8400, 2, 200, False)
    mh = 200
    infos = lv.extract_leaf_info_from_tree(tree.td, 2)
    host = lv.compute_turn_leaf_values(board, infos, mh, 1200, player_hands=ph)
    assert host.shape[0] == tree.num_leaves
    print(f"leaf into batch: turn 2p leaves={tree.num_leaves}", flush=True)
    _, ba = _solve_with(tree, hands, weights, host.reshape(tree.num_leaves, 2, mh))
    _, bb = _solve_with(tree, hands, weights, np.zeros((tree.num_leaves, 2, mh), np.float32))
    lv.load_leaf_values_into_batch(bb, 0, board, infos, ph)
    ba.run()
    bb.run()
    a, b = ba.output(0), bb.output(0)
    assert np.array_equal(a.avg, b.avg) and np.array_equal(a.root_ev, b.root_ev)
    ba.close()
    bb.close()


def test_leaf_values_into_resident_solve_tree_order(lib):
    """Tree order on a 3-player turn whose leaves have 2 or 3 active players
    (the reference layout does not fit such a tree): tree leaf l gets the
    producer's row L0 * 64 + sum_p s_p 4^p of the strategies its chain picks
    (players outside the chain at profile 0); the device-filled solve equals
    the solve given that gather made on the host."""
    board, ph, tree, hands, weights = _turn_problem(8401, 3, 60, False)
    mh = 60
    td = tree.td
    infos = lv.tree_leaf_infos(td, 3)
    prod = lv.compute_turn_leaf_values(board, infos, mh, 1200, player_hands=ph)
    parent, pos = {}, {}
    for i in range(td.num_nodes):
        n = td.nodes[i]
        if n.type == 0:
            for a in range(n.num_children):
                parent[td.children[n.first_child + a]] = i
                pos[td.children[n.first_child + a]] = a
    heads, rows, nactive = {}, {}, set()
    for i in range(td.num_nodes):
        n = td.nodes[i]
        if n.type != 4:
            continue
        act = [q for q in range(3) if n.active_players[q]]
        nactive.add(len(act))
        cur, combo = i, 0
        for q in reversed(act):
            assert td.nodes[parent[cur]].player == q
            combo += pos[cur] * 4 ** q
            cur = parent[cur]
        heads[n.leaf_idx], rows[n.leaf_idx] = cur, combo
    assert nactive == {2, 3}
    rank = {h: k for k, h in enumerate(sorted(set(heads.values())))}
    gathered = np.stack([prod[rank[heads[l]] * 64 + rows[l]] for l in range(tree.num_leaves)])
    print(f"leaf into batch: turn 3p tree order leaves={tree.num_leaves} heads={len(rank)}", flush=True)
    _, ba = _solve_with(tree, hands, weights, gathered)
    _, bb = _solve_with(tree, hands, weights, np.zeros_like(gathered))
    lv.load_leaf_values_into_batch(bb, 0, board, infos, ph, tree_order=True)
    ba.run()
    bb.run()
    a, b = ba.output(0), bb.output(0)
    assert np.array_equal(a.avg, b.avg) and np.array_equal(a.root_ev, b.root_ev)
    # the reference layout is refused for this tree
    with pytest.raises(RuntimeError):
        lv.load_leaf_values_into_batch(bb, 0, board, lv.extract_leaf_info_from_tree(td, 3), ph)
    ba.close()
    bb.close()
