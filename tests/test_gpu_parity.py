"""GPU parity: the HIP library against the CPU oracle and the golden vectors.

Tolerances. Integer work (hand ranks, trees) is bit-exact. FP32 CFR state is
compared after a few iterations, before rounding differences can flip the
sign of a near-zero regret (regret matching turns such flips into O(1)
strategy changes, and FP32-vs-FP64 runs of the same algorithm separate the
same way): regrets / strategy sums within 2e-5 relative to the array's scale,
average strategies within 1e-4 absolute. Long runs are compared through the
FP64 exploitability of the solution: within 1% relative at 50 iterations,
within the FP32 rounding envelope near convergence (see the test).
"""

import ctypes
import gzip
import json
import os

import numpy as np
import pytest

import oracle_bind as ob
import ss_abi
import ss_api
import workloads as wl

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _spot_problem(spot, use_cont=None):
    prob = ss_api.Problem.from_spot(spot, use_cont=use_cont)
    oprob = ob.Problem(prob.tree.td, spot.board, spot.hands, spot.weights, leaf_values=prob.leaf_values)
    return prob, oprob


def _close_state(gpu, ref, rel=2e-5, excused=None):
    """Every (node, hand) info-set row within rel * scale of the oracle, except
    the rows in `excused` (a boolean [N, max_h] mask from _tie_rows): there
    the FP64 regrets left regret matching at its discontinuity, so an FP32
    regret of +-1 ulp legitimately plays another strategy."""
    scale = max(1e-30, float(np.abs(ref).max()))
    d = np.abs(gpu.astype(np.float64) - ref)
    bad_rows = (d > rel * scale).any(axis=1)
    if excused is not None:
        bad_rows &= ~excused
    assert not bad_rows.any(), f"{int(bad_rows.sum())} rows off, max err {d.max()} vs scale {scale}"


def _tie_rows(prob, oprob, k, mode=None, tol=1e-6):
    """Rows (node, hand) whose FP64 regret row after some iteration j < k has
    its largest regret at 0 to within tol of the array's scale (the rounding
    of an FP32 regret, a difference of values of that scale): every regret is
    <= 0 there and one is 0, so regret matching plays uniform, and an FP32
    regret a few ulps above 0 plays a pure strategy instead. These are the
    only rows a state comparison after k >= 2 iterations may excuse."""
    mode = ob.TERM_BRUTE if mode is None else mode
    td = prob.tree.td
    N = prob.tree.num_nodes
    valid = np.zeros((N, 8), bool)
    for n in range(N):
        if td.nodes[n].type == ss_abi.SS_NODE_DECISION:
            valid[n, :td.nodes[n].num_children] = True
    mask = np.zeros((N, prob.max_h), bool)
    for j in range(1, k):
        R = ob.solve(oprob, j, fp64=True, mode=mode, want_state=True)["regrets"]
        scale = max(1e-300, float(np.abs(R).max()))
        top = np.where(valid[:, :, None], R, -np.inf).max(axis=1)
        mask |= np.abs(top) <= tol * scale
    return mask


# ---------------------------------------------------------------- evaluator
def test_device_combo_ranks_golden(lib):
    E = json.load(open(os.path.join(HERE, "golden", "eval_vectors.json")))
    lib.ssdbg_combo_ranks.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint32)]
    for rb in E["river_boards"]:
        bd = rb["board"]
        out = np.zeros(1326, np.uint32)
        assert lib.ssdbg_combo_ranks((ctypes.c_int * 5)(*bd), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))) == 0
        live = [ss for ss in range(1326)]
        got = []
        k = 0
        for a in range(52):
            for b in range(a + 1, 52):
                if a not in bd and b not in bd:
                    got.append(int(out[k]))
                else:
                    assert out[k] == 0
                k += 1
        assert got == rb["values"]


def test_device_eval7_golden(lib):
    E = json.load(open(os.path.join(HERE, "golden", "eval_vectors.json")))
    h = np.array(E["hands7"], dtype=np.int32).ravel()
    n = len(E["values"])
    out = np.zeros(n, np.uint32)
    lib.ssdbg_eval7.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int]
    assert lib.ssdbg_eval7(h.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n,
                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 1) == 0
    assert (out == np.array(E["values"], dtype=np.uint32)).all()


def test_device_eval_census(lib):
    C = json.load(open(os.path.join(HERE, "golden", "eval_census.json")))
    counts = (ctypes.c_ulonglong * 10)()
    s = ctypes.c_ulonglong()
    lib.ssdbg_eval_census.argtypes = [ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_ulonglong)]
    assert lib.ssdbg_eval_census(counts, ctypes.byref(s)) == 0
    assert list(counts) == C["category_counts"]
    assert s.value == C["rank_sum_mod_2_64"]


# ---------------------------------------------------------------- CFR state after k iterations
SMALL = [
    # (seed, players, hands per player, board cards, leaves)
    (0, 2, 12, 5, False), (1, 2, 9, 5, False), (2, 3, 6, 5, False), (3, 3, 7, 5, False),
    (4, 4, 4, 5, False), (5, 2, 10, 4, True), (6, 3, 5, 4, True),
    # three to five opponents: fold and multiway showdown values from the odometer
    (7, 5, 4, 5, False), (8, 6, 3, 5, False),
]


@pytest.mark.parametrize("seed,npl,hp,nb,leaves", SMALL)
@pytest.mark.parametrize("form", [0, 1])
def test_state_small_vs_brute_oracle(seed, npl, hp, nb, leaves, form):
    print(f"case seed={seed} np={npl} hands={hp} board={nb} leaves={leaves} form={form}", flush=True)
    sp = wl.small_spot(seed, num_players=npl, hands_per_player=hp, board_cards=nb, use_leaves=leaves,
                       bet_sizes=(0.5, 1.0), max_raises=2)
    prob, oprob = _spot_problem(sp)
    for k in (1, 2, 3):
        R, S, ev = ss_api.solve_state(prob, k, form)
        ref = ob.solve(oprob, k, fp64=True, mode=ob.TERM_BRUTE, want_state=True)
        ties = _tie_rows(prob, oprob, k)
        _close_state(R, ref["regrets"], excused=ties)
        _close_state(S, ref["strat_sum"], excused=ties)


def test_batched_equals_plain_bitwise():
    sp = wl.small_spot(11, num_players=3, hands_per_player=8)
    prob, _ = _spot_problem(sp)
    a = ss_api.solve_state(prob, 5, 0)
    b = ss_api.solve_state(prob, 5, 1)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("mode", ["persistent", "phased"])
def test_batch_session_modes_equal_plain(mode, monkeypatch):
    """Both batched strategies (persistent kernel, phased launches) give the
    plain form's results bit for bit, for every subgame of a mixed batch."""
    monkeypatch.setenv("SS_BATCH_MODE", mode)
    spots = [wl.small_spot(30 + i, num_players=2 + (i % 2), hands_per_player=9) for i in range(6)]
    spots.append(wl.river_2p(seed=77))
    probs = [ss_api.Problem.from_spot(s) for s in spots]
    b = ss_api.Batch(probs, 12, flags=1)
    reps, st = b.run()
    assert st.mode == (0 if mode == "persistent" else 1)
    for i, p in enumerate(probs):
        got = b.output(i)
        want = ss_api.solve_plain(p, 12, flags=1)
        assert np.array_equal(got.avg, want.avg), i
        assert np.array_equal(got.root_strategy, want.root_strategy), i
    b.close()


def test_batch_rivers_and_turns_equal_plain():
    """A batch mixing rivers (river tables built on host threads, uploaded
    as one image) with turns (continuation leaves, no river tables) gives the
    plain form's results bit for bit for every subgame, exploitability
    included."""
    spots = [wl.mix_river(wl.BASE_SEED + 4000 + 11, max_players=3), wl.river_2p(seed=95),
             wl.mix_turn(wl.BASE_SEED + 200000 + 3, max_players=3), wl.mix_turn(wl.BASE_SEED + 200000 + 4, max_players=3),
             wl.small_spot(41, num_players=3, hands_per_player=7, board_cards=4, use_leaves=True)]
    probs = [ss_api.Problem.from_spot(s) for s in spots]
    b = ss_api.Batch(probs, 20)
    reps, _ = b.run()
    for i, p in enumerate(probs):
        got = b.output(i, reps[i])
        want = ss_api.solve_plain(p, 20)
        assert np.array_equal(got.avg, want.avg), i
        assert np.array_equal(got.root_strategy, want.root_strategy), i
        assert np.array_equal(got.root_ev, want.root_ev), i
        assert got.exploitability == pytest.approx(want.exploitability, rel=1e-12), i
    b.close()


@pytest.mark.parametrize("shape", [(0, 1, -1), (0, 0, 1), (0, 0, 3), (1 << 30, 0, 4), (1 << 30, 0, 2)])
def test_phased_launch_shapes_equal_plain(shape, monkeypatch):
    """Every launch shape of the phased form gives the plain form's results
    bit for bit: forward tasks three to a workgroup or one, a pass's terminal
    classes as one launch or separate ones, two-opponent showdowns split over
    one to four workgroups (the automatic choice depends on the batch size,
    so the shapes are forced here)."""
    monkeypatch.setenv("SS_BATCH_MODE", "phased")
    lib = ss_abi.load_library()
    lib.ssdbg_set_phased_shape.restype = ctypes.c_int
    lib.ssdbg_set_phased_shape.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    assert lib.ssdbg_set_phased_shape(*shape) == 0
    try:
        spots = [wl.small_spot(60 + i, num_players=3, hands_per_player=12) for i in range(3)]
        spots += [wl.small_spot(70, num_players=2, hands_per_player=12), wl.river_3p(seed=79), wl.river_2p(seed=78)]
        probs = [ss_api.Problem.from_spot(s) for s in spots]
        b = ss_api.Batch(probs, 8, flags=1)
        reps, st = b.run()
        assert st.mode == 1
        for i, p in enumerate(probs):
            got = b.output(i)
            want = ss_api.solve_plain(p, 8, flags=1)
            assert np.array_equal(got.avg, want.avg), (shape, i)
            assert np.array_equal(got.root_strategy, want.root_strategy), (shape, i)
        b.close()
    finally:
        lib.ssdbg_set_phased_shape(-1, -1, -1)


def test_phased_batch_with_three_or_more_opponents_equals_plain(monkeypatch):
    """4-, 5- and 6-player spots take the odometer terminals (k_term_odo) in the
    phased form; results equal the plain form bit for bit and the state after
    one iteration matches the brute-force oracle."""
    monkeypatch.setenv("SS_BATCH_MODE", "phased")
    spots = [wl.small_spot(50, num_players=4, hands_per_player=5), wl.small_spot(51, num_players=5, hands_per_player=4),
             wl.small_spot(52, num_players=3, hands_per_player=6), wl.small_spot(53, num_players=6, hands_per_player=3)]
    probs = [ss_api.Problem.from_spot(s) for s in spots]
    b = ss_api.Batch(probs, 6, flags=1)
    reps, st = b.run()
    assert st.mode == 1
    for i, p in enumerate(probs):
        got = b.output(i)
        want = ss_api.solve_plain(p, 6, flags=1)
        assert np.array_equal(got.avg, want.avg), i
        assert np.array_equal(got.root_strategy, want.root_strategy), i
    b.close()
    prob, oprob = _spot_problem(spots[1])
    R, S, ev = ss_api.solve_state(prob, 1, 0)
    ref = ob.solve(oprob, 1, fp64=True, mode=ob.TERM_BRUTE, want_state=True)
    _close_state(R, ref["regrets"])
    _close_state(S, ref["strat_sum"])


def _frozen_for(tree, every=3):
    """Frozen action at every `every`-th decision node (the last action, so
    deeper lines with bets are frozen too); -1 elsewhere."""
    td = tree.td
    fz = np.full(tree.num_nodes, -1, np.int32)
    dec = [i for i in range(tree.num_nodes) if td.nodes[i].type == ss_abi.SS_NODE_DECISION]
    for j, i in enumerate(dec):
        if j % every == 0:
            fz[i] = td.nodes[i].num_children - 1
    return fz


def test_frozen_actions_vs_oracle_and_batched(monkeypatch):
    """A3 freezing (street_solve.cu:621-644): at a node with a frozen action
    the acting player's strategy is one-hot while its regrets keep updating.
    The plain form matches the FP64 oracle under the same freezing; both
    batched strategies and the phased form equal the plain form bit for bit."""
    sp = wl.small_spot(21, num_players=3, hands_per_player=7, bet_sizes=(0.5, 1.0), max_raises=2)
    base = ss_api.Problem.from_spot(sp)
    fz = _frozen_for(base.tree)
    assert (fz >= 0).sum() >= 3
    prob = ss_api.Problem(base.tree, sp.hands, sp.weights, leaf_values=base.leaf_values, frozen=fz)
    oprob = ob.Problem(prob.tree.td, sp.board, sp.hands, sp.weights, leaf_values=prob.leaf_values, frozen=fz)
    for k in (1, 2, 3):
        R, S, ev = ss_api.solve_state(prob, k, 0)
        ref = ob.solve(oprob, k, fp64=True, mode=ob.TERM_BRUTE, want_state=True)
        ties = _tie_rows(prob, oprob, k)
        _close_state(R, ref["regrets"], excused=ties)
        _close_state(S, ref["strat_sum"], excused=ties)
        # strategy sums at frozen nodes put all weight on the frozen action
        for i in np.nonzero(fz >= 0)[0]:
            others = np.delete(S[i, :base.tree.td.nodes[i].num_children], fz[i], axis=0)
            assert not others.any(), i
    want = ss_api.solve_plain(prob, 10, flags=1)
    for mode in ("persistent", "phased"):
        monkeypatch.setenv("SS_BATCH_MODE", mode)
        other = ss_api.Problem.from_spot(wl.small_spot(22, num_players=2, hands_per_player=9))
        b = ss_api.Batch([other, prob], 10, flags=1)
        reps, st = b.run()
        assert st.mode == (0 if mode == "persistent" else 1)
        got = b.output(1)
        assert np.array_equal(got.avg, want.avg), mode
        assert np.array_equal(got.root_strategy, want.root_strategy), mode
        b.close()


def test_state_river_2p_full_ranges():
    sp = wl.river_2p()
    prob, oprob = _spot_problem(sp)
    R, S, ev = ss_api.solve_state(prob, 3, 0)
    ref = ob.solve(oprob, 3, fp64=True, mode=ob.TERM_FAST, want_state=True)
    ties = _tie_rows(prob, oprob, 3, mode=ob.TERM_FAST)
    _close_state(R, ref["regrets"], rel=5e-5, excused=ties)
    _close_state(S, ref["strat_sum"], rel=5e-5, excused=ties)


def test_state_river_3p_sparse_ranges():
    sp = wl.river_3p()
    prob, oprob = _spot_problem(sp)
    R, S, ev = ss_api.solve_state(prob, 1, 0)
    ref = ob.solve(oprob, 1, fp64=True, mode=ob.TERM_FAST, want_state=True)
    _close_state(R, ref["regrets"], rel=5e-5)
    _close_state(S, ref["strat_sum"], rel=5e-5)


# ---------------------------------------------------------------- outputs and exploitability
def test_plain_avg_and_expl_vs_oracle_small():
    sp = wl.small_spot(21, num_players=2, hands_per_player=10)
    prob, oprob = _spot_problem(sp)
    res = ss_api.solve_plain(prob, 3)
    ref = ob.solve(oprob, 3, fp64=True, mode=ob.TERM_BRUTE)
    avg = res.avg_full(prob.tree.num_nodes)
    assert np.abs(avg - ref["avg"]).max() < 1e-4
    e_or, _ = ob.exploitability(oprob, avg, mode=ob.TERM_BRUTE)
    assert abs(res.exploitability - e_or) <= 1e-6 * max(1.0, abs(e_or))


@pytest.mark.parametrize("npl", [2, 3])
def test_gpu_exploitability_matches_oracle_on_gpu_solution(npl):
    sp = wl.river_2p() if npl == 2 else wl.river_3p()
    prob, oprob = _spot_problem(sp)
    res = ss_api.solve_plain(prob, 20)
    e_or, _ = ob.exploitability(oprob, res.avg_full(prob.tree.num_nodes), mode=ob.TERM_FAST)
    assert abs(res.exploitability - e_or) <= 1e-6 * max(1.0, abs(e_or)), (res.exploitability, e_or)


@pytest.mark.parametrize("npl,hp", [(4, 5), (5, 4), (6, 3)])
def test_multiway_avg_and_exploitability_vs_brute_oracle(npl, hp):
    """4-6 players (every fold node and the multiway showdowns have 3-5
    opponents, evaluated by the odometer): the average strategy after 3
    iterations within 1e-4 of the FP64 brute-force oracle on every info-set
    row except those whose FP64 regrets sat exactly at regret matching's
    discontinuity (all <= 0 with a 0; e.g. a hand tied with every live
    opponent hand where pot / 2 equals its bet, whose closed-form value is
    0 in exact arithmetic and +-1 ulp in FP32), and the GPU's FP64
    exploitability of its solution equal to the oracle's evaluation of the
    same strategy."""
    sp = wl.small_spot(70 + npl, num_players=npl, hands_per_player=hp)
    prob, oprob = _spot_problem(sp)
    print(f"multiway: {npl} players, {hp} hands, {prob.tree.num_nodes} nodes", flush=True)
    res = ss_api.solve_plain(prob, 3)
    ref = ob.solve(oprob, 3, fp64=True, mode=ob.TERM_BRUTE)
    avg = res.avg_full(prob.tree.num_nodes)
    err = np.abs(avg - ref["avg"]).max(axis=1)          # [N, max_h]
    ties = _tie_rows(prob, oprob, 3)
    assert not ((err > 1e-4) & ~ties).any(), float(np.where(ties, 0, err).max())
    e_or, _ = ob.exploitability(oprob, avg, mode=ob.TERM_BRUTE)
    assert abs(res.exploitability - e_or) <= 1e-6 * max(1.0, abs(e_or)), (res.exploitability, e_or)


@pytest.mark.parametrize("iters", [50, 200])
def test_long_run_exploitability_river_2p_vs_fp64_oracle(iters):
    """Exploitability of the GPU's FP32 solve vs the FP64 oracle at equal
    iterations. At 50 iterations within 1% relative. Near convergence the FP32
    rounding of any implementation moves exploitability by more than 1% (the
    reference's own FP32 arithmetic, restated in the oracle, lands 4.6% off
    the FP64 value at 200 iterations on this spot), so there the bound is the
    FP32 envelope: no further from FP64 than the oracle's FP32 run, and within
    0.02% of the pot."""
    sp = wl.river_2p()
    prob, oprob = _spot_problem(sp)
    res = ss_api.solve_plain(prob, iters)
    ref = ob.solve(oprob, iters, fp64=True, mode=ob.TERM_FAST)
    e_ref, _ = ob.exploitability(oprob, ref["avg"], mode=ob.TERM_FAST)
    if iters <= 50:
        assert abs(res.exploitability - e_ref) <= 0.01 * e_ref, (res.exploitability, e_ref)
    else:
        f32 = ob.solve(oprob, iters, fp64=False, mode=ob.TERM_FAST)
        e32, _ = ob.exploitability(oprob, f32["avg"], mode=ob.TERM_FAST)
        gap = abs(res.exploitability - e_ref)
        assert gap <= max(abs(e32 - e_ref), 0.01 * e_ref) and gap <= 0.02, (res.exploitability, e_ref, e32)


def test_exploitability_river_3p_vs_fp64_oracle():
    """Three players (1081 combos, 40% of each range live, 150 BB pot, sizes
    0.33 and 0.75 pot): the GPU's FP32 solve after 50 iterations is within 1%
    relative of the FP64 oracle's exploitability at equal iterations, through
    the plain and the batched form (two-opponent showdown and fold closed
    forms at full size)."""
    sp = wl.river_3p()
    prob, oprob = _spot_problem(sp)
    res = ss_api.solve_plain(prob, 50)
    ref = ob.solve(oprob, 50, fp64=True, mode=ob.TERM_FAST)
    e_ref, _ = ob.exploitability(oprob, ref["avg"], mode=ob.TERM_FAST)
    assert e_ref > 0
    assert abs(res.exploitability - e_ref) <= 0.01 * e_ref, (res.exploitability, e_ref)
    b = ss_api.Batch([prob, ss_api.Problem.from_spot(wl.river_2p(seed=5))], 50)
    reps, st = b.run()
    got = b.output(0, reps[0])
    b.close()
    assert got.exploitability == res.exploitability
    assert np.array_equal(got.avg, res.avg)


def test_legacy_abi_reference_spot(lib):
    # the spot of the reference's own hot-path test (tests/test_street_solve.py:178-190)
    from street_solver_gpu import StreetSolverGPU
    s = StreetSolverGPU(board=["Qs", "As", "2d", "7h", "4c"], oop_range="AhKh,QhQc,JhTh,6h5h,3c3d",
                        ip_range="AcKc,8c8d,Tc9c,KhJh,2h2c", pot_bb=10.0, stack_bb=50.0, bet_sizes=[0.33, 0.75])
    s.solve(iterations=200)
    strat = s.get_strategy("oop", hand_str="AhKh")
    assert len(strat) > 0 and abs(sum(strat.values()) - 1.0) < 0.05
    s.free()


# ---------------------------------------------------------------- root outputs (street_solve.cu:1563-1578)
def _legacy_solve(prob, iterations):
    """The legacy entry point ss_solve_gpu on a problem of at most 200 hands per
    player, with hands and weights written into the tree struct's fixed slots
    as the reference wrapper does (python/street_solver_gpu.py:286-291).
    Returns (root_strategy [na, nh_root], root_ev [max_h], root_player)."""
    lib = ss_abi.load_library()
    td = prob.tree.td
    for p in range(prob.num_players):
        n = len(prob.hands[p])
        assert n <= ss_abi.SS_MAX_HANDS
        td.num_hands[p] = n
        np.ctypeslib.as_array(td.hands[p])[:n] = prob.hands[p]
        np.ctypeslib.as_array(td.weights[p])[:n] = prob.weights[p]
    if prob.leaf_values is not None:
        td.leaf_values = prob.leaf_values.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    out = ss_abi.SSOutput()
    assert lib.ss_solve_gpu(ctypes.byref(td), int(iterations), ctypes.byref(out)) == 0, ss_abi.last_error()
    try:
        na, rp = out.root_num_actions, out.root_player
        nh = len(prob.hands[rp])
        rs = np.ctypeslib.as_array(out.root_strategy, shape=(na * nh,)).copy().reshape(na, nh)
        ev = np.ctypeslib.as_array(out.root_ev, shape=(prob.max_h,)).copy()
    finally:
        lib.ss_free_output(ctypes.byref(out))
    return rs, ev, rp


def _check_root(rs, ev, rp, prob, oprob, ref, k, mode):
    """root_strategy: the root actor's last-pass strategy, within 1e-4 for every
    hand whose FP64 root regrets are off regret matching's discontinuity;
    root_ev: node-0 counterfactual values of the last traverser (player NP-1,
    the reference's quirk), within 2e-5 of their scale."""
    na = rs.shape[0]
    want = ref["root_strategy"][:na, :rs.shape[1]]
    ties = _tie_rows(prob, oprob, k + 1, mode=mode)[0, :rs.shape[1]]
    off = (np.abs(rs - want) > 1e-4).any(axis=0) & ~ties
    assert not off.any(), f"root strategy off at hands {np.nonzero(off)[0][:8]}"
    assert np.allclose(rs.sum(axis=0), 1.0, atol=1e-5)
    nl = len(prob.hands[prob.num_players - 1])
    ref_ev = ref["root_ev"][:nl].astype(np.float64)
    scale = max(1e-30, float(np.abs(ref_ev).max()))
    assert np.abs(ev[:nl] - ref_ev).max() <= 2e-5 * scale, (np.abs(ev[:nl] - ref_ev).max(), scale)
    assert rp == prob.tree.td.nodes[0].player


@pytest.mark.parametrize("seed,npl,hp", [(40, 2, 11), (41, 3, 7), (42, 3, 9)])
def test_root_outputs_small_vs_brute_oracle(seed, npl, hp):
    """root_strategy and root_ev after k = 1..3 iterations, through the legacy
    ss_solve_gpu and through ss_solve_gpu_ex, against the FP64 brute-force
    oracle (street_solve.cu:1563-1578)."""
    sp = wl.small_spot(seed, num_players=npl, hands_per_player=hp)
    prob, oprob = _spot_problem(sp)
    for k in (1, 2, 3):
        print(f"root outputs seed={seed} np={npl} k={k}", flush=True)
        ref = ob.solve(oprob, k, fp64=True, mode=ob.TERM_BRUTE)
        rs, ev, rp = _legacy_solve(prob, k)
        _check_root(rs, ev, rp, prob, oprob, ref, k, ob.TERM_BRUTE)
        res = ss_api.solve_plain(prob, k, flags=ss_abi.SS_FLAG_NO_EXPL)
        _check_root(res.root_strategy, res.root_ev[:prob.max_h], res.root_player, prob, oprob, ref, k,
                    ob.TERM_BRUTE)


@pytest.mark.parametrize("npl", [2, 3])
def test_root_outputs_full_rivers_vs_fp64_oracle(npl):
    """The same on the 1081-combo 2- and 3-player rivers through
    ss_solve_gpu_ex (the legacy entry point is capped at 200 hands)."""
    sp = wl.river_2p() if npl == 2 else wl.river_3p()
    prob, oprob = _spot_problem(sp)
    for k in (1, 2, 3):
        print(f"root outputs full river np={npl} k={k}", flush=True)
        ref = ob.solve(oprob, k, fp64=True, mode=ob.TERM_FAST)
        res = ss_api.solve_plain(prob, k, flags=ss_abi.SS_FLAG_NO_EXPL)
        _check_root(res.root_strategy, res.root_ev[:prob.max_h], res.root_player, prob, oprob, ref, k,
                    ob.TERM_FAST)


def test_dropin_get_strategy_at_root_vs_oracle():
    """StreetSolverGPU.get_strategy with no action history returns the root's
    last-pass strategy (python/street_solver_gpu.py:398-410), labelled and
    cut at 0.001, for every hand of the reference's toy spot
    (tests/test_street_solve.py:178-190); checked against the oracle's
    root_strategy on the wrapper's own tree and ranges after 3 iterations."""
    from street_solver_gpu import StreetSolverGPU, _build_labels_at_node
    s = StreetSolverGPU(board=["Qs", "As", "2d", "7h", "4c"], oop_range="AhKh,QhQc,JhTh,6h5h,3c3d",
                        ip_range="AcKc,8c8d,Tc9c,KhJh,2h2c", pot_bb=10.0, stack_bb=50.0, bet_sizes=[0.33, 0.75])
    k = 3
    s.solve(iterations=k)
    oprob = ob.Problem(s._tree, s.board_cards, s._hand_arrs, s._weight_arrs)
    ref = ob.solve(oprob, k, fp64=True, mode=ob.TERM_BRUTE)
    labels = _build_labels_at_node(s._tree, 0)
    assert s._tree.nodes[0].player == 0
    for h, (c0, c1, _) in enumerate(s.player_hands[0]):
        got = s.get_strategy("oop", hand_idx=h)
        want = {labels[a]: float(ref["root_strategy"][a, h]) for a in range(len(labels))
                if float(ref["root_strategy"][a, h]) > 0.001}
        assert set(got) == set(want), (h, got, want)
        assert all(abs(got[x] - want[x]) < 1e-4 for x in got), (h, got, want)
    s.free()


@pytest.mark.parametrize("form", [0, 1])
def test_cfr_state_hand_derived_literals(form):
    """The GPU's regrets, strategy sums and root outputs on the hand-derived
    9-node river of test_oracle.py (P0 6h6s beats P1 AcAd, raw weights 2.0 and
    0.5) after one and two iterations, plain (form 0) and batched (form 1),
    equal the literals worked out by hand."""
    from test_oracle import TINY_BOARD, TINY_REGRETS, TINY_ROOT, TINY_STRAT_SUMS, _cards
    board = _cards(TINY_BOARD)
    tree = ss_api.Tree(board, 100, 50, [0.5], 2, max_raises=2)
    assert tree.num_nodes == 9
    prob = ss_api.Problem(tree, [np.array([_cards("6h 6s")]), np.array([_cards("Ac Ad")])],
                          [np.array([2.0], np.float32), np.array([0.5], np.float32)])
    for k in (1, 2):
        R, S, ev = ss_api.solve_state(prob, k, form)
        for (n, a), want in TINY_REGRETS[k].items():
            assert R[n, a, 0] == pytest.approx(want, rel=1e-6, abs=1e-6), (k, n, a)
        for (n, a), want in TINY_STRAT_SUMS[k].items():
            assert S[n, a, 0] == pytest.approx(want, rel=1e-6, abs=1e-6), (k, n, a)
        (s0, s1), root_ev = TINY_ROOT[k]
        assert ev[0] == pytest.approx(root_ev, rel=1e-6)
        res = ss_api.solve_plain(prob, k, flags=ss_abi.SS_FLAG_NO_EXPL)
        assert res.root_strategy[:, 0] == pytest.approx([s0, s1], abs=1e-6)
        assert res.root_ev[0] == pytest.approx(root_ev, rel=1e-6)


def test_autotuned_launch_shape_equals_plain_and_is_cached():
    """ss_batch_create times launch shapes of the phased form on the batch and
    keeps the fastest; a second batch of the same make-up reuses the choice
    without tuning. Every shape gives the plain form's results bit for bit."""
    spots = [wl.river_2p(seed=90 + i) for i in range(3)] + [wl.river_3p(seed=93)]
    spots += [wl.small_spot(94 + i, num_players=2 + i % 2, hands_per_player=11) for i in range(6)]
    probs = [ss_api.Problem.from_spot(s) for s in spots]   # 10 subgames: no other test's batch size
    tuned = []
    for rep in range(2):
        b = ss_api.Batch(probs, 9, flags=1)
        reps, st = b.run()
        assert st.mode == 1
        tuned.append((st.tune_ms, (st.shape_fwd_group_min, st.shape_term_small, st.shape_sd_parts,
                                   st.shape_sweep_grid)))
        print(f"autotune run {rep}: {tuned[-1]}", flush=True)
        for i, p in enumerate(probs):
            got = b.output(i)
            want = ss_api.solve_plain(p, 9, flags=1)
            assert np.array_equal(got.avg, want.avg), (rep, i)
            assert np.array_equal(got.root_strategy, want.root_strategy), (rep, i)
        b.close()
    assert tuned[0][0] > 0 and tuned[1][0] < tuned[0][0] and tuned[1][1] == tuned[0][1]
