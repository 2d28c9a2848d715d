"""Precision form (SS_PREC_BF16X2): the 2-player river showdown as a BF16
hi/lo MFMA contraction with FP32 accumulation, FP64 exploitability.

Error budget. The opponent reach is carried as hi = bf16(r), lo = bf16(r - hi):
16 significant bits, a relative error <= 2^-16 per element, and W (+1, -1, 0)
is exact, so a showdown value differs from the FP32 closed form by at most
about 2^-16 of pot * (opponent mass) plus FP32 accumulation rounding. State
after one iteration is checked against the FP64 oracle at 5e-5 of the array
scale (the FP32 closed form passes at 2e-5); after three iterations only
rows whose FP64 regrets sit at regret matching's discontinuity (largest
regret within the 5e-5 budget of 0) may differ.
"""

import numpy as np
import pytest

import oracle_bind as ob
import ss_abi
import ss_api
import workloads as wl

from test_gpu_parity import _close_state, _spot_problem, _tie_rows

pytestmark = pytest.mark.gpu
PREC = 2   # ssdbg_solve_state form: plain + BF16 showdown


@pytest.mark.parametrize("seed", [3, 4])
def test_precision_state_small_vs_brute_oracle(seed):
    sp = wl.small_spot(seed, num_players=2, hands_per_player=10)
    prob, oprob = _spot_problem(sp)
    R, S, ev = ss_api.solve_state(prob, 1, PREC)
    ref = ob.solve(oprob, 1, fp64=True, mode=ob.TERM_BRUTE, want_state=True)
    _close_state(R, ref["regrets"], rel=5e-5)
    _close_state(S, ref["strat_sum"], rel=5e-5)


def test_precision_state_river_2p_vs_fp64_oracle():
    sp = wl.river_2p()
    prob, oprob = _spot_problem(sp)
    R, S, ev = ss_api.solve_state(prob, 1, PREC)
    ref = ob.solve(oprob, 1, fp64=True, mode=ob.TERM_FAST, want_state=True)
    _close_state(R, ref["regrets"], rel=5e-5)
    _close_state(S, ref["strat_sum"], rel=5e-5)
    R3, S3, _ = ss_api.solve_state(prob, 3, PREC)
    ref3 = ob.solve(oprob, 3, fp64=True, mode=ob.TERM_FAST, want_state=True)
    ties = _tie_rows(prob, oprob, 3, mode=ob.TERM_FAST, tol=5e-5)   # the BF16 error budget
    _close_state(R3, ref3["regrets"], rel=5e-5, excused=ties)
    _close_state(S3, ref3["strat_sum"], rel=5e-5, excused=ties)


def test_precision_exploitability_river_2p_vs_fp64_oracle():
    """50 iterations: the solution's FP64 exploitability within 1% of the FP64
    oracle's at equal iterations (the FP32 closed form meets the same bound)."""
    sp = wl.river_2p()
    prob, oprob = _spot_problem(sp)
    res = ss_api.solve_plain(prob, 50, precision=ss_abi.SS_PREC_BF16X2)
    ref = ob.solve(oprob, 50, fp64=True, mode=ob.TERM_FAST)
    e_ref, _ = ob.exploitability(oprob, ref["avg"], mode=ob.TERM_FAST)
    assert abs(res.exploitability - e_ref) <= 0.01 * e_ref, (res.exploitability, e_ref)


def test_precision_batched_equals_plain():
    """The batched (phased) precision form gives the plain precision form's
    results bit for bit, for every board of a batch."""
    probs = [ss_api.Problem.from_spot(wl.river_2p(seed=20261203 + 1000 + b)) for b in range(3)]
    probs.append(ss_api.Problem.from_spot(wl.small_spot(12, num_players=2, hands_per_player=20)))
    b = ss_api.Batch(probs, 15, flags=1, precision=ss_abi.SS_PREC_BF16X2)
    b.run()
    for i, p in enumerate(probs):
        got = b.output(i)
        want = ss_api.solve_plain(p, 15, flags=1, precision=ss_abi.SS_PREC_BF16X2)
        assert np.array_equal(got.avg, want.avg), i
        assert np.array_equal(got.root_strategy, want.root_strategy), i
    b.close()


def test_precision_batched_exploitability_checks():
    """FP64 checks of a precision-form batch see every showdown: the reported
    exploitability equals the plain precision form's."""
    probs = [ss_api.Problem.from_spot(wl.river_2p(seed=20261203 + 1100 + b)) for b in range(2)]
    b = ss_api.Batch(probs, 40, precision=ss_abi.SS_PREC_BF16X2)
    reps, _ = b.run()
    for i, p in enumerate(probs):
        want = ss_api.solve_plain(p, 40, precision=ss_abi.SS_PREC_BF16X2)
        assert reps[i].exploitability == pytest.approx(want.exploitability, rel=1e-12), i
    b.close()


def test_precision_form_scope():
    """3+ players are FP32 only; the served form is FP32 only."""
    sp = wl.small_spot(5, num_players=3, hands_per_player=6)
    prob = ss_api.Problem.from_spot(sp)
    with pytest.raises(ss_api.SolveError):
        ss_api.solve_plain(prob, 2, precision=ss_abi.SS_PREC_BF16X2)
    with pytest.raises(ss_api.SolveError):
        ss_api.Batch([prob], 2, precision=ss_abi.SS_PREC_BF16X2)
