"""The allocation-free leaf-equity entry point (ss_leaf_equity_static) against
the allocating one (ss_leaf_equity), through the C-ABI: two calls in a row,
the second smaller than the first (the fixed scratch array then holds the
first call's contents), both equal bit for bit; a call needing more than the
fixed scratch is refused with SS_ERR_ARG before any device work."""

import ctypes

import numpy as np
import pytest

import ss_abi

pytestmark = pytest.mark.gpu

IP, DP = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)


def _case(seed, npl, nh, nrun, board_cards=4):
    rng = np.random.default_rng(seed)
    board = [int(c) for c in rng.choice(52, board_cards, replace=False)]
    free = [c for c in range(52) if c not in board]
    combos = [(a, b) for i, a in enumerate(free) for b in free[i + 1:]]
    mh = max(nh)
    hands = np.zeros((npl, mh, 2), np.int32)
    for p in range(npl):
        pick = rng.permutation(len(combos))[:nh[p]]
        hands[p, :nh[p]] = [combos[k] for k in pick]
    rest = [c for c in free]
    runs = []
    for _ in range(nrun):
        extra = rng.choice(rest, 5 - board_cards, replace=False).tolist()
        runs.append(board + [int(c) for c in extra])
    return hands, np.array(nh, np.int32), mh, np.array(runs, np.int32).reshape(-1, 5)


def _call(fn, npl, hands, nh, mh, runs):
    raw = np.full((npl, mh), -1.0)
    rc = fn(npl, hands.ctypes.data_as(IP), nh.ctypes.data_as(IP), mh, runs.ctypes.data_as(IP), len(runs),
            raw.ctypes.data_as(DP))
    return rc, raw


def test_static_equity_equals_allocating_twice_then_smaller():
    lib = ss_abi.load_library()
    cases = [(3, [1128, 1128, 1128], 12, 4), (2, [300, 250], 5, 3)]
    for i, (npl, nh, nrun, nb) in enumerate(cases):
        print(f"no-alloc equity call {i}: np={npl} hands={nh} runouts={nrun}", flush=True)
        hands, nh_a, mh, runs = _case(900 + i, npl, nh, nrun, nb)
        rc_s, raw_s = _call(lib.ss_leaf_equity_static, npl, hands, nh_a, mh, runs)
        assert rc_s == 0, ss_abi.last_error(lib)
        rc_a, raw_a = _call(lib.ss_leaf_equity, npl, hands, nh_a, mh, runs)
        assert rc_a == 0, ss_abi.last_error(lib)
        assert np.array_equal(raw_s.view(np.uint64), raw_a.view(np.uint64)), i
        assert (raw_s >= 0).all() and (raw_s <= 1).all()


def test_static_equity_refuses_oversized_call():
    lib = ss_abi.load_library()
    hands, nh_a, mh, runs = _case(990, 3, [1128, 1128, 1128], 1000, 4)
    print("no-alloc equity oversized call: np=3 hands=1128 runouts=1000", flush=True)
    rc, raw = _call(lib.ss_leaf_equity_static, 3, hands, nh_a, mh, runs)
    assert rc == -2
    assert "scratch" in ss_abi.last_error(lib)
    assert (raw == -1.0).all()   # nothing written
