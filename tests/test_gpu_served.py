"""Served form (ss_server_*): continuous batching on the GPU.

Every request's subgame is computed by its own workgroups with fixed-order
reductions, so its result must not depend on what else is resident: served
results equal the plain form's bit for bit, and a subgame solved alone equals
the same subgame solved inside a mix that arrives around it.
"""

import time

import numpy as np
import pytest

import ss_api
import workloads as wl

pytestmark = pytest.mark.gpu


def _wait(srv, tickets, timeout_s=120.0):
    out = {}
    t0 = time.time()
    while len(out) < len(tickets):
        for t in tickets:
            if t in out:
                continue
            r = srv.poll(t)
            if r is not None:
                out[t] = r
        assert time.time() - t0 < timeout_s, "served requests did not finish"
        time.sleep(0.002)
    return [out[t] for t in tickets]


def _same(a, b):
    assert np.array_equal(a.avg, b.avg)
    assert np.array_equal(a.root_strategy, b.root_strategy)
    assert np.array_equal(a.root_ev, b.root_ev)


def _mix_spots():
    spots = [wl.small_spot(60 + i, num_players=2 + (i % 2), hands_per_player=9) for i in range(5)]
    spots.append(wl.river_2p(seed=91))
    spots.append(wl.mix_river(wl.BASE_SEED + 4000 + 3, max_players=3))
    spots.append(wl.mix_turn(wl.BASE_SEED + 200000 + 1, max_players=3))
    return spots


def test_served_equals_plain():
    spots = _mix_spots()
    probs = [ss_api.Problem.from_spot(s) for s in spots]
    srv = ss_api.Server(max_slots=16, iters_per_step=5, check_every=10)
    try:
        tickets = [srv.submit(p, 20, target=0.0) for p in probs]
        got = _wait(srv, tickets)
    finally:
        srv.close()
    for p, g in zip(probs, got):
        want = ss_api.solve_plain(p, 20)
        assert g.iterations == 20
        _same(g, want)
        assert g.exploitability == pytest.approx(want.exploitability, rel=1e-12)


def test_served_alone_equals_in_mix():
    """Early stopping included: checks every 10 iterations against 0.5% pot."""
    target = wl.mix_river(wl.BASE_SEED + 4000 + 7, max_players=3)
    p = ss_api.Problem.from_spot(target)
    srv = ss_api.Server(max_slots=8, iters_per_step=10, check_every=10)
    try:
        alone = _wait(srv, [srv.submit(p, 300, target=0.5, check_every=10)])[0]
    finally:
        srv.close()

    others = [ss_api.Problem.from_spot(s) for s in _mix_spots()]
    srv = ss_api.Server(max_slots=8, iters_per_step=10, check_every=10)
    try:
        first = [srv.submit(q, 120, target=0.5, check_every=10) for q in others[:4]]
        mine = srv.submit(p, 300, target=0.5, check_every=10)
        time.sleep(0.05)
        later = [srv.submit(q, 120, target=0.5, check_every=10) for q in others[4:]]
        res = _wait(srv, first + [mine] + later)
    finally:
        srv.close()
    mixed = res[len(first)]
    assert mixed.iterations == alone.iterations
    assert mixed.converged == alone.converged
    assert mixed.exploitability == alone.exploitability
    _same(mixed, alone)


def test_served_admission_by_budget():
    """A budget that fits one subgame at a time: requests queue, then run."""
    spots = [wl.small_spot(80 + i, num_players=2, hands_per_player=8) for i in range(4)]
    probs = [ss_api.Problem.from_spot(s) for s in spots]
    fp = max(ss_api.plan_bytes(q)[1] for q in probs)
    srv = ss_api.Server(hbm_budget_bytes=int(fp * 1.5), max_slots=4, iters_per_step=5)
    try:
        tickets = [srv.submit(q, 15, target=0.0) for q in probs]
        got = _wait(srv, tickets)
    finally:
        srv.close()
    for q, g in zip(probs, got):
        _same(g, ss_api.solve_plain(q, 15))


def test_served_cancel_queued():
    spots = [wl.small_spot(90 + i, num_players=2, hands_per_player=8) for i in range(3)]
    probs = [ss_api.Problem.from_spot(s) for s in spots]
    fp = max(ss_api.plan_bytes(q)[1] for q in probs)
    srv = ss_api.Server(hbm_budget_bytes=int(fp * 1.5), max_slots=4, iters_per_step=5)
    try:
        t0 = srv.submit(probs[0], 400, target=0.0)
        t1 = srv.submit(probs[1], 10, target=0.0)
        srv.cancel(t1)
        with pytest.raises(ss_api.SolveError):
            _wait(srv, [t1])
        r0 = _wait(srv, [t0])[0]
        assert r0.iterations == 400
    finally:
        srv.close()


@pytest.mark.parametrize("sparse", [False, True])
def test_adaptive_checks_same_retirement_in_every_form(sparse):
    """SS_FLAG_ADAPTIVE_CHECKS: the first check at check_every, later checks
    where each subgame's exploitability would reach the target at t^-2 decay,
    on the 5-iteration step. The schedule depends only on the subgame's own
    exploitability, so the plain, batched and served forms retire each subgame
    at the same iteration with the same results bit for bit; each retires at or
    below the target, on the step. SS_FLAG_SPARSE_CHECKS keeps the checks on
    multiples of check_every (25) and skips those before the predicted
    crossing."""
    import ss_abi
    flags = ss_abi.SS_FLAG_SPARSE_CHECKS if sparse else ss_abi.SS_FLAG_ADAPTIVE_CHECKS
    step = 25 if sparse else 5
    spots = [wl.mix_river(wl.BASE_SEED + 4000 + i, max_players=3) for i in range(6)]
    probs = [ss_api.Problem.from_spot(s) for s in spots]
    b = ss_api.Batch(probs, 2000, target=0.5, check_every=25, flags=flags)
    reps, _ = b.run()
    batched = [b.output(i, reps[i]) for i in range(len(probs))]
    b.close()
    srv = ss_api.Server(max_slots=8, iters_per_step=5, check_every=25)
    try:
        served = _wait(srv, [srv.submit(p, 2000, target=0.5, check_every=25, flags=flags) for p in probs])
    finally:
        srv.close()
    for i, p in enumerate(probs):
        plain = ss_api.solve_plain(p, 2000, target=0.5, check_every=25, flags=flags)
        fixed = ss_api.solve_plain(p, 2000, target=0.5, check_every=25)
        print(f"subgame {i}: {p.num_players} players, adaptive {plain.iterations} / fixed {fixed.iterations} "
              f"iterations, exploitability {plain.exploitability:.4f}% pot", flush=True)
        assert plain.converged and plain.exploitability <= 0.5
        assert plain.iterations % step == 0 and plain.iterations >= 25
        for g in (batched[i], served[i]):
            assert g.iterations == plain.iterations, i
            assert g.exploitability == pytest.approx(plain.exploitability, rel=1e-12), i
            _same(g, plain)


def test_served_precision_form_equals_plain():
    """BF16 hi/lo precision requests (2-player rivers) served beside FP32
    requests: each retires at the plain precision form's iteration with its
    results bit for bit, the FP32 ones as their plain form; a 3-player
    request asking for the precision form is refused at admission."""
    import ss_abi
    bf = ss_abi.SS_PREC_BF16X2
    prec = [ss_api.Problem.from_spot(wl.river_2p(seed=300 + i)) for i in range(3)]
    fp32 = [ss_api.Problem.from_spot(wl.mix_river(wl.BASE_SEED + 4000 + i, max_players=3)) for i in range(3)]
    three = ss_api.Problem.from_spot(wl.river_3p(seed=301))
    srv = ss_api.Server(max_slots=8, iters_per_step=5, check_every=25)
    try:
        tp = [srv.submit(p, 2000, target=0.5, check_every=25, precision=bf) for p in prec]
        tf = [srv.submit(p, 2000, target=0.5, check_every=25) for p in fp32]
        got = _wait(srv, tp + tf)
        bad = srv.submit(three, 100, target=0.5, check_every=25, precision=bf)
        with pytest.raises(ss_api.SolveError):
            _wait(srv, [bad])
    finally:
        srv.close()
    for i, p in enumerate(prec + fp32):
        want = ss_api.solve_plain(p, 2000, target=0.5, check_every=25, precision=bf if i < 3 else ss_abi.SS_PREC_FP32)
        g = got[i]
        print(f"request {i}: {'bf16x2' if i < 3 else 'fp32'}, {g.iterations} iterations, "
              f"exploitability {g.exploitability:.4f}% pot", flush=True)
        assert g.iterations == want.iterations and g.converged
        assert g.exploitability == pytest.approx(want.exploitability, rel=1e-12)
        _same(g, want)


def test_served_concurrent_clients_drain():
    """Four client threads submit small requests at once and cancel every
    fifth; drain returns only once every request has finished, its outputs
    read back off the serving loop included, so every ticket then polls
    final. Requests not cancelled equal the plain form bit for bit."""
    import threading
    n = 24
    spots = [wl.small_spot(200 + i, num_players=2 + (i % 2), hands_per_player=7) for i in range(n)]
    probs = [ss_api.Problem.from_spot(s) for s in spots]
    srv = ss_api.Server(max_slots=6, iters_per_step=5)
    tickets, cancelled, errors = {}, set(), []
    lock = threading.Lock()

    def client(k):
        try:
            for i in range(k, n, 4):
                t = srv.submit(probs[i], 30, target=0.0)
                with lock:
                    tickets[i] = t
                if i % 5 == 4:
                    srv.cancel(t)
                    with lock:
                        cancelled.add(i)
        except Exception as e:   # surfaced below
            errors.append(e)

    try:
        th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errors, errors
        srv.drain()
        got = {}
        for i, t in tickets.items():
            if i in cancelled:
                try:
                    srv.poll(t)   # finished before the cancel took effect, or cancelled
                except ss_api.SolveError:
                    pass
                continue
            r = srv.poll(t)
            assert r is not None, f"request {i} not final after drain"
            got[i] = r
    finally:
        srv.close()
    assert len(got) == n - len(cancelled)
    for i in sorted(got)[::3]:
        want = ss_api.solve_plain(probs[i], 30)
        assert got[i].iterations == 30
        _same(got[i], want)
