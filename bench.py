#!/usr/bin/env python3
"""Subgame solves/s at <= 0.5% pot exploitability on one MI355X.

One step = one batched solve of a fixed batch of synthetic river subgames:
every subgame starts from zero regrets and is retired when its FP64
exploitability is <= target (% of its starting pot), checked every
`check_every` iterations, or at the iteration cap. All inputs are uploaded
to HBM before the timed region (ss_batch_create); each step resets the state
on the device and solves the whole batch (ss_batch_run).

Printed JSON line (rank 0, one GPU): value = subgames solved per second of
wall time over the timed steps. Also reported: the plain form (one subgame
per launch sequence, ss_solve_gpu_ex) on a sample of the same subgames, the
batched/plain ratio, the roofline of the persistent CFR kernel (algorithmic
bytes per launch / event-timed launch duration), and CPU baselines timed on
this host: the CPU oracle of the same semantics on a bounded sample, and
(when built) the reference's CPU solver src/solver_v2.c on the 2-player
river shape.
"""

import argparse
import concurrent.futures as cf
import json
import multiprocessing as mp
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "poker-solver_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_spots(workload, n, seed):
    import workloads as wl
    if workload == "batch_mix":
        return wl.batch_mix(n, seed=seed, max_players=3)
    if workload == "river_2p":
        return [wl.river_2p(seed=seed + i) for i in range(n)]
    if workload == "river_3p":
        return [wl.river_3p(seed=seed + i) for i in range(n)]
    raise SystemExit(f"unknown workload {workload}")


# ---------------------------------------------------------------- CPU baselines
def _oracle_iter_time(args):
    """Seconds per CFR iteration of the CPU oracle (FP32, closed-form terminals)
    on one subgame, averaged over `iters` iterations."""
    workload, seed, idx, iters = args
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_bind as ob
    import workloads as wl
    spot = make_spots(workload, idx + 1, seed)[idx]
    td = ob.build_tree(spot.board, spot.pot, spot.stack, spot.bet_sizes, spot.num_players,
                       spot.acting_order, max_raises=spot.max_raises)
    prob = ob.Problem(td, spot.board, spot.hands, spot.weights)
    t0 = time.perf_counter()
    ob.solve(prob, iters, fp64=False, mode=ob.TERM_FAST)
    return (time.perf_counter() - t0) / iters


def cpu_oracle_baseline(workload, seed, idxs, gpu_iters, cores, iters=2):
    with mp.get_context("spawn").Pool(cores) as pool:
        t_iter = pool.map(_oracle_iter_time, [(workload, seed, i, iters) for i in idxs])
    per_solve = [t * max(1, gpu_iters[i]) for t, i in zip(t_iter, idxs)]
    return cores / float(np.mean(per_solve)), float(np.mean(t_iter))


def _sv2_iter_time(seed):
    """Seconds for one iteration of the reference CPU solver (src/solver_v2.c,
    compiled in place into oracle/_ref/solver_v2.so) on a 2-player river with
    1081 combos per player, pot 100 BB, stacks 400 BB, sizes {0.5, 1.0}."""
    import ctypes
    import workloads as wl
    lib = ctypes.CDLL(os.path.join(REPO, "oracle", "_ref", "solver_v2.so"))
    spot = wl.river_2p(seed=seed)
    buf = ctypes.create_string_buffer(1 << 20)   # >= sizeof(SolverV2)
    h = [np.ascontiguousarray(x, dtype=np.int32) for x in spot.hands]
    w = [np.ascontiguousarray(x, dtype=np.float32) for x in spot.weights]
    b = (ctypes.c_int * 5)(*spot.board)
    sizes = (ctypes.c_float * 2)(0.5, 1.0)
    P = ctypes.c_void_p
    rc = lib.sv2_init(buf, b, 5, h[0].ctypes.data_as(P), w[0].ctypes.data_as(P), len(h[0]),
                      h[1].ctypes.data_as(P), w[1].ctypes.data_as(P), len(h[1]),
                      spot.pot, spot.stack, sizes, 2)
    if rc != 0:
        return None
    lib.sv2_solve.restype = ctypes.c_float
    t0 = time.perf_counter()
    lib.sv2_solve(buf, 1, ctypes.c_float(0.0))
    dt = time.perf_counter() - t0
    lib.sv2_free(buf)
    return dt


def cpu_reference_baseline(cores, iters_per_solve=500):
    path = os.path.join(REPO, "oracle", "_ref", "solver_v2.so")
    if not os.path.exists(path):
        return None
    with mp.get_context("spawn").Pool(cores) as pool:
        ts = [t for t in pool.map(_sv2_iter_time, [20261203 + 1000 + i for i in range(cores)]) if t]
    if not ts:
        return None
    t_iter = float(np.mean(ts))
    return {"value": cores / (t_iter * iters_per_solve), "unit": "solves/s", "cores": cores,
            "kind": "reference",
            "sample": f"src/solver_v2.c (compiled in place), 2-player river 1081 combos/player, "
                      f"pot 100 BB, stacks 400 BB, sizes {{0.5,1.0}}: 1 iteration per process on {cores} "
                      f"processes, {t_iter:.1f} s/iteration, extrapolated to {iters_per_solve} iterations "
                      f"(its cost per iteration is constant)",
            "seconds_per_iteration": t_iter}


# ---------------------------------------------------------------- served form
def served_bench(n, rate, seed, target, check_every, max_iters, iters_per_step, log_every=50, plain_fifo=False,
                 flags=0):
    """Poisson stream through ss_server_*: 70% rivers (2-3 players, as the
    batch mix) and 30% turns (2-3 players, continuation leaves N(0, 0.3 pot)).
    Requests are built before the clock starts and submitted at their arrival
    times from a separate thread; latency = arrival to retirement (measured
    inside the server), throughput = completed / (last retirement - first
    arrival).

    plain_fifo: the same stream through a plain-GPU request loop as well, one
    request per solve in arrival order: each request's plain-form solve is
    timed (setup, upload, solve, checks) and its latency follows the
    single-server queue, finish_i = max(arrival_i, finish_{i-1}) + service_i."""
    import threading
    import ss_api
    import workloads as wl
    stream = wl.served_stream(n, rate, seed=seed, river_max_players=3, turn_max_players=3)
    probs = [ss_api.Problem.from_spot(sp) for _, sp in stream]
    kinds = [sp.kind for _, sp in stream]
    srv = ss_api.Server(max_slots=256, iters_per_step=iters_per_step, check_every=check_every)
    tickets = [None] * n
    t_start = time.perf_counter()

    def submitter():
        for i, (t_arr, _) in enumerate(stream):
            d = t_start + t_arr - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = srv.submit(probs[i], max_iters, target=target, check_every=check_every, flags=flags)

    th = threading.Thread(target=submitter)
    th.start()
    done = {}
    t_last = t_start
    while len(done) < n:
        for i in range(n):
            if i in done or tickets[i] is None:
                continue
            r = srv.poll(tickets[i])
            if r is not None:
                done[i] = r
                t_last = time.perf_counter()
                if len(done) % log_every == 0:
                    log(f"[bench] served {len(done)}/{n}")
        time.sleep(0.001)
    th.join()
    srv.close()
    lat = np.array([done[i].elapsed_ms for i in range(n)])
    its = np.array([done[i].iterations for i in range(n)])
    conv = np.array([done[i].converged for i in range(n)])
    first_arrival = stream[0][0]
    span = (t_last - t_start) - first_arrival
    out = {
        "workload": f"Poisson arrivals at {rate}/s, {n} requests: 70% river (players U{{2,3}}, batch-mix "
                    f"shapes), 30% turn (4-card board, 1128 combos, players U{{2,3}}, 1-2 sizes, "
                    f"continuation leaves N(0, 0.3 pot)); retire at <= {target}% pot checked every "
                    f"{check_every} iterations, cap {max_iters}",
        "rate_per_s": rate, "requests": n,
        "river": kinds.count("river"), "turn": kinds.count("turn"),
        "solves_per_s": round(n / span, 3),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                       "p99": round(float(np.percentile(lat, 99)), 2),
                       "max": round(float(lat.max()), 2)},
        "iterations": {"mean": float(its.mean()), "max": int(its.max())},
        "converged_fraction": float(conv.mean()),
        "iters_per_step": iters_per_step,
    }
    if plain_fifo:
        service = []
        for i, pr in enumerate(probs):
            r = ss_api.solve_plain(pr, max_iters, target=target, check_every=check_every, flags=flags)
            service.append(r.elapsed_ms)
            if (i + 1) % log_every == 0:
                log(f"[bench] plain request loop {i + 1}/{n}")
        finish, lat_p = 0.0, []
        for (t_arr, _), sv in zip(stream, service):
            finish = max(1000.0 * t_arr, finish) + sv
            lat_p.append(finish - 1000.0 * t_arr)
        lat_p = np.array(lat_p)
        out["plain_request_loop"] = {
            "latency_ms": {"p50": round(float(np.percentile(lat_p, 50)), 2),
                           "p99": round(float(np.percentile(lat_p, 99)), 2),
                           "max": round(float(lat_p.max()), 2)},
            "solves_per_s": round(n / ((finish - 1000.0 * first_arrival) / 1000.0), 3),
            "mean_service_ms": round(float(np.mean(service)), 2),
        }
    return out


# ---------------------------------------------------------------- precision form
def _fp64_reference(args):
    """FP64 oracle solve of one 2-player 1081-combo river (pot 100 BB, stacks
    400 BB, sizes 0.5 and 1 pot): average strategy and root values after
    `iters` iterations."""
    seed, iters = args
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_bind as ob
    import workloads as wl
    spot = wl.river_2p(seed=seed)
    td = ob.build_tree(spot.board, spot.pot, spot.stack, spot.bet_sizes, spot.num_players,
                       spot.acting_order, max_raises=spot.max_raises)
    prob = ob.Problem(td, spot.board, spot.hands, spot.weights)
    r = ob.solve(prob, iters, fp64=True, mode=ob.TERM_FAST)
    return np.asarray(r["avg"]), np.asarray(r["root_ev"])


def precision_bench(boards, iters, ref_boards, cores):
    """Precision form: `boards` 2-player 1081-combo river subgames (pot 100 BB,
    stacks 400 BB, sizes 0.5 and 1 pot, seeds base + b), `iters` iterations
    each with the BF16 hi/lo MFMA showdown and FP32 accumulation. Errors of
    the average strategy and of the root values against the FP64 oracle on
    the first `ref_boards` boards, with the FP32 closed form's errors on the
    same boards beside them."""
    import ss_abi
    import ss_api
    import workloads as wl
    seeds = [20261203 + 1000 + b for b in range(boards)]
    probs = [ss_api.Problem.from_spot(wl.river_2p(seed=sd)) for sd in seeds]
    batch = ss_api.Batch(probs, iters, precision=ss_abi.SS_PREC_BF16X2)
    batch.run()   # warm-up
    t0 = time.perf_counter()
    reps, st = batch.run()
    wall = time.perf_counter() - t0
    res = [batch.output(i, reps[i]) for i in range(boards)]
    batch.close()
    t1 = time.perf_counter()
    plain = ss_api.solve_plain(probs[0], iters, precision=ss_abi.SS_PREC_BF16X2)
    plain_s = time.perf_counter() - t1
    # MFMA work per traverser pass: 2 Mp^2 16 ceil(2S / 16) flops as issued (hi and lo
    # columns packed into 16-wide tiles, hands padded to Mp); 2 nh^2 2S algorithmic
    S = [sum(1 for n in range(p.tree.num_nodes) if p.tree.td.nodes[n].type == 2) for p in probs]
    mpad = [(len(p.hands[0]) + 63) // 64 * 64 for p in probs]
    flops = sum(2.0 * m * m * 16 * ((2 * s + 15) // 16) * 2 * iters for m, s in zip(mpad, S))
    alg_flops = sum(2.0 * len(p.hands[0]) ** 2 * 2 * s * 2 * iters for p, s in zip(probs, S))
    out = {
        "workload": f"{boards} 2-player rivers (1081 combos each, pot 100 BB, stacks 400 BB, sizes {{0.5,1}}, "
                    f"max 2 raises), {iters} iterations each, one batched (phased) solve with the BF16 hi/lo "
                    f"MFMA showdown",
        "boards": boards, "iterations": iters,
        "solves_per_s": round(boards / wall, 3),
        "plain_form_solves_per_s": round(1.0 / plain_s, 3),
        "device_ms": {"cfr": round(st.cfr_ms, 2), "checks": round(st.check_ms, 2)},
        "mfma_tflop": round(flops / 1e12, 3),
        "mfma_tflops_over_cfr_time": round(flops / (st.cfr_ms / 1000.0) / 1e12, 3),
        "mfma_algorithmic_tflop": round(alg_flops / 1e12, 3),
        "exploitability_pct_pot": {"mean": float(np.mean([r.exploitability for r in reps])),
                                   "max": float(np.max([r.exploitability for r in reps]))},
    }
    if ref_boards > 0:
        nref = min(ref_boards, boards)
        with mp.get_context("spawn").Pool(min(cores, nref)) as pool:
            refs = pool.map(_fp64_reference, [(sd, iters) for sd in seeds[:nref]])
        f32 = [ss_api.solve_plain(p, iters) for p in probs[:nref]]

        def errs(rs):
            sa, sm, ra, rm = [], [], [], []
            for r, (avg, rev), p in zip(rs, refs, probs):
                g = r.avg_full(p.tree.num_nodes).astype(np.float64)
                live = np.zeros_like(avg, dtype=bool)
                live[r.node_ids] = True
                d = np.abs(g - avg)[live]
                sa.append(float(d.max()))
                sm.append(float(d.mean()))
                pot = float(p.tree.td.starting_pot)
                e = np.abs(r.root_ev[:len(rev)].astype(np.float64) - rev) / pot   # node-0 cfv, last traverser
                ra.append(float(e.max()))
                rm.append(float(e.mean()))
            return {"avg_strategy_abs": {"max": max(sa), "mean": float(np.mean(sm))},
                    "root_value_abs_per_pot": {"max": max(ra), "mean": float(np.mean(rm))}}
        out["error_vs_fp64"] = {"boards": nref, "bf16x2": errs(res[:nref]), "fp32_closed_form": errs(f32)}
    return out


# ---------------------------------------------------------------- roofline traffic
def leaf_bench(reps=5):
    """Depth-limit leaf values of a 3-player full-range turn subgame, and of a
    3-player full-range flop subgame (workloads.flop_leaf_workload, record
    "flop")
    (workloads.turn_leaf_workload: 1128 combos per player, 12 sampled rivers,
    2000 sampled opponent pairs per hero and river, 28 leaves x 64 bias
    profiles). GPU: wall time of compute_turn_leaf_values (host runout and
    tuple lists, upload, kernels, 24 MiB read back). CPU: the oracle's
    vectorised equity on the same inputs plus its value loop on one leaf,
    scaled to all leaves; the oracle's equity must equal the GPU's bit for
    bit."""
    import leaf_values as lv
    import ss_abi
    import workloads
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import leaf_oracle as lo
    lib = ss_abi.load_library()
    board, infos, hands, mh = workloads.turn_leaf_workload(lib)
    lv.compute_turn_leaf_values(board, infos, mh, 1500, player_hands=hands)
    t = time.perf_counter()
    for _ in range(reps):
        out = lv.compute_turn_leaf_values(board, infos, mh, 1500, player_hands=hands)
    gpu_ms = (time.perf_counter() - t) * 1000 / reps
    nh = [len(h) for h in hands]
    hc = [[(a, b) for a, b, _ in hs] for hs in hands]
    t = time.perf_counter()
    raw = lo.runout_equity(hc, nh, lo.river_runouts(board, nh), mh)
    eq_ms = (time.perf_counter() - t) * 1000
    t = time.perf_counter()
    one = lo.values_from_equity(raw, nh, mh, [(0, infos[0].pot)])
    val_ms = (time.perf_counter() - t) * 1000 * len(infos)
    same = bool(np.array_equal(one[:64], out[:64])) if infos[0].leaf_idx == 0 else None
    rec = {"workload": "3-player turn, 1128 combos per player, 12 sampled rivers, 2000 sampled opponent "
                       f"pairs, {len(infos)} leaves x 64 bias profiles",
           "gpu_ms_per_call": round(gpu_ms, 3), "output_mib": round(out.nbytes / 2**20, 1),
           "cpu_oracle_ms": round(eq_ms + val_ms, 1), "cpu_oracle_equity_ms": round(eq_ms, 1),
           "cpu_kind": "port (numpy-vectorised oracle, 1 core; value loop timed on one leaf and scaled)",
           "speedup_vs_cpu_oracle": round((eq_ms + val_ms) / gpu_ms, 1), "leaf0_equal_oracle": same}
    # Flop: 12 sampled turns x 12 sampled rivers; the oracle's equity is timed
    # on the first 12 runouts and scaled to all 144.
    board, infos, hands, mh = workloads.flop_leaf_workload(lib)
    lv.compute_flop_leaf_equity(board, infos, mh, 1500, player_hands=hands)
    t = time.perf_counter()
    for _ in range(reps):
        out = lv.compute_flop_leaf_equity(board, infos, mh, 1500, player_hands=hands)
    fgpu = (time.perf_counter() - t) * 1000 / reps
    nh = [len(h) for h in hands]
    hc = [[(a, b) for a, b, _ in hs] for hs in hands]
    runs = lo.flop_runouts(board, nh)
    t = time.perf_counter()
    lo.runout_equity(hc, nh, runs[:12], mh)
    feq = (time.perf_counter() - t) * 1000 * len(runs) / 12
    t = time.perf_counter()
    lo.values_from_equity(np.zeros((3, mh)), nh, mh, [(0, infos[0].pot)])
    fval = (time.perf_counter() - t) * 1000 * len(infos)
    rec["flop"] = {"workload": f"3-player flop, 1176 combos per player, {len(runs)} sampled runouts (12 turns x "
                               f"12 rivers), 2000 sampled opponent pairs, {len(infos)} leaves x 64 bias profiles",
                   "gpu_ms_per_call": round(fgpu, 3), "output_mib": round(out.nbytes / 2**20, 1),
                   "cpu_oracle_ms": round(feq + fval, 1),
                   "cpu_kind": "port (numpy-vectorised oracle, 1 core; equity timed on 12 of the runouts and "
                               "the value loop on one leaf, both scaled)",
                   "speedup_vs_cpu_oracle": round((feq + fval) / fgpu, 1)}
    return rec


FLOP_BOARD = [43, 51, 1]   # Qs As 2d


def flop_spot(n_hands, lib=None):
    """The flop-through-river workload of the reference's own flop benchmark
    (bench/bench_gpu_flop.py:50-66): flop Qs As 2d, pot 650, stacks 9750, one
    0.75-pot size with one raise per street; each player holds the first
    n_hands two-card combos (c0 < c1) off the board, weight 1."""
    import flop_solve
    t = flop_solve.FlopTree(FLOP_BOARD, 650, 9750, [0.75], lib=lib)
    combos = [(a, b) for a in range(52) for b in range(a + 1, 52) if a not in FLOP_BOARD and b not in FLOP_BOARD]
    t.set_hands(combos[:n_hands], combos[:n_hands])
    return t


def flop_bench(hands=(80, 200), iters=200, reps=3, cpu_iters=1):
    """Flop-through-river solves (include/flop_solve.h) on flop_spot: the wall
    time of one fs_solve_gpu call (tree upload, per-board tables, the
    captured iteration loop, read-back), the median of `reps` calls, at each
    hand count; the per-iteration cost from the difference of a full and a
    tenth-length call. CPU: the FP32 flop-solve oracle (one core) timed on
    `cpu_iters` iterations of the smallest case and scaled to `iters`."""
    import flop_solve
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_bind as ob
    rec = {"workload": f"2-player flop through river with chance nodes: flop Qs As 2d, pot 650, stacks 9750, "
                       f"0.75-pot size, 1 raise per street, {iters} iterations (the reference's "
                       f"bench/bench_gpu_flop.py spot)", "iterations": iters, "cases": []}
    for nh in hands:
        t = flop_spot(nh)
        flop_solve.solve(t, 2)
        walls = []
        for _ in range(reps):
            t0 = time.perf_counter()
            flop_solve.solve(t, iters)
            walls.append((time.perf_counter() - t0) * 1000)
        t0 = time.perf_counter()
        flop_solve.solve(t, iters // 10)
        short = (time.perf_counter() - t0) * 1000
        ms = float(np.median(walls))
        rec["cases"].append({"hands_per_player": nh, "nodes": t.num_nodes, "gpu_ms_per_solve": round(ms, 2),
                             "gpu_ms_per_iteration": round((ms - short) / (iters - iters // 10), 4)})
        log(f"[bench] flop solve: {nh} hands, {t.num_nodes} nodes: {ms:.1f} ms per {iters}-iteration solve")
        t.free()
    t = flop_spot(hands[0])
    t0 = time.perf_counter()
    ob.fs_solve(t.td, cpu_iters)
    cpu_ms = (time.perf_counter() - t0) * 1000 * iters / cpu_iters
    t.free()
    rec["cpu_oracle_ms_per_solve"] = round(cpu_ms, 1)
    rec["cpu_kind"] = (f"port (C flop-solve oracle, FP32, 1 core) at {hands[0]} hands per player, "
                       f"{cpu_iters} iteration(s) timed and scaled to {iters}")
    rec["speedup_vs_cpu_oracle"] = round(cpu_ms / rec["cases"][0]["gpu_ms_per_solve"], 1)
    return rec


def reference_latency_bench(reps=3):
    """The realtime re-solve call at the setting of the reference's published
    latencies (ARCHITECTURE.md:73, 116-117: RTX 3060, 200 hands, 200
    iterations, ~67 ms river / ~236 ms flop): the drop-in StreetSolverGPU
    (legacy ABI, ss_solve_gpu) on a 2-player spot, each player's range the
    200 heaviest of U[0,1]-weighted combos, pot 10 BB, stacks 50 BB, sizes
    0.33 / 0.75 / 1.5 pot (the class's default); turn and flop streets with
    the 4 continuation strategies per player at the depth limit and zero leaf
    values (the wrapper's fallback). Wall time of solve(200), median of
    `reps` calls; tree build and range installation timed apart."""
    import warnings
    import street_solver_gpu as sg
    import workloads as wl
    rec = {"setting": "2 players, 200 hands each, 200 iterations, pot 10 BB, stacks 50 BB, sizes 0.33/0.75/1.5 pot, "
                      "legacy ABI (ss_solve_gpu) through the drop-in StreetSolverGPU",
           "published_rtx3060_ms": {"river": 67, "flop": 236}, "streets": {}}
    for street, nb, seed in (("river", 5, 11), ("turn", 4, 12), ("flop", 3, 13)):
        r = wl.rng_for(20261203 + 9000 + seed)
        board = [int(c) for c in r.choice(52, nb, replace=False)]
        combos = wl.all_combos(board)
        ranges = [[(int(a), int(b), float(w)) for (a, b), w in zip(combos, r.uniform(0, 1, len(combos)))]
                  for _ in range(2)]
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            t0 = time.perf_counter()
            solver = sg.StreetSolverGPU(board=board, pot_bb=10.0, stack_bb=50.0, player_ranges=ranges)
            build_ms = (time.perf_counter() - t0) * 1000
            solver.solve(iterations=200)
            walls = []
            for _ in range(reps):
                t0 = time.perf_counter()
                solver.solve(iterations=200)
                walls.append((time.perf_counter() - t0) * 1000)
        nodes = solver._tree.num_nodes
        solver.free()
        rec["streets"][street] = {"nodes": nodes, "solve_ms": round(float(np.median(walls)), 2),
                                  "build_ms": round(build_ms, 2)}
        log(f"[bench] reference-setting re-solve, {street}: {nodes} nodes, {np.median(walls):.1f} ms per solve(200)")
    return rec


def traffic_per_launch(algo_per_launch, workload):
    """HBM bytes per CFR iteration from the committed counter profile
    (profiles/cfr_iteration_traffic.md: calibrated FETCH_SIZE + WRITE_SIZE of a
    full-batch iteration), carried over to this run's mean iteration by the
    measured traffic / algorithmic-bytes ratio. None when the profile is absent
    or was taken on another workload."""
    path = os.path.join(REPO, "profiles", "cfr_iteration_traffic.json")
    if workload != "batch_mix" or not os.path.exists(path):
        return None
    d = json.load(open(path))
    return int(d["traffic_over_algo"] * algo_per_launch)


# ---------------------------------------------------------------- profiler workloads
def profile_traffic(a, subgames=256, iters=20, copy_mib=256):
    """For `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE`: one plain device
    copy of a known size (ssdbg_copy, kernel k_copy: the dispatch that
    calibrates a byte counter), then `iters` CFR iterations of the mixed batch
    with every subgame live (no exploitability checks). HBM bytes per
    iteration = sum over the CFR dispatches of counter (KiB) * 1024 *
    (copy bytes / the copy dispatch's reading of the counter) / iterations."""
    import ctypes
    import ss_abi
    import ss_api
    import workloads as wl
    lib = ss_abi.load_library()
    lib.ssdbg_copy.restype = ctypes.c_int
    lib.ssdbg_copy.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    ms = ctypes.c_float()
    copy_bytes = copy_mib << 20
    assert lib.ssdbg_copy(copy_bytes, 1, ctypes.byref(ms)) == 0, ss_abi.last_error()
    probs = [ss_api.Problem.from_spot(s) for s in wl.batch_mix(subgames, max_players=3)]
    algo = sum(ss_api.plan_bytes(p)[0] for p in probs)
    # the phased form with its automatic launch shapes, no trial iterations:
    # the profiled dispatches are exactly the `iters` iterations
    os.environ["SS_BATCH_MODE"] = "phased"
    os.environ["SS_NO_AUTOTUNE"] = "1"
    b = ss_api.Batch(probs, iters, flags=ss_abi.SS_FLAG_NO_EXPL)
    reps, st = b.run()
    b.close()
    print(json.dumps({"copy_bytes": copy_bytes, "copy_ms": ms.value, "iterations": iters,
                      "algo_bytes_per_iteration": algo, "cfr_ms": st.cfr_ms, "mode": st.mode}), flush=True)
    return 0


def profile_plain(a, iters=20, copy_mib=256):
    """For rocprofv3 runs over the plain form's level sweeps (--kernel-trace,
    then --pmc FETCH_SIZE, then --pmc WRITE_SIZE): one plain device copy of a
    known size (ssdbg_copy, kernel k_copy: the dispatch that calibrates the
    byte counters), then the 3-player 1081-combo river of the plain
    benchmark configuration (pot 150 BB, stacks 500 BB, sizes {0.33, 0.75},
    40% of combos live, max 2 raises) for `iters` plain-form iterations with
    no exploitability evaluation: per traverser pass one k_forward launch per
    level, k_terminals, one k_backward launch per level."""
    import ctypes
    import ss_abi
    import ss_api
    import workloads as wl
    lib = ss_abi.load_library()
    lib.ssdbg_copy.restype = ctypes.c_int
    lib.ssdbg_copy.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    ms = ctypes.c_float()
    assert lib.ssdbg_copy(copy_mib << 20, 1, ctypes.byref(ms)) == 0, ss_abi.last_error()
    p = ss_api.Problem.from_spot(wl.river_3p())
    r = ss_api.solve_plain(p, iters, flags=ss_abi.SS_FLAG_NO_EXPL)
    td = p.tree.td
    print(json.dumps({"copy_bytes": copy_mib << 20, "iterations": iters, "nodes": p.tree.num_nodes,
                      "kernel_ms": r.kernel_ms, "levels": int(td.max_depth) + 1}), flush=True)
    return 0


def profile_precision(a, boards=64, iters=50):
    """64 two-player 1081-combo rivers (pot 100 BB, stacks 400 BB, sizes
    {0.5, 1}) for 50 iterations in one batched precision-form solve (BF16
    hi/lo MFMA showdown), then one of them in the plain precision form."""
    import ss_abi
    import ss_api
    import workloads as wl
    probs = [ss_api.Problem.from_spot(wl.river_2p(seed=20261203 + 1000 + b)) for b in range(boards)]
    b = ss_api.Batch(probs, iters, flags=ss_abi.SS_FLAG_NO_EXPL, precision=ss_abi.SS_PREC_BF16X2)
    reps, st = b.run()
    print(f"batched: {st.cfr_ms:.2f} ms for {boards} boards x {iters} iterations", flush=True)
    b.close()
    r = ss_api.solve_plain(probs[0], iters, precision=ss_abi.SS_PREC_BF16X2, flags=ss_abi.SS_FLAG_NO_EXPL)
    print(f"plain: {r.kernel_ms:.2f} ms for 1 board x {iters} iterations", flush=True)
    return 0


def profile_served(a):
    """The served workload of the bench (Poisson stream, 70% river / 30%
    turn, 2-3 players), 60 requests at 40/s."""
    r = served_bench(60, 40.0, 20261203 + 5000, 0.5, 25, 2000, 5, log_every=20)
    print(r["solves_per_s"], r["latency_ms"], flush=True)
    return 0


def profile_graph(a, subgames=256, iters=25):
    """The bench's mixed batch for `iters` iterations with every subgame live,
    replayed from the captured graph (5 iterations per graph): a kernel trace
    of the path the bench times, to compare with the direct launches
    (SS_NO_GRAPH=1) of the other profiles."""
    import ss_abi
    import ss_api
    import workloads as wl
    probs = [ss_api.Problem.from_spot(s) for s in wl.batch_mix(subgames, max_players=3)]
    os.environ["SS_BATCH_MODE"] = "phased"
    os.environ["SS_NO_AUTOTUNE"] = "1"
    b = ss_api.Batch(probs, iters, check_every=25, flags=ss_abi.SS_FLAG_NO_EXPL)
    reps, st = b.run()
    b.close()
    print(json.dumps({"iterations": iters, "cfr_ms": st.cfr_ms, "ms_per_iteration": st.cfr_ms / iters}), flush=True)
    return 0


def profile_served_sweep(a, rates=(10.0, 25.0, 50.0, 75.0, 100.0, 150.0), n=200):
    """Latency against load of the served form: the bench's Poisson stream
    (70% river / 30% turn, 2-3 players, adaptive checks) at several arrival
    rates, 200 requests each, with the plain-GPU request loop on the same
    stream beside it. One JSON line with a record per rate."""
    import ss_abi
    rows = []
    for i, rate in enumerate(rates):
        r = served_bench(n, rate, 20261203 + 5000, 0.5, 25, 2000, 5, log_every=100, plain_fifo=(i == 0),
                         flags=ss_abi.SS_FLAG_ADAPTIVE_CHECKS)
        log(f"[sweep] {rate}/s: {r['solves_per_s']} solves/s, p50 {r['latency_ms']['p50']} ms, "
            f"p99 {r['latency_ms']['p99']} ms")
        rows.append(r)
    print(json.dumps({"served_sweep": rows}), flush=True)
    return 0


def _profile_leaf(flop, reps=10):
    """`reps` calls of the turn (or flop) leaf-value producer on the 3-player
    full-range workload, then the two library calls alone on prepared
    arrays, each with its wall time per call."""
    import ctypes
    import leaf_values as lv
    import ss_abi
    import workloads
    lib = ss_abi.load_library()
    if flop:
        board, infos, hands, mh = workloads.flop_leaf_workload(lib)
        produce, runouts = lv.compute_flop_leaf_equity, lv.flop_runouts
    else:
        board, infos, hands, mh = workloads.turn_leaf_workload(lib)
        produce, runouts = lv.compute_turn_leaf_values, lv.turn_runouts
    print(f"{'flop' if flop else 'turn'} 3p leaves: {len(infos)} leaf infos, {mh} hands per player", flush=True)
    produce(board, infos, mh, 1500, player_hands=hands)
    t = time.perf_counter()
    for _ in range(reps):
        out = produce(board, infos, mh, 1500, player_hands=hands)
    print(f"{(time.perf_counter() - t) * 1000 / reps:.3f} ms per call, output {out.shape}", flush=True)
    nh = np.array([len(h) for h in hands], dtype=np.int32)
    hc = np.array([[(x, y) for x, y, _ in hs] for hs in hands], dtype=np.int32)
    run = np.array(runouts(board, list(nh)), dtype=np.int32)
    raw = np.zeros((len(hands), mh))
    ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    idx = np.array([li.leaf_idx for li in infos], dtype=np.int32)
    pot = np.array([li.pot for li in infos], dtype=np.int32)
    t = time.perf_counter()
    for _ in range(reps):
        lib.ss_leaf_equity(len(hands), hc.ctypes.data_as(ip), nh.ctypes.data_as(ip), mh, run.ctypes.data_as(ip),
                           len(run), raw.ctypes.data_as(dp))
    eq_ms = (time.perf_counter() - t) * 1000 / reps
    t = time.perf_counter()
    for _ in range(reps):
        lib.ss_leaf_values(len(hands), raw.ctypes.data_as(dp), nh.ctypes.data_as(ip), mh, idx.ctypes.data_as(ip),
                           pot.ctypes.data_as(ip), len(infos), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    val_ms = (time.perf_counter() - t) * 1000 / reps
    print(f"ss_leaf_equity {eq_ms:.3f} ms, ss_leaf_values {val_ms:.3f} ms per call", flush=True)
    return 0


# ---------------------------------------------------------------- error vs the FP64 oracle
def _fp64_solution(args):
    """FP64 oracle solve of a workloads spot: (average strategy [N][8][max_h],
    exploitability % pot). Runs in a worker process (one OpenMP thread)."""
    kind, seed, iters = args
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_bind as ob
    import workloads as wl
    spot = wl.river_2p(seed=seed) if kind == "river_2p" else wl.river_3p(seed=seed)
    td = ob.build_tree(spot.board, spot.pot, spot.stack, spot.bet_sizes, spot.num_players,
                       spot.acting_order, max_raises=spot.max_raises)
    prob = ob.Problem(td, spot.board, spot.hands, spot.weights)
    r = ob.solve(prob, iters, fp64=True, mode=ob.TERM_FAST)
    e, _ = ob.exploitability(prob, r["avg"], mode=ob.TERM_FAST)
    return np.asarray(r["avg"]), float(e)


def _fp32_oracle_solution(args):
    """FP32 oracle solve (a second FP32 implementation; mode = terminal mode)
    of a 2-player river: (average strategy, its FP64 exploitability % pot)."""
    seed, iters, mode = args
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_bind as ob
    import workloads as wl
    spot = wl.river_2p(seed=seed)
    td = ob.build_tree(spot.board, spot.pot, spot.stack, spot.bet_sizes, spot.num_players,
                       spot.acting_order, max_raises=spot.max_raises)
    prob = ob.Problem(td, spot.board, spot.hands, spot.weights)
    r = ob.solve(prob, iters, fp64=False, mode=mode)
    e, _ = ob.exploitability(prob, r["avg"], mode=ob.TERM_FAST)
    return np.asarray(r["avg"]), float(e)


def _own_reach(td, avg, weights, mh):
    """pi[p, n, h]: player p's own reach at node n for hand h under strategy
    avg ([N][8][mh]), starting from the raw range weights."""
    npl = len(weights)
    pi = np.zeros((npl, td.num_nodes, mh))
    for p in range(npl):
        pi[p, 0, :len(weights[p])] = weights[p]
    for i in range(td.num_nodes):
        n = td.level_order[i]
        nd = td.nodes[n]
        if nd.type != 0:
            continue
        for a in range(nd.num_children):
            c = td.children[nd.first_child + a]
            pi[:, c] = pi[:, n]
            pi[nd.player, c] *= avg[n, a]
    return pi


def strategy_error(td, avg, ref, weights, hands):
    """Average-strategy error of `avg` against `ref` (both [N][8][mh]) over
    the info-set rows (decision node, hand of its actor): the mean of
    max_a |avg - ref| weighted by the actor's own reach under ref, the
    fraction of rows with positive reach whose error exceeds 1e-4, and the
    largest error on such a row."""
    mh = ref.shape[2]
    pi = _own_reach(td, ref, weights, mh)
    num = den = 0.0
    reached = over = 0
    worst = 0.0
    for n in range(td.num_nodes):
        nd = td.nodes[n]
        if nd.type != 0:
            continue
        p, na, nh = nd.player, nd.num_children, len(hands[nd.player])
        err = np.abs(avg[n, :na, :nh].astype(np.float64) - ref[n, :na, :nh]).max(axis=0)
        w = pi[p, n, :nh]
        live = w > 0
        num += float((w * err).sum())
        den += float(w.sum())
        reached += int(live.sum())
        over += int((err[live] > 1e-4).sum())
        if live.any():
            worst = max(worst, float(err[live].max()))
    return {"reach_weighted_abs": num / den if den > 0 else 0.0, "reached_rows": reached,
            "frac_reached_rows_above_1e-4": over / max(1, reached), "max_abs_reached": worst}


def _dequant_reference(path, td, hands, mh):
    z = np.load(path)
    ref = np.zeros((td.num_nodes, 8, mh))
    off = 0
    q = z["avg_q16"].astype(np.float64) / 65535.0
    for n, na in zip(z["node_ids"], z["num_actions"]):
        nh = len(hands[td.nodes[int(n)].player])
        ref[n, :na, :nh] = q[off:off + na * nh].reshape(na, nh)
        off += na * nh
    return ref, float(z["exploitability"]), int(z["iterations"])


def error_report(a):
    """FP32 and BF16x2 solutions against the FP64 oracle at the configs' own
    iteration counts, per subgame: the 2-player 1081-combo river at 500
    iterations (8 seeds), the 64 precision-form boards at 1000 iterations
    (BF16 hi/lo MFMA showdown and the FP32 closed form), and the 3-player
    river at 1000 iterations (FP64 reference precomputed in
    tests/golden/river_3p_fp64_1000it.npz by tests/golden/make_fp64_reference.py).
    Exploitability is each solution's FP64 value (% pot); relative error =
    |e - e_fp64| / e_fp64."""
    import ss_abi
    import ss_api
    import workloads as wl
    cores = a.cpu_cores or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = "1"
    base = 20261203 + 1000
    rep = {"cores": cores}
    t0 = time.perf_counter()
    # GPU solutions
    seeds3 = [base + b for b in range(64)]
    probs = [ss_api.Problem.from_spot(wl.river_2p(seed=sd)) for sd in seeds3]
    gpu = {}
    for name, prec in (("bf16x2", ss_abi.SS_PREC_BF16X2), ("fp32", ss_abi.SS_PREC_FP32)):
        b = ss_api.Batch(probs, 1000, precision=prec)
        reps, _ = b.run()
        gpu[name] = [b.output(i, reps[i]) for i in range(len(probs))]
        b.close()
    gpu1 = [ss_api.solve_plain(probs[i], 500) for i in range(8)]
    log(f"[errors] GPU solves done ({time.perf_counter() - t0:.1f} s); FP64 oracle on {cores} processes")
    jobs = [("river_2p", sd, 1000) for sd in seeds3] + [("river_2p", seeds3[i], 500) for i in range(8)]
    rows3, rows1, refs = [], [], {}
    with mp.get_context("spawn").Pool(cores) as pool:
        for k, (ref, e_ref) in enumerate(pool.imap(_fp64_solution, jobs)):
            i = k if k < 64 else k - 64
            p = probs[i]
            td, hands, w = p.tree.td, p.hands, p.weights
            if k < 64:
                refs[i] = (ref, e_ref)
                row = {"board": i, "expl_fp64": e_ref}
                for name in ("bf16x2", "fp32"):
                    r = gpu[name][i]
                    row[name] = {"expl": r.exploitability,
                                 "expl_rel_err": abs(r.exploitability - e_ref) / e_ref,
                                 **strategy_error(td, r.avg_full(td.num_nodes), ref, w, hands)}
                rows3.append(row)
            else:
                r = gpu1[i]
                rows1.append({"subgame": i, "expl_fp64": e_ref, "expl": r.exploitability,
                              "expl_rel_err": abs(r.exploitability - e_ref) / e_ref,
                              **strategy_error(td, r.avg_full(td.num_nodes), ref, w, hands)})
            log(f"[errors] {'board' if k < 64 else 'river_2p 500-it subgame'} {i} done "
                f"({time.perf_counter() - t0:.1f} s)")
    # the FP32 envelope: the oracle's own FP32 solve of the same boards
    # (closed-form terminals) against the same FP64 solutions
    with mp.get_context("spawn").Pool(cores) as pool:
        jobs32 = [(sd, 1000, 1) for sd in seeds3]
        for i, (avg32, e32) in enumerate(pool.imap(_fp32_oracle_solution, jobs32)):
            ref, e_ref = refs[i]
            p = probs[i]
            rows3[i]["fp32_oracle_closed_form"] = {"expl": e32, "expl_rel_err": abs(e32 - e_ref) / e_ref,
                                                   **strategy_error(p.tree.td, avg32, ref, p.weights, p.hands)}
            log(f"[errors] FP32 oracle board {i} done ({time.perf_counter() - t0:.1f} s)")
    rep["precision_boards_1000it"] = rows3
    rep["river_2p_500it"] = rows1
    sp3 = wl.river_3p()
    p3 = ss_api.Problem.from_spot(sp3)
    r3 = ss_api.solve_plain(p3, 1000)
    ref3, e3, it3 = _dequant_reference(os.path.join(REPO, "tests", "golden", "river_3p_fp64_1000it.npz"),
                                       p3.tree.td, p3.hands, p3.max_h)
    rep["river_3p_1000it"] = {"iterations": it3, "expl_fp64": e3, "expl": r3.exploitability,
                              "expl_rel_err": abs(r3.exploitability - e3) / e3,
                              **strategy_error(p3.tree.td, r3.avg_full(p3.tree.num_nodes), ref3, p3.weights,
                                               p3.hands)}

    def summ(rows, key=None):
        get = (lambda r: r[key]) if key else (lambda r: r)
        out = {}
        for f in ("expl_rel_err", "reach_weighted_abs", "frac_reached_rows_above_1e-4", "max_abs_reached"):
            v = np.array([get(r)[f] for r in rows])
            out[f] = {"mean": float(v.mean()), "max": float(v.max())}
        return out
    rep["summary"] = {"precision_boards_bf16x2": summ(rows3, "bf16x2"), "precision_boards_fp32": summ(rows3, "fp32"),
                      "precision_boards_fp32_oracle_closed_form": summ(rows3, "fp32_oracle_closed_form"),
                      "river_2p_500it": summ(rows1), "river_3p_1000it": rep["river_3p_1000it"]}
    print(json.dumps(rep), flush=True)
    return 0


def _profile_flop(iters=50):
    """One flop_spot solve at 200 hands per player for a kernel trace."""
    import flop_solve
    t = flop_spot(200)
    flop_solve.solve(t, iters)
    log(f"[profile] flop solve: {t.num_nodes} nodes, {iters} iterations")
    return 0


PROFILE_TARGETS = {
    "traffic": profile_traffic,
    "precision": profile_precision,
    "served": profile_served,
    "served-sweep": profile_served_sweep,
    "graph": profile_graph,
    "reference-latency": lambda a: print(json.dumps(reference_latency_bench()), flush=True) or 0,
    "leaf": lambda a: _profile_leaf(False),
    "leaf-flop": lambda a: _profile_leaf(True),
    "error-report": error_report,
    "flop": lambda a: _profile_flop(),
    "plain": profile_plain,
}


# ---------------------------------------------------------------- main
def _read_output(batch, i):
    """One subgame's outputs through the C ABI: ss_batch_output, then
    ss_free_output (the reference wrapper's pattern, street_solver_gpu.py)."""
    import ctypes
    import ss_abi
    out = ss_abi.SSOutput()
    rc = batch.lib.ss_batch_output(batch.h, int(i), ctypes.byref(out))
    if rc == 0:
        batch.lib.ss_free_output(ctypes.byref(out))
    return rc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batch_mix", choices=["batch_mix", "river_2p", "river_3p"])
    ap.add_argument("--subgames", type=int, default=256)
    ap.add_argument("--seed", type=int, default=20261203 + 4000)
    ap.add_argument("--target", type=float, default=0.5, help="exploitability target, %% of pot")
    ap.add_argument("--check-every", type=int, default=25)
    ap.add_argument("--fixed-checks", action="store_true",
                    help="plain and served forms: check every --check-every iterations (default: the adaptive "
                         "schedule, SS_FLAG_ADAPTIVE_CHECKS)")
    ap.add_argument("--fixed-batched", action="store_true",
                    help="batched form: check every subgame every --check-every iterations (default: the sparse "
                         "schedule, SS_FLAG_SPARSE_CHECKS)")
    ap.add_argument("--max-iters", type=int, default=2000)
    ap.add_argument("--plain-sample", type=int, default=8)
    ap.add_argument("--cpu-cores", type=int, default=0, help="0 = min(16, host cores)")
    ap.add_argument("--cpu-sample", type=int, default=16, help="subgames timed by the CPU oracle")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-ref", action="store_true")
    ap.add_argument("--served-requests", type=int, default=200, help="0 skips the served-form run")
    ap.add_argument("--served-rate", type=float, default=25.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--served-iters-per-step", type=int, default=5)
    ap.add_argument("--precision-boards", type=int, default=64, help="0 skips the precision-form run")
    ap.add_argument("--precision-iters", type=int, default=1000)
    ap.add_argument("--precision-ref-boards", type=int, default=8, help="boards compared with the FP64 oracle")
    ap.add_argument("--no-leaf", action="store_true", help="skip the turn leaf-value producer run")
    ap.add_argument("--no-flop", action="store_true", help="skip the flop-through-river solve run")
    ap.add_argument("--fresh-steps", type=int, default=2,
                    help="steps of the setup-inclusive rate (fresh subgames each step); 0 skips it")
    ap.add_argument("--profile-target", default=None, choices=sorted(PROFILE_TARGETS),
                    help="run only one workload for a profiler (rocprofv3) and exit")
    a = ap.parse_args()

    if a.profile_target:
        return PROFILE_TARGETS[a.profile_target](a)

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: subgames are independent and shard across GPUs by request "
              "routing with no collectives, so no multi-GPU figure is produced (run with --gpus 1).",
              file=sys.stderr)
        return 2

    import ss_abi
    import ss_api
    spots = make_spots(a.workload, a.subgames, a.seed)
    log(f"[bench] building {len(spots)} subgames ({a.workload})")
    problems = [ss_api.Problem.from_spot(s) for s in spots]
    nodes = [p.tree.num_nodes for p in problems]
    players = [p.num_players for p in problems]
    # check schedules: the batched form keeps its check rounds on multiples of
    # check_every (one round serves the whole batch) and skips, per subgame,
    # the rounds before its predicted crossing; the plain and served forms,
    # whose checks are per subgame, use the fine adaptive schedule
    bflags = 0 if a.fixed_batched else ss_abi.SS_FLAG_SPARSE_CHECKS
    flags = 0 if a.fixed_checks else ss_abi.SS_FLAG_ADAPTIVE_CHECKS
    batch = ss_api.Batch(problems, a.max_iters, target=a.target, check_every=a.check_every, flags=bflags)
    log(f"[bench] uploaded; nodes mean {np.mean(nodes):.0f} max {max(nodes)}; players {np.bincount(players)[2:].tolist()}")

    for w in range(a.warmup):
        batch.run()
        log(f"[bench] warmup {w + 1}/{a.warmup} done")
    reps_all, stats_all = [], []
    t0 = time.perf_counter()
    for k in range(a.steps):
        reps, st = batch.run()
        reps_all.append(reps)
        stats_all.append(st)
        log(f"[bench] step {k + 1}/{a.steps}: {st.run_ms:.1f} ms, max iters {st.max_iterations}")
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / a.steps
    value = len(problems) * a.steps / (t1 - t0)

    # Setup-inclusive rate: fresh subgames every step (other seeds), with the
    # host tree build, upload, planning, mode choice, graph capture, the solve
    # and the read-back of every output inside the clock.
    setup_rec = None
    if a.fresh_steps > 0:
        fresh = [make_spots(a.workload, a.subgames, a.seed + 100000 * (k + 1)) for k in range(a.fresh_steps)]
        tf0 = time.perf_counter()
        for k, spots_k in enumerate(fresh):
            probs_k = [ss_api.Problem.from_spot(s) for s in spots_k]
            bk = ss_api.Batch(probs_k, a.max_iters, target=a.target, check_every=a.check_every, flags=bflags)
            reps_k, _ = bk.run()
            # every output read back through the C ABI (ss_batch_output fills
            # the caller's SSOutput, ss_free_output releases it), 8 host threads
            with cf.ThreadPoolExecutor(8) as ex:
                rcs = list(ex.map(lambda i: _read_output(bk, i), range(len(probs_k))))
            assert all(rc == 0 for rc in rcs), rcs
            bk.close()
            log(f"[bench] fresh-batch step {k + 1}/{a.fresh_steps} done")
        tf1 = time.perf_counter()
        setup_rec = {"value": round(len(problems) * a.fresh_steps / (tf1 - tf0), 3), "unit": "solves/s",
                     "steps": a.fresh_steps,
                     "ms_per_step": round((tf1 - tf0) * 1000.0 / a.fresh_steps, 3),
                     "what": "a fresh batch of new subgames per step: host tree build, upload, planning, "
                             "mode choice, graph capture, solve to target and read-back of every output "
                             "(ss_batch_output on 8 host threads) inside the clock (synthetic input "
                             "generation outside)"}

    reps = reps_all[-1]
    iters = np.array([r.iterations for r in reps])
    expl = np.array([r.exploitability for r in reps])
    conv = np.array([r.converged for r in reps])
    cfr_ms = sum(s.cfr_ms for s in stats_all)
    check_ms = sum(s.check_ms for s in stats_all)
    algo = sum(s.algo_bytes for s in stats_all)
    launches = sum(s.launches for s in stats_all)
    achieved = algo / (cfr_ms / 1000.0) / 1e9 if cfr_ms > 0 else 0.0

    # plain form on a sample of the same subgames
    plain_ms = []
    for i in range(min(a.plain_sample, len(problems))):
        r = ss_api.solve_plain(problems[i], a.max_iters, target=a.target, check_every=a.check_every, flags=flags)
        plain_ms.append(r.kernel_ms)
    plain_sps = 1000.0 / float(np.mean(plain_ms)) if plain_ms else None
    log(f"[bench] plain form: {np.mean(plain_ms):.1f} ms per solve on {len(plain_ms)} subgames")

    line = {
        "metric": "subgame solves/s at <=0.5% pot exploitability; p50/p99 latency vs plain GPU "
                  "(river subgames with players U{2,3}: 4-6 player full-range subgames are refused)",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"{a.workload}: {len(problems)} river subgames per batched solve"
                        + (", players U{2,3}, 1-4 sizes from {0.25,0.5,0.75,1,1.5} pot + all-in, range density "
                           "U[0.1,1], pot U[50,300] BB, stacks U[200,1000] BB, max 2 raises"
                           if a.workload == "batch_mix" else ""),
            "subgames": len(problems),
            "target_exploitability_pct_pot": a.target,
            "check_every": a.check_every,
            "check_schedule": {
                "batched": "every check_every iterations" if a.fixed_batched else
                           "sparse (SS_FLAG_SPARSE_CHECKS): rounds on multiples of check_every; a subgame skips "
                           "the rounds before the iteration where its exploitability would reach the target if it "
                           "fell as t^-2, at most 3 in a row",
                "plain_and_served": "every check_every iterations" if a.fixed_checks else
                                    "adaptive (SS_FLAG_ADAPTIVE_CHECKS): first at check_every, then where each "
                                    "subgame's exploitability would reach the target if it fell as t^-2, on the "
                                    "5-iteration step, at most 4 x check_every apart"},
            "max_iterations": a.max_iters,
            "form": ("batched, phased: per traverser pass one launch per tree level / terminal class over the "
                     "(subgame, node) tasks of every live subgame, check_every iterations per hipGraph replay"
                     if stats_all[-1].mode == 1 else
                     "batched, persistent: one workgroup per subgame, one launch per CFR iteration, "
                     "hipGraph-captured"),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic_per_launch(algo / max(1, launches), a.workload),
            "traffic_source": "profiles/cfr_iteration_traffic.md (calibrated FETCH_SIZE + WRITE_SIZE per "
                              "full-batch iteration) x this run's algorithmic bytes per iteration",
            "kernel": "cfr_iteration" if stats_all[-1].mode == 1 else "k_batched",
            "launches": launches,
            "avg_launch_us": round(cfr_ms * 1000.0 / max(1, launches), 3),
            "algo_bytes_per_launch": int(algo / max(1, launches)),
        },
        "setup_inclusive": setup_rec,
        "plain_solves_per_s": round(plain_sps, 3) if plain_sps else None,
        "batched_vs_plain": round(value / plain_sps, 2) if plain_sps else None,
        "iterations_to_target": {"mean": float(iters.mean()), "p50": float(np.percentile(iters, 50)),
                                 "p99": float(np.percentile(iters, 99)), "max": int(iters.max())},
        "converged_fraction": float(conv.mean()),
        "exploitability_pct_pot": {"mean": float(expl.mean()), "max": float(expl.max())},
        "device_ms_per_step": {"cfr": round(cfr_ms / a.steps, 3), "checks": round(check_ms / a.steps, 3)},
    }

    if a.served_requests > 0:
        log(f"[bench] served form: {a.served_requests} requests at {a.served_rate}/s")
        line["served"] = served_bench(a.served_requests, a.served_rate, 20261203 + 5000, a.target,
                                      a.check_every, a.max_iters, a.served_iters_per_step, plain_fifo=True,
                                      flags=flags)
        log(f"[bench] served: {line['served']['solves_per_s']} solves/s, "
            f"p50 {line['served']['latency_ms']['p50']} ms, p99 {line['served']['latency_ms']['p99']} ms")
        pl = line["served"]["plain_request_loop"]
        log(f"[bench] plain request loop on the same stream: {pl['solves_per_s']} solves/s, "
            f"p50 {pl['latency_ms']['p50']} ms, p99 {pl['latency_ms']['p99']} ms")
        log(f"[bench] served form capacity: the same {a.served_requests} requests submitted at once")
        cap = served_bench(a.served_requests, 1e9, 20261203 + 5000, a.target, a.check_every, a.max_iters,
                           a.served_iters_per_step, flags=flags)
        line["served"]["capacity_solves_per_s"] = cap["solves_per_s"]
        line["served"]["capacity_latency_ms"] = cap["latency_ms"]
        log(f"[bench] served capacity: {cap['solves_per_s']} solves/s")

    if a.precision_boards > 0:
        log(f"[bench] precision form: {a.precision_boards} boards x {a.precision_iters} iterations")
        line["precision_form"] = precision_bench(a.precision_boards, a.precision_iters, a.precision_ref_boards,
                                                 a.cpu_cores or min(16, os.cpu_count() or 1))
        log(f"[bench] precision: {line['precision_form']['solves_per_s']} solves/s")

    if not a.no_leaf:
        log("[bench] leaf values: 3-player full-range turn subgame")
        line["leaf_values"] = leaf_bench()
        log(f"[bench] leaf values: {line['leaf_values']['gpu_ms_per_call']} ms per call on the GPU, "
            f"{line['leaf_values']['cpu_oracle_ms']} ms CPU oracle")

    if not a.no_flop:
        log("[bench] flop-through-river solves")
        line["flop_solve"] = flop_bench()
        log("[bench] realtime re-solve at the reference's published setting")
        line["reference_setting_latency"] = reference_latency_bench()

    if not a.no_cpu:
        cores = a.cpu_cores or min(16, os.cpu_count() or 1)
        idxs = list(range(min(a.cpu_sample, len(problems))))
        log(f"[bench] CPU oracle baseline on {len(idxs)} subgames, {cores} processes")
        sps, t_iter = cpu_oracle_baseline(a.workload, a.seed, idxs, iters, cores)
        line["cpu_baseline"] = {
            "value": round(sps, 5), "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"CPU oracle (same semantics, FP32, closed-form terminals) on the first {len(idxs)} "
                      f"subgames of the batch, 2 iterations each ({t_iter * 1000:.0f} ms/iteration mean), "
                      f"extrapolated to each subgame's GPU iteration count to target"}
        if not a.no_cpu_ref:
            log("[bench] reference CPU solver (solver_v2) baseline")
            ref = cpu_reference_baseline(cores)
            if ref:
                line["cpu_reference_solver_v2"] = ref
    print(json.dumps(line), flush=True)
    batch.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
