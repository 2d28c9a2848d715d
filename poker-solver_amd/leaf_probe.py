#!/usr/bin/env python3
"""Leaf-value producer probe, run under rocprofv3 --kernel-trace --stats:
`--reps` calls of compute_turn_leaf_values on the 3-player full-range turn
workload (workloads.turn_leaf_workload), or with --flop of
compute_flop_leaf_equity on the 3-player full-range flop workload
(workloads.flop_leaf_workload), printing the wall time per call.

    rocprofv3 --kernel-trace --stats -d OUT -o run -- python3 poker-solver_amd/leaf_probe.py
"""

import argparse
import ctypes
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import numpy as np  # noqa: E402

import leaf_values as lv  # noqa: E402
import ss_abi  # noqa: E402
import workloads  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--flop", action="store_true")
    a = ap.parse_args()
    lib = ss_abi.load_library()
    if a.flop:
        board, infos, hands, mh = workloads.flop_leaf_workload(lib)
        produce, runouts = lv.compute_flop_leaf_equity, lv.flop_runouts
    else:
        board, infos, hands, mh = workloads.turn_leaf_workload(lib)
        produce, runouts = lv.compute_turn_leaf_values, lv.turn_runouts
    print(f"{'flop' if a.flop else 'turn'} 3p leaves: {len(infos)} leaf infos, {mh} hands per player", flush=True)
    produce(board, infos, mh, 1500, player_hands=hands)
    t = time.perf_counter()
    for _ in range(a.reps):
        out = produce(board, infos, mh, 1500, player_hands=hands)
    ms = (time.perf_counter() - t) * 1000 / a.reps
    print(f"{ms:.3f} ms per call, output {out.shape} ({out.nbytes / 2**20:.1f} MiB)", flush=True)
    # The two library calls alone, on prepared arrays.
    nh = np.array([len(h) for h in hands], dtype=np.int32)
    hc = np.array([[(a, b) for a, b, _ in hs] for hs in hands], dtype=np.int32)
    run = np.array(runouts(board, list(nh)), dtype=np.int32)
    raw = np.zeros((len(hands), mh))
    ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    idx = np.array([li.leaf_idx for li in infos], dtype=np.int32)
    pot = np.array([li.pot for li in infos], dtype=np.int32)
    t = time.perf_counter()
    for _ in range(a.reps):
        lib.ss_leaf_equity(len(hands), hc.ctypes.data_as(ip), nh.ctypes.data_as(ip), mh, run.ctypes.data_as(ip),
                           len(run), raw.ctypes.data_as(dp))
    eq_ms = (time.perf_counter() - t) * 1000 / a.reps
    t = time.perf_counter()
    for _ in range(a.reps):
        lib.ss_leaf_values(len(hands), raw.ctypes.data_as(dp), nh.ctypes.data_as(ip), mh, idx.ctypes.data_as(ip),
                           pot.ctypes.data_as(ip), len(infos), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    val_ms = (time.perf_counter() - t) * 1000 / a.reps
    print(f"ss_leaf_equity {eq_ms:.3f} ms, ss_leaf_values {val_ms:.3f} ms per call", flush=True)


if __name__ == "__main__":
    main()
