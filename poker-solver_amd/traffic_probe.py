#!/usr/bin/env python3
"""HBM traffic probe for the batched form, run under rocprofv3 --pmc.

Runs (1) a plain device copy of a known size (ssdbg_copy, kernel k_copy),
the dispatch that calibrates a byte counter, then (2) `--iters` CFR
iterations of the bench's mixed batch with every subgame live (no
exploitability checks), and prints the algorithmic bytes of one iteration of
the whole batch. Collect one counter per run:

    rocprofv3 --pmc FETCH_SIZE -d OUT_F -o run --output-format csv -- python3 poker-solver_amd/traffic_probe.py
    rocprofv3 --pmc WRITE_SIZE -d OUT_W -o run --output-format csv -- python3 poker-solver_amd/traffic_probe.py

HBM bytes per iteration = sum over the CFR dispatches of counter (KiB) *
1024 * (copy bytes / the copy dispatch's counter) / iterations.
"""

import argparse
import ctypes
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import ss_abi  # noqa: E402
import ss_api  # noqa: E402
import workloads as wl  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--subgames", type=int, default=256)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--copy-mib", type=int, default=256)
    a = ap.parse_args()
    lib = ss_abi.load_library()
    lib.ssdbg_copy.restype = ctypes.c_int
    lib.ssdbg_copy.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    ms = ctypes.c_float()
    copy_bytes = a.copy_mib << 20
    assert lib.ssdbg_copy(copy_bytes, 1, ctypes.byref(ms)) == 0, ss_abi.last_error()
    probs = [ss_api.Problem.from_spot(s) for s in wl.batch_mix(a.subgames, max_players=3)]
    algo = sum(ss_api.plan_bytes(p)[0] for p in probs)
    b = ss_api.Batch(probs, a.iters, flags=ss_abi.SS_FLAG_NO_EXPL)
    reps, st = b.run()
    b.close()
    print(json.dumps({"copy_bytes": copy_bytes, "copy_ms": ms.value, "iterations": a.iters,
                      "algo_bytes_per_iteration": algo, "cfr_ms": st.cfr_ms, "mode": st.mode}), flush=True)


if __name__ == "__main__":
    main()
