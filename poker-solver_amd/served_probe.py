#!/usr/bin/env python3
"""Served-form workload for profilers: the bench's Poisson stream (70% river,
30% turn, 2-3 players) through ss_server_*, 60 requests at 40/s.

    SS_NO_GRAPH=1 rocprofv3 --kernel-trace --stats -d OUT -o run -- python3 poker-solver_amd/served_probe.py
"""

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import bench  # noqa: E402


def main():
    r = bench.served_bench(60, 40.0, 20261203 + 5000, 0.5, 25, 2000, 5, log_every=20)
    print(r["solves_per_s"], r["latency_ms"], flush=True)


if __name__ == "__main__":
    main()
