#!/usr/bin/env python3
"""Precision-form workload for profilers: 64 two-player rivers (1081 combos,
pot 100 BB, stacks 400 BB, sizes {0.5, 1}, max 2 raises) solved for 50
iterations in one batched precision-form solve (BF16 hi/lo MFMA showdown),
then one of them in the plain precision form.

    SS_NO_GRAPH=1 rocprofv3 --kernel-trace --stats -d OUT -o run -- python3 poker-solver_amd/precision_probe.py
"""

import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import ss_abi  # noqa: E402
import ss_api  # noqa: E402
import workloads as wl  # noqa: E402


def main():
    probs = [ss_api.Problem.from_spot(wl.river_2p(seed=20261203 + 1000 + b)) for b in range(64)]
    b = ss_api.Batch(probs, 50, flags=ss_abi.SS_FLAG_NO_EXPL, precision=ss_abi.SS_PREC_BF16X2)
    reps, st = b.run()
    print(f"batched: {st.cfr_ms:.2f} ms for 64 boards x 50 iterations", flush=True)
    b.close()
    r = ss_api.solve_plain(probs[0], 50, precision=ss_abi.SS_PREC_BF16X2, flags=ss_abi.SS_FLAG_NO_EXPL)
    print(f"plain: {r.kernel_ms:.2f} ms for 1 board x 50 iterations", flush=True)


if __name__ == "__main__":
    main()
