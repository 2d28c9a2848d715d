"""Write golden vectors for the hot path from CPU builds of the reference.

Runs only in a container where the reference tree is mounted (default
/root/reference): it first runs `make -C oracle ref`, which compiles the
reference's own C sources where they lie into oracle/_ref/ (never committed),
then loads those builds and writes DATA ONLY (inputs and expected outputs)
into tests/golden/:

  eval_census.json   category counts + rank sum of src/hand_eval.h eval7 over
                     all C(52,7) hands (from oracle/_ref/ref_census)
  eval_vectors.json  seeded random 7-card hands and every river combo of a few
                     boards, with the reference eval7 value
  trees.json.gz      ss_build_tree (street_solve.cu:1192-1296, SS_MAX_RAISES=3)
                     layouts for a list of spots: nodes, children, BFS order,
                     depths and the per-type index lists

Usage:  python tests/golden/make_golden.py [--ref /root/reference]
"""

import argparse
import ctypes
import gzip
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "poker-solver_amd"))
import ss_abi  # noqa: E402

RANKS = "23456789TJQKA"
SUITS = "cdhs"


def card(s):
    return RANKS.index(s[0]) * 4 + SUITS.index(s[1])


# (name, board, pot, stack, bet_sizes, num_players, acting_order, use_cont)
TREE_SPOTS = [
    ("ref_test_river_2p", "Qs As 2d 7h 4c", 1000, 5000, [0.33, 0.75], 2, [0, 1], 0),
    ("river_2p_100bb_sizes_50_100", "Kd 7c 2h 9s 3d", 10000, 40000, [0.5, 1.0], 2, [0, 1], 0),
    ("river_3p_150bb_sizes_33_75", "Kd 7c 2h 9s 3d", 15000, 50000, [0.33, 0.75], 3, [0, 1, 2], 0),
    ("river_3p_order_201", "Ah Kh 2c 5d 9s", 3000, 9000, [0.5], 3, [2, 0, 1], 0),
    ("river_2p_order_10", "Ah Kh 2c 5d 9s", 1000, 5000, [0.75, 1.5], 2, [1, 0], 0),
    ("river_4p_one_size", "Ah Kh 2c 5d 9s", 2000, 6000, [1.0], 4, [0, 1, 2, 3], 0),
    ("river_2p_short_stack", "Ah Kh 2c 5d 9s", 1000, 700, [0.5, 1.0, 2.0], 2, [0, 1], 0),
    ("river_2p_tiny_sizes", "Ah Kh 2c 5d 9s", 100, 5000, [0.001, 0.25, 0.3], 2, [0, 1], 0),
    ("river_2p_three_sizes", "Qs As 2d 7h 4c", 650, 9750, [0.33, 0.75, 1.5], 2, [0, 1], 0),
    ("turn_2p_cont", "Qs As 2d 7h", 1000, 5000, [0.5, 1.0], 2, [0, 1], 1),
    ("turn_2p_nocont", "Qs As 2d 7h", 1000, 5000, [0.5, 1.0], 2, [0, 1], 0),
    ("turn_3p_cont_one_size", "Qs As 2d 7h", 1500, 5000, [0.75], 3, [0, 1, 2], 1),
    ("flop_2p_cont", "Qs As 2d", 650, 9750, [0.33, 0.75], 2, [0, 1], 1),
    ("river_5p_allin_only", "Ah Kh 2c 5d 9s", 1000, 1000, [1.0], 5, [0, 1, 2, 3, 4], 0),
]


def build_ref_tree(lib, spot):
    name, board, pot, stack, sizes, np_, order, cont = spot
    b = [card(c) for c in board.split()]
    is_river = 1 if len(b) == 5 else 0
    td = ss_abi.SSTreeData()
    barr = (ctypes.c_int * len(b))(*b)
    sarr = (ctypes.c_float * len(sizes))(*sizes)
    oarr = (ctypes.c_int * len(order))(*order)
    rc = lib.ss_build_tree(barr, len(b), pot, stack, sarr, len(sizes), np_, oarr,
                           is_river, cont, ctypes.byref(td))
    assert rc == 0
    d = ss_abi.tree_to_dict(td)
    d.update({"name": name, "board": b, "pot": pot, "stack": stack, "bet_sizes": sizes,
              "num_players": np_, "acting_order": order, "is_river": is_river,
              "use_cont_strats": cont, "max_raises": 3})
    return d


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--skip-census", action="store_true")
    a = ap.parse_args()
    oracle = os.path.join(REPO, "oracle")
    subprocess.check_call(["make", "-C", oracle, "ref", f"REF={a.ref}"])
    refdir = os.path.join(oracle, "_ref")

    # ---- evaluator vectors
    ev = ctypes.CDLL(os.path.join(refdir, "hand_eval_ref.so"))
    ev.ref_eval7.restype = ctypes.c_uint32
    ev.ref_eval7.argtypes = [ctypes.POINTER(ctypes.c_int)]
    rng = np.random.Generator(np.random.PCG64(20261203))
    hands7 = [sorted(rng.choice(52, 7, replace=False).tolist()) for _ in range(4000)]
    vals = [int(ev.ref_eval7((ctypes.c_int * 7)(*h))) for h in hands7]
    boards = [[card(c) for c in s.split()] for s in
              ["Qs As 2d 7h 4c", "Kd 7c 2h 9s 3d", "5h 6h 7h 8h 9h", "Ac Ad Ah 2s 2d", "2c 3c 4c 5d Ks"]]
    river = []
    for bd in boards:
        combos = [(x, y) for x in range(52) for y in range(x + 1, 52) if x not in bd and y not in bd]
        rv = [int(ev.ref_eval7((ctypes.c_int * 7)(*(bd + [x, y])))) for x, y in combos]
        river.append({"board": bd, "values": rv})
    with open(os.path.join(HERE, "eval_vectors.json"), "w") as f:
        json.dump({"source": "src/hand_eval.h eval7 (compiled in place)",
                   "hands7": hands7, "values": vals, "river_boards": river}, f)

    # ---- census
    if not a.skip_census:
        out = subprocess.check_output([os.path.join(refdir, "ref_census")]).decode().split()
        nums = [int(x) for x in out]
        with open(os.path.join(HERE, "eval_census.json"), "w") as f:
            json.dump({"source": "src/hand_eval.h eval7 over all C(52,7) hands",
                       "category_counts": nums[:10], "rank_sum_mod_2_64": nums[10]}, f, indent=1)

    # ---- trees
    lib = ctypes.CDLL(os.path.join(refdir, "ss_host_ref.so"))
    lib.ss_build_tree.restype = ctypes.c_int
    trees = []
    for spot in TREE_SPOTS:
        d = build_ref_tree(lib, spot)
        print(f"{d['name']}: {d['num_nodes']} nodes, {d['num_leaves']} leaves", file=sys.stderr)
        trees.append(d)
    with gzip.open(os.path.join(HERE, "trees.json.gz"), "wt") as f:
        json.dump({"source": "street_solve.cu ss_build_tree host half, SS_MAX_RAISES=3",
                   "trees": trees}, f)


if __name__ == "__main__":
    main()
