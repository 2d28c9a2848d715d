"""Write golden vectors for the leaf-value producers from the reference's own
python/leaf_values.py.

Runs only in a container where the reference tree is mounted (default
/root/reference). It imports that one module (numpy-only, no side effects),
calls compute_turn_leaf_values, compute_flop_leaf_equity and
extract_leaf_info_from_tree on seeded synthetic inputs, and writes DATA ONLY
into tests/golden/:

  leaf_values.npz   per case: board, hole cards [np][max_hands][2], hand
                    counts, max_hands, leaf infos (index, pot) and the
                    reference's float32 leaf-value array
  leaf_infos.json   extract_leaf_info_from_tree over the turn and flop trees
                    of trees.json.gz (which came from the reference builder)

The cases cover: two players with all 48 rivers (every range <= 50 hands) and
with 12 sampled rivers; three players with exact opponent products and with
the 2000 sampled tuples; four, five and six players (six: sampled tuples over
small lists, so samples repeat and collide); flop leaves over all 49 x 48
runouts and over 12 x 12 sampled ones; repeated and out-of-range leaf
indices. Hands may overlap between players.

Usage:  python tests/golden/make_leaf_golden.py [--ref /root/reference]
"""

import argparse
import gzip
import importlib.util
import json
import os
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
RANKS = "23456789TJQKA"
SUITS = "cdhs"


def card(s):
    return RANKS.index(s[0]) * 4 + SUITS.index(s[1])


# (name, kind, board, hands per player, max_hands, leaf infos (index, pot))
INFOS = [(0, 1000), (2, 2500), (2, 3100), (5, 700)]
CASES = [
    ("turn_2p_all_rivers", "turn", "Qs As 2d 7h", [20, 15], 20, INFOS),
    ("turn_2p_sampled_rivers", "turn", "Kd 7c 2h 9s", [60, 40], 64, INFOS),
    ("turn_3p_exact", "turn", "Qs As 2d 7h", [30, 25, 20], 30, INFOS),
    ("turn_3p_sampled", "turn", "Ah Kh 2c 5d", [55, 45, 30], 56, INFOS),
    ("turn_4p_exact", "turn", "Td 9d 4s 4c", [12, 10, 8, 6], 12, INFOS),
    ("turn_5p_exact", "turn", "Qs As 2d 7h", [4, 4, 4, 4, 4], 4, INFOS[:2]),
    ("turn_6p_sampled", "turn", "8h 8d 3c Js", [6, 6, 6, 6, 6, 6], 6, [(0, 1500)]),
    ("flop_2p_all_runouts", "flop", "Qs As 2d", [8, 6], 8, INFOS),
    ("flop_3p_sampled", "flop", "7c 8c 9d", [52, 45, 10], 52, INFOS[:2]),
]


def load_ref(ref):
    path = os.path.join(ref, "python", "leaf_values.py")
    spec = importlib.util.spec_from_file_location("ref_leaf_values", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def make_hands(rng, board, nh):
    free = [c for c in range(52) if c not in board]
    combos = [(a, b) for i, a in enumerate(free) for b in free[i + 1:]]
    out = []
    for n in nh:
        pick = rng.choice(len(combos), n, replace=False)
        out.append([(combos[k][0], combos[k][1], float(rng.uniform(0.1, 1.0))) for k in pick])
    return out


def tree_view(d):
    """Attribute view of a tree snapshot, shaped like the ctypes SSTreeData the
    reference function reads (nodes[i].type/.player/..., children[i], num_nodes)."""
    nodes = []
    for r in d["nodes"]:
        nodes.append(types.SimpleNamespace(type=r[0], player=r[1], num_children=r[2], first_child=r[3],
                                           pot=r[4], bets=r[5:11], leaf_idx=r[17]))
    return types.SimpleNamespace(nodes=nodes, children=d["children"], num_nodes=d["num_nodes"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    a = ap.parse_args()
    ref = load_ref(a.ref)
    rng = np.random.default_rng(20261203)
    arrays = {"names": np.array([c[0] for c in CASES])}
    for name, kind, board_s, nh, mh, infos in CASES:
        board = [card(c) for c in board_s.split()]
        hands = make_hands(rng, board, nh)
        lis = [ref.LeafInfo(leaf_idx=i, pot=p, bets=(p // 2,) * len(nh)) for i, p in infos]
        if kind == "turn":
            out = ref.compute_turn_leaf_values(board, lis, mh, 1000, player_hands=hands)
        else:
            out = ref.compute_flop_leaf_equity(board, lis, mh, 1000, player_hands=hands)
        hc = np.zeros((len(nh), mh, 2), dtype=np.int32)
        for p, hs in enumerate(hands):
            for h, (c0, c1, _) in enumerate(hs):
                hc[p, h] = (c0, c1)
        arrays[f"{name}__kind"] = np.array(kind)
        arrays[f"{name}__board"] = np.array(board, dtype=np.int32)
        arrays[f"{name}__hands"] = hc
        arrays[f"{name}__nh"] = np.array(nh, dtype=np.int32)
        arrays[f"{name}__max_hands"] = np.array(mh, dtype=np.int32)
        arrays[f"{name}__infos"] = np.array(infos, dtype=np.int64).reshape(-1, 2)
        arrays[f"{name}__out"] = np.asarray(out, dtype=np.float32)
        print(f"{name}: {out.shape}", flush=True)
    np.savez_compressed(os.path.join(HERE, "leaf_values.npz"), **arrays)

    with gzip.open(os.path.join(HERE, "trees.json.gz"), "rt") as f:
        trees = {t["name"]: t for t in json.load(f)["trees"]}
    ext = []
    for tname, npl in [("turn_2p_cont", 2), ("turn_3p_cont_one_size", 3), ("turn_3p_cont_one_size", 2),
                       ("flop_2p_cont", 2), ("turn_2p_nocont", 2)]:
        got = ref.extract_leaf_info_from_tree(tree_view(trees[tname]), num_players=npl)
        ext.append({"tree": tname, "num_players": npl,
                    "leaf_infos": [[li.leaf_idx, li.pot, list(li.bets)] for li in got]})
        print(f"extract {tname} np={npl}: {len(got)} leaves", flush=True)
    with open(os.path.join(HERE, "leaf_infos.json"), "w") as f:
        json.dump({"source": "python/leaf_values.py extract_leaf_info_from_tree", "cases": ext}, f, indent=1)


if __name__ == "__main__":
    main()
