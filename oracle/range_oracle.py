"""CPU restatement of the reference's Bayesian range update
(python/range_narrowing.py:76-122, RangeNarrower.update). TEST
INFRASTRUCTURE ONLY: tests/ use it as the checker; the library never
imports it.

ranges: [(card0, card1, weight)]; blueprint_probs: {(card0, card1): p}.
Hands missing from blueprint_probs use 1 / max(#distinct p values, 2); a hand
with weight > 0 keeps at least the floor; all weights are then divided by the
largest when it is positive.
"""

WEIGHT_FLOOR = 0.005


def range_update(ranges, blueprint_probs, floor=WEIGHT_FLOOR):
    n_approx = max(len(set(blueprint_probs.values())), 2)
    out = []
    for c0, c1, w in ranges:
        p = blueprint_probs.get((c0, c1))
        if p is None:
            p = 1.0 / n_approx
        v = w * p
        if v < floor and w > 0:
            v = floor
        out.append((c0, c1, v))
    m = max(v for _, _, v in out) if out else 1.0
    return [(c0, c1, v / m) for c0, c1, v in out] if m > 0 else out
