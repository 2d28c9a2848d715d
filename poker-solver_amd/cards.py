"""Card codec and range-string parser (same contract as python/solver.py:27-118).

card = rank * 4 + suit, rank 0 = '2' .. 12 = 'A', suit c d h s = 0..3
(src/hand_eval.h:17-24).
"""

RANKS = "23456789TJQKA"
SUITS = "cdhs"


def card_to_int(card_str):
    """'Ah' -> 50. Raises ValueError on a malformed card (python/solver.py:27-33)."""
    if len(card_str) < 2 or card_str[0] not in RANKS or card_str[1] not in SUITS:
        raise ValueError(f"Invalid card: {card_str}")
    return RANKS.index(card_str[0]) * 4 + SUITS.index(card_str[1])


def int_to_card(card_int):
    return RANKS[card_int >> 2] + SUITS[card_int & 3]


def parse_range_string(range_str):
    """PioSOLVER-style range -> [(c0, c1, weight)] with c0 < c1, first occurrence kept.

    Accepts specific combos ("AhKh", "AhKh:0.5"), suited / offsuit groups
    ("AKs", "AKo") and pairs ("QQ"), comma separated (python/solver.py:36-107).
    Unparseable weights fall back to 1.0; unknown tokens are skipped.
    """
    out, seen = [], set()

    def add(c0, c1, w):
        a, b = (c0, c1) if c0 < c1 else (c1, c0)
        if a != b and (a, b) not in seen:
            seen.add((a, b))
            out.append((a, b, w))

    if not range_str:
        return out
    for tok in range_str.split(","):
        tok = tok.strip()
        if not tok:
            continue
        w = 1.0
        if ":" in tok:
            tok, ws = tok.rsplit(":", 1)
            try:
                w = float(ws)
            except ValueError:
                pass
        if len(tok) == 4:
            add(card_to_int(tok[:2]), card_to_int(tok[2:]), w)
        elif len(tok) == 3 and tok[0] in RANKS and tok[1] in RANKS:
            r0, r1 = RANKS.index(tok[0]), RANKS.index(tok[1])
            if tok[2] == "s":
                for s in range(4):
                    add(r0 * 4 + s, r1 * 4 + s, w)
            elif tok[2] == "o":
                for s0 in range(4):
                    for s1 in range(4):
                        if s0 != s1:
                            add(r0 * 4 + s0, r1 * 4 + s1, w)
        elif len(tok) == 2 and tok[0] == tok[1] and tok[0] in RANKS:
            r = RANKS.index(tok[0])
            for s0 in range(4):
                for s1 in range(s0 + 1, 4):
                    add(r * 4 + s0, r * 4 + s1, w)
    return out
