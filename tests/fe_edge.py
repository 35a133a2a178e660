"""Edge cases of the 8 x 32-bit field representation, checked against Python
integers mod p.

The 8 x 32 layout keeps every value weakly reduced: any 256-bit integer
congruent to the element, so values in [2^255, 2^256) are ordinary working
values there, and the second-wrap branches of addition, subtraction and the
final x38 fold of multiplication and squaring only run near 2^256. The cases
below reach each of them:

- a + b >= 2^257 - 38 (both near 2^256): the sum wraps twice;
- a - b with b - a > 2^256 - 38: the difference borrows twice;
- products of values near 2^256 (all-ones words; 2^256 - 38, congruent to
  0 as 2^256 = 38 mod p; 2^256 - 39, congruent to -1): the high half folds
  to just below or above 2^256 again;
- the canonical encoding of values in [2^255, 2^256), including p, 2p,
  2^256 - 38 (congruent to 0) and 2^256 - 1 (congruent to 37).
"""

import random

P = 2**255 - 19
TOP = 2**256

EDGE = [0, 1, 2, 19, 37, 38, 39, P - 1, P, P + 1, 2 * P - 1, 2 * P, 2 * P + 1, 2**255 - 1, 2**255, 2**255 + 18,
        2**255 + 19, TOP - 1, TOP - 2, TOP - 19, TOP - 20, TOP - 37, TOP - 38, TOP - 39, TOP - 76, TOP - 77,
        TOP - 2**32, TOP - 2**224, 2**224 * (2**32 - 1), (TOP - 1) ^ (1 << 128)]

OPS = {0: lambda a, b: a * b, 1: lambda a, b: a * a, 2: lambda a, b: a + b, 3: lambda a, b: a - b,
       5: lambda a, b: a, 6: lambda a, b: a * 121665, 4: lambda a, b: pow(a, P - 2, P),
       7: lambda a, b: pow(a, P - 2, P)}


def words(v):
    return [(v >> (32 * i)) & 0xffffffff for i in range(8)]


def value(ws):
    return sum(w << (32 * i) for i, w in enumerate(ws))


def cases(seed=77, n_random=300):
    """Pairs (a, b) of 256-bit integers: every edge value against every other
    one, plus random values with bit 255 set, plus pairs built to wrap or
    borrow twice."""
    rng = random.Random(seed)
    pairs = [(a, b) for a in EDGE for b in EDGE]
    hi = [TOP - 1 - rng.getrandbits(rng.choice((8, 40, 200, 254))) for _ in range(n_random)]
    hi += [(1 << 255) | rng.getrandbits(255) for _ in range(n_random)]
    pairs += [(a, rng.choice(hi)) for a in hi]
    # sums at or above 2^257 - 38, differences with b - a above 2^256 - 38
    pairs += [(TOP - 1 - rng.getrandbits(5), TOP - 1 - rng.getrandbits(5)) for _ in range(64)]
    pairs += [(rng.getrandbits(5), TOP - 1 - rng.getrandbits(5)) for _ in range(64)]
    return pairs


def check(op, pairs, out_words, raw):
    """Mismatches of op on `pairs`: raw outputs need only be congruent mod p
    (and, being 8 words, below 2^256); canonical outputs (LXV_FE_CANON, or any
    op without raw limbs) must equal the residue."""
    bad = []
    for i, (a, b) in enumerate(pairs):
        if op in (4, 7) and a % P == 0:
            continue
        want = OPS[op](a, b) % P
        got = value(out_words[8 * i:8 * i + 8])
        ok = (got % P == want) if (raw and op != 5) else (got == want)
        if not ok:
            bad.append((a, b, got, want))
    return bad
