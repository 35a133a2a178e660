"""Edge cases of the 10 x 25.5-bit field representation (fe10.h), checked
against Python integers mod p.

The layout keeps unsigned limbs at bits ceil(25.5 i), 26 or 25 bits wide,
and lets them run over: a "tight" limb is below 2^w + 2^18, a "loose" one
(a sum of two tight values, or f + 2p - g) below tight + 2p's limb. Each
operation is fed limbs at the top of its documented input range:

- mul, sq and the x121665 scale: loose limbs at their maximum (the column
  sums reach their bound there), mixed with zeros, tight maxima and random
  loose limbs;
- sub: a loose minuend and a tight subtrahend at its maximum (the 2p bias
  must still cover it);
- the canonical encoding: limbs up to 2^32 - 2^12, and values at and around
  p, 2p and 2^255, including those whose top limb is 2^25 - 1 or above (the
  only ones that take the exact reduction path).

Raw outputs must be congruent to the result and keep to the output bound
the operation promises (tight for mul, sq and the scale).
"""

import random

P = 2**255 - 19
WIDTH = [26 if i % 2 == 0 else 25 for i in range(10)]
BIT = [(51 * i + 1) // 2 for i in range(10)]
TWO_P = [2**27 - 38] + [2 * (2**WIDTH[i] - 1) for i in range(1, 10)]
TIGHT_MAX = [2**WIDTH[i] - 1 + 2**18 for i in range(10)]
LOOSE_MAX = [TIGHT_MAX[i] + TWO_P[i] for i in range(10)]
ANY_MAX = 2**32 - 2**12 - 1

OPS = {0: lambda a, b: a * b, 1: lambda a, b: a * a, 2: lambda a, b: a + b, 3: lambda a, b: a - b,
       5: lambda a, b: a, 6: lambda a, b: a * 121665, 4: lambda a, b: pow(a, P - 2, P)}


def value(limbs):
    return sum(v << BIT[i] for i, v in enumerate(limbs))


def limbs_of(v):
    """The plain limbs of 0 <= v < 2^255."""
    return [(v >> BIT[i]) & (2**WIDTH[i] - 1) for i in range(10)]


def _rand_limbs(rng, top):
    return [rng.randrange(top[i] + 1) for i in range(10)]


def _near_max(rng, top):
    return [top[i] - rng.getrandbits(rng.choice((0, 3, 12))) for i in range(10)]


def cases(op, seed=91, n_random=200):
    """Pairs (a, b) of 10-limb lists inside op's input range."""
    rng = random.Random(seed * 16 + op)
    zero, one = [0] * 10, [1] + [0] * 9
    if op in (0, 1, 6):
        pool = [LOOSE_MAX, TIGHT_MAX, zero, one, [LOOSE_MAX[i] if i % 2 else 0 for i in range(10)],
                [0 if i % 2 else LOOSE_MAX[i] for i in range(10)], limbs_of(P - 1), limbs_of(2**255 - 1)]
        pool += [_near_max(rng, LOOSE_MAX) for _ in range(8)]
        pairs = [(a, b) for a in pool for b in pool]
        pairs += [(_rand_limbs(rng, LOOSE_MAX), _rand_limbs(rng, LOOSE_MAX)) for _ in range(n_random)]
        return pairs
    if op == 3:
        tight = [TIGHT_MAX, zero, one, limbs_of(P - 1), limbs_of(2**255 - 1)] + \
                [_near_max(rng, TIGHT_MAX) for _ in range(6)]
        loose = [LOOSE_MAX, zero, TIGHT_MAX] + [_near_max(rng, LOOSE_MAX) for _ in range(4)]
        pairs = [(a, b) for a in loose for b in tight]
        pairs += [(_rand_limbs(rng, LOOSE_MAX), _rand_limbs(rng, TIGHT_MAX)) for _ in range(n_random)]
        return pairs
    if op == 2:
        pool = [LOOSE_MAX, TIGHT_MAX, zero] + [_near_max(rng, LOOSE_MAX) for _ in range(4)]
        pairs = [(a, b) for a in pool for b in pool]
        pairs += [(_rand_limbs(rng, LOOSE_MAX), _rand_limbs(rng, LOOSE_MAX)) for _ in range(n_random)]
        return pairs
    # canonical encoding and inversion: any limbs below 2^32 - 2^12, and values
    # at and around p, 2p and 2^255 in plain and run-over limbs
    anymax = [ANY_MAX] * 10
    special = [P - 1, P, P + 1, P + 18, P + 19, 2**255 - 1, 2 * P - 1]
    pool = [anymax, zero, one, TIGHT_MAX, LOOSE_MAX] + [limbs_of(v) for v in special]
    # p and 2p with run-over limbs: p = (2^26 - 19) + sum (2^w - 1) 2^bit
    p_limbs = [2**26 - 19] + [2**WIDTH[i] - 1 for i in range(1, 10)]
    pool += [p_limbs, [2 * v for v in p_limbs], [v + 1 for v in p_limbs]]
    # top limb exactly 2^25 - 1 and just above, low limbs full or run over
    for top in (2**25 - 2, 2**25 - 1, 2**25, 2**25 + 2**8):
        pool.append([2**WIDTH[i] - 1 for i in range(9)] + [top])
        pool.append([2**WIDTH[i] - 1 + 2**10 for i in range(9)] + [top])
        pool.append([0] * 9 + [top])
    pool += [_rand_limbs(rng, anymax) for _ in range(n_random // 4)]
    pool += [_near_max(rng, TIGHT_MAX) for _ in range(n_random // 4)]
    return [(a, zero) for a in pool]


def flat(pairs):
    return sum((a for a, _ in pairs), []), sum((b for _, b in pairs), [])


def check(op, pairs, out):
    """Mismatches: raw results (10 limbs) congruent to the residue and within
    the operation's output bound; the canonical encoding (op 5, 8 words in a
    10-word slot) equal to it."""
    bad = []
    for i, (a, b) in enumerate(pairs):
        va, vb = value(a), value(b)
        if op == 4 and va % P == 0:
            continue
        want = OPS[op](va, vb) % P
        got = out[10 * i:10 * i + 10]
        if op == 5:
            ok = sum(w << (32 * k) for k, w in enumerate(got[:8])) == want and got[8:] == [0, 0]
        else:
            ok = value(got) % P == want
            if op in (0, 1, 6, 4):
                ok = ok and all(got[k] <= TIGHT_MAX[k] for k in range(10))
            if op == 3:
                ok = ok and all(got[k] <= a[k] + TWO_P[k] for k in range(10))
        if not ok:
            bad.append((a, b, got, want))
    return bad
