"""Writes tests/golden/walk_2p24_4hex.json: the small CPU workload's expected output.

Workload: seed 0 of the seeded generator gives the start scalar k, the
Ed25519 seed and a 4-hex-digit prefix; walk the 2^24 consecutive keys
k + 8g, g < 2^24. Expected output: every (g, address) whose address starts
with the prefix (about 256), and the XOR of all 2^24 addresses.

The 2^24-key result comes from the C oracle (oracle/liblxv_oracle.so, built by
oracle/Makefile). Before it is written it is cross-checked against the
independent Python big-integer walk of textbook.py on the first 2^16 keys:
the oracle's hits and XOR over that prefix of the range must equal Python's.

Run from the repository root:  python tests/golden/make_walk_2p24_4hex.py
"""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as orc  # noqa: E402
import textbook as tb  # noqa: E402

SEED = 0
DIGITS = 4
KEYS = 1 << 24
CHECK_KEYS = 1 << 16


def python_scan(start, ed_pub, prefix, n):
    acc = bytearray(16)
    hits = []
    for g, u in enumerate(tb.walk_us(start, n, 0)):
        a = tb.lxmf_address(u, ed_pub)
        for j in range(16):
            acc[j] ^= a[j]
        if tb.prefix_matches(prefix, a):
            hits.append((g, a))
    return hits, bytes(acc)


def main():
    start = tb.seeded_start(SEED)
    ed_seed = tb.seeded_ed_seed(SEED)
    ed_pub = tb.ed25519_pub(ed_seed)
    prefix = tb.seeded_prefix(SEED, DIGITS)
    if orc.ed25519_pub(ed_seed) != ed_pub:
        raise SystemExit("oracle and textbook disagree on the Ed25519 key")
    # cross-check on the first 2^16 keys
    py_hits, py_xor = python_scan(start, ed_pub, prefix, CHECK_KEYS)
    o_hits, o_xor = orc.scan_ranges(start, ed_pub, [(0, CHECK_KEYS)], prefix)
    if o_xor != py_xor or o_hits != py_hits:
        raise SystemExit("oracle and textbook disagree on the first 2^16 keys")
    hits, xor = orc.scan_ranges(start, ed_pub, [(0, KEYS)], prefix, cap=1 << 12)
    if [h for h in hits if h[0] < CHECK_KEYS] != py_hits:
        raise SystemExit("the 2^24 scan disagrees with its own 2^16 prefix")
    out = {"seed": SEED, "start": start.hex(), "ed_seed": ed_seed.hex(), "ed_pub": ed_pub.hex(),
           "prefix": prefix, "keys": KEYS, "xor": xor.hex(), "hits": [[g, a.hex()] for g, a in hits],
           "checked_prefix_keys": CHECK_KEYS, "checked_prefix_xor": py_xor.hex()}
    with open(os.path.join(HERE, "walk_2p24_4hex.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("walk_2p24_4hex.json: %d hits for %s over 2^24 keys" % (len(hits), prefix))


if __name__ == "__main__":
    main()
