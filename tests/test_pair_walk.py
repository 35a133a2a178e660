"""CPU checks of the pair walk (lxv_pairwalk.h), no GPU needed.

* The table of points T_i = i Q (Q = 8B) the walk steps with: every u against
  an independent x-only ladder, every v on the curve, and consecutive entries
  related by the affine group law with T_1, which fixes the signs of all the
  v's relative to each other (a wrong sign gives x((i - 2)Q) instead).
* The walk itself, compiled for the host from the device source: every
  public key of several batches equals the oracle's X25519(k + 8n, 9), for
  each batch size, both limb layouts and both ways of setting a lane up
  (from u(k - 8), u(k) as Backend::init hands them over, and the ladder of
  lxv_init_seeded).
"""

import os
import random
import re

import pytest

import oracle_lib as orc
import textbook as tb
import lxv

P = tb.P
A = 486662
TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lxmf-vanity_amd", "csrc",
                     "pair_table.inc")


def table_points():
    src = open(TABLE).read()
    body = src.split("LXV_PAIR_TABLE_INIT {", 1)[1].split("}", 1)[0]
    words = [int(w, 16) for w in re.findall(r"0x([0-9a-f]{8})u", body)]
    n = int(re.search(r"LXV_PAIR_TABLE_ENTRIES (\d+)", src).group(1))
    assert len(words) == 16 * n
    val = lambda ws: sum(w << (32 * i) for i, w in enumerate(ws))  # noqa: E731
    return [(val(words[16 * e:16 * e + 8]), val(words[16 * e + 8:16 * e + 16])) for e in range(n)]


def affine_add(p1, p2):
    (x1, y1), (x2, y2) = p1, p2
    if x1 == x2:
        lam = (3 * x1 * x1 + 2 * A * x1 + 1) * pow(2 * y1, P - 2, P) % P
    else:
        lam = (y2 - y1) * pow(x2 - x1, P - 2, P) % P
    x3 = (lam * lam - A - x1 - x2) % P
    return x3, (lam * (x1 - x3) - y1) % P


def test_pair_table_points():
    pts = table_points()
    assert len(pts) == 129
    index = list(range(1, 129)) + [256]
    for (x, y), i in zip(pts, index):
        assert x < P and y < P
        assert x == tb.le_int(tb.mont_mul(8 * i)), i
        assert y * y % P == (x ** 3 + A * x * x + x) % P, i
    q = pts[0]
    for e in range(1, 128):
        assert affine_add(pts[e - 1], q) == pts[e], e + 1
    assert affine_add(pts[127], pts[127]) == pts[128]


def u_words(k):
    u = orc.x25519(k)
    return [int.from_bytes(u[4 * i:4 * i + 4], "little") for i in range(8)]


@pytest.mark.parametrize("layout", [lxv.LAYOUT_8X32, lxv.LAYOUT_10X26], ids=["8x32", "10x26"])
@pytest.mark.parametrize("batch", [4, 16, 64, 256])
def test_host_pair_walk_matches_x25519(layout, batch):
    rng = random.Random(1000 + batch)
    batches = max(2, 512 // batch)
    for trial in range(2):
        start = tb.clamp(bytes(rng.getrandbits(8) for _ in range(32)))
        if not orc.start_scalar_ok(start):
            continue
        want = [orc.x25519(orc.step_scalar(start, n)) for n in range(batches * batch)]
        seeded = lxv.debug_pair_walk(batch, batches, start=start, layout=layout)
        assert seeded == want
        upc = u_words(orc.previous_scalar(start)) + u_words(start)
        from_u = lxv.debug_pair_walk(batch, batches, u_prev_cur=upc, layout=layout)
        assert from_u == want


def test_host_pair_walk_rejects_bad_arguments():
    start = lxv.seeded_search(3, 0)[0]
    for batch in (0, 2, 3, 48, 512):
        with pytest.raises(RuntimeError):
            lxv.debug_pair_walk(batch, 1, start=start)
    with pytest.raises(RuntimeError):
        lxv.debug_pair_walk(64, 1, start=bytes(32))  # not a clamped start scalar
