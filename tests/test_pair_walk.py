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
* lxv_init_seeded's lane set-up from two anchors (P0 + t S) against the full
  ladder, the oracle's public keys and the curve equation.
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


@pytest.mark.parametrize("layout", [lxv.LAYOUT_8X32, lxv.LAYOUT_10X26], ids=["8x32", "10x26"])
@pytest.mark.parametrize("batch", [64, 256])
def test_lane_setup_from_anchors_matches_the_ladder(layout, batch):
    # lxv_init_seeded's pair-walk set-up computes lane t's first centre as
    # P0 + t S from two anchors and a short ladder over t. It must give the
    # full ladder's point exactly; its x is the public key of the centre's key
    # (index t stride + B/2 - 1), the point is on the curve, and adding T_1
    # = Q gives the next key's public key (which fixes the sign of y).
    tab = table_points()
    rng = random.Random(batch + layout)
    start = lxv.seeded_search(11 + batch, 0)[0]
    H = batch // 2 - 1
    cases = [(1 << 40, t, 24) for t in (0, 1, 2, 3, 5, (1 << 23) - 1)]
    cases += [(1 << 38, (1 << 24) - 1, 24), (1, 0, 0), (1, 1, 1), (7, (1 << 32) - 1, 32), (1 << 20, 6, 40)]
    cases += [(1 << 30, rng.randrange(1 << 32), 32) for _ in range(3)]
    for stride, t, bits in cases:
        (x1, y1), (x2, y2) = lxv.debug_pair_centre(batch, start, stride, t, bits, layout=layout)
        assert (x2, y2) == (x1, y1), (stride, t)
        assert (y2 * y2 - (x2 ** 3 + A * x2 * x2 + x2)) % P == 0
        g = t * stride + H
        u = lambda k: int.from_bytes(orc.x25519(orc.step_scalar(start, k)), "little")  # noqa: E731
        assert x2 == u(g), (stride, t)
        assert affine_add((x2, y2), tab[0])[0] == u(g + 1), (stride, t)


def test_lane_setup_rejects_bad_arguments():
    start = lxv.seeded_search(3, 0)[0]
    for args in [(64, start, 0, 1, 1), (64, start, 1, 2, 1), (64, start, 1, 1, 64), (48, start, 1, 1, 1),
                 (64, bytes(32), 1, 1, 1), (64, start, 1 << 62, 2, 2)]:
        with pytest.raises(RuntimeError):
            lxv.debug_pair_centre(*args)
