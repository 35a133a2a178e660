"""CPU tests of the product library (no GPU needed).

* The C-ABI library loads and exports every function include/*.h declares.
* The host crypto of the search driver (used for setup and for re-deriving
  hits) matches the oracle, the golden fixtures and the reference's known
  answer.
* The device field and hash code, compiled for the host from the same source
  (lxv_debug_* with on_gpu=0), matches Python big integers and the oracle.
"""

import os
import random
import re

import pytest

import oracle_lib as orc
import textbook as tb
import lxv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = tb.P


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(lxv_[a-z0-9_]+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 25, names
    L = lxv.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_open_without_gpu_fails_cleanly():
    if os.path.exists("/dev/kfd") and os.environ.get("HIP_VISIBLE_DEVICES", "") != "-1":
        pytest.skip("a GPU may be present")
    with pytest.raises(lxv.GpuUnavailable):
        lxv.Backend()
    for kw in ({}, {"pipelined": True}):
        with pytest.raises(lxv.GpuUnavailable, match="no HIP device"):
            lxv.Multi(**kw)


def test_multi_options_are_checked_before_the_device():
    with pytest.raises(lxv.GpuUnavailable, match="lane_keys_log2"):
        lxv.Multi(lane_keys_log2=30)
    with pytest.raises(lxv.GpuUnavailable, match="no batched kernel"):
        lxv.Multi(batch=32)
    with pytest.raises(lxv.GpuUnavailable, match="no batched kernel"):
        lxv.Multi(batch=256, waves_per_cu=4)
    with pytest.raises(lxv.GpuUnavailable, match="no batched kernel"):
        lxv.Multi(batch=128, layout=lxv.LAYOUT_10X26)
    with pytest.raises(lxv.GpuUnavailable, match="oldest_share_pct"):
        lxv.Multi(oldest_share_pct=95)


def test_tuned_shapes_are_the_tables_picks():
    # the library is built with the autotuner's picks from tuned.json
    import json
    with open(os.path.join(ROOT, "lxmf-vanity_amd", "tuned.json")) as f:
        table = json.load(f)
    import gen_tuned
    picks = gen_tuned.picks(table)
    for w in ("single", "batched", "serving"):
        got = lxv.tuned_shape(w)
        for k, j in (("batch", "batch"), ("waves_per_cu", "waves_per_cu"), ("layout", "layout_id"),
                     ("walk", "walk_id"), ("step_ms", "step_ms"), ("lane_keys_log2", "lane_keys_log2")):
            assert got[k] == picks[w][j], (w, k)


def test_workloads_launch_at_the_tuned_shapes():
    # the serving loop and the batched workload open lxv_multi at the shape
    # the library was built with, unless the caller overrides a field
    import workloads
    for w in ("batched", "serving"):
        assert workloads.shape_for(w) == lxv.tuned_shape(w)
    sh = workloads.shape_for("serving", step_ms=50, batch=0)
    assert sh["step_ms"] == 50 and sh["batch"] == lxv.tuned_shape("serving")["batch"]


def test_tuned_header_follows_the_table(tmp_path):
    # a changed table gives a changed header (the build regenerates it)
    import json
    import gen_tuned
    table = {"pick": {"batch": 64, "waves_per_cu": 8, "layout_id": 1, "walk_id": 0, "keys_per_s": 1.0},
             "picks": {"serving": {"batch": 128, "waves_per_cu": 8, "layout_id": 0, "walk_id": 0, "step_ms": 15,
                                   "lane_keys_log2": 9, "keys_per_s": 2.0}}}
    src, dst = tmp_path / "t.json", tmp_path / "t.h"
    src.write_text(json.dumps(table))
    import sys
    old = sys.argv
    sys.argv = ["gen_tuned.py", str(src), str(dst)]
    try:
        gen_tuned.main()
    finally:
        sys.argv = old
    text = dst.read_text()
    assert "/* single */ {64, 8, 1, 0, 0, 0, 1.000000e+00}" in text
    assert "/* serving */ {128, 8, 0, 0, 15, 9, 2.000000e+00}" in text
    assert "/* batched */ {256, 16, 0, 0, 200, 14," in text


def test_host_crypto_known_answers():
    r = orc_kat = __import__("json").load(open(os.path.join(ROOT, "tests", "golden", "kat.json")))
    rns = r["rns"]
    ident = bytes.fromhex(rns["identity"])
    assert lxv.x25519_base(ident[:32]).hex() == rns["x_pub"]
    assert lxv.ed25519_public(ident[32:]).hex() == rns["ed_pub"]
    assert lxv.address_of_identity(ident).hex() == rns["address"]
    for seed, pub in orc_kat["ed25519"]:
        assert lxv.ed25519_public(bytes.fromhex(seed)).hex() == pub
    for k, u, out in orc_kat["x25519"]:
        o = ctypes_x25519(bytes.fromhex(k), bytes.fromhex(u))
        assert o is None or o.hex() == out


def ctypes_x25519(k, u):
    # the product exports X25519 only against the base point
    if u != tb.int_le(9):
        return None
    return lxv.x25519_base(k)


def test_host_crypto_matches_oracle_on_random_inputs():
    rng = random.Random(21)
    for _ in range(30):
        k = bytes(rng.getrandbits(8) for _ in range(32))
        s = bytes(rng.getrandbits(8) for _ in range(32))
        assert lxv.x25519_base(k) == orc.x25519(k)
        assert lxv.ed25519_public(s) == orc.ed25519_pub(s)
        assert lxv.address_of_identity(k + s) == orc.address_of_identity(k + s)
        n = rng.getrandbits(60)
        assert lxv.step_scalar(k, n) == orc.step_scalar(k, n)
        assert lxv.start_scalar_ok(k) == orc.start_scalar_ok(k)


def test_launch_args_constants():
    a = lxv.launch_args_for(bytes(range(32)), "b00b1e5")
    k = __import__("json").load(open(os.path.join(ROOT, "tests", "golden", "kat.json")))
    assert lxv.words_to_bytes_le(a.q).hex() == k["u_8B"]
    nh = bytes.fromhex(k["name_hash"])
    assert list(a.name_hash) == [int.from_bytes(nh[0:4], "big"), int.from_bytes(nh[4:8], "big"),
                                 nh[8] << 24 | nh[9] << 16]
    assert lxv.words_to_bytes_be(a.ed) == bytes(range(32))
    assert (list(a.want), list(a.mask)) == orc.prefix_words("b00b1e5")
    with pytest.raises(ValueError):
        lxv.prefix_words("BOOB")


def test_seeded_workloads_match_the_textbook_generator():
    for seed in (0, 1, 7, 2024, 2**63 + 5):
        s, e, p = lxv.seeded_search(seed, 9)
        assert s == tb.seeded_start(seed)
        assert e == tb.seeded_ed_seed(seed)
        assert p == tb.seeded_prefix(seed, 9)


def test_random_start_scalars_are_starts():
    for _ in range(20):
        assert orc.start_scalar_ok(lxv.random_start_scalar())


# ---- the device code, host build

EDGE = [0, 1, 2, 19, P - 1, P, P + 1, 2**255 - 1, 2**254, 2**255 - 20, (2**26 - 1) * 2**230]


def fe_cases(n, seed):
    rng = random.Random(seed)
    vals = list(EDGE) + [rng.getrandbits(255) for _ in range(n)]
    # Decoding drops bit 255 (RFC 7748 5), so inputs are taken below 2^255.
    return [v & (2**255 - 1) for v in vals]


def words(v):
    return [(v >> (32 * i)) & 0xffffffff for i in range(8)]


def from_words(ws):
    return sum(w << (32 * i) for i, w in enumerate(ws))


@pytest.mark.parametrize("layout", [lxv.LAYOUT_10X26, lxv.LAYOUT_8X32])
def test_device_field_code_on_host(layout):
    vals = fe_cases(200, 22)
    rng = random.Random(23)
    bs = [rng.choice(vals) for _ in vals]
    A = sum((words(v) for v in vals), [])
    B = sum((words(v) for v in bs), [])
    expect = {
        0: [a * b % P for a, b in zip(vals, bs)],
        1: [a * a % P for a in vals],
        2: [(a + b) % P for a, b in zip(vals, bs)],
        3: [(a - b) % P for a, b in zip(vals, bs)],
        5: [a % P for a in vals],
        6: [a * 121665 % P for a in vals],
    }
    for op, want in expect.items():
        got = lxv.debug_fe(op, A, B, on_gpu=False, layout=layout)
        got = [from_words(got[8 * i:8 * i + 8]) for i in range(len(vals))]
        assert got == want, op
    # inversion by divsteps (op 4) and by the reference's addition chain
    # (op 7) against Fermat's little theorem, 0 included (it maps to 0)
    for op in (4, 7):
        inv = lxv.debug_fe(op, A, B, on_gpu=False, layout=layout)
        for i, a in enumerate(vals):
            assert from_words(inv[8 * i:8 * i + 8]) == pow(a, P - 2, P), (op, i)


@pytest.mark.parametrize("layout", [lxv.LAYOUT_10X26, lxv.LAYOUT_8X32])
def test_device_field_code_chains_on_host(layout):
    # Outputs fed back as inputs many times: limb bounds hold across a chain.
    rng = random.Random(24)
    x = rng.getrandbits(255)
    y = rng.getrandbits(255)
    for _ in range(200):
        got = lxv.debug_fe(0, words(x), words(y), on_gpu=False, layout=layout)
        z = from_words(got)
        assert z == x * y % P
        x, y = z, (z + y) % P


@pytest.mark.parametrize("op", [0, 1, 2, 3, 5, 6, 4])
def test_fe8_raw_limbs_on_host(op):
    # The host build of the 8 x 32 field code on its own working range:
    # inputs anywhere below 2^256 (bit 255 kept), outputs congruent mod p.
    import fe_edge
    pairs = fe_edge.cases(n_random=60 if op == 4 else 300)
    if op == 4:
        pairs = pairs[::7]
    A = sum((fe_edge.words(a) for a, _ in pairs), [])
    B = sum((fe_edge.words(b) for _, b in pairs), [])
    out = lxv.debug_fe(op | lxv.FE_RAW_LIMBS, A, B, on_gpu=False, layout=lxv.LAYOUT_8X32)
    assert fe_edge.check(op, pairs, out, raw=True) == []


@pytest.mark.parametrize("op", [0, 1, 2, 3, 5, 6, 4])
def test_fe10_raw_limbs_on_host(op):
    # The host build of the 10 x 25.5 field code at the top of each
    # operation's input range (run-over limbs), outputs congruent mod p and
    # within their promised bounds.
    import fe10_edge
    pairs = fe10_edge.cases(op)
    if op == 4:
        pairs = pairs[::3]
    A, B = fe10_edge.flat(pairs)
    out = lxv.debug_fe(op | lxv.FE_RAW_LIMBS, A, B, on_gpu=False, layout=lxv.LAYOUT_10X26)
    assert fe10_edge.check(op, pairs, out) == []


def test_fe10_column_sums_fit_64_bits():
    # fe10.h's bound: with loose inputs every pre-scaled operand (2 f_i,
    # 4 f_i, 19 g_j) fits 32 bits and every column sum, with the largest
    # carry in, fits 64 bits; the carry out of a column fits two words.
    import fe10_edge as fe
    L = fe.LOOSE_MAX
    carry_in = 2**39
    worst = 0
    for k in range(10):
        s_mul, s_sq = carry_in, carry_in
        for i in range(10):
            j = (k - i) % 10
            a = L[i] * (2 if (i % 2 and j % 2) else 1)
            b = L[j] * (19 if i + j >= 10 else 1)
            assert a < 2**32 and b < 2**32, (k, i, j)
            s_mul += a * b
            if j >= i:
                dbl = (1 if i < j else 0) + (1 if (i % 2 and j % 2) else 0)
                a2 = L[i] << dbl
                assert a2 < 2**32, (k, i, j)
                s_sq += a2 * b
        worst = max(worst, s_mul, s_sq)
    assert worst < 2 ** 62.2, worst.bit_length()   # the bound fe10.h states
    assert worst >> 25 < carry_in
    # and the bound is not vacuous: doubling a wrapped even operand (the
    # doubling kept on the b side) would not fit 32 bits
    assert L[8] * 19 * 2 >= 2**32


def test_device_address_hash_on_host():
    rng = random.Random(25)
    ed_pub = bytes(rng.getrandbits(8) for _ in range(32))
    a = lxv.launch_args_for(ed_pub)
    us = [bytes(rng.getrandbits(8) for _ in range(32)) for _ in range(64)]
    U = sum(([int.from_bytes(u[4 * i:4 * i + 4], "little") for i in range(8)] for u in us), [])
    out = lxv.debug_address(list(a.ed), list(a.name_hash), U, on_gpu=False)
    for i, u in enumerate(us):
        assert lxv.words_to_bytes_be(out[4 * i:4 * i + 4]) == orc.lxmf_address(u, ed_pub)


def test_cli_cpu_walk_matches_x25519_across_batches():
    # the command-line tool's CPU worker walk (src/main.rs:527-539's check)
    import ctypes
    L = lxv.lib()
    L.lxv_debug_cpu_walk.argtypes = [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_void_p]
    start, _, _ = lxv.seeded_search(11, 0)
    out = ctypes.create_string_buffer(3 * 256 * 32)
    assert L.lxv_debug_cpu_walk(start, 3, out) == 0
    keys = out.raw
    for n in list(range(0, 768, 37)) + [255, 256, 511, 512, 767]:
        assert keys[32 * n:32 * n + 32] == orc.x25519(orc.step_scalar(start, n)), n
    w = orc.Walk(start)
    walk = []
    for _ in range(3):
        walk += w.next_batch()[1]
    assert [keys[32 * n:32 * n + 32] for n in range(768)] == walk


@pytest.mark.parametrize("script,generated", [("gen_fe8.py", "fe8_asm.inc"), ("gen_pair_table.py", "pair_table.inc")])
def test_generated_sources_match_their_generators(script, generated):
    # the committed inline asm and point table are exactly what their
    # generators print, so an edit to either side cannot drift unnoticed
    import subprocess
    import sys
    csrc = os.path.join(ROOT, "lxmf-vanity_amd", "csrc")
    out = subprocess.run([sys.executable, os.path.join(csrc, script)], capture_output=True, text=True,
                         check=True, timeout=120).stdout
    assert out == open(os.path.join(csrc, generated)).read()
