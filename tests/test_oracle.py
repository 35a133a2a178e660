"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

Each group is checked against published vectors, the reference's own test
cases (src/main.rs:501-563), Python big-integer arithmetic or hashlib, and the
committed fixtures of tests/golden/ (written by tests/golden/make_golden.py from
independent textbook code). CPU only.
"""

import hashlib
import json
import os
import random
import subprocess
import sys

import pytest

import oracle_lib as orc
import textbook as tb

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P = tb.P


def kat():
    with open(os.path.join(GOLDEN, "kat.json")) as f:
        return json.load(f)


def walk_fixture(seed):
    with open(os.path.join(GOLDEN, "walk_seed%d.json" % seed)) as f:
        return json.load(f)


# ---- hashes

def test_sha256_vectors_and_hashlib():
    for m, h in kat()["sha256"]:
        assert orc.sha256(bytes.fromhex(m)).hex() == h
    rng = random.Random(1)
    for n in list(range(0, 200)) + [1000, 4096, 10007]:
        m = bytes(rng.getrandbits(8) for _ in range(n))
        assert orc.sha256(m) == hashlib.sha256(m).digest(), n


def test_sha256_portable_path_matches_too():
    # The scan uses SHA-NI when the CPU has it; check the portable compression
    # in a process where it is forced.
    code = ("import sys; sys.path[:0] = %r; import oracle_lib as o, hashlib, random; "
            "r = random.Random(2); ms = [bytes(r.getrandbits(8) for _ in range(n)) for n in range(0, 300, 7)]; "
            "assert not o.shani(); assert all(o.sha256(m) == hashlib.sha256(m).digest() for m in ms); print('ok')"
            % [os.path.dirname(os.path.abspath(__file__))])
    env = dict(os.environ, LXV_ORACLE_NO_SHANI="1")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "ok" in out.stdout, out.stderr


def test_sha512_vectors_and_hashlib():
    for m, h in kat()["sha512"]:
        assert orc.sha512(bytes.fromhex(m)).hex() == h
    rng = random.Random(3)
    for n in list(range(0, 260, 3)) + [1000]:
        m = bytes(rng.getrandbits(8) for _ in range(n))
        assert orc.sha512(m) == hashlib.sha512(m).digest(), n


# ---- field arithmetic against Python big integers

EDGE = [0, 1, 2, 19, P - 1, P, P + 1, 2**255 - 1, 2**254, 2**51 - 1, 2**51, 2**102 + 5, 2**255 - 20]


def fe_inputs(n=300, seed=4):
    rng = random.Random(seed)
    vals = list(EDGE) + [rng.getrandbits(255) for _ in range(n)]
    # values with every 51-bit limb at its maximum, and single-limb values
    vals += [sum((2**51 - 1) << (51 * i) for i in range(5)) & (2**255 - 1)]
    vals += [(2**51 - 1) << (51 * i) for i in range(5)]
    return vals


def test_field_ops_match_big_integers():
    vals = fe_inputs()
    rng = random.Random(5)
    for a in vals:
        b = rng.choice(vals)
        A, B = tb.int_le(a), tb.int_le(b)
        assert orc.fe2("orc_fe_mul", A, B) == tb.int_le(a * b % P)
        assert orc.fe2("orc_fe_add", A, B) == tb.int_le((a + b) % P)
        assert orc.fe2("orc_fe_sub", A, B) == tb.int_le((a - b) % P)
        assert orc.fe1("orc_fe_sq", A) == tb.int_le(a * a % P)
        assert orc.fe1("orc_fe_canon", A) == tb.int_le(a % P)


def test_invert_chain():
    for a in fe_inputs(60, seed=6):
        got = tb.le_int(orc.fe1("orc_fe_invert", tb.int_le(a)))
        assert got == pow(a, P - 2, P)
        if a % P:
            assert got * a % P == 1


# ---- curve operations

def test_x25519_rfc7748():
    k = kat()
    for sc, u, out in k["x25519"]:
        assert orc.x25519(bytes.fromhex(sc), bytes.fromhex(u)).hex() == out
    kk = uu = tb.int_le(9)
    for i in range(1, 1001):
        kk, uu = orc.x25519(kk, uu), kk
        if str(i) in k["x25519_iterated"]:
            assert kk.hex() == k["x25519_iterated"][str(i)]
    dh = k["x25519_dh"]
    assert orc.x25519(bytes.fromhex(dh["alice_priv"])).hex() == dh["alice_pub"]
    assert orc.x25519(bytes.fromhex(dh["bob_priv"])).hex() == dh["bob_pub"]
    assert orc.x25519(bytes.fromhex(dh["alice_priv"]), bytes.fromhex(dh["bob_pub"])).hex() == dh["shared"]


def test_x25519_matches_textbook_on_random_scalars():
    rng = random.Random(8)
    for _ in range(20):
        k = bytes(rng.getrandbits(8) for _ in range(32))
        u = bytes(rng.getrandbits(8) for _ in range(32))
        assert orc.x25519(k, u) == tb.x25519(k, u)


def test_u_of_8B():
    assert orc.mont_ladder(tb.int_le(8)).hex() == kat()["u_8B"]


def test_ed25519_rfc8032():
    for seed, pub in kat()["ed25519"]:
        assert orc.ed25519_pub(bytes.fromhex(seed)).hex() == pub
    rng = random.Random(9)
    for _ in range(10):
        s = bytes(rng.getrandbits(8) for _ in range(32))
        assert orc.ed25519_pub(s) == tb.ed25519_pub(s)


# ---- the reference's own tests (src/main.rs:501-563)

def test_address_matches_reticulum_for_a_known_identity():
    r = kat()["rns"]
    ident = bytes.fromhex(r["identity"])
    assert orc.x25519(ident[:32]).hex() == r["x_pub"]
    assert orc.ed25519_pub(ident[32:]).hex() == r["ed_pub"]
    assert orc.name_hash().hex() == kat()["name_hash"]
    assert orc.lxmf_address(bytes.fromhex(r["x_pub"]), bytes.fromhex(r["ed_pub"])).hex() == r["address"]
    assert orc.address_of_identity(ident).hex() == r["address"]


def test_walked_public_keys_match_plain_x25519_across_batches():
    start = tb.seeded_start(11)
    w = orc.Walk(start)
    for _ in range(3):
        first, keys = w.next_batch()
        for i, key in enumerate(keys):
            assert key == orc.x25519(orc.step_scalar(start, first + i)), first + i


@pytest.mark.parametrize("batch", [1, 7, 64, 256])
def test_walk_any_batch_size(batch):
    start = tb.seeded_start(12)
    w = orc.Walk(start, batch)
    got = []
    while len(got) < 300:
        _, keys = w.next_batch()
        got += keys
    assert got[:300] == tb.walk_us(start, 300)


def test_stepped_scalars_stay_clamped():
    start = tb.seeded_start(13)
    for n in [0, 1, 1000, 2**32 - 1, 2**64 // 16]:
        k = orc.step_scalar(start, n)
        assert k[0] & 7 == 0 and k[31] & 0xC0 == 0x40
        assert k == tb.step_scalar(start, n)
    p = orc.previous_scalar(start)
    assert p[0] & 7 == 0 and p[31] & 0xC0 == 0x40
    assert orc.step_scalar(p, 1) == start
    assert p == tb.previous_scalar(start)


def test_start_scalar_rejection_rules():
    # bits 72..254 all set: k + 8n could carry into bit 255 (main.rs:164)
    k = bytearray(32); k[0] = 8; k[9:31] = b"\xff" * 22; k[31] = 127
    assert not orc.start_scalar_ok(bytes(k))
    k[20] = 0xfe
    assert orc.start_scalar_ok(bytes(k))
    # k = 2^254: k - 8 would lose the clamp (main.rs:165)
    k = bytearray(32); k[31] = 64
    assert not orc.start_scalar_ok(bytes(k))
    k[0] = 8
    assert orc.start_scalar_ok(bytes(k))
    # unclamped values are not starts
    k[0] = 9
    assert not orc.start_scalar_ok(bytes(k))
    for seed in range(50):
        s = tb.seeded_start(seed)
        assert orc.start_scalar_ok(s) and tb.start_scalar_ok(s)


def test_prefixes_match_whole_and_half_bytes():
    address = bytes([0xb0, 0x0b, 0x1e, 0x5f] + [0] * 12)
    assert orc.prefix_matches("B00B1E5", address)
    assert orc.prefix_matches("b00b1e5f", address)
    assert not orc.prefix_matches("b00b1e6", address)
    assert orc.prefix_parse("BOOB") is None
    assert orc.prefix_parse("") is None
    assert orc.prefix_parse("0" * 33) is None
    assert orc.prefix_parse("f" * 32) is not None


def test_prefix_words():
    # gpu/mod.rs:74-84: big-endian want/mask words of the fixed nibbles
    assert orc.prefix_words("b00b1e5") == ([0xb00b1e50, 0, 0, 0], [0xfffffff0, 0, 0, 0])
    assert orc.prefix_words("0123456789") == ([0x01234567, 0x89000000, 0, 0], [0xffffffff, 0xff000000, 0, 0])
    assert orc.prefix_words("a") == ([0xa0000000, 0, 0, 0], [0xf0000000, 0, 0, 0])


# ---- the walk and the scan against the golden fixtures

@pytest.mark.parametrize("seed", [1, 7, 2024])
def test_walk_matches_golden(seed):
    fx = walk_fixture(seed)
    start = bytes.fromhex(fx["start"])
    ed_pub = bytes.fromhex(fx["ed_pub"])
    assert orc.ed25519_pub(bytes.fromhex(fx["ed_seed"])) == ed_pub
    for key in ("entries_g0", "entries_far"):
        ents = fx[key]
        g0 = ents[0][0]
        w = orc.Walk(orc.step_scalar(start, g0))
        got = []
        while len(got) < len(ents):
            got += w.next_batch()[1]
        for (g, u, a), mine in zip(ents, got):
            assert mine.hex() == u, (key, g)
            assert orc.lxmf_address(mine, ed_pub).hex() == a, (key, g)


@pytest.mark.parametrize("seed", [1, 7, 2024])
def test_scan_matches_golden(seed):
    fx = walk_fixture(seed)
    start = bytes.fromhex(fx["start"])
    ed_pub = bytes.fromhex(fx["ed_pub"])
    # One range, and the same keys split into ragged ranges: same answer.
    for ranges in ([(0, fx["xor_keys"])], [(0, 1), (1, 999), (1000, 2**15 - 1000), (2**15, 2**15)]):
        hits, xor = orc.scan_ranges(start, ed_pub, ranges, fx["prefix"])
        assert xor.hex() == fx["xor"]
        assert [(g, a.hex()) for g, a in hits] == [tuple(h) for h in fx["hits"]]


def test_2p24_key_walk_matches_golden():
    # 2^24 consecutive keys of workload seed 0, 4-hex prefix: every hit and
    # the XOR of all addresses (tests/golden/walk_2p24_4hex.json, cross-checked
    # against the Python walk on its first 2^16 keys when it was written)
    with open(os.path.join(GOLDEN, "walk_2p24_4hex.json")) as f:
        fx = json.load(f)
    start = bytes.fromhex(fx["start"])
    assert start == tb.seeded_start(0)
    ed_pub = bytes.fromhex(fx["ed_pub"])
    hits, xor = orc.scan_ranges(start, ed_pub, [(0, fx["keys"])], fx["prefix"], cap=1 << 12)
    assert xor.hex() == fx["xor"]
    assert [(g, a.hex()) for g, a in hits] == [tuple(h) for h in fx["hits"]]
    assert 200 < len(hits) < 320


def test_scan_detects_a_planted_mistake():
    # The comparison above is sensitive: one key off, or a different Ed25519
    # half, gives a different checksum and hit list.
    fx = walk_fixture(1)
    start = bytes.fromhex(fx["start"])
    ed_pub = bytes.fromhex(fx["ed_pub"])
    n = fx["xor_keys"]
    want = [tuple(h) for h in fx["hits"]]
    hits, xor = orc.scan_ranges(start, ed_pub, [(1, n)], fx["prefix"])
    assert xor.hex() != fx["xor"]
    hits, xor = orc.scan_ranges(start, orc.ed25519_pub(bytes(32)), [(0, n)], fx["prefix"])
    assert xor.hex() != fx["xor"] and [(g, a.hex()) for g, a in hits] != want
    hits, xor = orc.scan_ranges(orc.step_scalar(start, n), ed_pub, [(0, n)], fx["prefix"])
    assert xor.hex() != fx["xor"] and [(g, a.hex()) for g, a in hits] != want


def test_scan_empty_and_prefixless():
    fx = walk_fixture(7)
    start = bytes.fromhex(fx["start"])
    ed_pub = bytes.fromhex(fx["ed_pub"])
    hits, xor = orc.scan_ranges(start, ed_pub, [(5, 0)], fx["prefix"])
    assert hits == [] and xor == bytes(16)
    # A 32-digit prefix equal to key 3's address finds exactly key 3.
    a3 = fx["entries_g0"][3][2]
    hits, _ = orc.scan_ranges(start, ed_pub, [(0, 256)], a3)
    assert [(g, a.hex()) for g, a in hits] == [(3, a3)]
