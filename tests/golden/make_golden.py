"""Writes the golden fixtures under tests/golden/ from textbook.py.

Run from the repository root:  python tests/golden/make_golden.py

The fixtures are small JSON files of inputs and expected outputs:
* kat.json: published vectors (FIPS 180-4, RFC 7748 5.2 and 6.1, RFC 8032
  7.1) with their expected values typed from the standards, each recomputed
  here from the textbook code and asserted before it is written; the RNS
  identity known answer of the reference (src/main.rs:509-525) with its
  intermediates; the name hash and u(8B).
* walk_seed<S>.json: for a few workload seeds, the start scalar, Ed25519
  seed and public key, a prefix, 1024 consecutive (g, u, address) entries
  from g = 0 and from a far g, the XOR of all addresses over the first 2^16
  keys, and every hit of a 2-digit prefix there.
"""

import hashlib
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import textbook as tb  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

SHA256_VECTORS = [
    ("", "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"),
    ("616263", "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"),
    ("6162636462636465636465666465666765666768666768696768696a68696a6b696a6b6c6a6b6c6d6b6c6d6e6c6d6e6f6d6e6f706e6f7071",
     "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1"),
]
SHA512_VECTORS = [
    ("616263", "ddaf35a193617abacc417349ae20413112e6fa4e89a97ea20a9eeee64b55d39a"
               "2192992a274fc1a836ba3c23a3feebbd454d4423643ce80e2a9ac94fa54ca49f"),
]
X25519_VECTORS = [  # RFC 7748 5.2
    ("a546e36bf0527c9d3b16154b82465edd62144c0ac1fc5a18506a2244ba449ac4",
     "e6db6867583030db3594c1a424b15f7c726624ec26b3353b10a903a6d0ab1c4c",
     "c3da55379de9c6908e94ea4df28d084f32eccf03491c71f754b4075577a28552"),
    ("4b66e9d4d1b4673c5ad22691957d6af5c11b6421e0ea01d42ca4169e7918ba0d",
     "e5210f12786811d3f4b7959d0538ae2c31dbe7106fc03c3efc4cd549c715a493",
     "95cbde9476e8907d7aade45cb4b873f88b595a68799fa152e6f8f7647aac7957"),
]
X25519_ITERATED = {1: "422c8e7a6227d7bca1350b3e2bb7279f7897b87bb6854b783c60e80311ae3079",
                   1000: "684cf59ba83309552800ef566f2f4d3c1c3887c49360e3875f2eb94d99532c51"}
X25519_DH = {  # RFC 7748 6.1
    "alice_priv": "77076d0a7318a57d3c16c17251b26645df4c2f87ebc0992ab177fba51db92c2a",
    "alice_pub": "8520f0098930a754748b7ddcb43ef75a0dbf3a0d26381af4eba4a98eaa9b4e6a",
    "bob_priv": "5dab087e624a8a4b79e17f8b83800ee66f3bb1292618b6fd1c2f8b27ff88e0eb",
    "bob_pub": "de9edb7d7b7dc1b4d35b61c2ece435373f8343c85b78674dadfc7e146f882b4f",
    "shared": "4a5d9d5ba4ce2de1728e3bf480350f25e07e21c947d19e3376f09b3c1e161742",
}
ED25519_VECTORS = [  # RFC 8032 7.1 tests 1-3
    ("9d61b19deffd5a60ba844af492ec2cc44449c5697b326919703bac031cae7f60",
     "d75a980182b10ab7d54bfed3c964073a0ee172f3daa62325af021a68f707511a"),
    ("4ccd089b28ff96da9db6c346ec114e0f5b8a319f35aba624da8cf6ed4fb8a6fb",
     "3d4017c3e843895a92b70aa74d1b7ebc9c982ccf2ec4968cc0cd55f12af4660c"),
    ("c5aa8df43f9f837bedb7442f31dcb7b166d38535076f094b85ce3a2e0b4458f7",
     "fc51cd8e6218a1a38da47ed00230f0580816ed13ba3303ac5deb911548908025"),
]
RNS = {  # src/main.rs:509-525 and its intermediates
    "identity": "e83d9ecdcdd012692583bd66f997bd0fd074d1ff05c62e892c3b5abbf1dd8678"
                "01cc1353accbf59a42a1b951c71c6454cda7e4b8252ae6a5cb7fa55bb0430442",
    "x_pub": "baf5e06908b2b4f8b714d808107411e6c852cb0f2fe762688b0c7dbc8e500b3a",
    "ed_pub": "503261d0694623ee7d475377cba53f34c008a313ea4fcde7e062b3d33e25b038",
    "identity_hash": "40eb27183deb9512caaa6fba058b93cf",
    "address": "b00b11fc764b58f813c3b3cc279f8b8a",
}
NAME_HASH = "6ec60bc318e2c0f0d908"
U_8B = "6899c948ed6f71f11ea974adcaed787aa15cb6c84d883bd82d5cd6aad7a65f27"


def check(cond, what):
    if not cond:
        raise SystemExit("fixture check failed: " + what)


def make_kat():
    for m, h in SHA256_VECTORS:
        check(hashlib.sha256(bytes.fromhex(m)).hexdigest() == h, "sha256 " + m)
    for m, h in SHA512_VECTORS:
        check(hashlib.sha512(bytes.fromhex(m)).hexdigest() == h, "sha512 " + m)
    for k, u, o in X25519_VECTORS:
        check(tb.x25519(bytes.fromhex(k), bytes.fromhex(u)).hex() == o, "x25519 " + k)
    k = u = tb.int_le(9)
    for i in range(1, 1001):
        k, u = tb.x25519(k, u), k
        if i in X25519_ITERATED:
            check(k.hex() == X25519_ITERATED[i], "x25519 iterated %d" % i)
    dh = X25519_DH
    check(tb.x25519(bytes.fromhex(dh["alice_priv"])).hex() == dh["alice_pub"], "alice pub")
    check(tb.x25519(bytes.fromhex(dh["bob_priv"])).hex() == dh["bob_pub"], "bob pub")
    check(tb.x25519(bytes.fromhex(dh["alice_priv"]), bytes.fromhex(dh["bob_pub"])).hex() == dh["shared"], "shared")
    for s, pk in ED25519_VECTORS:
        check(tb.ed25519_pub(bytes.fromhex(s)).hex() == pk, "ed25519 " + s)
    idb = bytes.fromhex(RNS["identity"])
    check(tb.x25519(idb[:32]).hex() == RNS["x_pub"], "rns x_pub")
    check(tb.ed25519_pub(idb[32:]).hex() == RNS["ed_pub"], "rns ed_pub")
    ih = hashlib.sha256(bytes.fromhex(RNS["x_pub"] + RNS["ed_pub"])).digest()[:16]
    check(ih.hex() == RNS["identity_hash"], "rns identity hash")
    check(tb.address_of_identity(idb).hex() == RNS["address"], "rns address")
    check(tb.NAME_HASH.hex() == NAME_HASH, "name hash")
    check(tb.mont_mul(8).hex() == U_8B, "u(8B)")
    return {
        "sha256": SHA256_VECTORS, "sha512": SHA512_VECTORS,
        "x25519": X25519_VECTORS, "x25519_iterated": {str(k): v for k, v in X25519_ITERATED.items()},
        "x25519_dh": X25519_DH, "ed25519": ED25519_VECTORS, "rns": RNS,
        "name_hash": NAME_HASH, "u_8B": U_8B,
    }


def make_walk(seed, n_keys=1024, far_g=(1 << 40) + 12345, xor_keys=1 << 16, prefix_digits=2):
    start = tb.seeded_start(seed)
    ed_seed = tb.seeded_ed_seed(seed)
    ed_pub = tb.ed25519_pub(ed_seed)
    prefix = tb.seeded_prefix(seed, prefix_digits)

    def entries(g0):
        us = tb.walk_us(start, n_keys, g0)
        # Every 64th walked key against an independent ladder.
        for i in range(0, n_keys, 64):
            check(us[i] == tb.x25519(tb.step_scalar(start, g0 + i)), "walk vs ladder g=%d" % (g0 + i))
        return [[g0 + i, u.hex(), tb.lxmf_address(u, ed_pub).hex()] for i, u in enumerate(us)]

    us = tb.walk_us(start, xor_keys, 0)
    for i in range(0, xor_keys, 4096):
        check(us[i] == tb.x25519(tb.step_scalar(start, i)), "xor walk vs ladder g=%d" % i)
    acc = bytearray(16)
    hits = []
    for g, u in enumerate(us):
        a = tb.lxmf_address(u, ed_pub)
        for j in range(16):
            acc[j] ^= a[j]
        if tb.prefix_matches(prefix, a):
            hits.append([g, a.hex()])
    return {
        "seed": seed, "start": start.hex(), "ed_seed": ed_seed.hex(), "ed_pub": ed_pub.hex(),
        "prefix": prefix, "entries_g0": entries(0), "entries_far": entries(far_g),
        "xor_keys": xor_keys, "xor": bytes(acc).hex(), "hits": hits,
    }


def main():
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump(make_kat(), f, indent=1)
    for seed in (1, 7, 2024):
        with open(os.path.join(HERE, "walk_seed%d.json" % seed), "w") as f:
            json.dump(make_walk(seed), f, indent=0)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
