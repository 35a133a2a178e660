"""Textbook big-integer restatements used to generate and check fixtures.

Test infrastructure only. Everything here is written from the published
standards with Python integers and hashlib, independently of the C oracle and
of the HIP code, so that agreement between them means something:

* X25519: RFC 7748 section 5 (the ladder pseudocode, a24 = 121665).
* Ed25519 public keys: RFC 8032 section 5.1 (extended coordinates, 5.1.4
  addition, 5.1.5 key generation).
* SHA-256 / SHA-512: hashlib.
* LXMF address: the derivation of the reference's src/main.rs:174-188.
* Walk values u(k + 8n): by an affine x-only step (each point normalised by
  pow(z, p - 2, p)), a different route to the same numbers than the reference's
  batched walk (src/main.rs:255-289), checked against the ladder.
* The seeded workload generator (a SHA-256 counter DRBG) shared with the
  product's host library, so tests can name workloads by seed.
"""

import hashlib
import struct

P = 2**255 - 19
A24 = 121665
ELL = 2**252 + 27742317777372353535851937790883648493


def le_int(b):
    return int.from_bytes(b, "little")


def int_le(v, n=32):
    return int(v).to_bytes(n, "little")


def decode_u(b):
    return (le_int(b) & ((1 << 255) - 1)) % P


def clamp(k_bytes):
    k = bytearray(k_bytes)
    k[0] &= 248
    k[31] &= 127
    k[31] |= 64
    return bytes(k)


def ladder(k, u, bits=255):
    """u([k]P) for P with u-coordinate u, RFC 7748 section 5 over `bits` bits of k."""
    x1, x2, z2, x3, z3, swap = u, 1, 0, u, 1, 0
    for t in reversed(range(bits)):
        kt = (k >> t) & 1
        swap ^= kt
        if swap:
            x2, x3, z2, z3 = x3, x2, z3, z2
        swap = kt
        a = (x2 + z2) % P
        aa = a * a % P
        b = (x2 - z2) % P
        bb = b * b % P
        e = (aa - bb) % P
        c = (x3 + z3) % P
        d = (x3 - z3) % P
        da = d * a % P
        cb = c * b % P
        x3 = (da + cb) ** 2 % P
        z3 = x1 * (da - cb) ** 2 % P
        x2 = aa * bb % P
        z2 = e * (aa + A24 * e) % P
    if swap:
        x2, x3, z2, z3 = x3, x2, z3, z2
    return x2 * pow(z2, P - 2, P) % P


def x25519(k_bytes, u_bytes=b"\x09" + b"\x00" * 31):
    return int_le(ladder(le_int(clamp(k_bytes)), decode_u(u_bytes), 255))


def mont_mul(k_int, u_bytes=b"\x09" + b"\x00" * 31):
    """Unclamped scalar multiple on all 256 bits (X25519_BASEPOINT * Scalar)."""
    return int_le(ladder(k_int, decode_u(u_bytes), 256))


# ---- Ed25519, RFC 8032 section 5.1

D = -121665 * pow(121666, P - 2, P) % P
SQRT_M1 = pow(2, (P - 1) // 4, P)


def _recover_x(y, sign):
    x2 = (y * y - 1) * pow(D * y * y + 1, P - 2, P) % P
    x = pow(x2, (P + 3) // 8, P)
    if (x * x - x2) % P != 0:
        x = x * SQRT_M1 % P
    if (x * x - x2) % P != 0:
        raise ValueError("not a square")
    if x % 2 != sign:
        x = P - x
    return x


_BY = 4 * pow(5, P - 2, P) % P
_BX = _recover_x(_BY, 0)
BASE = (_BX, _BY, 1, _BX * _BY % P)


def _pt_add(p, q):
    x1, y1, z1, t1 = p
    x2, y2, z2, t2 = q
    a = (y1 - x1) * (y2 - x2) % P
    b = (y1 + x1) * (y2 + x2) % P
    c = t1 * 2 * D * t2 % P
    d = z1 * 2 * z2 % P
    e, f, g, h = b - a, d - c, d + c, b + a
    return (e * f % P, g * h % P, f * g % P, e * h % P)


def _pt_mul(s, p):
    q = (0, 1, 1, 0)
    while s > 0:
        if s & 1:
            q = _pt_add(q, p)
        p = _pt_add(p, p)
        s >>= 1
    return q


def ed25519_pub(seed):
    h = hashlib.sha512(seed).digest()
    a = le_int(h[:32])
    a &= (1 << 254) - 8
    a |= 1 << 254
    x, y, z, _ = _pt_mul(a, BASE)
    zi = pow(z, P - 2, P)
    x, y = x * zi % P, y * zi % P
    return int_le(y | ((x & 1) << 255))


# ---- addresses and prefixes (src/main.rs:174-225)

NAME_HASH = hashlib.sha256(b"lxmf.delivery").digest()[:10]


def lxmf_address(x_pub, ed_pub):
    ih = hashlib.sha256(x_pub + ed_pub).digest()[:16]
    return hashlib.sha256(NAME_HASH + ih).digest()[:16]


def address_of_identity(identity):
    return lxmf_address(x25519(identity[:32]), ed25519_pub(identity[32:]))


def prefix_matches(prefix, addr):
    return addr.hex().startswith(prefix.lower())


def step_scalar(k_bytes, n):
    return int_le((le_int(k_bytes) + 8 * n) % (1 << 256))


def previous_scalar(k_bytes):
    return int_le((le_int(k_bytes) - 8) % (1 << 256))


def start_scalar_ok(k):
    if k[0] & 7 or (k[31] & 0xC0) != 0x40:
        return False
    high_all_set = all(b == 0xFF for b in k[9:31]) and k[31] == 127
    low_all_clear = all(b == 0 for b in k[:31]) and k[31] == 64
    return not high_all_set and not low_all_clear


# ---- walk values

def walk_us(k_bytes, count, g0=0):
    """u(k + 8g) for g in [g0, g0 + count), by affine x-only steps."""
    k = le_int(k_bytes) + 8 * g0
    q = decode_u(mont_mul(8))
    up = decode_u(int_le(ladder(k - 8, 9, 255)))
    uc = decode_u(int_le(ladder(k, 9, 255)))
    out = []
    for _ in range(count):
        out.append(int_le(uc))
        t1 = (uc - 1) * (q + 1) % P
        t2 = (uc + 1) * (q - 1) % P
        x = (t1 + t2) ** 2 % P
        z = up * (t1 - t2) ** 2 % P
        up, uc = uc, x * pow(z, P - 2, P) % P
    return out


# ---- the seeded workload generator (shared with the product's host library)

DRBG_TAG = b"lxmf-vanity-amd/v1\x00"


def drbg(seed, label, counter):
    msg = DRBG_TAG + struct.pack("<Q", seed) + label.encode() + b"\x00" + struct.pack("<I", counter)
    return hashlib.sha256(msg).digest()


def seeded_start(seed):
    c = 0
    while True:
        k = clamp(drbg(seed, "start", c))
        if start_scalar_ok(k):
            return k
        c += 1


def seeded_ed_seed(seed):
    return drbg(seed, "ed25519", 0)


def seeded_prefix(seed, nibbles):
    return drbg(seed, "prefix", 0).hex()[:nibbles]
