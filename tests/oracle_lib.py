"""ctypes binding of the CPU oracle (oracle/liblxv_oracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, as the checker.
"""

import ctypes
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liblxv_oracle.so")

_lib = None

u8p = ctypes.c_char_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name in ("orc_fe_mul", "orc_fe_add", "orc_fe_sub", "orc_x25519", "orc_mont_ladder",
                     "orc_lxmf_address"):
            getattr(L, name).argtypes = [u8p, u8p, ctypes.c_void_p]
        for name in ("orc_fe_sq", "orc_fe_invert", "orc_fe_canon", "orc_ed25519_pub",
                     "orc_previous_scalar", "orc_address_of_identity"):
            getattr(L, name).argtypes = [u8p, ctypes.c_void_p]
        L.orc_sha256.argtypes = [u8p, ctypes.c_size_t, ctypes.c_void_p]
        L.orc_sha512.argtypes = [u8p, ctypes.c_size_t, ctypes.c_void_p]
        L.orc_step_scalar.argtypes = [u8p, ctypes.c_uint64, ctypes.c_void_p]
        L.orc_start_scalar_ok.argtypes = [u8p]
        L.orc_name_hash.argtypes = [ctypes.c_void_p]
        L.orc_prefix_parse.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.orc_prefix_matches.argtypes = [u8p, ctypes.c_int, u8p]
        L.orc_prefix_words.argtypes = [u8p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_walk_new_batch.argtypes = [u8p, ctypes.c_size_t]
        L.orc_walk_new_batch.restype = ctypes.c_void_p
        L.orc_walk_next_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.orc_walk_next_batch.restype = ctypes.c_uint64
        L.orc_walk_free.argtypes = [ctypes.c_void_p]
        L.orc_scan_ranges.argtypes = [u8p, u8p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                      u8p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), ctypes.c_void_p,
                                      ctypes.c_int]
        _lib = L
    return _lib


def _out(n):
    return ctypes.create_string_buffer(n)


def sha256(m):
    o = _out(32); lib().orc_sha256(m, len(m), o); return o.raw


def sha512(m):
    o = _out(64); lib().orc_sha512(m, len(m), o); return o.raw


def shani():
    return bool(lib().orc_sha256_uses_shani())


def fe2(name, a, b):
    o = _out(32); getattr(lib(), name)(a, b, o); return o.raw


def fe1(name, a):
    o = _out(32); getattr(lib(), name)(a, o); return o.raw


def x25519(k, u=b"\x09" + b"\x00" * 31):
    return fe2("orc_x25519", k, u)


def mont_ladder(k, u=b"\x09" + b"\x00" * 31):
    return fe2("orc_mont_ladder", k, u)


def ed25519_pub(seed):
    return fe1("orc_ed25519_pub", seed)


def step_scalar(k, n):
    o = _out(32); lib().orc_step_scalar(k, n, o); return o.raw


def previous_scalar(k):
    return fe1("orc_previous_scalar", k)


def start_scalar_ok(k):
    return bool(lib().orc_start_scalar_ok(k))


def name_hash():
    o = _out(10); lib().orc_name_hash(o); return o.raw


def lxmf_address(x_pub, ed_pub):
    o = _out(16); lib().orc_lxmf_address(x_pub, ed_pub, o); return o.raw


def address_of_identity(identity):
    o = _out(16); lib().orc_address_of_identity(identity, o); return o.raw


def prefix_parse(s):
    b = _out(16); n = ctypes.c_int(0)
    rc = lib().orc_prefix_parse(s.encode(), b, ctypes.byref(n))
    return None if rc != 0 else (b.raw, n.value)


def prefix_matches(s, addr):
    p = prefix_parse(s)
    return bool(lib().orc_prefix_matches(p[0], p[1], addr))


def prefix_words(s):
    p = prefix_parse(s)
    want = (ctypes.c_uint32 * 4)(); mask = (ctypes.c_uint32 * 4)()
    lib().orc_prefix_words(p[0], p[1], want, mask)
    return list(want), list(mask)


class Walk:
    """The reference CPU walk (src/main.rs:229-290)."""

    def __init__(self, start, batch=256):
        self.batch = batch
        self.h = lib().orc_walk_new_batch(start, batch)
        if not self.h:
            raise MemoryError

    def next_batch(self, n=None):
        n = self.batch if n is None else n
        buf = _out(32 * n)
        first = lib().orc_walk_next_batch(self.h, buf, n)
        return first, [buf.raw[32 * i:32 * i + 32] for i in range(n)]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_walk_free(self.h)
            self.h = None


def scan_ranges(start, ed_pub, ranges, prefix, cap=1 << 16, threads=0):
    """Walks k + 8g over each (g0, length) range; returns (hits, xor16).

    hits is the sorted list of (g, address) whose address matches `prefix`."""
    n = len(ranges)
    g0 = (ctypes.c_uint64 * n)(*[r[0] for r in ranges])
    ln = (ctypes.c_uint64 * n)(*[r[1] for r in ranges])
    pb, pn = prefix_parse(prefix) if prefix else (b"\x00" * 16, 0)
    hg = (ctypes.c_uint64 * cap)()
    ha = _out(16 * cap)
    nh = ctypes.c_uint32(0)
    xo = _out(16)
    rc = lib().orc_scan_ranges(start, ed_pub, g0, ln, n, pb, pn, hg, ha, cap, ctypes.byref(nh), xo, threads)
    if rc != 0:
        raise RuntimeError("orc_scan_ranges failed")
    hits = [(hg[i], ha.raw[16 * i:16 * i + 16]) for i in range(min(nh.value, cap))]
    return hits, xo.raw
