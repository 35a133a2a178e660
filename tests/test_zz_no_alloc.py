"""GPU test of lxv_launch_noalloc, the launch that allocates nothing (its
scratch is a static device array), against lxv_launch on a twin context.

Both contexts are opened with the same shape and set up from the same seeded
start; one walks with lxv_launch (+ lxv_poll), the other with
lxv_launch_noalloc into device buffers this test owns. Every hit, the hit
count, the XOR checksum of all walked addresses and every lane's key counter
must agree exactly, on two calls in a row, the second with a smaller context
(fewer lanes and grid lanes, smaller batch), so leftovers of the first call in
the shared scratch would show. Both are also compared with the oracle.
"""

import ctypes
import os

import pytest

import oracle_lib as orc
import lxv

pytestmark = pytest.mark.gpu

STRIDE = 1 << 20
HIP_H2D, HIP_D2H = 1, 2


def hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            H = ctypes.CDLL(name)
            break
        except OSError:
            H = None
    assert H is not None, "libamdhip64.so not found"
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipDeviceSynchronize.argtypes = []
    return H


def noalloc(b, args, H):
    """lxv_launch_noalloc on backend b into device buffers; returns (hits,
    hit count)."""
    L = lxv.lib()
    L.lxv_launch_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(lxv.LaunchArgs), ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p]
    cap = args.max_hits
    d_hits, d_cnt = ctypes.c_void_p(), ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(d_hits), max(cap, 1) * ctypes.sizeof(lxv.Hit)) == 0
    assert H.hipMalloc(ctypes.byref(d_cnt), 4) == 0
    try:
        junk = (ctypes.c_uint32 * 1)(0xdeadbeef)   # the count must be cleared by the call
        assert H.hipMemcpy(d_cnt, junk, 4, HIP_H2D) == 0
        rc = L.lxv_launch_noalloc(b.h, ctypes.byref(args), d_hits, d_cnt, None)
        assert rc == 0, L.lxv_last_error(b.h).decode()
        assert H.hipDeviceSynchronize() == 0
        cnt = (ctypes.c_uint32 * 1)()
        assert H.hipMemcpy(cnt, d_cnt, 4, HIP_D2H) == 0
        n = min(cnt[0], cap)
        hits = (lxv.Hit * max(n, 1))()
        if n:
            assert H.hipMemcpy(hits, d_hits, n * ctypes.sizeof(lxv.Hit), HIP_D2H) == 0
        return [(hits[i].thread, hits[i].n, lxv.words_to_bytes_be(hits[i].address)) for i in range(n)], cnt[0]
    finally:
        H.hipFree(d_hits)
        H.hipFree(d_cnt)


@pytest.mark.parametrize("order", ["large-then-small"])
def test_noalloc_launch_matches_lxv_launch(order):
    H = hip()
    shapes = [dict(threads=8192, batch=64, resident_lanes=2048, seed=970, iters=3, digits=2),
              dict(threads=1024, batch=16, resident_lanes=256, seed=971, iters=5, digits=2)]
    for sh in shapes:
        start, ed_seed, prefix = lxv.seeded_search(sh["seed"], sh["digits"])
        ed_pub = lxv.ed25519_public(ed_seed)
        print("shape", sh, flush=True)
        ref = lxv.Backend(threads=sh["threads"], batch=sh["batch"], resident_lanes=sh["resident_lanes"])
        new = lxv.Backend(threads=sh["threads"], batch=sh["batch"], resident_lanes=sh["resident_lanes"])
        try:
            assert ref.walk_kernel() == new.walk_kernel()
            assert new.walk_kernel()[0].startswith("walk_pair_queue_kernel")
            ref.init_seeded(start, STRIDE)
            new.init_seeded(start, STRIDE)
            args = lxv.launch_args_for(ed_pub, prefix, iters=sh["iters"], max_hits=1 << 14)
            ref.launch_async(args)
            want_hits, _ = ref.poll(True, cap=args.max_hits)
            want_total = ref.stats().hits_total
            assert want_total == len(want_hits) < args.max_hits
            got_hits, got_total = noalloc(new, args, H)
            assert got_total == want_total
            assert sorted(got_hits) == sorted(want_hits)
            assert new.checksum() == ref.checksum()
            per_lane = sh["iters"] * sh["batch"]
            assert new.read_ns() == ref.read_ns() == [per_lane] * sh["threads"]
            # and both equal the oracle walking the same lanes
            oh, ox = orc.scan_ranges(start, ed_pub, [(t * STRIDE, per_lane) for t in range(sh["threads"])], prefix,
                                     cap=1 << 14)
            assert sorted((t * STRIDE + n, a) for t, n, a in got_hits) == oh
            assert new.checksum() == ox
            print("hits", len(got_hits), "match", flush=True)
        finally:
            ref.close()
            new.close()


def test_noalloc_refuses_a_grid_beyond_its_scratch():
    # the default context's grid (every resident wave, B = 64) needs far more
    # than the 32 MiB of static scratch: refused with a clear error, nothing run
    H = hip()
    b = lxv.Backend(batch=64)
    try:
        start, ed_seed, prefix = lxv.seeded_search(972, 2)
        b.init_seeded(start, STRIDE)
        L = lxv.lib()
        L.lxv_launch_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(lxv.LaunchArgs), ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]
        d = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(d), 1 << 16) == 0
        try:
            args = lxv.launch_args_for(lxv.ed25519_public(ed_seed), prefix, iters=1, max_hits=16)
            rc = L.lxv_launch_noalloc(b.h, ctypes.byref(args), d, d, None)
            assert rc == -1
            assert "static scratch" in L.lxv_last_error(b.h).decode()
        finally:
            H.hipFree(d)
    finally:
        b.close()
