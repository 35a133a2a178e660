"""GPU parity of the walk kernel against the CPU oracle, through the C ABI.

Every (lane, n, address) hit the GPU reports, every dumped public key and the
XOR checksum of every walked address must match the oracle bit for bit on the
same seeded inputs. Lane t walks keys start + 8 (t x stride + n).
"""

import json
import os
import random
import time

import pytest

import oracle_lib as orc
import textbook as tb
import lxv

pytestmark = pytest.mark.gpu

P = tb.P
STRIDE = 1 << 20


def words(v):
    return [(v >> (32 * i)) & 0xffffffff for i in range(8)]


def from_words(ws):
    return sum(w << (32 * i) for i, w in enumerate(ws))


def seeded(seed, nibbles):
    start, ed_seed, prefix = lxv.seeded_search(seed, nibbles)
    return start, ed_seed, lxv.ed25519_public(ed_seed), prefix


@pytest.mark.parametrize("layout", [lxv.LAYOUT_10X26, lxv.LAYOUT_8X32])
def test_gpu_field_ops_match_big_integers(layout):
    rng = random.Random(31)
    vals = [0, 1, 2, 19, P - 1, P, P + 1, 2**255 - 1, 2**254, 2**255 - 20, 2**256 - 1, 2**224 * (2**32 - 1)]
    vals += [rng.getrandbits(255) for _ in range(500)]
    # 8x32 limbs see inputs with every word at its maximum too (values mod 2^255)
    vals += [(2**255 - 1) - rng.getrandbits(20) for _ in range(50)]
    vals = [v & (2**255 - 1) for v in vals]
    bs = [rng.choice(vals) for _ in vals]
    A = sum((words(v) for v in vals), [])
    B = sum((words(v) for v in bs), [])
    expect = {0: [a * b % P for a, b in zip(vals, bs)], 1: [a * a % P for a in vals],
              2: [(a + b) % P for a, b in zip(vals, bs)], 3: [(a - b) % P for a, b in zip(vals, bs)],
              5: [a % P for a in vals], 6: [a * 121665 % P for a in vals],
              4: [pow(a, P - 2, P) for a in vals], 7: [pow(a, P - 2, P) for a in vals]}
    for op, want in expect.items():
        print("field op", op, flush=True)
        got = lxv.debug_fe(op, A, B, on_gpu=True, layout=layout)
        got = [from_words(got[8 * i:8 * i + 8]) for i in range(len(vals))]
        assert got == want, op


@pytest.mark.parametrize("op", [0, 1, 2, 3, 5, 6, 4])
def test_gpu_fe10_raw_limbs_at_their_bounds(op):
    # The 10 x 25.5 code on the GPU (column scans on v_mad_u64_u32) at the top
    # of each operation's input range: results congruent mod p and within
    # their output bounds, and limb for limb equal to the host build.
    import fe10_edge
    pairs = fe10_edge.cases(op)
    if op == 4:
        pairs = pairs[::3]
    A, B = fe10_edge.flat(pairs)
    print("fe10 raw op", op, "cases", len(pairs), flush=True)
    out = lxv.debug_fe(op | lxv.FE_RAW_LIMBS, A, B, on_gpu=True, layout=lxv.LAYOUT_10X26)
    assert fe10_edge.check(op, pairs, out) == []
    host = lxv.debug_fe(op | lxv.FE_RAW_LIMBS, A, B, on_gpu=False, layout=lxv.LAYOUT_10X26)
    assert out == host


@pytest.mark.parametrize("op", [0, 1, 2, 3, 5, 6, 4, 7])
def test_gpu_fe8_raw_limbs_weakly_reduced_inputs(op):
    # The 8 x 32 asm on its own working range: raw limbs in [0, 2^256) with
    # bit 255 kept (values near 2^256, all-ones words, sums that wrap twice,
    # differences that borrow twice), results congruent mod p; the canonical
    # encoding (op 5) of every such value is exact.
    import fe_edge
    pairs = fe_edge.cases()
    if op in (4, 7):
        pairs = pairs[::5]
    A = sum((fe_edge.words(a) for a, _ in pairs), [])
    B = sum((fe_edge.words(b) for _, b in pairs), [])
    print("raw op", op, "cases", len(pairs), flush=True)
    out = lxv.debug_fe(op | lxv.FE_RAW_LIMBS, A, B, on_gpu=True, layout=lxv.LAYOUT_8X32)
    assert fe_edge.check(op, pairs, out, raw=True) == []
    # the same limbs through the host build give the same residues
    host = lxv.debug_fe(op | lxv.FE_RAW_LIMBS, A, B, on_gpu=False, layout=lxv.LAYOUT_8X32)
    assert [fe_edge.value(out[8 * i:8 * i + 8]) % P for i in range(len(pairs))] == \
        [fe_edge.value(host[8 * i:8 * i + 8]) % P for i in range(len(pairs))]


def test_gpu_address_hash_matches_oracle():
    rng = random.Random(32)
    ed_pub = bytes(rng.getrandbits(8) for _ in range(32))
    a = lxv.launch_args_for(ed_pub)
    us = [bytes(rng.getrandbits(8) for _ in range(32)) for _ in range(300)]
    U = sum(([int.from_bytes(u[4 * i:4 * i + 4], "little") for i in range(8)] for u in us), [])
    out = lxv.debug_address(list(a.ed), list(a.name_hash), U, on_gpu=True)
    for i, u in enumerate(us):
        assert lxv.words_to_bytes_be(out[4 * i:4 * i + 4]) == orc.lxmf_address(u, ed_pub)


# One walk per lane, 64 keys per inversion, in both limb layouts and both
# walk forms (the pair walk, the default, and the reference's projective
# walk), and the sub-walk form (64 affine walks per lane: a thread is one of
# them, one key per thread per iteration). The "-queue" forms run the pair
# walk's persistent grid with a few resident waves (4 and 3) taking the 64
# groups of 64 lanes from the queue; with 4096 lanes the others launch one
# thread per lane.
FORMS = {"10x26": dict(threads=4096, batch=64, layout=lxv.LAYOUT_10X26),
         "8x32": dict(threads=4096, batch=64, layout=lxv.LAYOUT_8X32),
         "8x32-queue": dict(threads=4096, batch=64, layout=lxv.LAYOUT_8X32, resident_lanes=256),
         "10x26-queue": dict(threads=4096, batch=64, layout=lxv.LAYOUT_10X26, resident_lanes=192),
         "10x26-proj": dict(threads=4096, batch=64, layout=lxv.LAYOUT_10X26, walk=lxv.WALK_PROJECTIVE),
         "8x32-proj": dict(threads=4096, batch=64, layout=lxv.LAYOUT_8X32, walk=lxv.WALK_PROJECTIVE),
         "8x32-b16": dict(threads=4096, batch=16, layout=lxv.LAYOUT_8X32),
         "8x32-sw64": dict(threads=512, subwalks=64, layout=lxv.LAYOUT_8X32)}


@pytest.fixture(scope="module", params=sorted(FORMS))
def small_backend(request):
    print("backend", request.param, flush=True)
    b = lxv.Backend(**FORMS[request.param])
    yield b
    b.close()


def test_walk_kernel_and_grid_of_each_form():
    def kernel(**kw):
        b = lxv.Backend(**kw)
        try:
            return b.walk_kernel(), b.threads
        finally:
            b.close()
    assert kernel(**FORMS["8x32-queue"]) == (("walk_pair_queue_kernel<lxv::Field8, 64, 4>", 256), 4096)
    assert kernel(**FORMS["10x26-queue"]) == (("walk_pair_queue_kernel<lxv::Field10, 64, 4>", 192), 4096)
    assert kernel(**FORMS["8x32"]) == (("walk_pair_kernel<lxv::Field8, 64, 4>", 4096), 4096)
    assert kernel(**FORMS["8x32-proj"]) == (("walk_kernel<lxv::Field8, 64, 4>", 4096), 4096)
    # the defaults: a persistent grid of the resident waves (16 per CU at
    # 128 VGPRs) over 64 times as many lanes
    (name, grid), lanes = kernel(batch=64)
    assert name == "walk_pair_queue_kernel<lxv::Field8, 64, 4>"
    assert grid % (16 * 64) == 0 and lanes == 64 * grid
    (name, grid), lanes = kernel(batch=64, walk=lxv.WALK_PROJECTIVE)
    assert name == "walk_kernel<lxv::Field8, 64, 4>" and grid == lanes


def test_walk_dump_matches_x25519(small_backend):
    b = small_backend
    start, _, ed_pub, _ = seeded(41, 2)
    b.init_seeded(start, STRIDE)
    hits, dump = b.launch(lxv.launch_args_for(ed_pub, "00", iters=1, dump_threads=256))
    B = b.batch
    assert len(dump) == 256 * B * 8
    rng = random.Random(42)
    lanes = [0, 1, 63, 64, 255] + rng.sample(range(256), 20)
    for t in lanes:
        # every key of the batch on five lanes (every position of a pair
        # walk's batch), a sample on the others
        keys = range(B) if t in (0, 1, 63, 64, 255) else [i for i in list(range(8)) + [31, 32, 62, 63] if i < B]
        for i in keys:
            k = orc.step_scalar(start, t * STRIDE + i)
            got = lxv.words_to_bytes_le(dump[(t * B + i) * 8:(t * B + i) * 8 + 8])
            assert got == orc.x25519(k), (t, i)


@pytest.mark.parametrize("iters,prefix_len", [(1, 2), (4, 3)])
def test_walk_hits_and_checksum_match_oracle(small_backend, iters, prefix_len):
    b = small_backend
    start, _, ed_pub, prefix = seeded(43 + iters, prefix_len)
    b.init_seeded(start, STRIDE)
    hits, _ = b.launch(lxv.launch_args_for(ed_pub, prefix, iters=iters, max_hits=1 << 16))
    per_lane = iters * b.batch
    st = b.stats()
    assert st.keys == b.threads * per_lane and st.stopped == 0
    assert st.hits_total == len(hits)
    got = sorted((t * STRIDE + n, addr) for t, n, addr in hits)
    ranges = [(t * STRIDE, per_lane) for t in range(b.threads)]
    want, xor = orc.scan_ranges(start, ed_pub, ranges, prefix, cap=1 << 16)
    print("hits", len(got), "expected", len(want), flush=True)
    assert len(want) > 10
    assert got == [(g, a) for g, a in want]
    assert b.checksum() == xor
    assert b.read_ns() == [per_lane] * b.threads


def test_walk_continues_across_launches(small_backend):
    # Two launches of 2 batches walk the same keys as one of 4.
    b = small_backend
    start, _, ed_pub, prefix = seeded(47, 3)
    b.init_seeded(start, STRIDE)
    a = lxv.launch_args_for(ed_pub, prefix, iters=2, max_hits=1 << 16)
    h1, _ = b.launch(a)
    h2, _ = b.launch(a)
    x_two = b.checksum()
    b.init_seeded(start, STRIDE)
    h4, _ = b.launch(lxv.launch_args_for(ed_pub, prefix, iters=4, max_hits=1 << 16))
    assert sorted(h1 + h2) == sorted(h4)
    assert b.checksum() == x_two


@pytest.mark.parametrize("form", [dict(threads=256, batch=64), dict(threads=256, batch=64, walk=lxv.WALK_PROJECTIVE),
                                  dict(threads=256, batch=256, layout=lxv.LAYOUT_10X26),
                                  dict(threads=64, subwalks=64)], ids=["pairs", "proj", "pairs-10x26-b256", "sw64"])
def test_reference_style_init_from_host_start_points(form):
    # Backend::init with u(k-8), u(k) per thread from independent random
    # starts (src/gpu/mod.rs:210-228), computed with the product's host X25519.
    b = lxv.Backend(**form)
    try:
        rng = random.Random(48)
        starts = [tb.clamp(bytes(rng.getrandbits(8) for _ in range(32))) for _ in range(b.threads)]
        starts = [s for s in starts if orc.start_scalar_ok(s)]
        assert len(starts) == b.threads
        upc = []
        for s in starts:
            prev = orc.previous_scalar(s)
            for k in (prev, s):
                u = lxv.x25519_base(k)
                upc += [int.from_bytes(u[4 * i:4 * i + 4], "little") for i in range(8)]
        b.init(upc)
        ed_pub = lxv.ed25519_public(bytes(32))
        B = b.batch
        hits, dump = b.launch(lxv.launch_args_for(ed_pub, "0", iters=max(1, 64 // B), dump_threads=b.threads,
                                                  max_hits=1 << 14))
        for t in range(0, b.threads, 7):
            for i in [i for i in (0, 1, 33, 63) if i < B]:
                got = lxv.words_to_bytes_le(dump[(t * B + i) * 8:(t * B + i) * 8 + 8])
                assert got == orc.x25519(orc.step_scalar(starts[t], i)), (t, i)
        for t, n, addr in hits:
            assert addr == orc.lxmf_address(orc.x25519(orc.step_scalar(starts[t], n)), ed_pub)
            assert addr.hex().startswith("0")
        assert len(hits) > 0
    finally:
        b.close()


@pytest.mark.parametrize("batch", [32, 128, 256])
@pytest.mark.parametrize("layout", [lxv.LAYOUT_8X32, lxv.LAYOUT_10X26], ids=["8x32", "10x26"])
def test_pair_walk_every_batch_size(layout, batch):
    # The pair walk at the batch sizes the small backends above do not use:
    # every key of the first batch on four lanes, and every hit and the
    # checksum of two batches on all lanes, against the oracle.
    b = lxv.Backend(threads=1024, batch=batch, layout=layout)
    try:
        start, _, ed_pub, prefix = seeded(60 + batch, 2)
        b.init_seeded(start, STRIDE)
        hits, dump = b.launch(lxv.launch_args_for(ed_pub, prefix, iters=2, dump_threads=4, max_hits=1 << 14))
        for t in range(4):
            for i in range(batch):
                k = orc.step_scalar(start, t * STRIDE + i)
                assert lxv.words_to_bytes_le(dump[(t * batch + i) * 8:(t * batch + i) * 8 + 8]) == orc.x25519(k), (t, i)
        want, xor = orc.scan_ranges(start, ed_pub, [(t * STRIDE, 2 * batch) for t in range(b.threads)], prefix,
                                    cap=1 << 14)
        assert sorted((t * STRIDE + n, a) for t, n, a in hits) == want
        assert b.checksum() == xor
    finally:
        b.close()


# Every occupancy the library builds: waves_per_cu 4, 8 and 16 select the
# kernels compiled for 1, 2 and 4 waves per SIMD (different register budgets
# and spills), in both layouts, both walk forms and at B = 16, 64 and 256;
# the pair walk also on its persistent grid.
OCCUPANCY = [(walk, layout, batch, wpc, queue)
             for walk in (lxv.WALK_PAIRS, lxv.WALK_PROJECTIVE)
             for layout in (lxv.LAYOUT_8X32, lxv.LAYOUT_10X26)
             for batch in (16, 64, 256)
             for wpc in (4, 8, 16)
             for queue in ((False, True) if walk == lxv.WALK_PAIRS else (False,))]


@pytest.mark.parametrize("walk,layout,batch,wpc,queue", OCCUPANCY,
                         ids=["%s-%s-b%d-w%d%s" % ("pairs" if w == lxv.WALK_PAIRS else "proj",
                                                   "8x32" if l == lxv.LAYOUT_8X32 else "10x26", b, c,
                                                   "-queue" if q else "")
                              for w, l, b, c, q in OCCUPANCY])
def test_walk_at_every_occupancy(walk, layout, batch, wpc, queue):
    lanes = 512
    b = lxv.Backend(threads=lanes, batch=batch, layout=layout, walk=walk, waves_per_cu=wpc,
                    resident_lanes=128 if queue else 0)
    try:
        W = {4: 1, 8: 2, 16: 4}[wpc]
        name, grid = b.walk_kernel()
        assert name.endswith(", %d, %d>" % (batch, W)), name
        assert grid == (128 if queue else lanes)
        start, _, ed_pub, prefix = seeded(900 + batch + wpc, 2)
        b.init_seeded(start, STRIDE)
        iters = max(1, 512 // batch)
        hits, dump = b.launch(lxv.launch_args_for(ed_pub, prefix, iters=iters, dump_threads=3, max_hits=1 << 14))
        for t in range(3):
            for i in range(batch):
                k = orc.step_scalar(start, t * STRIDE + i)
                assert lxv.words_to_bytes_le(dump[(t * batch + i) * 8:(t * batch + i) * 8 + 8]) == orc.x25519(k), \
                    (t, i)
        per_lane = iters * batch
        want, xor = orc.scan_ranges(start, ed_pub, [(t * STRIDE, per_lane) for t in range(lanes)], prefix,
                                    cap=1 << 14)
        assert len(want) > 100
        assert sorted((t * STRIDE + n, a) for t, n, a in hits) == want
        assert b.checksum() == xor
        assert b.read_ns() == [per_lane] * lanes
    finally:
        b.close()


def test_full_address_prefix_finds_exactly_that_key(small_backend):
    b = small_backend
    start, _, ed_pub, _ = seeded(49, 0)
    t, n = 1234, 37
    want = orc.lxmf_address(orc.x25519(orc.step_scalar(start, t * STRIDE + n)), ed_pub)
    b.init_seeded(start, STRIDE)
    hits, _ = b.launch(lxv.launch_args_for(ed_pub, want.hex(), iters=n // b.batch + 1))
    assert hits == [(t, n, want)]


def test_raw_want_mask_words_keep_the_and_compare(small_backend):
    """LaunchArgs carries raw words and a hit is (address & mask) == want
    (common.h:344-348): a want bit outside the mask, in the first word or in
    a word the mask leaves empty, matches nothing; a mask in a later word
    alone selects by that word, as the oracle's addresses say."""
    b = small_backend
    start, _, ed_pub, _ = seeded(55, 0)
    b.init_seeded(start, STRIDE)
    b.launch(lxv.launch_args_for(ed_pub, "0"))
    assert b.stats().hits_total > 1000
    for word, bit in ((0, 1), (2, 1 << 7)):
        a = lxv.launch_args_for(ed_pub, "0")
        a.want[word] |= bit
        b.init_seeded(start, STRIDE)
        got, _ = b.launch(a)
        assert got == [] and b.stats().hits_total == 0, (word, bit)
    # address byte 8 == 0xa5, nothing else fixed
    a = lxv.launch_args_for(ed_pub, None)
    a.want[:] = [0, 0, 0xA5000000, 0]
    a.mask[:] = [0, 0, 0xFF000000, 0]
    b.init_seeded(start, STRIDE)
    got, _ = b.launch(a)
    assert b.stats().hits_total == len(got)
    per_lane = b.batch
    every, _ = orc.scan_ranges(start, ed_pub, [(t * STRIDE, per_lane) for t in range(b.threads)], None,
                               cap=b.threads * per_lane)
    assert len(every) == b.threads * per_lane
    want = [(g, addr) for g, addr in every if addr[8] == 0xA5]
    assert len(want) > 50
    assert sorted((t * STRIDE + n, addr) for t, n, addr in got) == want


def test_hits_beyond_max_hits_are_counted_not_stored(small_backend):
    b = small_backend
    start, _, ed_pub, _ = seeded(50, 0)
    b.init_seeded(start, STRIDE)
    iters = max(1, (1 << 18) // (b.threads * b.batch))
    hits, _ = b.launch(lxv.launch_args_for(ed_pub, "0", iters=iters, max_hits=100))
    st = b.stats()
    assert len(hits) == 100 and st.hits_total > 10000


def test_stop_ends_a_long_launch_and_the_walk_resumes(small_backend):
    b = small_backend
    start, _, ed_pub, prefix = seeded(51, 3)
    b.init_seeded(start, STRIDE)
    b.launch_async(lxv.launch_args_for(ed_pub, prefix, iters=20000, max_hits=1 << 16))
    time.sleep(0.3)
    b.stop()
    hits, _ = b.poll(True)
    st = b.stats()
    B = b.batch
    assert st.stopped == 1 and 0 < st.keys < b.threads * B * 20000
    ns = b.read_ns()
    assert all(n % B == 0 for n in ns) and sum(ns) == st.keys
    # The walk resumes where each lane stopped: the next batch's keys match.
    h2, dump = b.launch(lxv.launch_args_for(ed_pub, prefix, iters=1, dump_threads=16))
    for t in range(16):
        for i in sorted({0, B - 1}):
            k = orc.step_scalar(start, t * STRIDE + ns[t] + i)
            assert lxv.words_to_bytes_le(dump[(t * B + i) * 8:(t * B + i) * 8 + 8]) == orc.x25519(k)
    for t, n, addr in hits + h2:
        assert addr == orc.lxmf_address(orc.x25519(orc.step_scalar(start, t * STRIDE + n)), ed_pub)


@pytest.mark.parametrize("form", [dict(threads=65536, batch=256, layout=lxv.LAYOUT_8X32),
                                  dict(threads=65536, batch=256, layout=lxv.LAYOUT_10X26),
                                  dict(threads=65536, batch=256, layout=lxv.LAYOUT_8X32, resident_lanes=16384),
                                  dict(threads=65536, batch=256, layout=lxv.LAYOUT_10X26, resident_lanes=12288),
                                  dict(threads=65536, batch=256, layout=lxv.LAYOUT_8X32, walk=lxv.WALK_PROJECTIVE),
                                  dict(threads=65536, batch=256, layout=lxv.LAYOUT_10X26, walk=lxv.WALK_PROJECTIVE),
                                  dict(threads=1024, subwalks=64)],
                         ids=["8x32", "10x26", "8x32-queue", "10x26-queue", "8x32-proj", "10x26-proj", "8x32-sw64"])
def test_2p24_key_walk_matches_golden(form):
    # The small workload at full size: 2^24 consecutive keys of workload
    # seed 0 split evenly over the threads (thread t walks g in [m t, m t + m)
    # with m = 2^24 / threads), every hit of the 4-hex prefix and the XOR of
    # all 2^24 addresses against the committed fixture.
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "walk_2p24_4hex.json")) as f:
        fx = json.load(f)
    start, ed_pub = bytes.fromhex(fx["start"]), bytes.fromhex(fx["ed_pub"])
    b = lxv.Backend(**form)
    try:
        per_thread = fx["keys"] // b.threads
        assert per_thread * b.threads == fx["keys"]
        b.init_seeded(start, per_thread)
        hits, _ = b.launch(lxv.launch_args_for(ed_pub, fx["prefix"], iters=per_thread // b.batch, max_hits=1 << 12))
        assert b.stats().keys == fx["keys"]
        got = sorted((t * per_thread + n, a.hex()) for t, n, a in hits)
        assert got == [tuple(h) for h in fx["hits"]]
        assert b.checksum().hex() == fx["xor"]
    finally:
        b.close()


def test_headline_shape_2p28_subrange_matches_oracle():
    # The headline workload's shape (workload seed 1, 7-hex prefix, B = 64,
    # 2^22 lanes at 2^40 keys apart on the persistent grid of every resident
    # wave): a 2^28-key sub-range, lanes 0..L-1 x n < 2^28 / L, walked on the
    # GPU and by the oracle; every hit and the XOR of all 2^28 addresses must
    # agree.
    start, _, ed_pub, prefix = seeded(1, 7)
    b = lxv.Backend(batch=64, threads=1 << 22)
    try:
        assert b.walk_kernel()[0].startswith("walk_pair_queue_kernel")
        per_lane = (1 << 28) // b.threads
        assert per_lane * b.threads == 1 << 28 and per_lane % 64 == 0
        b.init_seeded(start, lxv.GpuSearch.LANE_STRIDE)
        hits, _ = b.launch(lxv.launch_args_for(ed_pub, prefix, iters=per_lane // 64, max_hits=1 << 12))
        gpu_xor = b.checksum()
        ranges = [(t * lxv.GpuSearch.LANE_STRIDE, per_lane) for t in range(b.threads)]
        print("lanes", b.threads, "keys per lane", per_lane, "hits", len(hits), flush=True)
        want, xor = orc.scan_ranges(start, ed_pub, ranges, prefix, cap=1 << 12)
        assert sorted((t * lxv.GpuSearch.LANE_STRIDE + n, a) for t, n, a in hits) == want
        assert gpu_xor == xor
    finally:
        b.close()


@pytest.mark.parametrize("op", range(16))
def test_valu_probe_rates_are_sane(op):
    # The issue-rate probes the VALU roofline is priced against: each runs,
    # reports a positive rate no more than twice one wave64 instruction per
    # SIMD per 4 clocks at the clock it measured, and a clock within the
    # range the device can run at.
    rate, clk, nominal, cus = lxv.debug_valu_rate(op, 16)
    print("probe", op, "%.3e lane-ops/s at %.3f GHz (nominal %.3f GHz), %d CUs" % (rate, clk / 1e9, nominal / 1e9,
                                                                                  cus), flush=True)
    assert cus > 0 and 0.3e9 < clk < 1.2 * nominal
    assert 0 < rate < 2.0 * cus * 64 * clk


def test_poll_with_a_small_dump_buffer_keeps_the_launch():
    # lxv_poll refuses a dump buffer that is too small before collecting
    # anything: the launch stays in flight and a second poll gets everything
    import ctypes
    b = lxv.Backend(threads=1024, batch=64)
    try:
        start, _, ed_pub, prefix = seeded(62, 2)
        b.init_seeded(start, STRIDE)
        b.launch_async(lxv.launch_args_for(ed_pub, prefix, iters=1, dump_threads=4, max_hits=1 << 12))
        L = lxv.lib()
        small = (ctypes.c_uint32 * 8)()
        n = ctypes.c_uint32(0)
        hits = (lxv.Hit * 4096)()
        assert L.lxv_poll(b.h, 1, hits, 4096, ctypes.byref(n), small, 8) == -1
        assert "dump buffer too small" in L.lxv_last_error(b.h).decode()
        got, dump = b.poll(True, dump_threads=4, cap=4096)
        want, _ = orc.scan_ranges(start, ed_pub, [(t * STRIDE, 64) for t in range(1024)], prefix, cap=1 << 12)
        assert sorted((t * STRIDE + n_, a) for t, n_, a in got) == want
        assert lxv.words_to_bytes_le(dump[:8]) == orc.x25519(start)
    finally:
        b.close()


def test_self_test_and_run():
    start, ed_seed, _, _ = seeded(52, 0)
    s = lxv.GpuSearch(start=start, ed_seed=ed_seed, threads=16384)
    try:
        report = s.self_test()
        print(report, flush=True)
        identity, tries = s.run("abc")
        addr = orc.address_of_identity(identity)
        assert addr.hex().startswith("abc") and tries > 0
    finally:
        s.close()


@pytest.mark.parametrize("resident", [0, 64], ids=["thread-per-lane", "queue"])
def test_ragged_lane_count_and_empty_launch(resident):
    # 100 lanes asked for: the library rounds up to whole waves (128), walks
    # every one of them, and a launch of zero batches walks nothing and
    # leaves every lane where it was.
    b = lxv.Backend(threads=100, batch=64, resident_lanes=resident)
    try:
        assert b.threads == 128
        start, _, ed_pub, prefix = seeded(61, 2)
        b.init_seeded(start, STRIDE)
        hits, _ = b.launch(lxv.launch_args_for(ed_pub, prefix, iters=0))
        assert hits == [] and b.stats().keys == 0 and b.read_ns() == [0] * 128
        hits, _ = b.launch(lxv.launch_args_for(ed_pub, prefix, iters=2, max_hits=1 << 12))
        want, xor = orc.scan_ranges(start, ed_pub, [(t * STRIDE, 128) for t in range(128)], prefix, cap=1 << 12)
        assert sorted((t * STRIDE + n, a) for t, n, a in hits) == want
        assert b.checksum() == xor and b.read_ns() == [128] * 128
    finally:
        b.close()


def test_headline_shape_launch_at_full_size():
    # The headline workload's own launch: every lane of the default grid (2^24
    # lanes on MI355X), 8 batches of 64 keys each, 2^33 keys in one launch.
    # Every hit of the 7-hex prefix is re-derived on the CPU, every lane has
    # walked exactly its 512 keys (sampled at both ends of the lane range and
    # in the middle), and the launch counted every key.
    start, _, ed_pub, prefix = seeded(1, 7)
    b = lxv.Backend(batch=64)
    try:
        stride = lxv.lane_stride(b.threads)
        b.init_seeded(start, stride)
        hits, _ = b.launch(lxv.launch_args_for(ed_pub, prefix, iters=8, max_hits=1 << 12))
        st = b.stats()
        print("lanes", b.threads, "kernel", b.walk_kernel(), "hits", len(hits), "ms %.1f" % st.kernel_ms, flush=True)
        assert st.keys == b.threads * 512 and st.stopped == 0 and st.hits_total == len(hits)
        assert 8 <= len(hits) <= 96  # 2^33 / 16^7 = 32 expected
        for t, n, addr in hits:
            assert addr == orc.lxmf_address(orc.x25519(orc.step_scalar(start, t * stride + n)), ed_pub)
            assert addr.hex().startswith(prefix)
        for first in (0, b.threads // 2, b.threads - 1000):
            assert b.read_ns(first, 1000) == [512] * 1000
    finally:
        b.close()
