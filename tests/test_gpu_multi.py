"""GPU tests of the batched search (lxv_multi_*), against the CPU oracle.

A search with V virtual lanes walks keys start + 8 (v * 2^40 + n); its answer
must be its first hit in walk order (lowest block of B keys, then lowest
(n, v)), bit-exact with the oracle walking the same lanes, and the same
whether the search runs alone or among others, eagerly or by graph replay,
one step at a time or pipelined (the next step enqueued before the previous
one is collected).
"""

import random
import time

import pytest

import oracle_lib as orc
import lxv

pytestmark = pytest.mark.gpu

STRIDE = lxv.STRIDE_MULTI
B = 256
MIX_LENGTHS_4_TO_10 = ((4, .10), (5, .15), (6, .20), (7, .20), (8, .15), (9, .12), (10, .08))


def workload(seeds, digits):
    out = []
    for s, d in zip(seeds, digits):
        start, ed_seed, prefix = lxv.seeded_search(s, d)
        out.append((start, lxv.ed25519_public(ed_seed), prefix))
    return out


def run(searches, budget_ms=50, priorities=None, **opts):
    opts.setdefault("batch", B)
    opts.setdefault("waves_per_cu", 16)
    m = lxv.Multi(**opts)
    try:
        prio = priorities or [0] * len(searches)
        ids = [m.submit(st, ed, pf, tag=i, priority=prio[i]) for i, (st, ed, pf) in enumerate(searches)]
        print("searches", len(ids), opts, flush=True)
        res = {r["tag"]: r for r in m.run_all(budget_ms=budget_ms)}
        assert sorted(res) == list(range(len(ids)))
        return res, m.stats()
    finally:
        m.close()


def oracle_first_hit(start, ed_pub, prefix, vlanes, upto_block, batch=B):
    ranges = [(v * STRIDE, (upto_block + 1) * batch) for v in range(vlanes)]
    hits, _ = orc.scan_ranges(start, ed_pub, ranges, prefix, cap=1 << 16)
    best = min(((g % STRIDE) // batch, g % STRIDE, g // STRIDE, g, a) for g, a in hits)
    return best[3], best[4]


def check_first_hit(r, start, ed_pub, prefix, batch=B):
    """r is the search's first hit in walk order, as the oracle walks it."""
    g, addr = r["g"], r["address"]
    assert addr == orc.lxmf_address(orc.x25519(orc.step_scalar(start, g)), ed_pub)
    assert addr.hex().startswith(prefix)
    block = (g % STRIDE) // batch
    assert (g, addr) == oracle_first_hit(start, ed_pub, prefix, r["vlanes"], block, batch)
    assert r["keys"] == (block + 1) * r["vlanes"] * batch


# Every batched kernel the library builds: the pair walk at B = 64, 128, 256
# (8 x 32) and 256 (10 x 25.5), each for 2 and 4 waves per SIMD (8 and 16
# waves per CU); the projective walk at B = 64, 256 (8 x 32) and 256
# (10 x 25.5); and the pair walk with pipelined steps.
MULTI_KERNELS = [dict(walk=lxv.WALK_PAIRS, layout=lxv.LAYOUT_8X32, batch=b, waves_per_cu=w)
                 for b in (64, 128, 256) for w in (8, 16)]
MULTI_KERNELS += [dict(walk=lxv.WALK_PAIRS, layout=lxv.LAYOUT_10X26, batch=256, waves_per_cu=w) for w in (8, 16)]
MULTI_KERNELS += [dict(walk=lxv.WALK_PROJECTIVE, layout=lxv.LAYOUT_8X32, batch=b, waves_per_cu=16) for b in (64, 256)]
MULTI_KERNELS += [dict(walk=lxv.WALK_PROJECTIVE, layout=lxv.LAYOUT_10X26, batch=256, waves_per_cu=16)]
MULTI_KERNELS += [dict(walk=lxv.WALK_PAIRS, layout=lxv.LAYOUT_8X32, batch=256, waves_per_cu=16, pipelined=True)]


def kernel_id(k):
    return "%s-%s-b%d-w%d%s" % ("pairs" if k["walk"] == lxv.WALK_PAIRS else "proj",
                                "8x32" if k["layout"] == lxv.LAYOUT_8X32 else "10x26", k["batch"],
                                k["waves_per_cu"], "-pipelined" if k.get("pipelined") else "")


@pytest.mark.parametrize("kernel", MULTI_KERNELS, ids=[kernel_id(k) for k in MULTI_KERNELS])
def test_answers_are_first_hits_in_walk_order(kernel):
    searches = workload(range(100, 112), [2, 3] * 6)
    res, st = run(searches, **kernel)
    for i, (start, ed_pub, prefix) in enumerate(searches):
        assert res[i]["status"] == 1
        check_first_hit(res[i], start, ed_pub, prefix, kernel["batch"])
    assert st["done"] == len(searches) and st["keys"] > 0


def test_mix_of_256_searches_of_4_to_10_digits():
    # The batched workload's shape: 256 seeded searches with prefix lengths
    # drawn from its 4-10 digit table, each with its own start and Ed25519
    # key, in one persistent kernel. Searches give up after about 2^24 keys,
    # so the 4-6 digit ones mostly finish and the longer ones give way; every
    # finished answer must be the oracle's first hit in walk order.
    rng = random.Random(3)
    digits = []
    for _ in range(256):
        x, acc = rng.random(), 0.0
        for d, pr in MIX_LENGTHS_4_TO_10:
            acc += pr
            if x < acc:
                break
        digits.append(d)
    searches = workload(range(5_000_000, 5_000_256), digits)
    m = lxv.Multi(batch=B, waves_per_cu=16, pool_bytes=4 << 30, pipelined=True)
    try:
        for i, (st, ed, pf) in enumerate(searches):
            m.submit(st, ed, pf, tag=i, max_keys=1 << 24)
        res = {r["tag"]: r for r in m.run_all(budget_ms=100, timeout_s=300)}
    finally:
        m.close()
    assert sorted(res) == list(range(256))
    found = [i for i in res if res[i]["status"] == 1]
    print("found", len(found), "of 256; by length",
          {d: sum(1 for i in found if digits[i] == d) for d in range(4, 11)}, flush=True)
    assert sum(1 for i in found if digits[i] == 4) == digits.count(4)
    assert len(found) >= digits.count(4) + digits.count(5) // 2
    for i in found:
        check_first_hit(res[i], *searches[i])
    for i, r in res.items():
        if r["status"] == 0:
            assert r["keys"] >= 1 << 24


def test_serving_admits_arrivals_while_others_run():
    # Requests arrive as a Poisson process (60/s for 2 s, 3-5 hex digits:
    # each needs several steps of 5 ms) and are submitted between steps,
    # while earlier ones are resident on the GPU. Each answer must be the
    # oracle's first hit in walk order and the same answer the request gets
    # alone.
    rng = random.Random(11)
    t, trace = 0.0, []
    while True:
        t += rng.expovariate(60.0)
        if t >= 2.0:
            break
        trace.append((t, 6_000_000 + len(trace), rng.choice((3, 4, 4, 5))))
    m = lxv.Multi(batch=B, waves_per_cu=16, lane_keys_log2=10, oldest_share_pct=20)
    reqs, res, overlapped = {}, {}, 0
    try:
        t0, nxt = time.perf_counter(), 0
        while nxt < len(trace) or m.stats()["resident"] or m.stats()["queued"]:
            now = time.perf_counter() - t0
            while nxt < len(trace) and trace[nxt][0] <= now:
                _, seed, d = trace[nxt]
                (st, ed, pf), = workload([seed], [d])
                if m.stats()["resident"]:
                    overlapped += 1
                sid = m.submit(st, ed, pf, tag=nxt, priority=d)
                reqs[sid] = (st, ed, pf)
                nxt += 1
            for r in m.step(5):
                res[r["id"]] = r
            if time.perf_counter() - t0 > 120:
                raise TimeoutError("serving loop did not drain")
    finally:
        m.close()
    print("requests", len(trace), "submitted while others were resident", overlapped, flush=True)
    assert len(res) == len(trace) and overlapped >= len(trace) // 4
    for sid, r in res.items():
        assert r["status"] == 1
        check_first_hit(r, *reqs[sid])
    # the same answer alone, for a sample of the requests
    for sid in sorted(res)[::4]:
        alone = run([reqs[sid]], budget_ms=5, lane_keys_log2=10)[0][0]
        assert (alone["g"], alone["address"]) == (res[sid]["g"], res[sid]["address"])


def test_oldest_share_bounds_the_wait_of_a_low_priority_search():
    # A 4-digit search at priority 9, then a 10-digit search at priority 0
    # whose lanes can take the whole GPU every step. With strict priority the
    # first gets no work for as long as the second runs (40 steps of 5 ms
    # here); with a 30% share for the longest-waiting search it finishes
    # within them (its first hit is in its 13th block of 64 x 256 keys), with
    # the answer the oracle gives.
    (slow, hog) = workload([7_000_000, 7_000_001], [4, 10])

    def serve(share):
        m = lxv.Multi(batch=B, waves_per_cu=16, lane_keys_log2=10, oldest_share_pct=share)
        try:
            sid = m.submit(*slow, priority=9)
            m.submit(*hog, priority=0, max_keys=1 << 40)
            for step in range(40):
                for r in m.step(5):
                    if r["id"] == sid:
                        return step, r
            return None
        finally:
            m.close()

    assert serve(0) is None
    got = serve(30)
    print("finished at step", got and got[0], flush=True)
    assert got is not None
    check_first_hit(got[1], *slow)


@pytest.mark.parametrize("walk,lane_bytes", [(lxv.WALK_PAIRS, 64), (lxv.WALK_PROJECTIVE, 128)],
                         ids=["pairs", "projective"])
def test_preempted_search_resumes_with_its_answer(walk, lane_bytes):
    # Pool room for one 64-lane search. A 4-digit search at priority 9 (its
    # first hit in block 9) walks a block per step; after 3 steps a 3-digit
    # search at priority 0 arrives. With preemption it suspends the first,
    # finishes first, and the first resumes at the block it reached and
    # finishes with the oracle's first hit and a key count of an
    # uninterrupted walk; without, the newcomer waits.
    low, high = workload([7_100_007, 7_100_010], [4, 3])

    def serve(preempt):
        m = lxv.Multi(batch=B, waves_per_cu=16, lane_keys_log2=10, pool_bytes=64 * lane_bytes, walk=walk,
                      preempt=preempt)
        try:
            lid = m.submit(*low, priority=9)
            done = []
            for _ in range(3):
                done += m.step(5)
            assert done == [] and m.stats()["resident"] == 1
            hid = m.submit(*high, priority=0)
            for _ in range(200):
                done += m.step(5)
                assert m.stats()["resident"] <= 1
                if len(done) == 2:
                    break
            by = {r["id"]: r for r in done}
            return m.suspensions(), by[lid], by[hid]
        finally:
            m.close()

    suspended, lo, hi = serve(False)
    assert suspended == 0 and lo["done_s"] < hi["done_s"]
    suspended, lo2, hi2 = serve(True)
    assert suspended == 1 and hi2["done_s"] < lo2["done_s"]
    for r, spec in ((lo2, low), (hi2, high), (lo, low)):
        assert r["status"] == 1
        check_first_hit(r, *spec)
    assert (lo2["g"], lo2["address"], lo2["keys"]) == (lo["g"], lo["address"], lo["keys"])
    assert (lo2["g"] % STRIDE) // B == 9


@pytest.mark.parametrize("walk,lane_bytes,pipelined", [(lxv.WALK_PAIRS, 64, False), (lxv.WALK_PROJECTIVE, 128, False),
                                                       (lxv.WALK_PAIRS, 64, True)],
                         ids=["pairs", "projective", "pairs-pipelined"])
def test_admission_is_bounded_by_device_memory(walk, lane_bytes, pipelined):
    # room for three 64-lane searches (8 x 32-bit limbs: 2 or 4 field
    # elements of state per lane); finished searches hand their slots and
    # lanes to queued ones, also while a pipelined step is still running
    m = lxv.Multi(batch=B, waves_per_cu=16, pool_bytes=3 * 64 * lane_bytes, walk=walk, pipelined=pipelined)
    try:
        searches = workload(range(500, 510), [3] * 10)
        for i, (st, ed, pf) in enumerate(searches):
            m.submit(st, ed, pf, tag=i)
        res, peak = [], 0
        while True:
            res += m.step(20)
            s = m.stats()
            peak = max(peak, s["resident"])
            if s["resident"] == 0 and s["queued"] == 0:
                break
        assert peak <= 3 and len(res) == 10 and all(r["status"] == 1 for r in res)
        for r in res:
            st, ed, pf = searches[r["tag"]]
            assert r["address"] == orc.lxmf_address(orc.x25519(orc.step_scalar(st, r["g"])), ed)
    finally:
        m.close()


@pytest.mark.parametrize("pipelined", [False, True], ids=["stepped", "pipelined"])
def test_give_up_and_cancel(pipelined):
    # b is cancelled after its first step was launched (pipelined: while it
    # still runs); a gives up after about 2^22 keys
    m = lxv.Multi(batch=B, waves_per_cu=16, pipelined=pipelined)
    try:
        (s1, e1, _), (s2, e2, _) = workload([600, 601], [0, 0])
        a = m.submit(s1, e1, "0123456789", max_keys=1 << 22)
        b = m.submit(s2, e2, "fedcba9876")
        res = m.step(20)
        m.cancel(b)
        for _ in range(200):
            res += m.step(20)
            if m.stats()["resident"] == 0 and m.stats()["queued"] == 0:
                break
        by = {r["id"]: r for r in res}
        assert by[b]["status"] == -1
        assert by[a]["status"] == 0 and by[a]["keys"] >= 1 << 22
    finally:
        m.close()
