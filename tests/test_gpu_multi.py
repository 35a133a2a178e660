"""GPU tests of the batched search (lxv_multi_*), against the CPU oracle.

A search with V virtual lanes walks keys start + 8 (v * 2^40 + n); its answer
must be its first hit in walk order (lowest block of B keys, then lowest
(n, v)), bit-exact with the oracle walking the same lanes, and the same
whether the search runs alone or among others, eagerly or by graph replay,
one step at a time or pipelined (the next step enqueued before the previous
one is collected).
"""

import pytest

import oracle_lib as orc
import lxv

pytestmark = pytest.mark.gpu

STRIDE = lxv.STRIDE_MULTI
B = 256


def workload(seeds, digits):
    out = []
    for s, d in zip(seeds, digits):
        start, ed_seed, prefix = lxv.seeded_search(s, d)
        out.append((start, lxv.ed25519_public(ed_seed), prefix))
    return out


def run(searches, budget_ms=50, priorities=None, **opts):
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


def oracle_first_hit(start, ed_pub, prefix, vlanes, upto_block):
    ranges = [(v * STRIDE, (upto_block + 1) * B) for v in range(vlanes)]
    hits, _ = orc.scan_ranges(start, ed_pub, ranges, prefix, cap=1 << 16)
    best = min(((g % STRIDE) // B, g % STRIDE, g // STRIDE, g, a) for g, a in hits)
    return best[3], best[4]


@pytest.mark.parametrize("walk,pipelined", [(lxv.WALK_PAIRS, False), (lxv.WALK_PROJECTIVE, False),
                                             (lxv.WALK_PAIRS, True)], ids=["pairs", "projective", "pairs-pipelined"])
def test_answers_are_first_hits_in_walk_order(walk, pipelined):
    searches = workload(range(100, 112), [2, 3] * 6)
    res, st = run(searches, walk=walk, pipelined=pipelined)
    for i, (start, ed_pub, prefix) in enumerate(searches):
        r = res[i]
        assert r["status"] == 1
        g, addr = r["g"], r["address"]
        assert addr == orc.lxmf_address(orc.x25519(orc.step_scalar(start, g)), ed_pub)
        assert addr.hex().startswith(prefix)
        block = (g % STRIDE) // B
        want_g, want_addr = oracle_first_hit(start, ed_pub, prefix, r["vlanes"], block)
        assert (g, addr) == (want_g, want_addr), i
    assert st["done"] == len(searches) and st["keys"] > 0


def test_same_answer_alone_and_in_a_mix():
    targets = workload([200, 201, 202], [3, 4, 3])
    others = workload(range(300, 320), [2, 3, 4, 5] * 5)
    alone = [run([t])[0][0] for t in targets]
    mixed, _ = run(others[:10] + targets + others[10:])
    # priorities change who is admitted and served first, never an answer
    prio = [i % 4 for i in range(10)] + [5, 5, 5] + [i % 3 for i in range(10)]
    ranked, _ = run(others[:10] + targets + others[10:], budget_ms=5, priorities=prio)
    piped, _ = run(others[:10] + targets + others[10:], budget_ms=5, priorities=prio, pipelined=True)
    for i, a in enumerate(alone):
        for res in (mixed, ranked, piped):
            m = res[10 + i]
            assert (m["g"], m["address"]) == (a["g"], a["address"])


def test_blocks_split_across_steps_keep_the_answer():
    # 6 digits at 2^8 expected keys per lane: 65536 virtual lanes, so a block
    # is 2^24 keys, far above a 1 ms step's share: each block is walked a few
    # virtual waves per step. The answer must match whole-block steps and the
    # oracle, and `keys` must count every block up to the answer's.
    searches = workload(range(700, 704), [6] * 4)
    split, st = run(searches, budget_ms=1, lane_keys_log2=8)
    piped, st_p = run(searches, budget_ms=1, lane_keys_log2=8, pipelined=True)
    whole, _ = run(searches, budget_ms=2000, lane_keys_log2=8)
    assert st["steps"] > 8 and st_p["steps"] > 8
    for i, (start, ed_pub, prefix) in enumerate(searches):
        r = split[i]
        assert r["vlanes"] == 65536 and r["status"] == 1
        assert (r["g"], r["address"]) == (whole[i]["g"], whole[i]["address"])
        assert (piped[i]["g"], piped[i]["address"], piped[i]["keys"]) == (r["g"], r["address"], r["keys"])
        block = (r["g"] % STRIDE) // B
        assert r["keys"] == (block + 1) * r["vlanes"] * B
        assert (r["g"], r["address"]) == oracle_first_hit(start, ed_pub, prefix, r["vlanes"], block)


def test_graph_replay_matches_eager_launches():
    searches = workload(range(400, 416), [3, 4] * 8)
    g_res, g_st = run(searches)
    e_res, e_st = run(searches, no_graph=True)
    assert g_st["steps"] > 1
    for i in range(len(searches)):
        assert (g_res[i]["g"], g_res[i]["address"]) == (e_res[i]["g"], e_res[i]["address"])


@pytest.mark.parametrize("walk,lane_bytes,pipelined", [(lxv.WALK_PAIRS, 64, False), (lxv.WALK_PROJECTIVE, 128, False),
                                                       (lxv.WALK_PAIRS, 64, True)],
                         ids=["pairs", "projective", "pairs-pipelined"])
def test_admission_is_bounded_by_device_memory(walk, lane_bytes, pipelined):
    # room for three 64-lane searches (8 x 32-bit limbs: 2 or 4 field
    # elements of state per lane); finished searches hand their slots and
    # lanes to queued ones, also while a pipelined step is still running
    m = lxv.Multi(pool_bytes=3 * 64 * lane_bytes, walk=walk, pipelined=pipelined)
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
    m = lxv.Multi(pipelined=pipelined)
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
