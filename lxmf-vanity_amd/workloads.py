"""Many-search workloads on one MI355X over the batched API (lxv_multi_*).

batched(): 256 seeded searches submitted at once, prefix lengths 4-10 hex
digits drawn with P = .10/.15/.20/.20/.15/.12/.08, stepped for a wall budget;
reports aggregate keys/s (to compare with the plain kernel on one search),
completions per length, and re-derives every answer on the CPU.

serving(): requests arrive as a Poisson process (rate 20/s by default) with
prefix lengths 5-9 drawn with P = .30/.30/.20/.15/.05, each a fresh identity
(its own start scalar and Ed25519 key); the loop admits them into the
persistent kernel as slots and HBM pool memory allow, in a chosen order
(shortest expected work first, optionally with a share of every step for the
longest-waiting request, or arrival order), and reports p50/p99 time-to-hit
from arrival (over finished requests, and over all with the unfinished
counted at their age), queue depth over time, aggregate keys/s and the
requests still unfinished at the end.

Both run at the launch shape the library was tuned for (lxv.tuned_shape:
keys per inversion, waves per CU, layout, step budget, keys per virtual
lane), unless the caller overrides it. one_by_one() replays the same arrivals
through the plain kernel, one search at a time in arrival order, as the
reference's binary would serve them (src/gpu/mod.rs:146-174).

Workload identities come from the seeded generator (lxv.seeded_search):
request i of the batched workload uses seed 3_000_000 + i, of the serving
workload 4_000_000 + i; lengths and arrival gaps come from Python's
random.Random(seed), which is stable across Python versions.
"""

import math
import random
import time

import lxv

BATCHED_LENGTHS = ((4, .10), (5, .15), (6, .20), (7, .20), (8, .15), (9, .12), (10, .08))
SERVING_LENGTHS = ((5, .30), (6, .30), (7, .20), (8, .15), (9, .05))


def draw_length(rng, table):
    x, acc = rng.random(), 0.0
    for digits, p in table:
        acc += p
        if x < acc:
            return digits
    return table[-1][0]


class Request:
    """One search: a fresh identity from workload seed `seed` and a prefix."""

    def __init__(self, seed, digits, arrival=0.0):
        self.seed, self.digits, self.arrival = seed, digits, arrival
        self.start, self.ed_seed, self.prefix = lxv.seeded_search(seed, digits)
        self.ed_pub = lxv.ed25519_public(self.ed_seed)
        self.done = None      # seconds from the workload's start, when the answer was on the host
        self.g = None
        self.keys = 0

    def recheck(self, g, address):
        """The identity for step index g, re-derived from scratch on the CPU
        (src/gpu/mod.rs:157-160): its address must be `address` and match."""
        ident = lxv.step_scalar(self.start, g) + self.ed_seed
        return lxv.address_of_identity(ident) == address and address.hex().startswith(self.prefix)


def quantile(xs, q):
    if not xs:
        return None
    xs = sorted(xs)
    i = min(len(xs) - 1, max(0, int(math.ceil(q * len(xs))) - 1))
    return xs[i]


def _log(progress, msg):
    if progress:
        progress(msg)


def shape_for(workload, **override):
    """The tuned launch shape of `workload` with the caller's overrides (keys
    whose value is None or 0 keep the tuned value)."""
    sh = lxv.tuned_shape(workload)
    sh.update({k: v for k, v in override.items() if v})
    return sh


def open_multi(sh, **kw):
    return lxv.Multi(batch=sh["batch"], waves_per_cu=sh["waves_per_cu"], layout=sh["layout"], walk=sh["walk"],
                     lane_keys_log2=sh["lane_keys_log2"], **kw)


def batched(n=256, wall_s=60.0, budget_ms=0, seed=3, pool_bytes=8 << 30, lane_keys_log2=0, progress=print,
            shape=None):
    ov = {"step_ms": budget_ms, "lane_keys_log2": lane_keys_log2}
    ov.update(shape or {})
    sh = shape_for("batched", **ov)
    budget_ms = sh["step_ms"]
    rng = random.Random(seed)
    reqs = [Request(3_000_000 + i, draw_length(rng, BATCHED_LENGTHS)) for i in range(n)]
    m = open_multi(sh, pool_bytes=pool_bytes, pipelined=True)
    try:
        by_id = {}
        for i, r in enumerate(reqs):
            by_id[m.submit(r.start, r.ed_pub, r.prefix, tag=i)] = r
        t0 = time.perf_counter()
        last = t0
        found = 0
        while True:
            # budget 0 once the time is up: collect the step still in flight
            done_stepping = time.perf_counter() - t0 >= wall_s
            for res in m.step(0 if done_stepping else budget_ms):
                r = by_id[res["id"]]
                if res["status"] != 1:
                    continue
                if not r.recheck(res["g"], res["address"]):
                    raise SystemExit("search %d (step %d) failed its CPU recheck" % (r.seed, res["g"]))
                r.g, r.keys, r.done = res["g"], res["keys"], time.perf_counter() - t0
                found += 1
            now = time.perf_counter()
            st = m.stats()
            if now - last > 5:
                _log(progress, "batched: %.0f s, %d/%d found, %.3e keys, resident %d, queued %d"
                     % (now - t0, found, n, st["keys"], st["resident"], st["queued"]))
                last = now
            if (st["resident"] == 0 and st["queued"] == 0) or done_stepping:
                break
        wall = time.perf_counter() - t0
        st = m.stats()
    finally:
        m.close()
    per_len = {}
    for r in reqs:
        d = per_len.setdefault(r.digits, {"searches": 0, "found": 0, "mean_keys_found": 0.0})
        d["searches"] += 1
        if r.done is not None:
            d["found"] += 1
            d["mean_keys_found"] += r.keys
    for d in per_len.values():
        d["mean_keys_found"] = d["mean_keys_found"] / d["found"] if d["found"] else None
    return {"searches": n, "found": found, "rechecked": found, "wall_s": wall, "keys": st["keys"], "shape": sh,
            "busy_s": st["busy_s"], "steps": st["steps"], "keys_per_s": st["keys"] / wall,
            "kernel_keys_per_s": st["keys"] / st["busy_s"] if st["busy_s"] else None,
            "per_length": {str(k): v for k, v in sorted(per_len.items())}}


def arrivals(rate, duration_s, seed, table):
    rng = random.Random(seed)
    t, out, i = 0.0, [], 0
    while True:
        t += rng.expovariate(rate)
        if t >= duration_s:
            return out
        out.append((t, 4_000_000 + i, draw_length(rng, table)))
        i += 1


def _latency_report(reqs, end):
    done = [r.done - r.arrival for r in reqs if r.done is not None]
    open_ages = [end - r.arrival for r in reqs if r.done is None]
    censored = done + open_ages   # unfinished counted at their age: a lower bound
    per_len = {}
    for r in reqs:
        d = per_len.setdefault(r.digits, {"requests": 0, "done": [], "unfinished": 0})
        d["requests"] += 1
        if r.done is not None:
            d["done"].append(r.done - r.arrival)
        else:
            d["unfinished"] += 1
    table = {str(k): {"requests": v["requests"], "finished": len(v["done"]), "unfinished": v["unfinished"],
                      "p50_s": quantile(v["done"], .5), "p99_s": quantile(v["done"], .99)}
             for k, v in sorted(per_len.items())}
    return {"requests": len(reqs), "finished": len(done), "unfinished": len(open_ages),
            "p50_s": quantile(done, .5), "p99_s": quantile(done, .99),
            "p50_s_with_unfinished": quantile(censored, .5), "p99_s_with_unfinished": quantile(censored, .99),
            "unfinished_age_p50_s": quantile(open_ages, .5), "unfinished_age_max_s": max(open_ages) if open_ages else None,
            "per_length": table}


POLICIES = ("short-first", "short-first-oldest", "fifo")


def serving(rate=20.0, duration_s=60.0, policy="short-first-oldest", budget_ms=0, seed=4, lane_keys_log2=0,
            pool_bytes=32 << 30, max_searches=4096, progress=print, oldest_share_pct=20, shape=None, preempt=True):
    """The serving loop. Requests are admitted into lxv_multi in `policy`
    order: "short-first" (priority = digits: fewest expected keys, 16^digits,
    then arrival; strict, so under overload the longest requests starve),
    "short-first-oldest" (the same, but every step gives oldest_share_pct of
    its work to the longest-waiting resident request, which is also admitted
    first) or "fifo" (one priority: processor sharing in arrival order).
    The device-memory budget is 32 GiB of walk state (a 7-9 digit request
    holds 16 MiB: 2^18 virtual lanes of 64 B) and 4096 slots, about two
    thousand long requests, so under the 9x overload of 20 requests/s the
    pool fills only late in a 600 s run. With preempt (the default), a
    request that finds the HBM pool or the slots full suspends resident
    requests of a lower priority (the library keeps their progress and sets
    their lanes up again when they return), so
    long requests that wait for service cannot hold the memory that new short
    ones need. Each step walks about budget_ms of work over every resident request;
    time-to-hit is from arrival to the answer being on the host. Measured on
    MI355X over 60 s (8x32, B = 256, 16 waves/CU): 15/25/35/50/100 ms steps
    gave p50 0.053/0.050/0.060/0.076/0.117 s at 87/92/94/92/95% of the plain
    kernel's keys/s."""
    if policy not in POLICIES:
        raise ValueError("policy must be one of %s" % (POLICIES,))
    ov = {"step_ms": budget_ms, "lane_keys_log2": lane_keys_log2}
    ov.update(shape or {})
    sh = shape_for("serving", **ov)
    budget_ms, lane_keys_log2 = sh["step_ms"], sh["lane_keys_log2"]
    trace = arrivals(rate, duration_s, seed, SERVING_LENGTHS)
    reqs = []
    # One step at a time: pipelining the steps keeps the GPU busier (measured
    # over 60 s: 93.1% of the plain kernel's keys/s against 91.7%) but makes a
    # new request wait for the step already in flight (p50 0.112 s against
    # 0.076 s), and time-to-hit is what this workload is for.
    m = open_multi(sh, pool_bytes=pool_bytes, max_searches=max_searches, preempt=preempt,
                   oldest_share_pct=oldest_share_pct if policy == "short-first-oldest" else 0)
    try:
        by_id = {}
        t0 = time.perf_counter()
        nxt, last, depth = 0, t0, []

        def take(results):
            for res in results:
                r = by_id.pop(res["id"])
                if res["status"] != 1:
                    continue
                if not r.recheck(res["g"], res["address"]):
                    raise SystemExit("request %d (step %d) failed its CPU recheck" % (r.seed, res["g"]))
                r.g, r.keys, r.done = res["g"], res["keys"], time.perf_counter() - t0

        while True:
            now = time.perf_counter() - t0
            if now >= duration_s:
                break
            while nxt < len(trace) and trace[nxt][0] <= now:
                t_arr, s, digits = trace[nxt]
                r = Request(s, digits, t_arr)
                reqs.append(r)
                # admission order inside the library: (priority, submission)
                prio = 0 if policy == "fifo" else digits
                by_id[m.submit(r.start, r.ed_pub, r.prefix, priority=prio)] = r
                nxt += 1
            st = m.stats()
            if st["resident"] == 0 and st["queued"] == 0:
                if nxt < len(trace):
                    time.sleep(max(0.0, min(trace[nxt][0] - now, budget_ms / 1e3)))
                continue
            take(m.step(budget_ms))
            now = time.perf_counter() - t0
            st = m.stats()
            if not depth or now - depth[-1][0] >= 1.0:
                depth.append((round(now, 2), st["queued"], st["resident"]))
            if time.perf_counter() - last > 5:
                last = time.perf_counter()
                fin = sum(1 for r in reqs if r.done is not None)
                _log(progress, "serving: %.0f s, %d arrived, %d finished, queued %d, resident %d"
                     % (now, len(reqs), fin, st["queued"], st["resident"]))
        end = time.perf_counter() - t0
        st = m.stats()
        suspensions = m.suspensions()
    finally:
        m.close()
    out = _latency_report(reqs, end)
    out.update({"policy": policy, "oldest_share_pct": oldest_share_pct if policy == "short-first-oldest" else 0,
                "preempt": preempt, "suspensions": suspensions, "pool_bytes": pool_bytes,
                "max_searches": max_searches,
                "shape": sh, "rate_per_s": rate, "duration_s": duration_s, "wall_s": end,
                "budget_ms": budget_ms, "lane_keys_log2": lane_keys_log2,
                "keys": st["keys"], "busy_s": st["busy_s"], "steps": st["steps"],
                "keys_per_s": st["keys"] / end, "rechecked": out["finished"],
                "queue_depth": [{"t_s": t, "queued": w, "resident": k} for t, w, k in depth[::5]]})
    return out


def one_by_one(rate=20.0, duration_s=60.0, seed=4, batch=0, waves_per_cu=0, progress=print):
    """The same arrivals through the plain kernel, one search at a time in
    arrival order, every lane on the current search; launches grow while a
    launch is short (src/gpu/mod.rs:146-174). Every request sets up all
    lanes anew, so the lanes are eight rounds of the resident waves launched
    one thread per lane rather than the default 64 rounds on the persistent
    grid: an eighth of the set-up per request."""
    trace = arrivals(rate, duration_s, seed, SERVING_LENGTHS)
    single = lxv.tuned_shape("single")
    batch, waves_per_cu = batch or single["batch"], waves_per_cu or single["waves_per_cu"]
    probe = lxv.Backend(batch=batch, waves_per_cu=waves_per_cu)
    grid, lanes = probe.walk_kernel()[1], probe.threads
    probe.close()
    resident = grid if grid < lanes else lanes // 8
    b = lxv.Backend(batch=batch, waves_per_cu=waves_per_cu, threads=8 * resident, resident_lanes=8 * resident)
    stride = lxv.lane_stride(b.threads)
    reqs, keys, busy = [], 0, 0.0
    try:
        per_iter = b.threads * b.batch
        t0 = time.perf_counter()
        last = t0
        for t_arr, s, digits in trace:
            now = time.perf_counter() - t0
            if now >= duration_s:
                break
            if t_arr > now:
                time.sleep(t_arr - now)
            r = Request(s, digits, t_arr)
            reqs.append(r)
            b.init_seeded(r.start, stride)
            iters = 1
            while time.perf_counter() - t0 < duration_s:
                t1 = time.perf_counter()
                hits, _ = b.launch(lxv.launch_args_for(r.ed_pub, r.prefix, iters=iters))
                took = time.perf_counter() - t1
                st = b.stats()
                keys += st.keys
                busy += st.kernel_ms / 1e3
                if hits:
                    t, n, addr = hits[0]
                    g = t * stride + n
                    if not r.recheck(g, addr):
                        raise SystemExit("request %d (lane %d, key %d) failed its CPU recheck" % (s, t, n))
                    r.g, r.done = g, time.perf_counter() - t0
                    break
                if took < lxv.TARGET_LAUNCH_S / 2 and iters < (1 << 16):
                    iters *= 2
            if time.perf_counter() - last > 5:
                last = time.perf_counter()
                _log(progress, "one-by-one: %.0f s, %d started, %d finished"
                     % (last - t0, len(reqs), sum(1 for x in reqs if x.done is not None)))
        end = time.perf_counter() - t0
    finally:
        b.close()
    # requests that arrived but never started are unfinished too
    for t_arr, s, digits in trace[len(reqs):]:
        if t_arr < end:
            r = Request.__new__(Request)
            r.seed, r.digits, r.arrival, r.done = s, digits, t_arr, None
            reqs.append(r)
    out = _latency_report(reqs, end)
    out.update({"policy": "one-by-one, arrival order, plain kernel B=%d" % batch, "wall_s": end, "keys": keys,
                "busy_s": busy, "keys_per_s": keys / end})
    return out
