"""Benchmark of the MI355X walk kernel on the headline workload.

Workload (the plain HIP kernel, single search on one MI355X): one seeded
search (start scalar, Ed25519 key, 7-hex-digit prefix from workload seed 1),
64 keys per batched inversion, one walk per lane over every CU, a budget of
2^36 keys split evenly over the timed steps. A step is one launch of the walk
kernel over every lane (inputs resident in HBM; launch and completion timed
with HIP events on the kernel's own stream), and every hit is re-derived on
the CPU before it counts.

Prints one JSON line: keys/s, the kernel's roofline against the measured
v_mad_u64_u32 issue peak, and the CPU oracle (the C restatement of the
reference's CPU walk) timed on this host's cores on a bounded sample.

    python bench.py [--steps K] [--warmup W] [--budget-log2 36] [--no-cpu]

Other workloads (one JSON line each, same keys where they apply):

    python bench.py --workload batched [--wall-s 60]
        256 concurrent seeded searches (4-10 hex digits) in the persistent
        kernel; value = aggregate keys/s, compared with the plain kernel on
        one search measured in the same process.
    python bench.py --workload serving [--duration-s 60] [--policy short-first|fifo] [--one-by-one]
        Poisson arrivals (20/s, 5-9 hex digits); p50/p99 time-to-hit from
        arrival, queue depth, unfinished requests; --one-by-one also replays
        the same arrivals through the plain kernel one search at a time.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lxmf-vanity_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import lxv  # noqa: E402

METRIC = "keys/s on 1 MI355X vs plain HIP kernel; p50/p99 time-to-hit, mixed prefixes"


def mad_work_per_key(batch, affine=False):
    """32x32->64 products per key on the 8x32-bit schoolbook count of the
    reference's walk: M(B) = 8 + 11/B multiplications and S(B) = 2 + 254/B
    squarings per key for B keys per inversion (64 M + 36 S products). This
    is the algorithmic work the roofline is priced in, whatever form computes
    the keys; the affine sub-walk form's own count is 6 + 11/B."""
    return 64 * ((6 if affine else 8) + 11 / batch) + 36 * (2 + 254 / batch)


def pair_mads_per_key(batch):
    """Products per key the pair walk executes: per batch of B keys, B/2
    forward and B backward multiplications of Montgomery's trick, 2M + 2S
    per pair of keys (B/2 - 1 pairs), 1M + 1S for the single key, 2M + 1S for
    the next centre and the inversion's 11M + 254S: M = 2.5 + 12/B and
    S = 1 + 254/B."""
    return 64 * (2.5 + 12 / batch) + 36 * (1 + 254 / batch)


def measure_mad_peak():
    """The chip's v_mad_u64_u32 issue peak: the mad probe (8 independent
    chains per lane, unrolled, lxv_debug_valu_rate) at 16 and 32 waves per
    CU, the higher rate in lane-level mads/s; with the clock the probe
    measured and the CU count."""
    best = None
    for wpc in (16, 32):
        rate, clk, _, cus = lxv.debug_valu_rate(7, wpc)
        if best is None or rate > best[0]:
            best = (rate, clk, cus, wpc)
    return best


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_cores():
    """(threads to use, CPUs visible, why). One thread per core of the lease,
    as the reference runs one search thread per core (available_parallelism,
    src/main.rs:474-477). A shared GPU host may lease a share of its CPUs per
    GPU and say so in OMP_NUM_THREADS; sched_getaffinity may show the whole
    host there, so the lease's figure wins when it is smaller."""
    visible = os.cpu_count() or 1
    try:
        visible = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    lease = os.environ.get("OMP_NUM_THREADS", "")
    if lease.isdigit() and 0 < int(lease) < visible:
        return int(lease), visible, "OMP_NUM_THREADS=%s: the host CPUs leased with this GPU" % lease
    return visible, visible, "every CPU in this process's affinity mask"


def cpu_baseline(seconds=12.0):
    """The oracle walk (the reference's CPU path restated in C, one thread per
    core as the reference runs, SHA-NI when present) on a bounded sample."""
    import oracle_lib as orc
    start, ed_seed, prefix = lxv.seeded_search(1, 7)
    ed_pub = orc.ed25519_pub(ed_seed)
    cores, visible, why = host_cores()
    # calibrate with a short run, then size the sample to `seconds`
    n = 1 << 18
    t0 = time.perf_counter()
    orc.scan_ranges(start, ed_pub, [(0, n)], prefix, threads=cores)
    rate = n / (time.perf_counter() - t0)
    n = max(1 << 20, int(rate * seconds))
    t0 = time.perf_counter()
    orc.scan_ranges(start, ed_pub, [(0, n)], prefix, threads=cores)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "keys/s", "cores": cores, "kind": "port",
            "sample": "%d consecutive keys of the same seeded search (7-hex prefix), %.1f s, SHA-NI=%s"
                      % (n, dt, orc.shani()),
            "per_core": n / dt / cores, "cpu_model": cpu_model(), "cpus_visible": visible, "cores_reason": why}


def split_iters(budget, per_iter, steps):
    """The key budget in whole launch iterations (per_iter keys each), spread
    over the timed steps so that exactly budget // per_iter iterations are
    walked (at least one per step): each step gets the floor or the ceiling
    of the mean. 2^36 keys at the default shape are 64 iterations of 2^30."""
    total = max(steps, budget // per_iter)
    return [total // steps + (1 if i < total % steps else 0) for i in range(steps)]


def keys_text(k):
    return "2^%d" % (k.bit_length() - 1) if k & (k - 1) == 0 else "%.4g" % k


def valu_issue_roofline(keys_per_s, insts_per_key, kname, busy_pct):
    """The walk's VALU issue rate (lane-level instructions per second: keys/s
    x VALU instructions per key from the PMC summary) against measured issue
    ceilings at the kernel's occupancy: the probe with the pair walk's own
    instruction mix (the peak), the SHA-256 round mix, and the fastest single
    instruction class measured; plus the shader clock the probes ran at."""
    waves = int(kname.rstrip(">").split(",")[-1]) if kname.endswith(">") else 4
    wpc = 4 * waves
    achieved = keys_per_s * insts_per_key
    walk_mix, clk, nominal, cus = lxv.debug_valu_rate(11, wpc)
    sha_mix = lxv.debug_valu_rate(8, wpc)[0]
    singles = {lxv.VALU_PROBES[op]: lxv.debug_valu_rate(op, wpc)[0] for op in (0, 1, 2, 3, 7)}
    best_name = max(singles, key=singles.get)
    per4 = cus * 64 * clk
    return {"valu_insts_per_key": insts_per_key, "achieved": achieved / 1e12, "peak": walk_mix / 1e12,
            "unit": "T lane-instr/s", "frac": achieved / walk_mix,
            "peak_source": "issue-rate probe with the pair walk's per-key instruction mix at %d waves/CU" % wpc,
            "busy_pct_measured": busy_pct, "cus": cus, "clock_hz_measured": clk, "clock_hz_nominal": nominal,
            "achieved_per_simd_4clk": achieved / per4, "walk_mix_per_simd_4clk": walk_mix / per4,
            "sha_mix_per_simd_4clk": sha_mix / per4, "frac_vs_sha_mix": achieved / sha_mix,
            "fastest_single_class": best_name, "fastest_single_per_simd_4clk": singles[best_name] / per4,
            "frac_vs_fastest_single": achieved / singles[best_name],
            "frac_vs_nominal_clock": achieved / (cus * 64 * nominal)}


def load_traffic():
    path = os.path.join(ROOT, "profiles", "walk_pmc.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def pmc_summary(out_dir, dest=os.path.join(ROOT, "profiles", "walk_pmc.json")):
    """profiles/walk_pmc.json from three PMC runs of this script (one launch
    each, --copy-probe-mib 1024) under out_dir/pmc_sq, pmc_fetch and
    pmc_write, with the first run's log as out_dir/pmc_sq.log: VALU
    instructions and HBM bytes per key of the walk kernel the run launched,
    FETCH_SIZE and WRITE_SIZE scaled by what they read for the 1 GiB copy."""
    import csv
    import glob

    def find(pattern):
        g = glob.glob(os.path.join(out_dir, pattern), recursive=True)
        if not g:
            raise SystemExit("missing %s under %s" % (pattern, out_dir))
        return g[0]

    line = [ln for ln in open(os.path.join(out_dir, "pmc_sq.log")) if ln.startswith('{"metric"')][-1]
    run = json.loads(line)
    kernel, keys = run["config"]["kernel"], run["keys"]

    def counters(path):
        walk, other = {}, {}
        for r in csv.DictReader(open(path)):
            tgt = walk if kernel in r["Kernel_Name"] else other.setdefault(r["Kernel_Name"], {})
            tgt[r["Counter_Name"]] = tgt.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        return walk, other

    sq, _ = counters(find("pmc_sq/**/*_counter_collection.csv"))
    fe, fo = counters(find("pmc_fetch/**/*_counter_collection.csv"))
    wr, wo = counters(find("pmc_write/**/*_counter_collection.csv"))
    copy = [k for k in fo if "copy_kernel" in k][0]
    out = {"kernel": kernel, "keys_per_launch": keys,
           "valu_insts_per_key": sq["SQ_INSTS_VALU"] * 64 / keys,
           "salu_insts_per_key": sq["SQ_INSTS_SALU"] * 64 / keys,
           "vmem_rd_insts_per_key": sq["SQ_INSTS_VMEM_RD"] * 64 / keys,
           "vmem_wr_insts_per_key": sq["SQ_INSTS_VMEM_WR"] * 64 / keys,
           "valu_busy_pct": 100.0 * sq["SQ_ACTIVE_INST_VALU"] * 4 / (1024 * sq["GRBM_GUI_ACTIVE"] / 8),
           "fetch_size_kib": fe["FETCH_SIZE"], "write_size_kib": wr["WRITE_SIZE"],
           "copy_probe_bytes": 1 << 30, "copy_fetch_kib": fo[copy]["FETCH_SIZE"],
           "copy_write_kib": wo[copy]["WRITE_SIZE"]}
    out["fetch_scale"] = (1 << 30) / 1024 / out["copy_fetch_kib"]
    out["write_scale"] = (1 << 30) / 1024 / out["copy_write_kib"]
    out["hbm_read_bytes_per_key"] = out["fetch_size_kib"] * 1024 * out["fetch_scale"] / keys
    out["hbm_write_bytes_per_key"] = out["write_size_kib"] * 1024 * out["write_scale"] / keys
    out["hbm_bytes_per_key"] = out["hbm_read_bytes_per_key"] + out["hbm_write_bytes_per_key"]
    out["note"] = ("FETCH_SIZE/WRITE_SIZE of the walk kernel scaled by the factor that makes the same counters read "
                   "a 1 GiB plain copy as 1 GiB; per-key figure x keys per launch = traffic per launch")
    with open(dest, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))
    return out


def plain_rate(batch=0, waves_per_cu=0, budget_log2=33):
    """keys/s of the plain kernel on one search (workload seed 2, 8 hex
    digits) at the tuned single-search shape, kernel time only: the yardstick
    for the many-search workloads."""
    single = lxv.tuned_shape("single")
    batch, waves_per_cu = batch or single["batch"], waves_per_cu or single["waves_per_cu"]
    start, ed_seed, prefix = lxv.seeded_search(2, 8)
    s = lxv.GpuSearch(start=start, ed_seed=ed_seed, batch=batch, waves_per_cu=waves_per_cu)
    try:
        b = s.backend
        iters = max(1, (1 << budget_log2) // (b.threads * b.batch))
        a = lxv.launch_args_for(s.ed_pub, prefix, iters=iters)
        b.launch(lxv.launch_args_for(s.ed_pub, prefix, iters=1))
        keys, ms = 0, 0.0
        for _ in range(2):
            b.launch(a)
            st = b.stats()
            keys += st.keys
            ms += st.kernel_ms
        return keys / (ms / 1e3)
    finally:
        s.close()


def batched_pass(args, plain_keys_per_s):
    """The batched mix in the same run: 256 seeded searches (4-10 hex digits)
    in the persistent kernel at the tuned batched shape for --batched-s
    seconds; aggregate keys/s against this run's plain-kernel keys/s and
    against the plain kernel at the batched kernel's own shape."""
    import workloads
    log = lambda msg: print(msg, file=sys.stderr, flush=True)  # noqa: E731
    r = workloads.batched(wall_s=args.batched_s, progress=log)
    sh = r["shape"]
    same_shape = plain_rate(batch=sh["batch"], waves_per_cu=sh["waves_per_cu"])
    return {"workload": "batched: 256 seeded searches, 4-10 hex digits (P=.10/.15/.20/.20/.15/.12/.08), "
                        "%.0f s wall" % args.batched_s,
            "keys_per_s": r["keys_per_s"], "kernel_keys_per_s": r["kernel_keys_per_s"], "found": r["found"],
            "rechecked": r["rechecked"], "shape": sh, "keys_per_s_vs_plain": r["keys_per_s"] / plain_keys_per_s,
            "plain_same_shape_keys_per_s": same_shape, "keys_per_s_vs_plain_same_shape": r["keys_per_s"] / same_shape}


def serving_pass(args, plain_keys_per_s):
    """The metric's second half, in the same run: the serving loop over the
    Poisson request stream (20/s, 5-9 hex digits) for --serving-s seconds,
    p50/p99 time-to-hit from arrival over the finished requests and over all
    of them (unfinished counted at their age), and the same arrivals replayed
    one by one on the plain kernel for --one-by-one-s seconds."""
    import workloads
    log = lambda msg: print(msg, file=sys.stderr, flush=True)  # noqa: E731
    r = workloads.serving(duration_s=args.serving_s, policy=args.policy, oldest_share_pct=args.oldest_share_pct,
                          preempt=not args.no_preempt, progress=log)
    keep = ("requests", "finished", "unfinished", "p50_s", "p99_s", "p50_s_with_unfinished", "p99_s_with_unfinished",
            "policy", "oldest_share_pct", "preempt", "suspensions", "pool_bytes", "max_searches", "shape", "keys_per_s",
            "busy_s", "steps", "rechecked", "per_length")
    out = {k: r[k] for k in keep}
    out["workload"] = ("serving: Poisson 20/s for %.0f s, 5-9 hex digits (P=.30/.30/.20/.15/.05), a fresh seeded "
                       "identity per request, admission %s" % (args.serving_s, args.policy))
    out["keys_per_s_vs_plain"] = r["keys_per_s"] / plain_keys_per_s
    out["queue_depth_end"] = r["queue_depth"][-1] if r["queue_depth"] else None
    if args.one_by_one_s > 0:
        o = workloads.one_by_one(duration_s=args.one_by_one_s, progress=log)
        out["one_by_one"] = {k: o[k] for k in ("requests", "finished", "unfinished", "p50_s", "p99_s",
                                               "p50_s_with_unfinished", "p99_s_with_unfinished", "keys_per_s",
                                               "policy")}
    return out


def many_search_main(args):
    import workloads
    log = lambda msg: print(msg, file=sys.stderr, flush=True)  # noqa: E731
    base = plain_rate()
    single = lxv.tuned_shape("single")
    log("plain kernel, one search (B=%d, %d waves/CU): %.3e keys/s" % (single["batch"], single["waves_per_cu"], base))
    common = {"n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
              "dtype": "u32 (exact integer limbs)",
              "data": "synthetic (seeded start scalars, Ed25519 keys and prefixes)", "plain_keys_per_s": base}
    if args.workload == "batched":
        kw = {"lane_keys_log2": args.lane_keys_log2} if args.lane_keys_log2 else {}
        if args.budget_ms:
            kw["budget_ms"] = args.budget_ms
        r = workloads.batched(wall_s=args.wall_s, progress=log, **kw)
        out = {"metric": "aggregate keys/s, 256 concurrent searches", "value": r["keys_per_s"], "unit": "keys/s",
               "steps": r["steps"], "warmup": 0, "ms_per_step": 1e3 * r["wall_s"] / max(1, r["steps"]),
               "config": {"workload": "batched: 256 seeded searches, 4-10 hex digits "
                                      "(P=.10/.15/.20/.20/.15/.12/.08), %.0f s wall" % args.wall_s,
                          "shape": r["shape"]},
               "vs_plain": r["keys_per_s"] / base, "kernel_vs_plain": (r["kernel_keys_per_s"] or 0) / base}
        out.update(common)
        out["result"] = r
    elif args.workload == "one-by-one":
        r = workloads.one_by_one(duration_s=args.duration_s, progress=log)
        out = {"metric": "p50/p99 time-to-hit, mixed prefixes (one search at a time, plain kernel)",
               "value": r["keys_per_s"], "unit": "keys/s", "steps": 0, "warmup": 0, "ms_per_step": None,
               "p50_s": r["p50_s"], "p99_s": r["p99_s"], "unfinished": r["unfinished"],
               "config": {"workload": "serving arrivals: Poisson 20/s for %.0f s, 5-9 hex digits "
                                      "(P=.30/.30/.20/.15/.05), one search at a time in arrival order"
                                      % args.duration_s, "shape": lxv.tuned_shape("single")}}
        out.update(common)
        out["result"] = r
    else:
        kw = {"lane_keys_log2": args.lane_keys_log2} if args.lane_keys_log2 else {}
        if args.budget_ms:
            kw["budget_ms"] = args.budget_ms
        r = workloads.serving(duration_s=args.duration_s, policy=args.policy, progress=log,
                              oldest_share_pct=args.oldest_share_pct, preempt=not args.no_preempt, **kw)
        out = {"metric": "p50/p99 time-to-hit, mixed prefixes (serving)", "value": r["keys_per_s"],
               "unit": "keys/s", "steps": r["steps"], "warmup": 0,
               "ms_per_step": 1e3 * r["wall_s"] / max(1, r["steps"]),
               "p50_s": r["p50_s"], "p99_s": r["p99_s"], "unfinished": r["unfinished"],
               "config": {"workload": "serving: Poisson 20/s for %.0f s, 5-9 hex digits (P=.30/.30/.20/.15/.05), "
                                      "admission %s%s, %d ms steps, 2^%d keys per lane"
                                      % (args.duration_s, args.policy, "" if args.no_preempt else " with preemption",
                                         r["budget_ms"], r["lane_keys_log2"]),
                          "shape": r["shape"]}}
        out.update(common)
        out["result"] = r
        if args.one_by_one:
            out["one_by_one"] = workloads.one_by_one(duration_s=args.duration_s, progress=log)
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("plain", "batched", "serving", "one-by-one"), default="plain",
                    help="one-by-one: the serving arrivals through the plain kernel alone, one search at a time")
    ap.add_argument("--wall-s", type=float, default=60.0)
    ap.add_argument("--duration-s", type=float, default=60.0)
    ap.add_argument("--policy", choices=("short-first", "short-first-oldest", "fifo"), default="short-first-oldest")
    ap.add_argument("--oldest-share-pct", type=int, default=20,
                    help="short-first-oldest: share of every step for the longest-waiting request")
    ap.add_argument("--no-preempt", action="store_true",
                    help="serving: admit only into free room (no suspension of lower-priority requests)")
    ap.add_argument("--batched-s", type=float, default=20.0,
                    help="plain workload: also run the 256-search batched mix for this long (0: skip)")
    ap.add_argument("--serving-s", type=float, default=60.0,
                    help="plain workload: also run the serving loop for this long and report time-to-hit (0: skip)")
    ap.add_argument("--one-by-one-s", type=float, default=60.0,
                    help="plain workload: also replay the serving arrivals one by one on the plain kernel (0: skip)")
    ap.add_argument("--one-by-one", action="store_true")
    ap.add_argument("--lane-keys-log2", type=int, default=0,
                    help="serving/batched: expected keys per virtual lane, log2 (0: the workload's default)")
    ap.add_argument("--budget-ms", type=int, default=0, help="serving/batched: step budget (0: the workload's default)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--budget-log2", type=int, default=36)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--layout", type=int, default=lxv.LAYOUT_8X32)
    ap.add_argument("--walk", choices=("pairs", "projective"), default="pairs",
                    help="how a lane computes its batch (same keys): around an affine centre, or the reference's "
                         "chained x-only differential additions")
    ap.add_argument("--subwalks", type=int, default=0,
                    help="K affine sub-walks per lane, inverted across the K (0: one walk per lane)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run one plain copy kernel of this many MiB (counter calibration)")
    ap.add_argument("--pmc-summary", metavar="DIR", default=None,
                    help="write profiles/walk_pmc.json from the three PMC runs under DIR, then exit (no GPU)")
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(args.pmc_summary)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the library walks on a single MI355X and has no multi-GPU path "
              "(a search would shard by disjoint lane ranges with no exchange).", file=sys.stderr)
        sys.exit(2)
    if args.workload != "plain":
        many_search_main(args)
        return

    start, ed_seed, prefix = lxv.seeded_search(1, 7)
    ed_pub = lxv.ed25519_public(ed_seed)
    walk = lxv.WALK_PAIRS if args.walk == "pairs" else lxv.WALK_PROJECTIVE
    pairs = walk == lxv.WALK_PAIRS and not args.subwalks
    search = lxv.GpuSearch(start=start, ed_seed=ed_seed, batch=args.batch, waves_per_cu=args.waves_per_cu,
                           layout=args.layout, subwalks=args.subwalks, walk=walk)
    b = search.backend
    per_iter = b.threads * b.batch
    budget = 1 << args.budget_log2
    step_iters = split_iters(budget, per_iter, args.steps)
    total_iters = sum(step_iters)
    iters = min(step_iters)   # warm-up launches: the common size, so profiles average close to the timed ones
    mean_iters = total_iters / args.steps
    a = lxv.launch_args_for(ed_pub, prefix, iters=iters, max_hits=lxv.MAX_HITS)

    for _ in range(args.warmup):
        b.launch(a)

    kernel_ms, keys, wall, hits = [], 0, 0.0, []
    t0 = time.perf_counter()
    for it in step_iters:
        a.iters = it
        h, _ = b.launch(a)
        st = b.stats()
        kernel_ms.append(st.kernel_ms)
        keys += st.keys
        hits += h
    wall = time.perf_counter() - t0
    # every hit re-derived from scratch on the CPU (src/gpu/mod.rs:157-160)
    rechecked = 0
    for t, n, addr in hits:
        ident = search.identity(t, n)
        if lxv.address_of_identity(ident) != addr or not addr.hex().startswith(prefix):
            raise SystemExit("hit (lane %d, key %d) failed its CPU recheck" % (t, n))
        rechecked += 1
    gpu_s = sum(kernel_ms) / 1e3
    value = keys / gpu_s

    mad_peak, clock, cus, mad_wpc = measure_mad_peak()
    kpi = args.subwalks or b.batch  # keys per inversion
    work = mad_work_per_key(kpi, affine=bool(args.subwalks))
    achieved = value * work
    roofline = {"bound": "valu-imad", "achieved": achieved / 1e12, "peak": (mad_peak or 0) / 1e12,
                "unit": "T mad/s", "frac": (achieved / mad_peak) if mad_peak else None,
                "traffic": None, "mads_per_key": work,
                "peak_source": "v_mad_u64_u32 issue-rate probe, 8 chains per lane, best of 16 and 32 waves/CU "
                               "(%d), %.3f GHz measured" % (mad_wpc, clock / 1e9)}
    if pairs:
        roofline["mads_per_key_executed"] = pair_mads_per_key(kpi)
    # HBM: the algorithmic bytes per key are the batch arrays (a_i and Z_i,
    # written in the forward pass and read back in the backward pass) plus
    # the lane state read and written once per launch; in the sub-walk form,
    # the affine state (u_n, u_{n-1}) read and u_{n+1} written every key.
    limbs = 10 if args.layout == lxv.LAYOUT_10X26 else 8
    fe_bytes = 4 * limbs
    if args.subwalks:
        alg_bytes_per_key = 2 * 2 * fe_bytes + 3 * 32
    elif pairs:
        # B/2 prefix products written and read back per batch of B keys, plus
        # the centre (x, y) read and written once per launch
        alg_bytes_per_key = 2 * fe_bytes / 2 + 2 * 2 * fe_bytes / (mean_iters * b.batch)
    else:
        alg_bytes_per_key = 2 * 2 * fe_bytes + 2 * 4 * fe_bytes / (mean_iters * b.batch)
    roofline["hbm"] = {"algorithmic_bytes_per_key": alg_bytes_per_key,
                       "achieved_GBps": value * alg_bytes_per_key / 1e9, "peak_GBps": 8000.0,
                       "frac": value * alg_bytes_per_key / 8e12}
    pmc = load_traffic()
    kname, grid_lanes = b.walk_kernel()
    if pmc and kname in pmc.get("kernel", "") and "hbm_bytes_per_key" in pmc:
        # bytes per launch of this run = measured bytes per key x keys per launch
        roofline["traffic"] = pmc["hbm_bytes_per_key"] * keys / len(step_iters)
        roofline["traffic_source"] = "profiles/walk_pmc.json (FETCH_SIZE/WRITE_SIZE, copy-calibrated, same kernel)"
        v = pmc.get("valu_insts_per_key")
        if v:
            roofline["valu_issue"] = valu_issue_roofline(value, v, kname, pmc.get("valu_busy_pct"))
    if mad_peak and "valu_issue" in roofline and roofline["valu_issue"].get("clock_hz_measured"):
        # the mad probe against one wave64 instruction per SIMD per 4 clocks
        # at the clock the probes measured, not the nominal one
        vi = roofline["valu_issue"]
        roofline["mad_probe_per_simd_4clk"] = mad_peak / (vi["cus"] * 64 * vi["clock_hz_measured"])

    out = {
        "metric": METRIC, "value": value, "unit": "keys/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": sum(kernel_ms) / len(kernel_ms), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32 (exact integer limbs)",
        "data": "synthetic (seeded start scalar, Ed25519 key and prefix)",
        "config": {"workload": ("plain HIP walk (%s form), single search, 7-hex prefix, B=%d, %d keys (%s) in %d "
                                "launches of %s iterations" % (args.walk, b.batch, keys, keys_text(keys), args.steps,
                                                               "/".join(str(x) for x in sorted(set(step_iters)))))
                   if not args.subwalks else
                               ("sub-walk form: %d affine walks per lane inverted together, single search, 7-hex "
                                "prefix, 2^%d-key budget" % (args.subwalks, args.budget_log2)),
                   "threads": b.threads, "kernel": kname, "grid_lanes": grid_lanes,
                   "keys_per_inversion": kpi, "iters_per_step": step_iters,
                   "prefix": prefix, "layout": "10x25.5" if args.layout == lxv.LAYOUT_10X26 else "8x32",
                   "device": b.describe().strip()},
        "keys": keys, "wall_s": wall, "hits": len(hits), "hits_rechecked": rechecked,
        "roofline": roofline,
        "reference_gb10_keys_per_s": 630e6,
    }
    if args.copy_probe_mib:
        import ctypes
        L = lxv.lib()
        L.lxv_debug_copy.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
        ms = ctypes.c_double(0)
        L.lxv_debug_copy(0, args.copy_probe_mib << 20, ctypes.byref(ms))
        out["copy_probe"] = {"bytes": args.copy_probe_mib << 20, "ms": ms.value}
    search.close()
    if args.batched_s > 0:
        out["batched"] = batched_pass(args, value)
    if args.serving_s > 0:
        out["time_to_hit"] = serving_pass(args, value)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
