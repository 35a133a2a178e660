"""Autotuner of the walk kernels for three workloads: a single search on the
plain kernel (keys per inversion x waves per CU x limb layout x walk form),
the batched mix of 256 searches, and the serving stream (the batched kernel's
keys per inversion x waves per CU x layout, and for serving the step budget).
The picks go to tuned.json's "picks"; the library is built with them
(gen_tuned.py), so they are the default shapes of lxv_open, lxv_multi_open,
the command-line tool and the serving loop.

Each point opens the backend with that shape at the library's own launch
shape (--shape grid, the default): the pair walk on a persistent grid of
exactly `waves per CU` resident waves per CU taking 64 groups of 64 lanes
each, the projective walk one thread per lane over 8 rounds of that many
waves. --shape round launches one resident round (lanes = CUs x waves per CU
x 64, one thread per lane) instead. It sets every lane up on the GPU
and walks a fixed budget of keys of one seeded search (an 8-hex-digit prefix
from workload seed 2), timed with HIP events on the kernel's stream; every hit
is re-derived on the CPU. Per point it records keys/s, the kernel's VGPRs and
scratch bytes per lane (from its code object), and, when a PMC pass has been
merged, VALU busy %. The fastest point is the pick; the table goes to JSON.

    python autotune.py [--budget-log2 34] [--out tuned.json]
    # VALU busy %: a second, short pass under the profiler, then a merge
    rocprofv3 --pmc SQ_ACTIVE_INST_VALU SQ_INSTS_VALU GRBM_GUI_ACTIVE --output-format csv \\
        -d OUT -o pmc -- python3 autotune.py --budget-log2 28 --out /dev/null
    python autotune.py --merge-pmc OUT/pmc_counter_collection.csv --out tuned.json
    # separate sweeps (say one per walk form) joined, picked anew and merged
    python autotune.py --join a.json b.json [--merge-pmc CSV] --out tuned.json
    # the batched mix and the serving stream, each point run for --wall-s
    # seconds on its workload; picks merged into --out
    python autotune.py --mix batched [--wall-s 20] --out tuned.json
    python autotune.py --mix serving [--wall-s 20] --out tuned.json
    # a partial sweep (here 10 ms steps at 8 waves/CU) merged into the
    # earlier one, the pick made over all points
    python autotune.py --mix serving --steps-ms 10 --waves-per-cu 8 --extend

Mix picks: batched, the highest aggregate keys/s; serving, the lowest p50
time-to-hit over all requests (unfinished counted at their age) among the
points within 10% of the best keys/s, then the highest keys/s.

VALU busy % = SQ_ACTIVE_INST_VALU x 4 / (SIMDs x GRBM_GUI_ACTIVE / XCDs): the
share of SIMD cycles issuing VALU work (a wave64 instruction occupies a
16-lane SIMD for 4 cycles; GRBM_GUI_ACTIVE is summed over the 8 XCDs).
"""

import argparse
import csv
import ctypes
import functools
import json
import os
import re
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import lxv  # noqa: E402

BATCHES = (16, 32, 64, 128, 256)
WAVES = (2, 4, 8, 16)
LAYOUTS = ((lxv.LAYOUT_8X32, "8x32"), (lxv.LAYOUT_10X26, "10x25.5"))
WALKS = ((lxv.WALK_PAIRS, "pairs"), (lxv.WALK_PROJECTIVE, "projective"))
SIMDS_PER_CU = 4
XCDS = 8


def kernel_attrs(layout, batch, wpc, walk):
    L = lxv.lib()
    L.lxv_debug_kernel_attrs.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32] + \
        [ctypes.POINTER(ctypes.c_uint32)] * 4
    v, s, sc, w = (ctypes.c_uint32() for _ in range(4))
    rc = L.lxv_debug_kernel_attrs(layout, batch, wpc, walk, ctypes.byref(v), ctypes.byref(s), ctypes.byref(sc),
                                  ctypes.byref(w))
    if rc != 0:
        return None
    return {"vgprs": v.value, "scratch_bytes": sc.value, "waves_per_simd": w.value}


@functools.lru_cache(maxsize=None)
def device_cus():
    b = lxv.Backend(threads=64)
    try:
        return int(re.search(r"(\d+) CUs", b.describe()).group(1))
    finally:
        b.close()


def run_point(layout, batch, wpc, walk, budget, start, ed_seed, prefix, shape="grid"):
    # one resident round = CUs x waves per CU x 64 lanes: "waves per CU" is
    # the occupancy of the point in either shape
    resident = device_cus() * wpc * 64
    if shape == "round":
        lanes, grid = resident, 0
    elif walk == lxv.WALK_PAIRS:
        lanes, grid = 64 * resident, resident
    else:
        lanes, grid = 8 * resident, 0
    s = lxv.GpuSearch(start=start, ed_seed=ed_seed, batch=batch, waves_per_cu=wpc, layout=layout, walk=walk,
                      threads=lanes, resident_lanes=grid)
    try:
        b = s.backend
        per_iter = b.threads * b.batch
        iters = max(1, budget // (per_iter * 2))
        a = lxv.launch_args_for(s.ed_pub, prefix, iters=iters)
        b.launch(lxv.launch_args_for(s.ed_pub, prefix, iters=1))   # warm-up
        ms, keys, hits = 0.0, 0, []
        for _ in range(2):
            h, _ = b.launch(a)
            st = b.stats()
            ms += st.kernel_ms
            keys += st.keys
            hits += h
        for t, n, addr in hits:
            if lxv.address_of_identity(s.identity(t, n)) != addr or not addr.hex().startswith(prefix):
                raise SystemExit("autotune: hit (lane %d, key %d) failed its CPU recheck" % (t, n))
        kname, grid_lanes = b.walk_kernel()
        return {"lanes": b.threads, "grid_lanes": grid_lanes, "cus": resident // (wpc * 64), "kernel": kname,
                "keys": keys, "kernel_ms": ms,
                "keys_per_s": keys / (ms / 1e3), "hits": len(hits), "device": b.describe().strip()}
    finally:
        s.close()


def sweep(budget, walks=WALKS, shape="grid"):
    start, ed_seed, prefix = lxv.seeded_search(2, 8)
    points = []
    for walk, wname in walks:
        for layout, lname in LAYOUTS:
            for batch in BATCHES:
                for wpc in WAVES:
                    t0 = time.time()
                    r = run_point(layout, batch, wpc, walk, budget, start, ed_seed, prefix, shape)
                    r.update({"walk": wname, "walk_id": walk, "layout": lname, "layout_id": layout, "batch": batch,
                              "waves_per_cu": wpc})
                    attrs = kernel_attrs(layout, batch, wpc, walk)
                    if attrs:
                        r.update(attrs)
                    points.append(r)
                    print("%-10s %-8s B=%-3d waves/CU=%-2d  %.3e keys/s  vgprs=%s scratch=%s  (%.1f s)"
                          % (wname, lname, batch, wpc, r["keys_per_s"], r.get("vgprs"), r.get("scratch_bytes"),
                             time.time() - t0), flush=True)
    return {"workload": "autotune sweep: single seeded search (seed 2), 8-hex prefix %s, 2^%d keys per point, "
                        "%s launch shape" % (prefix, budget.bit_length() - 1, shape),
            "points": points}


def kernel_key(name, grid):
    # the kernel's name (walk_pair_queue_kernel<lxv::Field8, 64, 4>, ...) and
    # its grid identify a point
    for kname, wname in (("walk_pair_queue_kernel<", "pairs"), ("walk_pair_kernel<", "pairs"),
                         ("walk_kernel<", "projective")):
        if kname in name:
            inner = name.split(kname, 1)[1].split(">", 1)[0]
            field, batch, waves = [x.strip() for x in inner.split(",")]
            return (wname, "8x32" if "Field8" in field else "10x25.5", int(batch), int(waves), int(grid))
    return None


def pmc_pass_keys(point, budget_log2):
    """Keys the profiled pass (run_point at 2^budget_log2) walked at a point:
    one warm-up launch of one iteration, then two of `iters`."""
    per_iter = point["lanes"] * point["batch"]
    iters = max(1, (1 << budget_log2) // (per_iter * 2))
    return per_iter * (1 + 2 * iters)


def merge_pmc(table, csv_path, budget_log2=28):
    acc = {}
    for r in csv.DictReader(open(csv_path)):
        k = kernel_key(r["Kernel_Name"], r["Grid_Size"])
        if k is None:
            continue
        d = acc.setdefault(k, {})
        d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    for p in table["points"]:
        grid = p.get("grid_lanes", p["lanes"])
        k = (p.get("walk", "projective"), p["layout"], p["batch"], p.get("waves_per_simd", 0), grid)
        d = acc.get(k)
        if not d or not d.get("GRBM_GUI_ACTIVE"):
            continue
        cus = p.get("cus") or p["lanes"] // (p["waves_per_cu"] * 64)
        cycles = d["GRBM_GUI_ACTIVE"] / XCDS
        p["valu_busy_pct"] = 100.0 * d["SQ_ACTIVE_INST_VALU"] * 4 / (cus * SIMDS_PER_CU * cycles)
        if "SQ_INSTS_VALU" in d:
            # wave instructions x 64 lanes over the keys of the profiled pass
            p["valu_insts_per_key"] = d["SQ_INSTS_VALU"] * 64 / pmc_pass_keys(p, budget_log2)
            p.pop("valu_insts_per_wave_key", None)
    return table


def pick(table):
    best = max(table["points"], key=lambda p: p["keys_per_s"])
    by_batch = {}
    for p in table["points"]:
        if p["batch"] not in by_batch or p["keys_per_s"] > by_batch[p["batch"]]["keys_per_s"]:
            by_batch[p["batch"]] = p
    table["pick"] = {k: best[k] for k in ("walk", "walk_id", "layout", "layout_id", "batch", "waves_per_cu",
                                          "keys_per_s")}
    table["pick_by_batch"] = {str(b): {k: p[k] for k in ("walk", "walk_id", "layout", "layout_id", "waves_per_cu",
                                                         "keys_per_s")}
                              for b, p in sorted(by_batch.items())}
    return table


MIX_SHAPES = ([dict(walk=lxv.WALK_PAIRS, layout=lxv.LAYOUT_8X32, batch=b, waves_per_cu=w)
               for b in (64, 128, 256) for w in (8, 16)] +
              [dict(walk=lxv.WALK_PAIRS, layout=lxv.LAYOUT_10X26, batch=256, waves_per_cu=w) for w in (8, 16)])
SERVING_STEPS_MS = (15, 25, 50)
# expected keys per virtual lane (log2): fewer keys per lane means more lanes
# per search and shorter units of work
BATCHED_LANE_KEYS_LOG2 = (10, 12, 14)
SERVING_LANE_KEYS_LOG2 = (8, 10)


def mix_sweep(kind, wall_s, progress=print, steps_ms=None, lane_keys_log2=None, waves_per_cu=None):
    """Runs the batched mix or the serving stream at every batched shape and
    every expected keys per virtual lane (and, for serving, every step budget)
    for wall_s seconds each. steps_ms, lane_keys_log2 and waves_per_cu, when
    given, replace the default lists (a partial sweep, for --extend)."""
    import workloads
    points = []
    steps = tuple(steps_ms or SERVING_STEPS_MS)
    lks = tuple(lane_keys_log2 or (SERVING_LANE_KEYS_LOG2 if kind == "serving" else BATCHED_LANE_KEYS_LOG2))
    for sh in MIX_SHAPES:
        if waves_per_cu and sh["waves_per_cu"] not in waves_per_cu:
            continue
        if kind == "serving":
            variants = [(step, lk) for step in steps for lk in lks]
        else:
            variants = [(0, lk) for lk in lks]
        for step_ms, lk in variants:
            shape = {"batch": sh["batch"], "waves_per_cu": sh["waves_per_cu"], "layout": sh["layout"],
                     "walk": sh["walk"], "lane_keys_log2": lk}
            t0 = time.time()
            if kind == "batched":
                r = workloads.batched(wall_s=wall_s, shape=shape, progress=None)
                pt = {"keys_per_s": r["keys_per_s"], "found": r["found"]}
            else:
                shape["step_ms"] = step_ms
                r = workloads.serving(duration_s=wall_s, shape=shape, progress=None)
                pt = {k: r[k] for k in ("keys_per_s", "p50_s", "p99_s", "p50_s_with_unfinished",
                                        "p99_s_with_unfinished", "finished", "unfinished", "policy")}
            pt.update({"batch": sh["batch"], "waves_per_cu": sh["waves_per_cu"], "layout_id": sh["layout"],
                       "walk_id": sh["walk"], "layout": "8x32" if sh["layout"] == lxv.LAYOUT_8X32 else "10x25.5",
                       "step_ms": r["shape"]["step_ms"], "lane_keys_log2": r["shape"]["lane_keys_log2"]})
            points.append(pt)
            if progress:
                progress("%s %-7s B=%-3d waves/CU=%-2d step %3d ms 2^%d keys/lane: %.3e keys/s%s  (%.0f s)" % (
                    kind, pt["layout"], pt["batch"], pt["waves_per_cu"], pt["step_ms"], pt["lane_keys_log2"],
                    pt["keys_per_s"],
                    "" if kind == "batched" else ", p50 %.3f s (all %.3f s)" % (pt["p50_s"] or -1,
                                                                              pt["p50_s_with_unfinished"] or -1),
                    time.time() - t0))
    return points


def merge_points(old, new):
    """The points of an earlier mix sweep with those of a newer (partial) one;
    a newer point replaces an older one of the same shape, step budget and
    keys per lane."""
    def key(p):
        return (p["batch"], p["waves_per_cu"], p["layout_id"], p["walk_id"], p["step_ms"], p["lane_keys_log2"])
    fresh = {key(p) for p in new}
    return [p for p in old if key(p) not in fresh] + list(new)


def mix_pick(kind, points):
    if kind == "batched":
        best = max(points, key=lambda p: p["keys_per_s"])
    else:
        top = max(p["keys_per_s"] for p in points)
        near = [p for p in points if p["keys_per_s"] >= 0.9 * top and p["p50_s_with_unfinished"] is not None]
        best = min(near, key=lambda p: (p["p50_s_with_unfinished"], -p["keys_per_s"]))
    keys = ("batch", "waves_per_cu", "layout_id", "walk_id", "layout", "step_ms", "lane_keys_log2", "keys_per_s")
    return {k: best[k] for k in keys}


def load(path=os.path.join(HERE, "tuned.json")):
    """The tuned table, or None."""
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--budget-log2", type=int, default=34)
    ap.add_argument("--out", default=os.path.join(HERE, "tuned.json"))
    ap.add_argument("--merge-pmc", default=None, help="PMC csv of a profiled pass to merge into --out")
    ap.add_argument("--pmc-budget-log2", type=int, default=28, help="the --budget-log2 of the profiled pass")
    ap.add_argument("--mix", choices=("batched", "serving"), default=None,
                    help="sweep the batched kernel's shapes on this workload and merge its pick into --out")
    ap.add_argument("--wall-s", type=float, default=20.0, help="--mix: seconds per point")
    ap.add_argument("--steps-ms", type=int, nargs="+", default=None, help="--mix serving: step budgets to run")
    ap.add_argument("--lane-keys-log2", type=int, nargs="+", default=None,
                    help="--mix: expected keys per virtual lane (log2) to run")
    ap.add_argument("--waves-per-cu", type=int, nargs="+", default=None, help="--mix: only these waves per CU")
    ap.add_argument("--extend", action="store_true",
                    help="--mix: merge the points into the table's earlier sweep and pick over all of them")
    ap.add_argument("--join", nargs="+", default=None,
                    help="tables of separate sweeps (e.g. one per walk form) to join into --out, picking anew")
    ap.add_argument("--walk", choices=("all", "pairs", "projective"), default="all")
    ap.add_argument("--shape", choices=("grid", "round"), default="grid",
                    help="grid: the library's launch shape at each point's occupancy; round: one resident round")
    args = ap.parse_args()
    if args.mix:
        table = load(args.out) or {"points": []}
        pts = mix_sweep(args.mix, args.wall_s, steps_ms=args.steps_ms, lane_keys_log2=args.lane_keys_log2,
                        waves_per_cu=args.waves_per_cu)
        earlier = table.get("mix_points", {}).get(args.mix)
        if args.extend and earlier:
            pts = merge_points(earlier["points"], pts)
        table.setdefault("mix_points", {})[args.mix] = {
            "workload": ("batched: 256 seeded searches (4-10 hex digits)" if args.mix == "batched" else
                         "serving: Poisson 20/s, 5-9 hex digits, short-first-oldest admission") +
                        ", %.0f s per point" % args.wall_s, "points": pts}
        table.setdefault("picks", {})[args.mix] = mix_pick(args.mix, pts)
        if "pick" in table:
            table["picks"]["single"] = dict(table["pick"])
        if args.out != "/dev/null":
            with open(args.out, "w") as f:
                json.dump(table, f, indent=1)
        print(json.dumps(table["picks"][args.mix]))
        return
    if args.join:
        tables = [load(p) for p in args.join]
        if any(t is None for t in tables):
            raise SystemExit("cannot read every table of %s" % args.join)
        table = pick({"workload": tables[0]["workload"], "points": [p for t in tables for p in t["points"]]})
        if args.merge_pmc:
            table = merge_pmc(table, args.merge_pmc, args.pmc_budget_log2)
    elif args.merge_pmc:
        table = load(args.out)
        if table is None:
            raise SystemExit("no table at %s to merge into" % args.out)
        table = merge_pmc(table, args.merge_pmc, args.pmc_budget_log2)
    else:
        walks = [w for w in WALKS if args.walk in ("all", w[1])]
        table = pick(sweep(1 << args.budget_log2, walks, args.shape))
    if args.out != "/dev/null":
        with open(args.out, "w") as f:
            json.dump(table, f, indent=1)
    print(json.dumps(table.get("pick")))


if __name__ == "__main__":
    main()
