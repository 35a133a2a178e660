"""CPU checks of the autotuner's bookkeeping: kernel names from the profiler
mapped back to sweep points, the VALU-busy merge and the pick."""

import csv
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "lxmf-vanity_amd"))
import autotune  # noqa: E402


def test_kernel_names_map_to_points():
    k = autotune.kernel_key("void lxv::walk_pair_queue_kernel<lxv::Field8, 64, 4>(lxv::WalkParams)", "262144")
    assert k == ("pairs", "8x32", 64, 4, 262144)
    k = autotune.kernel_key("void lxv::walk_pair_kernel<lxv::Field10, 256, 2>(lxv::WalkParams)", "131072")
    assert k == ("pairs", "10x25.5", 256, 2, 131072)
    k = autotune.kernel_key("void lxv::walk_kernel<lxv::Field8, 16, 1>(lxv::WalkParams)", "65536")
    assert k == ("projective", "8x32", 16, 1, 65536)
    assert autotune.kernel_key("lxv::clear_kernel(unsigned int*, unsigned int)", "256") is None


def test_merge_and_pick(tmp_path):
    pts = [dict(walk="pairs", walk_id=0, layout="8x32", layout_id=0, batch=64, waves_per_cu=16, waves_per_simd=4,
                lanes=1 << 24, grid_lanes=1 << 18, cus=256, keys_per_s=9.2e9),
           dict(walk="pairs", walk_id=0, layout="8x32", layout_id=0, batch=256, waves_per_cu=16, waves_per_simd=4,
                lanes=1 << 24, grid_lanes=1 << 18, cus=256, keys_per_s=9.6e9),
           dict(walk="projective", walk_id=1, layout="8x32", layout_id=0, batch=64, waves_per_cu=16,
                waves_per_simd=4, lanes=1 << 21, grid_lanes=1 << 21, cus=256, keys_per_s=6.9e9)]
    table = autotune.pick({"points": pts})
    assert table["pick"]["batch"] == 256 and table["pick"]["walk"] == "pairs"
    assert table["pick_by_batch"]["64"]["keys_per_s"] == 9.2e9
    path = tmp_path / "pmc.csv"
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"])
        w.writeheader()
        name = "void lxv::walk_pair_queue_kernel<lxv::Field8, 256, 4>(lxv::WalkParams)"
        # one SIMD-cycle in four issuing VALU work on 256 CUs x 4 SIMDs, per XCD clock
        for c, v in (("GRBM_GUI_ACTIVE", 8e6), ("SQ_ACTIVE_INST_VALU", 0.25 * 1024 * 1e6 / 4),
                     ("SQ_INSTS_VALU", 123.0)):
            w.writerow({"Kernel_Name": name, "Grid_Size": str(1 << 18), "Counter_Name": c, "Counter_Value": v})
    table = autotune.merge_pmc(table, str(path), budget_log2=28)
    merged = [p for p in table["points"] if "valu_busy_pct" in p]
    assert len(merged) == 1 and merged[0]["batch"] == 256
    assert abs(merged[0]["valu_busy_pct"] - 25.0) < 1e-9
    # 2^24 lanes x 256 keys = 2^32 keys per iteration, above the 2^28 budget:
    # the pass walked one warm-up iteration and two of one, 3 x 2^32 keys
    assert autotune.pmc_pass_keys(merged[0], 28) == 3 << 32
    assert merged[0]["valu_insts_per_key"] == 123.0 * 64 / (3 << 32)


def test_mix_pick():
    pts = [dict(batch=64, waves_per_cu=16, layout_id=0, walk_id=0, layout="8x32", step_ms=25, lane_keys_log2=10,
                keys_per_s=8.0e9, p50_s_with_unfinished=0.9),
           dict(batch=256, waves_per_cu=16, layout_id=0, walk_id=0, layout="8x32", step_ms=25, lane_keys_log2=10,
                keys_per_s=9.0e9, p50_s_with_unfinished=1.5),
           dict(batch=128, waves_per_cu=8, layout_id=0, walk_id=0, layout="8x32", step_ms=15, lane_keys_log2=10,
                keys_per_s=8.5e9, p50_s_with_unfinished=1.0)]
    assert autotune.mix_pick("batched", pts)["batch"] == 256
    # serving: lowest p50 among points within 10% of the best rate (8.0e9 is
    # out, 8.5e9 is in)
    assert autotune.mix_pick("serving", pts)["batch"] == 128


def test_mix_sweep_runs_every_variant(monkeypatch):
    # the sweep's bookkeeping without a GPU: each point records the shape,
    # step budget and keys per lane it was run with, and every combination the
    # sweep names is run once
    import workloads
    calls = []

    def fake(kind):
        def run(**kw):
            sh = dict(kw["shape"])
            calls.append((kind, sh["batch"], sh["waves_per_cu"], sh["layout"], sh.get("step_ms", 0),
                          sh["lane_keys_log2"]))
            # a rate that rises with B and keys per lane, a p50 that rises with
            # the step budget and falls with keys per lane
            rate = 8e9 + 1e6 * sh["batch"] + 1e7 * sh["lane_keys_log2"]
            p50 = 0.01 * (sh.get("step_ms") or 1) - 0.001 * sh["lane_keys_log2"]
            r = {"keys_per_s": rate, "found": 1, "p50_s": p50, "p99_s": 1.0, "p50_s_with_unfinished": p50,
                 "p99_s_with_unfinished": 2.0, "finished": 1, "unfinished": 0, "policy": "x",
                 "shape": {"step_ms": sh.get("step_ms") or 200, "lane_keys_log2": sh["lane_keys_log2"]}}
            return r
        return run

    monkeypatch.setattr(workloads, "batched", fake("batched"))
    monkeypatch.setattr(workloads, "serving", fake("serving"))
    pts = autotune.mix_sweep("batched", 1.0, progress=None)
    assert len(pts) == len(autotune.MIX_SHAPES) * len(autotune.BATCHED_LANE_KEYS_LOG2)
    assert sorted({p["lane_keys_log2"] for p in pts}) == sorted(autotune.BATCHED_LANE_KEYS_LOG2)
    assert autotune.mix_pick("batched", pts)["lane_keys_log2"] == max(autotune.BATCHED_LANE_KEYS_LOG2)
    pts = autotune.mix_sweep("serving", 1.0, progress=None)
    n = len(autotune.MIX_SHAPES) * len(autotune.SERVING_STEPS_MS) * len(autotune.SERVING_LANE_KEYS_LOG2)
    assert len(pts) == n and len(set(calls[-n:])) == n
    assert {(p["step_ms"], p["lane_keys_log2"]) for p in pts} == \
        {(s, k) for s in autotune.SERVING_STEPS_MS for k in autotune.SERVING_LANE_KEYS_LOG2}
    pick = autotune.mix_pick("serving", pts)
    assert pick["step_ms"] == min(autotune.SERVING_STEPS_MS)
    # a partial sweep: 5 ms steps at 8 waves/CU only, merged into the full one;
    # its points replace nothing else, and the pick is made over all points
    part = autotune.mix_sweep("serving", 1.0, progress=None, steps_ms=[5], lane_keys_log2=[8], waves_per_cu=[8])
    shapes8 = [s for s in autotune.MIX_SHAPES if s["waves_per_cu"] == 8]
    assert len(part) == len(shapes8) and {(p["step_ms"], p["lane_keys_log2"], p["waves_per_cu"]) for p in part} == \
        {(5, 8, 8)}
    merged = autotune.merge_points(pts, part)
    assert len(merged) == len(pts) + len(part)
    assert autotune.mix_pick("serving", merged)["step_ms"] == 5
    # a point run again replaces the older one of the same shape and variant
    again = [dict(p, keys_per_s=1.0) for p in part]
    assert len(autotune.merge_points(merged, again)) == len(merged)
    assert all(p["keys_per_s"] == 1.0 for p in autotune.merge_points(merged, again) if p["step_ms"] == 5)


def test_committed_picks_follow_from_the_committed_points():
    # the picks in tuned.json (compiled into the library) are what the pick
    # rules give on the points recorded beside them
    table = autotune.load()
    assert table is not None
    for kind in ("batched", "serving"):
        pts = table["mix_points"][kind]["points"]
        assert autotune.mix_pick(kind, pts) == table["picks"][kind], kind
    assert autotune.pick({"workload": table["workload"], "points": table["points"]})["pick"] == table["pick"]
