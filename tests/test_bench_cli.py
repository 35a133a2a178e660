"""CPU checks of bench.py's command line: it measures one GPU and says so."""

import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(args, env=None):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, capture_output=True, text=True,
                          timeout=120, env=env)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = run(["--gpus", "2"])
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_a_multi_rank_launch_is_refused():
    env = dict(os.environ, WORLD_SIZE="8", RANK="0", LOCAL_RANK="0")
    r = run(["--gpus", "8"], env=env)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_pmc_summary_scales_the_counters_by_the_copy_probe(tmp_path):
    # Three synthetic PMC runs: the walk kernel and a 1 GiB copy whose
    # FETCH_SIZE reads half the bytes and WRITE_SIZE all of them.
    sys.path.insert(0, ROOT)
    import bench
    kernel = "walk_pair_queue_kernel<lxv::Field8, 64, 4>"
    keys = 1 << 30
    (tmp_path / "pmc_sq.log").write_text(
        "profiler noise\n" + json.dumps({"metric": "m", "keys": keys, "config": {"kernel": kernel}}) + "\n")
    rows = {"pmc_sq": [(kernel, "SQ_INSTS_VALU", 4554.0 * keys / 64), (kernel, "SQ_INSTS_SALU", 145.0 * keys / 64),
                       (kernel, "SQ_INSTS_VMEM_RD", 7.0 * keys / 64), (kernel, "SQ_INSTS_VMEM_WR", 2.0 * keys / 64),
                       (kernel, "SQ_ACTIVE_INST_VALU", 1e9), (kernel, "GRBM_GUI_ACTIVE", 8e6)],
            "pmc_fetch": [(kernel, "FETCH_SIZE", 20.0 * keys / 1024 / 2), ("copy_kernel(...)", "FETCH_SIZE", 524288.0)],
            "pmc_write": [(kernel, "WRITE_SIZE", 16.0 * keys / 1024), ("copy_kernel(...)", "WRITE_SIZE", 1048576.0)]}
    for d, rs in rows.items():
        (tmp_path / d / "host").mkdir(parents=True)
        with open(tmp_path / d / "host" / "x_counter_collection.csv", "w", newline="") as f:
            w = csv.writer(f)
            w.writerow(["Kernel_Name", "Counter_Name", "Counter_Value"])
            w.writerows(rs)
    out = bench.pmc_summary(str(tmp_path), dest=str(tmp_path / "walk_pmc.json"))
    assert out["kernel"] == kernel and abs(out["valu_insts_per_key"] - 4554.0) < 1e-9
    assert out["fetch_scale"] == 2.0 and out["write_scale"] == 1.0
    assert abs(out["hbm_read_bytes_per_key"] - 20.0) < 1e-9 and abs(out["hbm_write_bytes_per_key"] - 16.0) < 1e-9
    assert abs(out["valu_busy_pct"] - 100.0 * 1e9 * 4 / (1024 * 1e6)) < 1e-9
    assert json.load(open(tmp_path / "walk_pmc.json")) == out


def test_the_key_budget_is_walked_exactly():
    sys.path.insert(0, ROOT)
    import bench
    per_iter = (1 << 24) * 64          # the default grid's lanes x B = 64
    for steps in (1, 8, 20, 64):
        it = bench.split_iters(1 << 36, per_iter, steps)
        assert len(it) == steps and sum(it) * per_iter == 1 << 36 and max(it) - min(it) <= 1
    assert bench.split_iters(1 << 36, per_iter, 20)[:5] == [4, 4, 4, 4, 3]
    assert bench.keys_text(1 << 36) == "2^36"


def test_host_cores_follow_the_lease(monkeypatch):
    sys.path.insert(0, ROOT)
    import bench
    visible = len(os.sched_getaffinity(0))
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    cores, vis, why = bench.host_cores()
    assert vis == visible and cores == (1 if visible > 1 else visible)
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert bench.host_cores()[0] == visible
