"""Tests of the command-line tool (lxmf-vanity_amd/lxmf-vanity), the host
side above the C ABI: usage and error texts, exit codes and the identity
file of src/main.rs:325-498, and searches whose saved identity is
re-derived here by the oracle.

CPU tests run the CPU-only mode (no flag); the GPU tests (marked gpu) run
--gpu, --gpu --cpu and --gpu-self-test on the device.
"""

import os
import re
import stat
import subprocess

import pytest

import oracle_lib as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "lxmf-vanity_amd", "lxmf-vanity")

USAGE_HEAD = "usage: lxmf-vanity [--gpu [--cpu]] <hex prefix>\n       lxmf-vanity --gpu-self-test\n"


def run(args, cwd, timeout=300):
    if not os.path.exists(CLI):
        pytest.fail("%s is not built: run __graft_entry__.build()" % CLI)
    print("cli", args, flush=True)
    return subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=timeout)


def saved_identity_ok(path, prefix):
    st = os.stat(path)
    assert stat.S_IMODE(st.st_mode) == 0o600
    data = open(path, "rb").read()
    assert len(data) == 64
    addr = orc.address_of_identity(data)
    assert addr.hex().startswith(prefix.lower())
    return addr


@pytest.mark.parametrize("args", [[], ["--cpu", "ab"], ["--bogus", "ab"], ["ab", "cd"], ["--gpu"]])
def test_usage_errors(tmp_path, args):
    r = run(args, tmp_path)
    assert r.returncode == 1
    assert r.stderr.startswith(USAGE_HEAD)
    assert r.stdout == ""


def test_bad_prefix_is_refused_with_the_reference_message(tmp_path):
    r = run(["BOOB"], tmp_path)
    assert r.returncode == 1
    assert r.stderr == "'boob' isn't 1 to 32 hex digits: use 0-9 and a-f only (zero, not the letter O).\n"
    r = run(["0123456789abcdef0123456789abcdef0"], tmp_path)
    assert r.returncode == 1 and "isn't 1 to 32 hex digits" in r.stderr


def test_cpu_search_saves_a_private_identity(tmp_path):
    r = run(["A1B"], tmp_path)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == "Looking for an LXMF address starting a1b."
    assert re.fullmatch(r"Searching on \d+ CPU threads; about 4096 tries on average\.", lines[1])
    addr = saved_identity_ok(os.path.join(tmp_path, "a1b.identity"), "a1b")
    assert lines[-2] == "Found " + addr.hex()
    assert re.fullmatch(r"About \d+ tries in \d+\.\d s \(\d+/s\)\. Saved to a1b\.identity", lines[-1])


def test_progress_line_every_ten_seconds_parses_as_fleet_sh_reads_it(tmp_path):
    # An unreachable prefix (16 hex digits) on the CPU for about 11 s: the
    # 10-second report (src/main.rs:407-421) must be on stdout by then, in
    # the form fleet.sh reads (fleet.sh:100-103): the last line containing
    # "tries, ", its first word the tries, and after "tries, " the rate up to
    # "/s", both integers.
    import signal
    import time
    p = subprocess.Popen([CLI, "0123456789abcdef"], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                         text=True)
    try:
        time.sleep(11.5)
    finally:
        p.send_signal(signal.SIGTERM)
        out, err = p.communicate(timeout=30)
    lines = [ln for ln in out.splitlines() if "tries, " in ln]
    assert len(lines) == 1, out
    line = lines[-1]
    assert re.fullmatch(r"\d+ tries, \d+/s; the average is about \d+ min away", line), line
    tried = int(line.split(" ", 1)[0])             # ${line%% *}
    rate = int(line.split("tries, ", 1)[1].split("/s", 1)[0])   # r=${line#*tries, }; ${r%%/s*}
    assert tried > 0 and rate > 0 and 9 * rate <= tried <= 11 * rate
    assert not os.path.exists(os.path.join(tmp_path, "0123456789abcdef.identity"))


def test_an_existing_identity_file_is_never_overwritten(tmp_path):
    p = os.path.join(tmp_path, "c0.identity")
    with open(p, "wb") as f:
        f.write(b"keep")
    r = run(["C0"], tmp_path)
    assert r.returncode == 1
    assert r.stderr == "c0.identity already exists; not overwriting it.\n"
    assert open(p, "rb").read() == b"keep"


@pytest.mark.gpu
def test_gpu_self_test_passes():
    r = run(["--gpu-self-test"], ROOT)
    assert r.returncode == 0, r.stderr
    out = r.stdout.splitlines()
    assert re.fullmatch(r".*: \d+ GPU threads, ready in \d+\.\d s\.", out[0])
    assert re.fullmatch(r"Self-test passed: 65536 public keys match X25519, \d+ hits for 'b0' all recheck on the CPU\.",
                        out[1])


@pytest.mark.gpu
@pytest.mark.parametrize("args,prefix", [(["--gpu"], "5eed5"), (["--gpu", "--cpu"], "c0de")])
def test_gpu_search_saves_a_private_identity(tmp_path, args, prefix):
    r = run(args + [prefix], tmp_path)
    assert r.returncode == 0, r.stderr
    out = r.stdout.splitlines()
    assert "GPU threads, ready in" in r.stdout
    on = "the GPU" if args == ["--gpu"] else "the GPU and"
    assert out[1].startswith("Searching on " + on)
    addr = saved_identity_ok(os.path.join(tmp_path, prefix + ".identity"), prefix)
    assert out[-2] == "Found " + addr.hex()
