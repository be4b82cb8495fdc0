"""bench.py's command-line contract that holds without a device: a request for
more than one GPU exits with a message instead of running (every box here has
one GPU; the multi-GPU exchange is design only), and the committed profile
summaries bench.py reads are present and well-formed."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
    assert not any(line.startswith("{") for line in r.stdout.splitlines())


def test_committed_profiles_bench_reads():
    prof = os.path.join(ROOT, "profiles")
    traffic = json.load(open(os.path.join(prof, "walk_6max_100bb_traffic.json")))
    assert traffic["hbm_bytes_per_launch"] > 0 and os.path.exists(os.path.join(ROOT, traffic["source"]))
    issue = json.load(open(os.path.join(prof, "walk_6max_100bb_issue.json")))
    assert issue["valu_instructions_per_launch"] > 0 and 0 < issue["lanes_active_per_instruction"] <= 1
    assert os.path.exists(os.path.join(ROOT, issue["source"]))
    l2 = json.load(open(os.path.join(prof, "large_table_l2.json")))
    assert 0 < l2["l2_hit_rate"] < 1 and l2["table_atomics_per_launch"] == 0
