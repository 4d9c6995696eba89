"""bench.py's contract on the CPU: the one-GPU guard, and the committed HBM
traffic summaries it reports beside the roofline."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""          # no JSON line for a run that did not happen


def test_traffic_summary_missing_is_none():
    assert bench.traffic_from_profile("no_such_kernel") == (None, None)


def test_committed_traffic_summaries_parse_and_name_their_command():
    """Every committed PMC summary gives the counters bench.py reads and the
    full command that produced it."""
    pdir = os.path.join(ROOT, "profiles")
    for fn in sorted(os.listdir(pdir)):
        if not fn.startswith("traffic_"):
            continue
        name = fn[len("traffic_"):-len(".txt")]
        nbytes, src = bench.traffic_from_profile(name)
        assert src == f"profiles/{fn}" and nbytes > 0, fn
        assert "rocprofv3" in open(os.path.join(pdir, fn)).read(), fn


def test_committed_profiles_name_commands_bench_accepts():
    """Every bench.py command line quoted in profiles/ parses with the current
    bench.py arguments, so each summary can be reproduced from the tree."""
    import re
    import shlex
    pdir = os.path.join(ROOT, "profiles")
    seen = 0
    for fn in sorted(os.listdir(pdir)):
        text = open(os.path.join(pdir, fn)).read().replace("\\\n", " ")
        for line in text.splitlines():
            m = re.search(r"bench\.py(.*)$", line)
            if not m:
                continue
            args = [a for a in shlex.split(m.group(1)) if a not in (">", "2>&1")]
            bench.parser().parse_args(args)
            seen += 1
    assert seen > 0
