"""bench.py's contract on the CPU: the one-GPU guard, and the committed HBM
traffic summaries it reports beside the roofline (one per swept operator
shape, with and without the coefficient class table)."""
import os
import subprocess
import sys

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""          # no JSON line for a run that did not happen


@pytest.mark.parametrize("classes", [True, False], ids=["class_table", "streamed"])
@pytest.mark.parametrize("cw,kz", bench.SHAPES)
def test_traffic_summary_committed_for_every_swept_shape(cw, kz, classes):
    kernel = "k_operator_apply_config3_128" + ("_classes" if classes else "")
    nbytes, src = bench.traffic_from_profile(kernel, f"{cw}x{kz}")
    assert src == f"profiles/traffic_{kernel}_{cw}x{kz}.txt"
    kv = {}
    for line in open(os.path.join(ROOT, src)):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    # the summary belongs to the kernel variant of that shape
    assert kv["kernel"] == f"k_operator<double, double, 1, {cw}, {'true' if classes else 'false'}>"
    # the same command shape is recorded in the header
    text = open(os.path.join(ROOT, src)).read()
    assert f"--zchunk {kz} --colwidth {cw}" in text
    # y (8 * 4 * 128^3 bytes) is written once; reads at least x, at most a few times the arrays
    n1 = 128 ** 3
    assert float(kv["write_kib"]) * float(kv["write_scale"]) * 1024 == pytest.approx(32 * n1, rel=0.01)
    reads = float(kv["fetch_kib"]) * float(kv["fetch_scale"]) * 1024
    own = 32 * n1 + (2 * n1 if classes else 48 * n1)
    assert own <= reads <= 1.5 * own
    assert nbytes == pytest.approx(reads + 32 * n1, rel=0.01)


def test_traffic_summary_missing_shape_is_none():
    assert bench.traffic_from_profile("k_operator_apply_config3_128_classes", "64x5") == (None, None)
