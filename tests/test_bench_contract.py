"""bench.py's contract pieces that need no GPU: the traffic figure it takes from
the committed rocprofv3 summary of its workload, and the refusal of more than one
GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def summary_figures():
    bytes_per_build = launches = traffic = None
    with open(bench.TREE_PROFILE) as f:
        for ln in f:
            w = ln.split()
            if "bytes_per_build" in w and "(algorithmic)," in w:
                bytes_per_build = float(w[w.index("bytes_per_build") + 1])
            if "launches_per_build" in w:
                launches = int(w[w.index("launches_per_build") + 1].rstrip(","))
            if ln.startswith("Per dispatch (the bench's traffic figure):"):
                traffic = float(ln.split(":", 1)[1].split()[0])
    return bytes_per_build, launches, traffic


def test_traffic_comes_from_the_committed_summary():
    b, n, t = summary_figures()
    assert None not in (b, n, t) and t > b / n  # traffic above the algorithmic bytes per launch
    assert bench.profiled_traffic(b, n) == t
    assert bench.profiled_traffic(b * 1.01, n) == t  # within 2%: the same work
    assert bench.profiled_traffic(b * 1.05, n) is None  # other work: no figure
    assert bench.profiled_traffic(b, n + 1) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)
