"""bench.py's contract pieces that need no GPU: the traffic figure it takes from
the committed rocprofv3 summary of its workload (only for the same kernel sources
and the same work), and the refusal of more than one GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def summary_figures(path):
    bytes_per_build = launches = traffic = src = None
    with open(path) as f:
        for ln in f:
            w = ln.split()
            if "bytes_per_build" in w and "(algorithmic)," in w:
                bytes_per_build = float(w[w.index("bytes_per_build") + 1])
            if "launches_per_build" in w:
                launches = int(w[w.index("launches_per_build") + 1].rstrip(","))
            if ln.startswith("Per dispatch (the bench's traffic figure):"):
                traffic = float(ln.split(":", 1)[1].split()[0])
            if ln.startswith("Kernel sources sha256 (tree.hip, dev.hpp):"):
                src = ln.split(":", 1)[1].strip()
    return bytes_per_build, launches, traffic, src


@pytest.mark.parametrize("path", [bench.TREE_PROFILE, bench.TREE_PROFILE_C4])
def test_traffic_comes_from_the_committed_summary(path, monkeypatch):
    b, n, t, src = summary_figures(path)
    assert None not in (b, n, t, src) and t > b / n  # traffic above the algorithmic bytes per launch
    monkeypatch.setattr(bench, "kernel_source_hash", lambda: src)
    assert bench.profiled_traffic(b, n, path) == t
    assert bench.profiled_traffic(b * 1.01, n, path) == t  # within 2%: the same work
    assert bench.profiled_traffic(b * 1.05, n, path) is None  # other work: no figure
    assert bench.profiled_traffic(b, n + 1, path) is None
    monkeypatch.setattr(bench, "kernel_source_hash", lambda: "0" * 64)
    assert bench.profiled_traffic(b, n, path) is None  # other kernel sources: no figure


def test_source_hash_covers_the_kernel_files():
    import hashlib
    h = hashlib.sha256()
    for f in ("tree.hip", "dev.hpp"):
        h.update(open(os.path.join(ROOT, "plant_amd", "csrc", f), "rb").read())
    assert bench.kernel_source_hash() == h.hexdigest()


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)
