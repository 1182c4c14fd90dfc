"""Whole-build CPU timing of the 1M-vertex road-like workload, for the CPU
baseline that bench.py reports (test infrastructure: it times the oracle, the
thing bench.py's cpu_baseline leg runs, and is never part of the product).

bench.py times the oracle's GLL (run_paraPLL restated) on the GPU host's cores
over the rank prefix it completes in about 20 s, because the whole build takes
minutes. This script measures, on one host, both that same bounded prefix and
the whole build (lt + st + ct, as shared_memory/akiba.hpp:214 times it), and
writes oracle/c2_whole_build.json with the labels/s of each and their ratio:
bench.py scales its prefix rate by that ratio to a whole-build rate and states
so. The reference's own whole-build time on the same inputs is in
tests/golden/fullsize.json (reference_stdout).

    python oracle/calibrate_cpu.py [threads]
"""
import json
import os
import platform
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
# the repository root, not this folder (whose oracle.py would shadow the package)
sys.path = [os.path.dirname(HERE)] + [p for p in sys.path if os.path.abspath(p or ".") != HERE]

from oracle import oracle as orc  # noqa: E402
from plant_amd import chl  # noqa: E402


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 8)
    g = chl.Graph.knn(1_000_000, 3, seed=1)
    order = g.order_betweenness(256, seed=7)
    src, dst, w = g.arc_list(1)
    og = orc.OGraph.from_arcs(g.n, src, dst, w)
    Lp, s = orc.gll_bounded(og, order, threads, 20.0)
    prefix_s = float(s[0] + s[1] + s[2])
    prefix_labels = int(Lp.total())
    del Lp
    t0 = time.time()
    L, t = orc.gll(og, order, threads=threads)
    whole_s = float(t[0] + t[1] + t[2])
    whole_labels = int(L.total())
    out = {
        "workload": "knn(1000000, 3, seed=1), order_betweenness(256, seed=7)",
        "threads": threads, "cpu": platform.processor() or platform.machine(),
        "prefix": {"roots": int(s[4]), "labels": prefix_labels, "seconds": prefix_s,
                   "labels_per_s": prefix_labels / prefix_s},
        "whole": {"labels": whole_labels, "seconds": whole_s, "lt": float(t[0]), "st": float(t[1]),
                  "ct": float(t[2]), "labels_per_s": whole_labels / whole_s, "wall_s": time.time() - t0},
    }
    out["whole_over_prefix_rate"] = out["whole"]["labels_per_s"] / out["prefix"]["labels_per_s"]
    with open(os.path.join(HERE, "c2_whole_build.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
