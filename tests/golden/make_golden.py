"""Writes the golden fixtures of tests/golden/ from the reference's own GLL.

The reference's shared_memory/hhl.cpp is compiled from its sources by
oracle/Makefile into oracle/_ref/hhl. This script generates small seeded inputs
with this project's generators, writes them as DIMACS graph + order files (the
formats of shared_memory/graph.hpp:168-189 and ordering.hpp:37-49), runs the
reference hhl on them with -l, and stores inputs and label files here, plus a
manifest with sha256 sums. For the 100x100 grid workload only the sha256 and
sizes of the reference's label file are kept (the file is ~12 MB).

The reference's degree-order tool (distributed_memory/degree.cpp, built by
oracle/Makefile into oracle/_ref/degree) is run on every golden graph file; its
order goes to <name>.degree and its stdout line into the manifest
("degree_stdout").

Run from the repository root:  python tests/golden/make_golden.py
(python tests/golden/make_golden.py degree  redoes only the degree fixtures)
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from plant_amd import chl  # noqa: E402

HHL = os.path.join(ROOT, "oracle", "_ref", "hhl")
DEGREE = os.path.join(ROOT, "oracle", "_ref", "degree")


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def cases():
    yield "grid12", chl.Graph.grid(12, 12, 1, 100, seed=21), dict(kind="betweenness", roots=16, seed=4)
    yield "knn300", chl.Graph.knn(300, 3, seed=8), dict(kind="betweenness", roots=32, seed=6)
    rng = np.random.default_rng(13)
    n, m = 120, 600
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    w = rng.integers(1, 30, m).astype(np.int32)
    yield "directed120", chl.Graph.from_arcs(n, src, dst, w), dict(kind="random", seed=5)
    yield "rmat8", chl.Graph.rmat(8, 3 * 256, seed=3), dict(kind="degree", seed=2)
    yield "rmat8w", chl.Graph.rmat(8, 3 * 256, wmax=16, seed=3), dict(kind="degree", seed=2)


def make_order(g, spec):
    if spec["kind"] == "betweenness":
        return g.order_betweenness(spec["roots"], seed=spec["seed"])
    if spec["kind"] == "degree":
        return g.order_degree(seed=spec["seed"])
    return np.random.default_rng(spec["seed"]).permutation(g.n).astype(np.uint32)


def run_ref(gpath, opath, lpath):
    subprocess.run([HHL, "-g", gpath, "-o", opath, "-p", "4", "-l", lpath], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


def main():
    if not os.path.exists(HHL):
        raise SystemExit(f"{HHL} missing: build it with make -C oracle (needs the reference tree)")
    manifest = {}
    for name, g, spec in cases():
        gp, op, lp = (os.path.join(HERE, f"{name}.{ext}") for ext in ("gr", "order", "labels"))
        g.write(gp)
        chl.order_write(op, make_order(g, spec))
        run_ref(gp, op, lp)
        manifest[name] = {"n": g.n, "graph_sha256": sha(gp), "order_sha256": sha(op), "labels_sha256": sha(lp),
                          "order": spec}
    # 100x100 grid, w in [1,100], 64-root sampled-betweenness order: hashes only
    g = chl.Graph.grid(100, 100, 1, 100, seed=1)
    order = g.order_betweenness(64, seed=1)
    tmp = os.path.join(HERE, "_tmp_grid100")
    os.makedirs(tmp, exist_ok=True)
    gp, op, lp = (os.path.join(tmp, f"grid100.{ext}") for ext in ("gr", "order", "labels"))
    g.write(gp)
    chl.order_write(op, order)
    run_ref(gp, op, lp)
    manifest["grid100"] = {"n": g.n, "graph_sha256": sha(gp), "order_sha256": sha(op), "labels_sha256": sha(lp),
                           "labels_bytes": os.path.getsize(lp), "order": {"kind": "betweenness", "roots": 64,
                                                                          "seed": 1}}
    for p in (gp, op, lp):
        os.remove(p)
    os.rmdir(tmp)
    degree_fixtures(manifest)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print(json.dumps(manifest, indent=1))


def degree_fixtures(manifest):
    """<name>.degree and the stdout line of the reference degree tool, per golden graph."""
    if not os.path.exists(DEGREE):
        raise SystemExit(f"{DEGREE} missing: build it with make -C oracle (needs the reference tree)")
    for name in sorted(manifest):
        gp = os.path.join(HERE, f"{name}.gr")
        if not os.path.exists(gp):
            continue
        dp = os.path.join(HERE, f"{name}.degree")
        r = subprocess.run([DEGREE, "-o", dp, gp], check=True, capture_output=True, text=True)
        manifest[name]["degree_stdout"] = r.stdout
        manifest[name]["degree_sha256"] = sha(dp)


if __name__ == "__main__":
    if sys.argv[1:] == ["degree"]:
        with open(os.path.join(HERE, "manifest.json")) as f:
            m = json.load(f)
        degree_fixtures(m)
        with open(os.path.join(HERE, "manifest.json"), "w") as f:
            json.dump(m, f, indent=1, sort_keys=True)
    else:
        main()
