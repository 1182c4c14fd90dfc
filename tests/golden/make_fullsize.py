"""Whole-label-set fixtures at the full 1M-vertex sizes, from the reference's own GLL.

For the road-like 1M graph (3-NN, 256-root sampled-betweenness order) and the two
R-MAT scale-20 graphs (unit and [1, 16] lengths, seeded degree order) this script
writes the graph as DIMACS and the order file (into a scratch folder), runs the
reference's shared_memory/hhl (compiled from its sources by oracle/Makefile into
oracle/_ref/hhl) with -l into a FIFO, and streams the label text through
tests/golden/label_digest (compiled here with g++), which prints the entry count
and the digest of chl_labels_checksum per side. Only those numbers, the input
sha256 sums and the reference's own printed timings (akiba.hpp:214-215) go into
tests/golden/fullsize.json; the label files themselves (12 GB of text at 1M
road-like) are never written to disk.

Run from the repository root on a host with the reference tree:
    python tests/golden/make_fullsize.py [c2 c3u c3w]
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from plant_amd import chl  # noqa: E402

HHL = os.path.join(ROOT, "oracle", "_ref", "hhl")
DIGEST_SRC = os.path.join(HERE, "label_digest.cpp")
OUT = os.path.join(HERE, "fullsize.json")


def cases():
    yield "c2", lambda: chl.Graph.knn(1_000_000, 3, seed=1), lambda g: g.order_betweenness(256, seed=7), \
        "knn(1000000, 3, seed=1); order_betweenness(256, seed=7)"
    yield "c3u", lambda: chl.Graph.rmat(20, 3 << 20, wmax=1, seed=1), lambda g: g.order_degree(seed=2), \
        "rmat(20, 3<<20, wmax=1, seed=1); order_degree(seed=2)"
    yield "c3w", lambda: chl.Graph.rmat(20, 3 << 20, wmax=16, seed=1), lambda g: g.order_degree(seed=2), \
        "rmat(20, 3<<20, wmax=16, seed=1); order_degree(seed=2)"


def sha_file(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 24), b""):
            h.update(chunk)
    return h.hexdigest()


def main(names):
    if not os.path.exists(HHL):
        raise SystemExit(f"{HHL} missing: build it with make -C oracle (needs the reference tree)")
    scratch = tempfile.mkdtemp(prefix="chl_fullsize_", dir=os.environ.get("CHL_SCRATCH"))
    digest = os.path.join(scratch, "label_digest")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", digest, DIGEST_SRC], check=True)
    threads = str(os.cpu_count() or 8)
    fixtures = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name, make_graph, make_order, spec in cases():
        if names and name not in names:
            continue
        g = make_graph()
        order = make_order(g)
        gp, op, fifo = (os.path.join(scratch, f"{name}.{ext}") for ext in ("gr", "order", "fifo"))
        g.write(gp)
        chl.order_write(op, order)
        os.mkfifo(fifo)
        t0 = time.time()
        reader = subprocess.Popen([digest, fifo], stdout=subprocess.PIPE, text=True)
        ref = subprocess.run([HHL, "-g", gp, "-o", op, "-p", threads, "-l", fifo], check=True, capture_output=True,
                             text=True)
        out, _ = reader.communicate()
        if reader.returncode != 0:
            raise SystemExit(f"label_digest failed on {name}")
        d = json.loads(out)
        fixtures[name] = {
            "inputs": spec, "n": g.n, "graph_sha256": sha_file(gp), "order_sha256": sha_file(op),
            "count": d["count"], "digest": d["digest"],
            "reference_stdout": ref.stdout.strip().splitlines()[-2:], "reference_threads": int(threads),
            "wall_s": round(time.time() - t0, 1),
        }
        print(name, json.dumps(fixtures[name]), flush=True)
        for p in (gp, op, fifo):
            os.remove(p)
        with open(OUT, "w") as f:
            json.dump(fixtures, f, indent=1, sort_keys=True)
    os.remove(digest)
    os.rmdir(scratch)


if __name__ == "__main__":
    main(sys.argv[1:])
