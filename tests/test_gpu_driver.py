"""The drop-in driver plant_amd/bin/hhl (shared_memory/hhl.cpp:16-79 plus the query
flags of distributed_memory/hhl.cpp:47-48,82-91) on a committed golden graph: the
output lines, the label file (byte-identical to the reference's own, written by
the reference binary when the fixture was made) and the query results against
the oracle's query over the same labels."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
HHL = os.path.join(HERE, "..", "plant_amd", "bin", "hhl")


def run_hhl(args, timeout=120):
    print("hhl", " ".join(args), flush=True)
    return subprocess.run([HHL] + args, capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("name", ["grid12", "directed120"])
def test_driver_labels_and_lines(chl, tmp_path, name):
    out = tmp_path / "l.txt"
    args = ["-g", os.path.join(GOLD, name + ".gr"), "-o", os.path.join(GOLD, name + ".order"), "-p", "4",
            "-l", str(out)]
    r = run_hhl(args)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0].split() == [HHL] + args  # the argv echo
    assert re.fullmatch(r"Average label size after gll is \S+, labeling time = \S+, sorting time = \S+, "
                        r"cleaning time = \S+, total time = \S+", lines[1])
    avg = float(lines[1].split()[6].rstrip(","))
    assert lines[2].split()[0] == lines[1].split()[6].rstrip(",")
    n = chl.Labels.read(os.path.join(GOLD, name + ".labels")).n
    assert re.search(rf"Order size: {n} \|V\|: {n}", r.stderr)
    assert out.read_bytes() == open(os.path.join(GOLD, name + ".labels"), "rb").read()
    assert abs(avg - chl.Labels.read(str(out)).avg) < 1e-3 * max(1.0, avg)


def test_driver_queries(chl, orc, tmp_path):
    name = "knn300"
    ref = orc.OLabels.read(os.path.join(GOLD, name + ".labels"))
    n = ref.n
    rng = np.random.default_rng(5)
    pairs = rng.integers(0, n, (2000, 2)).astype(np.uint32)
    qf = tmp_path / "q.txt"
    qf.write_text("".join(f"{u} {v}\n" for u, v in pairs))
    res = tmp_path / "r.txt"
    base = ["-g", os.path.join(GOLD, name + ".gr"), "-o", os.path.join(GOLD, name + ".order")]
    r = run_hhl(base + ["-rq", str(qf), "-qo", str(res)])
    assert r.returncode == 0, r.stderr
    got = np.array([int(x) for x in res.read_text().split()], np.int32)
    assert np.array_equal(got, ref.query(pairs))
    # generated queries (unseeded rand() % n) go to stdout after the result lines
    r = run_hhl(base + ["-nq", "300"])
    assert r.returncode == 0, r.stderr
    got = np.array([int(x) for x in r.stdout.splitlines()[3:]], np.int32)
    assert np.array_equal(got, ref.query(chl.query_generate(300, n)))
