"""The device sampled-betweenness pipeline's own source (the tree kernel of
plant_amd/csrc/tree.hip in TREE_KEEP mode and the four passes of
plant_amd/csrc/order.hip), compiled for the host with g++ and AddressSanitizer
against tests/emu/hip/hip_runtime.h, every workgroup of every launch run, over
scratch arrays sized exactly as plant_amd/csrc/order_device.cpp sizes them,
with several groups of roots per slot set. With every vertex as a root the
summed scores do not depend on the sampled root set, so they must equal the host
order's scores (chl_order_betweenness_scored, itself pinned against scipy in
tests/test_host.py), and every search cell must come back to its rest state.

Not a model of the GPU memory system: it checks indexing, bounds and the
barrier and counting logic without a GPU."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "emu", "order_emu.cpp")


@pytest.fixture(scope="module")
def emulator(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("emu") / "order_emu")
    cmd = ["g++", "-std=c++20", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-pthread",
           "-I", os.path.join(HERE, "emu"), "-x", "c++", SRC, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return exe


def run_order_emulator(exe, tmp_path, g, slots, delta, seed=7, threads=64):
    n = g.n
    off, head, ln = g.csr(1)
    ioff, ihead, iln = g.csr(0)
    arc = (head.astype(np.uint64) << np.uint64(32)) | ln.astype(np.uint32).astype(np.uint64)
    iarc = (ihead.astype(np.uint64) << np.uint64(32)) | iln.astype(np.uint32).astype(np.uint64)
    roots = np.random.default_rng(seed).permutation(n).astype(np.uint32)  # every vertex
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.array([n, arc.size, iarc.size, n, delta, slots, threads], np.uint32).tobytes())
        f.write(off.astype(np.uint32).tobytes())
        f.write(arc.tobytes())
        f.write(ioff.astype(np.uint32).tobytes())
        f.write(iarc.tobytes())
        f.write(np.arange(n, dtype=np.uint32).tobytes())  # device id = vertex id
        f.write(roots.tobytes())
    env = dict(os.environ, EMU_SEED=str(seed), ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0, (r.returncode, r.stderr[-3000:])
    data = np.fromfile(out, np.uint8)
    score = data[:8 * n].view(np.uint64)
    restored = int(data[8 * n:].view(np.uint32)[0])
    return score, restored


def cases(chl):
    yield "grid", chl.Graph.grid(7, 7, 1, 100, seed=1)
    yield "grid_ties", chl.Graph.grid(6, 6, 1, 3, seed=2)  # many equal-length paths
    yield "knn", chl.Graph.knn(60, 3, seed=8)
    rng = np.random.default_rng(13)
    n, m = 40, 150
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    yield "directed", chl.Graph.from_arcs(n, src, dst, rng.integers(1, 30, m).astype(np.int32))
    yield "rmat_unit", chl.Graph.rmat(6, 3 * 64, seed=3)


@pytest.mark.parametrize("slots,threads", [(7, 64), (5, 256)])
def test_betweenness_pipeline_emulated_equals_host_scores(chl, emulator, tmp_path, slots, threads):
    for name, g in cases(chl):
        _, want = g.order_betweenness(g.n, seed=1, with_score=True)
        for delta in (2, 200):
            score, restored = run_order_emulator(emulator, tmp_path, g, slots, delta, threads=threads)
            assert restored == 1, (name, slots, delta)
            assert np.array_equal(score.astype(np.float64), np.asarray(want, np.float64)), (name, slots, delta)
