"""The segment build of the label store (plant_amd/csrc/labels.hip: row count,
scan, cursor copy, scatter, per-row sort, launched as engine.cpp's build_segment
launches them), compiled for the host with g++ against tests/emu/hip/
hip_runtime.h and run with every workgroup's lanes as cooperative fibers. A
batch's labels arrive in the tree kernel's emission order (any order); the
segment must hold each row's labels hub-ascending, with the row offsets of the
counts, in the 5-byte (hub offset, distance) and the 8-byte entry formats.

Rows of every length class are present: up to 16 entries (sorted by one lane),
up to 4,096 (a bitonic sort in LDS) and longer (the bitonic network in global
memory). The expected segment is numpy's sort of the same labels; this pins the
restatement of labeling.hpp:110-120 (rows hub-sorted) that the GPU tests check
against the oracle."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "emu", "segment_emu.cpp")
SEG_F5, SEG_F8 = 5, 8


@pytest.fixture(scope="module")
def emulator(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("emu") / "segment_emu")
    cmd = ["g++", "-std=c++20", "-O2", "-pthread", "-I", os.path.join(HERE, "emu"), "-x", "c++", SRC, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return exe


def batch(rng, n, lengths, h0, nhub, maxd):
    """Labels of one batch: row r gets lengths[r] distinct hubs of [h0, h0 + nhub)
    (a root labels a vertex at most once), shuffled."""
    rows, hubs = [], []
    for r, k in enumerate(lengths):
        rows.append(np.full(k, r, np.uint32))
        hubs.append(h0 + rng.choice(nhub, size=k, replace=False).astype(np.uint32))
    row = np.concatenate(rows) if rows else np.zeros(0, np.uint32)
    hub = np.concatenate(hubs) if hubs else np.zeros(0, np.uint32)
    dist = rng.integers(0, maxd, row.size).astype(np.uint32)
    perm = rng.permutation(row.size)
    return row[perm], hub[perm], dist[perm]


def run(exe, tmp_path, n, fmt, h0, row, hub, dist, seed=3):
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    hd = (hub.astype(np.uint64) << np.uint64(32)) | dist.astype(np.uint64)
    with open(inp, "wb") as f:
        f.write(np.array([n, fmt, h0], np.uint32).tobytes())
        f.write(np.array([row.size], np.uint64).tobytes())
        f.write(row.astype(np.uint32).tobytes())
        f.write(hd.tobytes())
    r = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, timeout=900,
                       env=dict(os.environ, EMU_SEED=str(seed)))
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    data = np.fromfile(out, np.uint8)
    off = data[:8 * (n + 1)].view(np.uint64)
    ent = data[8 * (n + 1):].view(np.uint32).reshape(-1, 2)
    return off, ent[:, 0], ent[:, 1]


@pytest.mark.parametrize("fmt", [SEG_F5, SEG_F8])
def test_segment_build_sorts_rows(emulator, tmp_path, fmt):
    rng = np.random.default_rng(17 + fmt)
    # empty rows, rows of 1-16, 17-4096 and one past 4,096 entries
    lengths = np.concatenate([np.zeros(40, int), rng.integers(1, 17, 300), rng.integers(17, 600, 12),
                              [4096, 4097, 5000]])
    rng.shuffle(lengths)
    n = lengths.size
    h0 = 1000
    maxd = 1 << 16 if fmt == SEG_F5 else 1 << 31
    row, hub, dist = batch(rng, n, lengths, h0, 6000, maxd)
    off, got_hub, got_dist = run(emulator, tmp_path, n, fmt, h0, row, hub, dist)
    # expected: the labels sorted by (row, hub), the offsets their running counts
    order = np.lexsort((hub, row))
    want_off = np.concatenate([[0], np.cumsum(np.bincount(row, minlength=n))]).astype(np.uint64)
    assert np.array_equal(off, want_off)
    assert np.array_equal(got_hub, hub[order])
    assert np.array_equal(got_dist, dist[order])


def test_segment_build_empty_batch(emulator, tmp_path):
    z = np.zeros(0, np.uint32)
    off, h, d = run(emulator, tmp_path, 50, SEG_F5, 7, z, z, z)
    assert np.array_equal(off, np.zeros(51, np.uint64)) and h.size == 0
