"""The label-store kernels of plant_amd/csrc/labels.hip, compiled for the host
with g++ against tests/emu/hip/hip_runtime.h and run with every workgroup's
lanes as cooperative fibers: the segment build (row count, scan, cursor copy,
scatter, per-row sort, launched as engine.cpp's build_segment launches them),
the merge of adjacent segments (store.cpp's merge_range), the GLL cleaning of
a batch's candidates (clean_side) and the compaction after it (compact_side). A batch's labels arrive in the tree kernel's
emission order (any order); every segment must hold each row's labels
hub-ascending, with the row offsets of the counts, in the 5-byte (hub offset,
distance) and the 8-byte entry formats.

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


def run(exe, tmp_path, n, batches, mode=0, out_fmt=SEG_F8, seed=3):
    """batches: (fmt, h0, row, hub, dist) each; mode 0 returns the first batch's
    segment, 1 the merge of all, 2 the first compacted by keep_flags, 3 the
    first segment with its keep flags after cleaning against the second."""
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.array([n, len(batches), mode, out_fmt], np.uint32).tobytes())
        for fmt, h0, row, hub, dist in batches:
            hd = (hub.astype(np.uint64) << np.uint64(32)) | dist.astype(np.uint64)
            f.write(np.array([fmt, h0], np.uint32).tobytes())
            f.write(np.array([row.size], np.uint64).tobytes())
            f.write(row.astype(np.uint32).tobytes())
            f.write(hd.tobytes())
    r = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, timeout=900,
                       env=dict(os.environ, EMU_SEED=str(seed)))
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    data = np.fromfile(out, np.uint8)
    off = data[:8 * (n + 1)].view(np.uint64)
    m = int(off[-1])
    ent = data[8 * (n + 1):8 * (n + 1) + 8 * m].view(np.uint32).reshape(-1, 2)
    if mode == 3:
        return off, ent[:, 0], ent[:, 1], data[8 * (n + 1) + 8 * m:]
    return off, ent[:, 0], ent[:, 1]


def keep_flags(row, hub):
    """The compaction test's keep rule (the harness computes the same)."""
    return ((hub.astype(np.uint64) * np.uint64(2654435761)) % np.uint64(1 << 32)
            ^ row.astype(np.uint64)) % np.uint64(3) != 0


def segment_of(n, row, hub, dist):
    """Expected segment: the labels sorted by (row, hub), the offsets their counts."""
    order = np.lexsort((hub, row))
    off = np.concatenate([[0], np.cumsum(np.bincount(row, minlength=n))]).astype(np.uint64)
    return off, hub[order], dist[order]


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
    got = run(emulator, tmp_path, n, [(fmt, h0, row, hub, dist)])
    for a, b in zip(got, segment_of(n, row, hub, dist)):
        assert np.array_equal(a, b)


def test_segment_build_empty_batch(emulator, tmp_path):
    z = np.zeros(0, np.uint32)
    off, h, d = run(emulator, tmp_path, 50, [(SEG_F5, 7, z, z, z)])
    assert np.array_equal(off, np.zeros(51, np.uint64)) and h.size == 0


# (formats of three adjacent batches' segments, format of the merged segment)
@pytest.mark.parametrize("fmts,out_fmt", [((SEG_F5, SEG_F5, SEG_F5), SEG_F5), ((SEG_F5, SEG_F8, SEG_F5), SEG_F8)])
def test_segments_merge_into_one(emulator, tmp_path, fmts, out_fmt):
    """Segments of adjacent hub ranges concatenate row by row into one segment
    whose rows are still hub-ascending (store.cpp's merge_range)."""
    rng = np.random.default_rng(5)
    n = 120
    parts, h0 = [], 100
    for fmt in fmts:
        lengths = rng.integers(0, 40, n)
        lengths[rng.integers(0, n, 3)] = rng.integers(100, 300, 3)
        # (a batch gets the 5-byte form only when its distances fit 16 bits)
        maxd = 1 << 16 if fmt == SEG_F5 else 1 << 31
        row, hub, dist = batch(rng, n, lengths, h0, 400, maxd)
        parts.append((fmt, h0, row, hub, dist))
        h0 += 400  # the next batch's hubs all rank below
    got = run(emulator, tmp_path, n, parts, mode=1, out_fmt=out_fmt)
    row = np.concatenate([p[2] for p in parts])
    hub = np.concatenate([p[3] for p in parts])
    dist = np.concatenate([p[4] for p in parts])
    for a, b in zip(got, segment_of(n, row, hub, dist)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("fmt", [SEG_F5, SEG_F8])
def test_segment_compaction_keeps_order(emulator, tmp_path, fmt):
    """Dropping the covered candidates (compact_side) keeps the kept entries of
    every row, in row and hub order, with the new row offsets."""
    rng = np.random.default_rng(9 + fmt)
    n = 150
    lengths = rng.integers(0, 70, n)
    lengths[rng.integers(0, n, 2)] = 200  # rows spanning several 64-entry chunks
    row, hub, dist = batch(rng, n, lengths, 50, 300, 1 << 16)
    got = run(emulator, tmp_path, n, [(fmt, 50, row, hub, dist)], mode=2)
    k = keep_flags(row, hub)
    assert 0 < k.sum() < k.size
    for a, b in zip(got, segment_of(n, row[k], hub[k], dist[k])):
        assert np.array_equal(a, b)


def clean_cover(off, hub, dist, ooff, ohub, odist):
    """shared_memory/labeling.hpp:78-88 by brute force, one candidate at a time:
    candidate (h, d) of row v is covered when some hub x < h is in both L(v) and
    the other side's L(h) with L(v)[x] + L'(h)[x] <= d. Rows are taken as sets, so
    neither the entry order nor a merge is assumed."""
    keep = np.ones(hub.size, bool)
    for v in range(off.size - 1):
        b, e = int(off[v]), int(off[v + 1])
        mine = dict(zip(hub[b:e].tolist(), dist[b:e].tolist()))
        for c in range(b, e):
            h, d = int(hub[c]), int(dist[c])
            ob, oe = int(ooff[h]), int(ooff[h + 1])
            other = dict(zip(ohub[ob:oe].tolist(), odist[ob:oe].tolist()))
            keep[c] = not any(x < h and x in other and dx + other[x] <= d for x, dx in mine.items())
    return keep


@pytest.mark.parametrize("fmt", [SEG_F5, SEG_F8])
def test_cleaning_matches_clean_cover(emulator, tmp_path, fmt):
    """The batch's candidates of one side are cleaned against the other side's
    batch rows (clean_side): the keep flags equal clean_cover's, with rows staged
    in LDS (up to 1,024 entries) and read from the segment (longer)."""
    rng = np.random.default_rng(21 + fmt)
    n, h0, nhub = 1600, 0, 1300  # hubs are ranks in [0, 1300), rows are ranks in [0, 1600)
    lengths = rng.integers(0, 25, n)
    lengths[rng.integers(0, n, 2)] = [1100, 1200]  # past the LDS stage
    lengths[rng.integers(0, nhub, 40)] = rng.integers(60, 200, 40)  # long other rows of some hubs
    row, hub, dist = batch(rng, n, lengths, h0, nhub, 200)
    lengths2 = rng.integers(0, 25, n)
    lengths2[rng.integers(0, nhub, 40)] = rng.integers(60, 200, 40)
    row2, hub2, dist2 = batch(rng, n, lengths2, h0, nhub, 200)
    off, got_hub, got_dist, keep = run(emulator, tmp_path, n, [(fmt, h0, row, hub, dist), (fmt, h0, row2, hub2, dist2)],
                                       mode=3)
    want = segment_of(n, row, hub, dist)
    for a, b in zip((off, got_hub, got_dist), want):
        assert np.array_equal(a, b)
    o2 = segment_of(n, row2, hub2, dist2)
    want_keep = clean_cover(*want, *o2)
    assert 0 < (~want_keep).sum() < want_keep.size  # some covered, some kept
    assert np.array_equal(keep.astype(bool), want_keep)
