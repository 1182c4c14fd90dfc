"""The query kernel's own source (plant_amd/csrc/query.hip), compiled for the host
with g++ against tests/emu/hip/hip_runtime.h, on label sets made here: random
hub-ascending rows over several segments (disjoint hub ranges), in the 5-byte
and the 8-byte entry forms and mixed, with rows long enough to span several tiles
and chunks, and pairs including out-of-range ids. Every distance must equal the
min-plus merge-join of the two rows (Labeling::query, shared_memory/labeling.hpp:
56-65): the smallest L1(u)[h] + L0(v)[h] over common hubs, INT_MAX when there is
none, -1 for an id out of range.

This checks the kernel's tiling, the resumption of the second row's scan across
tiles, chunk layouts and segment handling without a GPU; the GPU tests stay the
parity tests proper."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "emu", "query_emu.cpp")
INF = 2147483647


def build(tmp_path_factory, tile):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("qemu") / f"query_emu_{tile}")
    cmd = ["g++", "-std=c++20", "-O2", "-pthread", f"-DCHL_QUERY_TILE={tile}", "-I", os.path.join(HERE, "emu"),
           "-x", "c++", SRC, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return exe


@pytest.fixture(scope="module")
def emulators(tmp_path_factory):
    return {t: build(tmp_path_factory, t) for t in (256, 512)}


def make_side(rng, n, bounds, fmts, long_rows, dmax):
    """Rows of one side: per segment s, per vertex, a sorted random hub subset of
    [bounds[s], bounds[s+1]) with distances in [0, dmax]."""
    segs = []
    for s, fmt in enumerate(fmts):
        h0, h1 = bounds[s], bounds[s + 1]
        rows = []
        for v in range(n):
            if v in long_rows:
                k = min(h1 - h0, int(rng.integers(900, 1700)))
            else:
                k = int(rng.integers(0, min(h1 - h0, 40) + 1))
            hubs = np.sort(rng.choice(np.arange(h0, h1), size=k, replace=False)).astype(np.uint32)
            dist = rng.integers(0, dmax + 1, size=k).astype(np.uint32)
            rows.append((hubs, dist))
        segs.append((fmt, h0, rows))
    return segs


def encode(f, n, segs):
    for fmt, h0, rows in segs:
        off = np.zeros(n + 1, np.uint64)
        off[1:] = np.cumsum([len(h) for h, _ in rows])
        hubs = np.concatenate([h for h, _ in rows]) if rows else np.zeros(0, np.uint32)
        dist = np.concatenate([d for _, d in rows]) if rows else np.zeros(0, np.uint32)
        f.write(np.array([fmt, h0], np.uint32).tobytes())
        f.write(np.array([hubs.size], np.uint64).tobytes())
        f.write(off.tobytes())
        if fmt == 5:
            assert np.all(dist <= 65535) and np.all(hubs - h0 < (1 << 24))
            key = ((hubs - np.uint32(h0)) << np.uint32(8)) | (dist >> np.uint32(8))
            f.write(key.astype(np.uint32).tobytes())
            f.write((dist & np.uint32(0xFF)).astype(np.uint8).tobytes())
        else:
            ent = (hubs.astype(np.uint64) << np.uint64(32)) | dist.astype(np.uint64)
            f.write(ent.tobytes())


def expected(n, side0, side1, pairs):
    row0 = [dict() for _ in range(n)]
    row1 = [dict() for _ in range(n)]
    for segs, rows_out in ((side0, row0), (side1, row1)):
        for _, _, rows in segs:
            for v, (h, d) in enumerate(rows):
                rows_out[v].update(zip(h.tolist(), d.tolist()))
    out = []
    for u, v in pairs.tolist():
        if u >= n or v >= n:
            out.append(-1)
            continue
        a, b = row1[u], row0[v]
        if len(a) > len(b):
            a, b = b, a
        best = INF
        for h, d in a.items():
            if h in b:
                best = min(best, d + b[h])
        out.append(best)
    return np.array(out, np.int64)


@pytest.mark.parametrize("tile", [256, 512])
@pytest.mark.parametrize("fmts,seed", [((5,), 1), ((5, 5, 5, 5), 2), ((8, 8), 3), ((5, 8, 5), 4), ((8,), 5)])
def test_query_kernel_matches_merge_join(emulators, tmp_path, tile, fmts, seed):
    rng = np.random.default_rng(seed)
    n = 24
    ks = len(fmts)
    bounds = np.linspace(0, 6000 * ks, ks + 1).astype(np.int64).tolist()
    long_rows = set(rng.choice(n, 6, replace=False).tolist())
    dmax = 65535 if all(f == 5 for f in fmts) else (1 << 30)
    dmax5 = 65535
    side0 = make_side(rng, n, bounds, fmts, long_rows, dmax5 if 5 in fmts else dmax)
    side1 = make_side(rng, n, bounds, fmts, long_rows, dmax5 if 5 in fmts else dmax)
    pairs = np.array([(u, v) for u in range(n) for v in range(n)], np.uint32)
    pairs = np.concatenate([pairs, np.array([[n, 0], [0, n + 5], [2 ** 32 - 1, 1]], np.uint32)])
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.array([n, ks, pairs.shape[0]], np.uint32).tobytes())
        encode(f, n, side0)
        encode(f, n, side1)
        f.write(pairs.tobytes())
    r = subprocess.run([emulators[tile], str(inp), str(outp)], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, EMU_SEED=str(seed)))
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    got = np.fromfile(outp, np.int32).astype(np.int64)
    want = expected(n, side0, side1, pairs)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(pairs[i].tolist(), int(got[i]), int(want[i])) for i in bad[:5]]
    assert np.any((want != INF) & (want >= 0))  # some pairs share hubs
