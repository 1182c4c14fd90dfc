"""The tree kernel's own source (plant_amd/csrc/tree.hip), compiled for the host
with g++ against tests/emu/hip/hip_runtime.h and run as one persistent workgroup
of 64 or 256 lanes (cooperative fibers that switch at barriers and, at
random, at atomics). Its PLaNT labels must equal the oracle's canonical labels, and
every search cell must be consistent with its slot's counts afterwards: the
vertex's rank kept, no state newer than the slot's last tree, no stamp at or
above the slot's next stamp base.

This checks the kernel's indexing, bucket bookkeeping, deferred pile, queueing
stamps, epoch keys (including the in-place reset of a slot whose counts reach
their bounds), early stop and emission without a GPU. It is not a model of the GPU
memory system: the GPU tests remain the parity tests proper."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "emu", "tree_emu.cpp")
TREE_FLAGS = 1


def compile_emulator(tmp_path_factory, src, name, defines=()):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("emu") / name)
    cmd = ["g++", "-std=c++20", "-O2", "-pthread", *defines, "-I", os.path.join(HERE, "emu"), "-x", "c++", src,
           "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return exe


@pytest.fixture(scope="module")
def emulator(tmp_path_factory):
    return compile_emulator(tmp_path_factory, SRC, "tree_emu")


def run_emulator(exe, tmp_path, g, order, threads, delta, slots=2, seed=12345, env_extra=None):
    off, head, ln = g.csr(1)
    n = g.n
    arc = (head.astype(np.uint64) << np.uint64(32)) | ln.astype(np.uint32).astype(np.uint64)
    rank = np.empty(n, np.uint32)
    rank[order] = np.arange(n, dtype=np.uint32)
    roots = np.asarray(order, np.uint32)  # every root, in rank order
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.array([n, arc.size, n, delta, threads, slots, TREE_FLAGS], np.uint32).tobytes())
        f.write(off.astype(np.uint32).tobytes())
        f.write(arc.tobytes())
        f.write(rank.tobytes())
        f.write(roots.tobytes())
    env = dict(os.environ, EMU_SEED=str(seed), **(env_extra or {}))
    r = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    data = np.fromfile(out, np.uint8)
    count = int(data[:8].view(np.uint64)[0])
    pairs = data[8:8 + 16 * count].view(np.uint64).reshape(-1, 2)
    restored = int(data[8 + 16 * count:].view(np.uint32)[0])
    # key = (vertex rank << 32) | hub rank, ent = (hub rank << 32) | dist
    vrank = (pairs[:, 0] >> np.uint64(32)).astype(np.int64)
    hub = (pairs[:, 1] >> np.uint64(32)).astype(np.uint32)
    dist = (pairs[:, 1] & np.uint64(0xFFFFFFFF)).astype(np.int32)
    vert = np.asarray(order, np.int64)[vrank]
    srt = np.lexsort((hub, vert))
    cnt = np.bincount(vert, minlength=n)
    offs = np.concatenate([[0], np.cumsum(cnt)]).astype(np.uint64)
    return (offs, hub[srt], dist[srt]), restored


def cases(chl):
    g = chl.Graph.grid(6, 6, 1, 30, seed=4)  # small weights: equal-length paths
    yield "grid", g, g.order_betweenness(16, seed=2)
    rng = np.random.default_rng(9)
    n, m = 40, 160
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    d = chl.Graph.from_arcs(n, src, dst, rng.integers(1, 12, m).astype(np.int32))
    yield "directed", d, rng.permutation(n).astype(np.uint32)


@pytest.mark.parametrize("threads", [64, 256])
def test_tree_kernel_emulated_equals_oracle(chl, orc, emulator, tmp_path, threads):
    for name, g, order in cases(chl):
        src, dst, w = g.arc_list(1)
        ref, _ = orc.gll(orc.OGraph.from_arcs(g.n, src, dst, w), order, threads=4)
        want = ref.copy(0)
        # a narrow bucket (many rounds, deferrals and pile entries) and a wide one
        for delta in (3, 64):
            got, restored = run_emulator(emulator, tmp_path, g, order, threads, delta)
            assert restored == 1, (name, delta)
            for a, b, what in zip(got, want, ("off", "hub", "dist")):
                assert np.array_equal(np.asarray(a), np.asarray(b)), (name, delta, what)


def wide_star(chl, leaves=300, seed=5):
    """A centre joined to every leaf and the leaves joined in a ring: a tree that
    reaches the centre queues every leaf in one round."""
    rng = np.random.default_rng(seed)
    n = leaves + 1
    lv = np.arange(1, n, dtype=np.uint32)
    nxt = np.roll(lv, -1)
    src = np.concatenate([np.zeros(leaves, np.uint32), lv, lv, nxt])
    dst = np.concatenate([lv, np.zeros(leaves, np.uint32), nxt, lv])
    w = rng.integers(1, 9, src.size).astype(np.int32)
    g = chl.Graph.from_arcs(n, src, dst, w)
    return g, rng.permutation(n).astype(np.uint32)


# (LDS bound, smallest frontier sorted): every frontier (the default bound);
# no frontier; every one up to 128 entries (the star's
# widest rounds run unsorted); every one that fits one wavefront's block; only
# the star's widest
@pytest.mark.parametrize("sort,threads", [((None, 2), 256), ((None, 0), 256), ((128, 2), 256), ((64, 2), 256),
                                          ((None, 129), 256), ((None, 2), 1024)])
def test_frontier_sort_wide_frontiers(chl, orc, tmp_path_factory, tmp_path, sort, threads):
    """Frontiers of up to 300 entries (sorts of up to 512 entries: bitonic stages
    across wavefronts behind workgroup barriers, within one wavefront's block
    behind a wavefront sync), sorted in LDS before they are processed or left
    unsorted outside the bounds: the labels must be the oracle's."""
    bound, smin = sort
    defines = () if bound is None else (f"-DCHL_FRONT_SORT={bound}",)
    exe = compile_emulator(tmp_path_factory, SRC, f"tree_emu_sort_{bound}", defines)
    # (the defaults on a 300-leaf star, whose 512-entry sorts caught a missing
    # barrier after a cross-wavefront stage; the others on 150 leaves, to stay quick)
    full = sort == (None, 2) and threads == 256
    g, order = wide_star(chl, leaves=300 if full else 150)
    src, dst, w = g.arc_list(1)
    ref, _ = orc.gll(orc.OGraph.from_arcs(g.n, src, dst, w), order, threads=4)
    # both bucket widths for the defaults, a wide one (larger frontiers) otherwise
    for delta in (2, 64) if full else (64,):
        got, restored = run_emulator(exe, tmp_path, g, order, threads, delta, env_extra={"EMU_SORT_MIN": str(smin)})
        assert restored == 1, (sort, delta)
        for a, b, what in zip(got, ref.copy(0), ("off", "hub", "dist")):
            assert np.array_equal(np.asarray(a), np.asarray(b)), (sort, delta, what)


# (grid seed, bucket width, schedule seed): schedules in which a vertex put on the
# deferred pile is improved below the bucket bound in the same round. A kernel
# that used one stamp for piling and queueing lost that improvement here (one
# extra label); these three were found among 72 schedules of that kernel.
PILE_RACES = [(2, 100, 2), (2, 150, 1), (8, 100, 3)]


@pytest.mark.parametrize("gseed,delta,seed", PILE_RACES)
def test_pile_race_schedules(chl, orc, emulator, tmp_path, gseed, delta, seed):
    g = chl.Graph.grid(10, 10, 1, 100, seed=gseed)
    order = np.random.default_rng(gseed).permutation(g.n).astype(np.uint32)
    src, dst, w = g.arc_list(1)
    ref, _ = orc.gll(orc.OGraph.from_arcs(g.n, src, dst, w), order, threads=4)
    got, restored = run_emulator(emulator, tmp_path, g, order, 64, delta, seed=seed)
    assert restored == 1
    for a, b in zip(got, ref.copy(0)):
        assert np.array_equal(np.asarray(a), np.asarray(b))


EPOCH_MAX = (1 << 30) - 1
STAMP_MAX = 0x7FFFFFFF


@pytest.mark.parametrize("start", [{"EMU_EPOCH0": str(EPOCH_MAX - 3)}, {"EMU_SBASE0": str(STAMP_MAX - 40)},
                                   {"EMU_EPOCH0": str(EPOCH_MAX), "EMU_SBASE0": str(STAMP_MAX + 1)}])
def test_slot_counts_at_their_bounds(chl, orc, emulator, tmp_path, start):
    """A slot whose tree count or stamp base reaches its bound resets its cells in
    place and restarts both: the labels stay canonical."""
    for name, g, order in cases(chl):
        src, dst, w = g.arc_list(1)
        ref, _ = orc.gll(orc.OGraph.from_arcs(g.n, src, dst, w), order, threads=4)
        got, restored = run_emulator(emulator, tmp_path, g, order, 64, 3, env_extra=start)
        assert restored == 1, (name, start)
        for a, b, what in zip(got, ref.copy(0), ("off", "hub", "dist")):
            assert np.array_equal(np.asarray(a), np.asarray(b)), (name, start, what)


SMALL_SRC = os.path.join(HERE, "emu", "small_emu.cpp")


@pytest.fixture(scope="module")
def small_emulator(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("emu") / "small_emu")
    cmd = ["g++", "-std=c++20", "-O2", "-pthread", "-I", os.path.join(HERE, "emu"), "-x", "c++", SMALL_SRC, "-o",
           exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return exe


def run_small(exe, tmp_path, g, order, delta, cap, slots, bigstep, seed=7, medium=(0, 0, 0)):
    off, head, ln = g.csr(1)
    n = g.n
    arc = (head.astype(np.uint64) << np.uint64(32)) | ln.astype(np.uint32).astype(np.uint64)
    rank = np.empty(n, np.uint32)
    rank[order] = np.arange(n, dtype=np.uint32)
    roots = np.asarray(order, np.uint32)
    inp, out = tmp_path / "sin.bin", tmp_path / "sout.bin"
    with open(inp, "wb") as f:
        f.write(np.array([n, arc.size, n, delta, cap, slots, bigstep, *medium], np.uint32).tobytes())
        f.write(off.astype(np.uint32).tobytes())
        f.write(arc.tobytes())
        f.write(rank.tobytes())
        f.write(roots.tobytes())
    r = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, EMU_SEED=str(seed)))
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    data = np.fromfile(out, np.uint8)
    count = int(data[:8].view(np.uint64)[0])
    pairs = data[8:8 + 16 * count].view(np.uint64).reshape(-1, 2)
    done_small, done_medium, empty = data[8 + 16 * count:].view(np.uint32)
    vrank = (pairs[:, 0] >> np.uint64(32)).astype(np.int64)
    hub = (pairs[:, 1] >> np.uint64(32)).astype(np.uint32)
    dist = (pairs[:, 1] & np.uint64(0xFFFFFFFF)).astype(np.int32)
    vert = np.asarray(order, np.int64)[vrank]
    srt = np.lexsort((hub, vert))
    cnt = np.bincount(vert, minlength=n)
    offs = np.concatenate([[0], np.cumsum(cnt)]).astype(np.uint64)
    return (offs, hub[srt], dist[srt]), int(done_small), int(done_medium), int(empty)


# (table entries, slots, every k-th root marked big): a table too small for the
# top roots' trees (they are abandoned and handed over), and big roots skipped
@pytest.mark.parametrize("cap,slots,bigstep", [(64, 2, 0), (256, 3, 5), (4096, 1, 0)])
def test_small_tree_kernel_then_tree_kernel_equal_oracle(chl, orc, small_emulator, tmp_path, cap, slots, bigstep):
    """The small-tree kernel (small.hip) over every root, then the tree kernel over
    the roots it handed over (marked big, or trees that outgrew half a table), as
    the label build runs them: the labels must be the oracle's canonical labels,
    and every small table must be empty afterwards."""
    for name, g, order in cases(chl):
        src, dst, w = g.arc_list(1)
        ref, _ = orc.gll(orc.OGraph.from_arcs(g.n, src, dst, w), order, threads=4)
        for delta in (3, 64):
            got, done_small, _, empty = run_small(small_emulator, tmp_path, g, order, delta, cap, slots, bigstep)
            assert empty == 1, (name, cap, delta)
            assert done_small > 0, (name, cap, delta)
            if cap == 64:
                assert done_small < g.n  # some trees outgrew the table and went to the tree kernel
            for a, b, what in zip(got, ref.copy(0), ("off", "hub", "dist")):
                assert np.array_equal(np.asarray(a), np.asarray(b)), (name, cap, delta, what)


# (small table, medium table, every k-th root big for the small tables, for the
# medium ones): small tables that hand most trees over, medium tables too small
# for the top trees (abandoned there as well), and roots kept from both
@pytest.mark.parametrize("cap,cap2,bigstep,bigstep2", [(64, 128, 0, 0), (64, 4096, 3, 0), (64, 512, 2, 6)])
def test_small_then_medium_then_tree_kernel_equal_oracle(chl, orc, small_emulator, tmp_path, cap, cap2, bigstep,
                                                         bigstep2):
    """Small tables one wavefront per root, then medium tables 256 threads per root
    over the roots handed over, then the tree kernel over the rest: the labels
    must be the oracle's canonical labels and every table empty afterwards."""
    for name, g, order in cases(chl):
        src, dst, w = g.arc_list(1)
        ref, _ = orc.gll(orc.OGraph.from_arcs(g.n, src, dst, w), order, threads=4)
        for delta in (3, 64):
            got, done_small, done_medium, empty = run_small(small_emulator, tmp_path, g, order, delta, cap, 2,
                                                            bigstep, medium=(cap2, 2, bigstep2))
            assert empty == 1, (name, cap2, delta)
            assert done_medium > 0, (name, cap2, delta)
            for a, b, what in zip(got, ref.copy(0), ("off", "hub", "dist")):
                assert np.array_equal(np.asarray(a), np.asarray(b)), (name, cap2, delta, what)
