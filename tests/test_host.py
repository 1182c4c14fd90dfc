"""Host-side parts of libchl.so that need no GPU: graph readers (against the
oracle's restatement of shared_memory/graph.hpp, itself checked through the
reference's own hhl), generators, order files, and the exported C ABI."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc_mod

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DEGREE_TOOL = os.path.join(os.path.dirname(GOLD), "..", "plant_amd", "bin", "degree")
REF_DEGREE = os.path.join(os.path.dirname(orc_mod.REF_HHL), "degree")


def golden_manifest():
    with open(os.path.join(GOLD, "manifest.json")) as f:
        return json.load(f)

FIXTURES = {
    # duplicate arcs (min length kept), unequal reverse lengths, a self-loop, comments
    "dimacs": "c tiny\np sp 5 8\na 1 2 7\na 1 2 3\na 2 1 3\na 2 3 4\na 3 2 9\na 3 3 2\nc mid\na 4 5 1\na 5 4 1\n",
    # weighted METIS (fmt 1): vertex lines of (neighbour, length)
    "metis": "% comment\n4 4 1\n2 5 3 7\n1 5\n1 7 4 2\n3 2\n",
    "metis_unweighted": "3 2\n2\n1 3\n2\n",
    "snap": "# Directed graph\n# Nodes: 4 Edges: 4\n0 1\n1 2\n2 3\n3 0\n",
    "edgelist": "0 1 5\n1 2 7\n2 0 1\n",
}


def write(tmp_path, name, text):
    p = tmp_path / f"{name}.txt"
    p.write_text(text)
    return str(p)


@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_reader_matches_oracle(chl, orc, tmp_path, name):
    path = write(tmp_path, name, FIXTURES[name])
    g = chl.Graph.load(path)
    og, fmt = orc.OGraph.read(path)
    assert og is not None and fmt == g.format
    assert g.n == og.n
    for rng in (0, 1):
        for a, b in zip(g.csr(rng), og.csr(rng)):
            assert np.array_equal(a, b)


def test_dimacs_keeps_min_and_unequal_reverse(chl, tmp_path):
    g = chl.Graph.load(write(tmp_path, "d", FIXTURES["dimacs"]))
    off, head, ln = g.csr(1)
    row0 = list(zip(head[off[0]:off[1]], ln[off[0]:off[1]]))
    assert row0 == [(1, 3)]  # a 1 2 7 and a 1 2 3: min kept
    row2 = list(zip(head[off[2]:off[3]], ln[off[2]:off[3]]))
    assert row2 == [(1, 9), (2, 2)]  # 3->2 has length 9 while 2->3 has 4; self-loop kept
    assert not g.symmetric


def test_unterminated_dimacs_rejected(chl, orc, tmp_path):
    path = write(tmp_path, "bad", "p sp 2 1\na 1 2 3")
    with pytest.raises(chl.ChlError):
        chl.Graph.load(path)
    og, r = orc.OGraph.read(path)
    assert og is None and r < 0


DIMACS_LAYOUTS = {
    # the parallel reader's plain layout, and every way out of it, against the oracle
    "plain": "c x\np sp 3 3\na 1 2 4\na 2 3 5\nc y\na 3 1 6\n",
    "trailing_junk": "p sp 3 2\na 1 2 4 junk\na 2 3 5\n",
    "tabs_crlf": "p sp 3 2\r\na\t1\t2\t4\r\na 2 3 5\r\n",
    "no_space_after_a": "p sp 3 2\na1 2 4\na 2 3 5\n",
    "long_comment": "c " + "x" * 900 + "\np sp 3 2\na 1 2 4\nc " + "y" * 700 + "\na 2 3 5\n",
    "long_arc_line": "p sp 3 2\na 1 2 4" + " " * 600 + "\na 2 3 5\n",
    "line_511": "p sp 3 2\na 1 2 4" + " " * (511 - 8 - 1) + "\na 2 3 5\n",
    "line_512": "p sp 3 2\na 1 2 4" + " " * (512 - 8 - 1) + "\na 2 3 5\n",
    "blank_line": "p sp 3 2\na 1 2 4\n\na 2 3 5\n",
    "second_p": "p sp 3 2\na 1 2 4\np sp 3 2\na 2 3 5\n",
    "arc_before_p": "a 1 2 4\np sp 3 2\na 2 3 5\n",
    "missing_length": "p sp 3 2\na 1 2\na 2 3 5\n",
    "out_of_range": "p sp 3 2\na 1 4 4\na 2 3 5\n",
    "signed": "p sp 3 2\na +1 2 4\na 2 3 +5\n",
    "huge_number": "p sp 3 2\na 1 2 12345678901234567890\na 2 3 5\n",
}


@pytest.mark.parametrize("name", sorted(DIMACS_LAYOUTS))
def test_dimacs_layouts_match_oracle(chl, orc, tmp_path, name):
    path = write(tmp_path, name, DIMACS_LAYOUTS[name])
    og, r = orc.OGraph.read(path)
    try:
        g = chl.Graph.load(path)
    except chl.ChlError as e:
        if e.code == -4:
            # a length the reference keeps as a non-positive int and asserts on in
            # Dijkstra (dijkstra.hpp:91) is rejected at load time
            assert og is not None and min(og.csr(1)[2].min(), og.csr(0)[2].min()) <= 0, name
        else:
            assert og is None, name
        return
    assert og is not None and g.n == og.n, name
    for rng in (0, 1):
        for a, b in zip(g.csr(rng), og.csr(rng)):
            assert np.array_equal(a, b), name


SNAP_HEAD = "# Directed graph\n# Nodes: 4 Edges: 4\n"
LINE_LAYOUTS = {
    # the parallel METIS, SNAP and edge-list readers' plain layouts, and ways out of them,
    # against the oracle (format detected, n, arcs, or rejection)
    "snap_plain": SNAP_HEAD + "0 1\n1 2\n2 3\n3 0\n",
    "snap_junk_after_ids": SNAP_HEAD + "0 1 junk\n1 2\t7\n2 3\n3 0\n",
    "snap_crlf": "# Nodes: 4 Edges: 2\r\n0 1\r\n2 3\r\n",
    "snap_comment_in_body": SNAP_HEAD + "0 1\n# note\n2 3\n",
    "snap_second_nodes_line": SNAP_HEAD + "0 1\n# Nodes: 4 Edges: 4\n2 3\n",
    "snap_nodes_after_comments": "# a\n# b\n# Nodes: 3 Edges: 1\n# c\n0 2\n",
    "snap_no_nodes_line": "# Directed graph\n0 1\n1 2\n",
    "snap_out_of_range": SNAP_HEAD + "0 4\n",
    "snap_negative": SNAP_HEAD + "0 -1\n",
    "snap_blank_line": SNAP_HEAD + "0 1\n\n2 3\n",
    "snap_no_final_newline": SNAP_HEAD + "0 1\n2 3",
    "snap_long_comment": "# " + "x" * 700 + "\n# Nodes: 4 Edges: 1\n0 1\n",
    "snap_long_line": SNAP_HEAD + "0 1" + " " * 600 + "\n2 3\n",
    "snap_line_511": SNAP_HEAD + "0 1" + " " * (511 - 4) + "\n2 3\n",
    "snap_line_512": SNAP_HEAD + "0 1" + " " * (512 - 4) + "\n2 3\n",
    "snap_huge_id": SNAP_HEAD + "0 12345678901234567890\n",
    "snap_header_only": "# Nodes: 3 Edges: 0\n",
    "el_plain": "0 1 5\n1 2 7\n2 0 3\n",
    "el_fmt_like_first_line": "0 1 1\n1 2 7\n2 0 3\n",
    "el_fmt_like_10": "2 0 10\n0 1 4\n1 2 5\n",
    "el_signed_first": "+0 1 1\n1 2 7\n2 0 3\n",
    "el_junk_after": "0 1 5 x\n1 2 7 y\n2 0 3\n",
    "el_gap_in_ids": "0 1 5\n1 3 7\n3 0 3\n",
    "el_zero_weight": "0 1 5\n1 2 0\n2 0 3\n",
    "el_negative_id": "0 1 5\n-1 2 7\n2 0 3\n",
    "el_two_numbers": "0 1 5\n1 2\n2 0 3\n",
    "el_no_final_newline": "0 1 5\n1 2 7\n2 0 3",
    "el_long_line": "0 1 5" + " " * 600 + "\n1 2 7\n2 0 3\n",
    "el_line_512": "0 1 5" + " " * (512 - 6) + "\n1 2 7\n2 0 3\n",
    "el_huge_weight": "0 1 5\n1 2 3000000000\n2 0 3\n",
    "el_tabs": "0\t1\t5\n1\t2\t7\n2\t0\t3\n",
    "el_big_id": "0 1 5\n1 3000000000 7\n",
    # METIS (the parallel reader's plain layout and ways out of it)
    "metis_plain": "% c\n4 4 1\n2 5 3 7\n1 5\n1 7 4 2\n3 2\n",
    "metis_unweighted": "3 2\n2\n1 3\n2\n",
    "metis_comments_between": "3 2\n% x\n2\n  % y\n1 3\n2\n",
    "metis_blank_vertex_line": "3 1\n2\n1\n\n",
    "metis_fewer_lines": "5 1\n2\n1\n",
    "metis_extra_blank_lines": "2 1\n2\n1\n\n\n",
    "metis_extra_vertex_line": "2 1\n2\n1\n1\n",
    "metis_vertex_weights": "3 2 10\n9 2\n9 1 3\n9 2\n",
    "metis_sizes_weights_edges": "3 2 111\n4 9 2 6\n4 9 1 6 3 8\n4 9 2 8\n",
    "metis_ncon": "3 2 10 2\n9 9 2\n9 9 1 3\n9 9 2\n",
    "metis_ncon_without_flag": "3 2 1 2\n2 5\n1 5\n",
    "metis_bad_fmt": "3 2 2\n2\n1 3\n2\n",
    "metis_odd_pairs": "3 2 1\n2 5 3\n1 5\n2 4\n",
    "metis_zero_weight": "3 2 1\n2 0\n1 0 3 4\n2 4\n",
    "metis_out_of_range": "3 2\n4\n1 3\n2\n",
    "metis_zero_id": "3 2\n0\n1 3\n2\n",
    "metis_tabs_crlf": "3 2\r\n2\r\n1\t3\r\n2\r\n",
    "metis_percent_mid_line": "3 2\n2 % x\n1 3\n2\n",  # the next line continues vertex 1
    "metis_percent_mid_token": "3 2\n2\n1%x\n3\n2\n",
    "metis_percent_weighted": "3 2 1\n2 % x\n5\n1 5 3 4\n2 4\n",
    "metis_odd_pairs_at_eof": "3 2 1\n2 5\n1 5 3 4\n2 4 1",
    "metis_too_few_at_eof": "3 2 10\n9 2\n9 1 3\n",
    "metis_header_comment": "3 2 % c\n2\n1 3\n2\n",
    "metis_letter": "3 2\n2\n1 x3\n2\n",
    "metis_no_final_newline": "3 2\n2\n1 3\n2",
    "metis_blank_before_header": "\n3 2\n2\n1 3\n2\n",
    "metis_huge": "3 2\n2\n1 123456789012345678901\n2\n",
    "metis_n_zero": "0 0\n",
}


def check_against_oracle(chl, orc, path, name):
    og, r = orc.OGraph.read(path)
    try:
        g = chl.Graph.load(path)
    except chl.ChlError as e:
        if e.code == -4:
            assert og is not None and min(og.csr(1)[2].min(), og.csr(0)[2].min()) <= 0, name
        else:
            assert og is None or og.n == 0, name
        return
    assert og is not None and g.n == og.n and g.format == r, name
    for rng in (0, 1):
        for a, b in zip(g.csr(rng), og.csr(rng)):
            assert np.array_equal(a, b), name


@pytest.mark.parametrize("name", sorted(LINE_LAYOUTS))
def test_line_format_layouts_match_oracle(chl, orc, tmp_path, name):
    check_against_oracle(chl, orc, write(tmp_path, name, LINE_LAYOUTS[name]), name)


def test_line_formats_at_size_match_oracle(chl, orc, tmp_path):
    """Files of 20K to 300K lines, read in ranges on every host thread."""
    rng = np.random.default_rng(3)
    n, m = 50000, 300000
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    u[:n] = np.arange(n)  # every id present
    w = rng.integers(2, 1000, m)
    w[0] = 7  # a first line the METIS header check rejects
    el = tmp_path / "big.el"
    el.write_text("".join(f"{a} {b} {c}\n" for a, b, c in zip(u, v, w)))
    check_against_oracle(chl, orc, str(el), "edgelist")
    assert chl.Graph.load(str(el)).format == 4
    sn = tmp_path / "big.snap"
    sn.write_text(f"# big\n# Nodes: {n} Edges: {m}\n" + "".join(f"{a}\t{b}\n" for a, b in zip(u, v)))
    check_against_oracle(chl, orc, str(sn), "snap")
    assert chl.Graph.load(str(sn)).format == 3
    # METIS: weighted adjacency lines of a symmetric graph, comments sprinkled
    g = chl.Graph.knn(20000, 3, seed=2)
    off, head, ln = g.csr(1)
    lines = [f"{g.n} {g.arcs(1) // 2} 1"]
    for x in range(g.n):
        if x % 997 == 0:
            lines.append("% every 997th line a comment")
        lines.append(" ".join(f"{h + 1} {w}" for h, w in zip(head[off[x]:off[x + 1]], ln[off[x]:off[x + 1]])))
    mt = tmp_path / "big.metis"
    mt.write_text("\n".join(lines) + "\n")
    check_against_oracle(chl, orc, str(mt), "metis")
    gm = chl.Graph.load(str(mt))
    assert gm.format == 2
    for a, b in zip(gm.csr(1), g.csr(1)):
        assert np.array_equal(a, b)


def test_zero_weight_rejected(chl, tmp_path):
    with pytest.raises(chl.ChlError) as e:
        chl.Graph.load(write(tmp_path, "z", "p sp 2 1\na 1 2 0\n"))
    assert e.value.code == -4


def test_oracle_readers_agree_with_reference_binary(chl, orc, tmp_path):
    """The oracle's readers against the reference's own hhl on every reader case of
    this file: a file the oracle rejects makes hhl fail, with "Unable to read graph"
    (shared_memory/hhl.cpp:46-48) or, for an edge list whose ids are not 0 .. n-1, by
    a crash (read_edgelist has no bounds check, graph.hpp:285-305); a file it accepts
    (n > 0, lengths > 0, where hhl runs to the end) gives the same labels."""
    if not os.path.exists(orc_mod.REF_HHL):
        pytest.skip("reference hhl not built (reference tree absent)")
    cases = dict(FIXTURES)
    cases.update(DIMACS_LAYOUTS)
    cases.update(LINE_LAYOUTS)
    for name, text in cases.items():
        path = write(tmp_path, name, text)
        og, _ = orc.OGraph.read(path)
        if og is None:
            r = subprocess.run([orc_mod.REF_HHL, "-g", path, "-o", "none", "-p", "2"], capture_output=True,
                               text=True, timeout=60)
            assert r.returncode != 0, name
            assert r.returncode < 0 or "Unable to read graph" in r.stderr, name
            continue
        if og.n == 0 or min(og.csr(1)[2].min(initial=1), og.csr(0)[2].min(initial=1)) <= 0:
            continue  # the reference asserts on such a graph (dijkstra.hpp:91)
        order = np.arange(og.n, dtype=np.uint32)[::-1].copy()
        op = str(tmp_path / f"{name}.order")
        chl.order_write(op, order)
        lp = str(tmp_path / f"{name}.labels")
        subprocess.run([orc_mod.REF_HHL, "-g", path, "-o", op, "-p", "2", "-l", lp], check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        L, _ = orc.gll(og, order, threads=2)
        mine = str(tmp_path / f"{name}.mine")
        L.write(mine)
        assert open(mine, "rb").read() == open(lp, "rb").read(), name


def test_generators_deterministic_and_shaped(chl):
    g1, g2 = chl.Graph.grid(20, 30, 1, 100, seed=4), chl.Graph.grid(20, 30, 1, 100, seed=4)
    assert g1.n == 600 and g1.arcs(1) == 2 * (20 * 29 + 30 * 19) and g1.symmetric
    for a, b in zip(g1.csr(1), g2.csr(1)):
        assert np.array_equal(a, b)
    _, _, ln = g1.csr(1)
    assert ln.min() >= 1 and ln.max() <= 100
    k = chl.Graph.knn(2000, 3, seed=2)
    assert k.symmetric
    off, head, ln = k.csr(1)
    assert np.diff(off).min() >= 3  # every point keeps its 3 nearest neighbours
    src = np.repeat(np.arange(k.n), np.diff(off).astype(np.int64))
    assert not np.any(src == head)
    r = chl.Graph.rmat(10, 3 * 1024, seed=1)
    off, head, ln = r.csr(1)
    src = np.repeat(np.arange(r.n), np.diff(off).astype(np.int64))
    assert r.symmetric and not np.any(src == head) and set(np.unique(ln)) == {1}
    rw = chl.Graph.rmat(10, 3 * 1024, wmax=16, seed=1)
    assert rw.symmetric and rw.csr(1)[2].max() <= 16


def test_knn_is_exact_on_small_input(chl):
    # brute-force 3-NN on the same points: the generator's edge count is the union size
    k = chl.Graph.knn(300, 3, seed=9)
    off, head, ln = k.csr(1)
    assert np.all(ln >= 1)
    assert k.arcs(1) % 2 == 0


def test_orders(chl, tmp_path):
    g = chl.Graph.grid(10, 10, 1, 10, seed=3)
    ob = g.order_betweenness(16, seed=1)
    assert sorted(ob.tolist()) == list(range(100))
    od = g.order_degree(seed=5)
    off, _, _ = g.csr(1)
    deg = np.diff(off).astype(np.int64)
    assert np.all(np.diff(deg[od]) <= 0)  # degree descending
    ref = g.order_degree(reference_ties=True)
    # distributed_memory/degree.cpp:62-67: ties by higher id first
    keys = [(-int(deg[v]), -int(v)) for v in ref]
    assert keys == sorted(keys)
    p = str(tmp_path / "o.txt")
    chl.order_write(p, ob)
    assert np.array_equal(chl.order_load(p), ob)
    assert open(p).read().split("\n")[0] == "100"


def _betweenness_all_roots(n, src, dst, w):
    """Sampled betweenness with every vertex a root, restated with scipy's Dijkstra:
    parent of v = smallest-id u with d(u) + w(u, v) == d(v); score = proper
    descendants summed over the trees."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import dijkstra
    A = csr_matrix((w.astype(np.float64), (src, dst)), shape=(n, n))
    D = dijkstra(A, directed=True)
    score = np.zeros(n, np.int64)
    for r in range(n):
        d = D[r]
        par = np.full(n, n, np.int64)
        for u, v, ww in zip(src, dst, w):
            if np.isfinite(d[u]) and d[u] + ww == d[v] and v != r and u < par[v]:
                par[v] = u
        sub = np.ones(n, np.int64)
        reach = np.flatnonzero(np.isfinite(d))
        for v in reach[np.argsort(-d[reach], kind="stable")]:
            if v != r:
                sub[par[v]] += sub[v]
        score[reach] += sub[reach] - 1
    order = np.lexsort((np.arange(n), -score))
    return order, score


def test_betweenness_order_pinned(chl):
    # closed form: on a unit path with every vertex a root, vertex k has k (n-1-k)
    # proper descendants in each tree rooted on either side, and n-1 in its own
    n = 9
    src = np.arange(n - 1, dtype=np.uint32)
    g = chl.Graph.from_arcs(n, src, src + 1, np.ones(n - 1, np.int32), undirected=True)
    order, score = g.order_betweenness(n, seed=3, with_score=True)
    k = np.arange(n)
    assert np.array_equal(score, 2 * k * (n - 1 - k) + n - 1)
    assert order.tolist() == [4, 3, 5, 2, 6, 1, 7, 0, 8]
    # against scipy's Dijkstra and the tie rule, on weighted graphs with ties,
    # a directed graph and a disconnected one
    rng = np.random.default_rng(5)
    cases = []
    gr = chl.Graph.grid(6, 7, 1, 3, seed=2)  # small weights: many equal-length paths
    cases.append((gr, *gr.arc_list(1)))
    m = 160
    s2, d2 = rng.integers(0, 40, m).astype(np.uint32), rng.integers(0, 40, m).astype(np.uint32)
    w2 = rng.integers(1, 4, m).astype(np.int32)
    gd = chl.Graph.from_arcs(40, s2, d2, w2)
    cases.append((gd, *gd.arc_list(1)))
    s3 = np.array([0, 1, 2, 4, 5], np.uint32)
    d3 = np.array([1, 2, 0, 5, 6], np.uint32)
    gu = chl.Graph.from_arcs(8, s3, d3, np.array([3, 4, 5, 1, 2], np.int32), undirected=True)
    cases.append((gu, *gu.arc_list(1)))
    for g, src, dst, w in cases:
        want_order, want_score = _betweenness_all_roots(g.n, src, dst, w)
        order, score = g.order_betweenness(g.n, seed=11, with_score=True)
        assert np.array_equal(score.astype(np.int64), want_score)
        assert np.array_equal(order, want_order)


def _read_status(chl, path, reader):
    """0 when the file parses (the labels then reach the device, or the call stops
    at the missing device on a host without one); else the error code."""
    try:
        reader(str(path))
        return 0
    except chl.ChlError as e:
        return 0 if e.code == -7 else e.code


LABEL_TEXT = "3\n2 0 0 1 4\n2 0 0 1 4\n1 1 0\n1 1 0\n2 0 7 2 0\n2 0 7 2 0\n"


def test_label_text_reader_rules(chl, tmp_path):
    # labeling.hpp:185-204 reads whitespace-separated tokens in any layout and
    # accepts only whitespace after the last label
    cases = {
        "rows": (LABEL_TEXT, 0),
        "one_line": (LABEL_TEXT.replace("\n", " "), 0),
        "split_rows": (LABEL_TEXT.replace(" 4\n", "\n4\n").replace("2 0 7", "2\n0\t7"), 0),
        "blank_lines": (LABEL_TEXT.replace("\n", "\n\n"), 0),
        "crlf": (LABEL_TEXT.replace("\n", "\r\n"), 0),
        "no_final_newline": (LABEL_TEXT.rstrip("\n"), 0),
        "trailing_blank": (LABEL_TEXT + "  \n\n", 0),
        "trailing_garbage": (LABEL_TEXT + "x\n", -3),
        "extra_token": (LABEL_TEXT + "5\n", -3),
        "negative_count": (LABEL_TEXT.replace("1 1 0\n1 1 0", "-1 1 0\n1 1 0"), -3),
        "truncated": (LABEL_TEXT[:-6], -3),
        "letter": (LABEL_TEXT.replace("0 1 4", "0 a 4", 1), -3),
        "empty": ("", -3),
    }
    for name, (text, want) in cases.items():
        p = tmp_path / f"{name}.txt"
        p.write_bytes(text.encode())
        assert _read_status(chl, p, chl.Labels.read) == want, name


def test_label_binary_checks(chl, tmp_path):
    # a well-formed sidecar (shared sides) and damaged copies of it
    n = 3
    off = np.array([0, 2, 3, 5], np.uint64)
    ent = np.array([(0 << 32) | 0, (1 << 32) | 4, (1 << 32) | 0, (0 << 32) | 7, (2 << 32) | 0], np.uint64)
    row_of = np.array([0, 1, 2], np.uint32)
    head = b"CHLLABv1" + np.array([n], np.uint64).tobytes() + np.array([1, 0], np.uint32).tobytes() + \
        np.array([5, 0], np.uint64).tobytes()
    good = head + off.tobytes() + ent.tobytes() + row_of.tobytes()
    bad = {
        "good": (good, 0),
        "magic": (b"CHLLABv2" + good[8:], -3),
        "short": (good[:-4], -3),
        "long": (good + b"\0", -3),
        "offsets": (head + np.array([0, 2, 1, 5], np.uint64).tobytes() + ent.tobytes() + row_of.tobytes(), -3),
        "hub_order": (head + off.tobytes() + ent[[1, 0, 2, 3, 4]].tobytes() + row_of.tobytes(), -3),
        "row_map": (head + off.tobytes() + ent.tobytes() + np.array([0, 0, 2], np.uint32).tobytes(), -3),
        # counts whose byte total wraps 2^64 to the file's size
        "wrap": (b"CHLLABv1" + np.array([n], np.uint64).tobytes() + np.array([1, 0], np.uint32).tobytes() +
                 np.array([(1 << 61) + 5, 0], np.uint64).tobytes() + off.tobytes() + ent.tobytes() +
                 row_of.tobytes(), -3),
    }
    for name, (data, want) in bad.items():
        p = tmp_path / f"{name}.bin"
        p.write_bytes(data)
        assert _read_status(chl, p, chl.Labels.load) == want, name


def test_library_exports_header(chl):
    lib = chl.lib()
    names = chl.header_symbols()
    assert len(names) > 30
    missing = [s for s in names if not hasattr(lib, s)]
    assert not missing, missing


def test_no_device_is_an_error_not_a_fallback(chl):
    if chl.device_count() > 0:
        pytest.skip("a device is visible")
    g = chl.Graph.grid(4, 4)
    with pytest.raises(chl.ChlError) as e:
        chl.build(g, np.arange(16, dtype=np.uint32))
    assert e.value.code == -7


def test_build_rejects_bad_order(chl):
    g = chl.Graph.grid(4, 4)
    with pytest.raises(chl.ChlError) as e:
        chl.build(g, np.zeros(16, np.uint32))
    assert e.value.code == -5
    with pytest.raises(chl.ChlError):
        chl.build(g, np.arange(15, dtype=np.uint32))


def test_query_generate_is_the_unseeded_c_rand_stream(chl):
    # generate_query (distributed_memory/query.hpp:44-51) draws rand() % range
    # from an unseeded generator: the C library's stream for seed 1
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    want = np.array([libc.rand() % 977 for _ in range(2 * 500)], np.uint32).reshape(-1, 2)
    got = chl.query_generate(500, 977)
    assert np.array_equal(got, want)
    # the library keeps its own state: a second call gives the same stream
    assert np.array_equal(chl.query_generate(500, 977), want)


def test_query_files_roundtrip_and_reject_malformed(chl, tmp_path):
    p = tmp_path / "q.txt"
    p.write_text("0 5\n17 3\n4294967295 0\n")
    pairs = chl.query_load(p)
    assert pairs.tolist() == [[0, 5], [17, 3], [4294967295, 0]]
    bad = tmp_path / "bad.txt"
    bad.write_text("0 5\n17\n")
    with pytest.raises(chl.ChlError):
        chl.query_load(bad)
    out = tmp_path / "r.txt"
    chl.query_write(out, np.array([3, chl.INFTY, 0], np.int32))
    assert out.read_text() == "3\n2147483647\n0\n"


def test_local_pairs_stay_within_the_hop_ball(chl):
    # a path: the targets of a 3-hop ball are within 3 positions of the source,
    # and every offset in [-3, 3] occurs
    n = 200
    src = np.arange(n - 1, dtype=np.uint32)
    g = chl.Graph.from_arcs(n, src, src + 1, np.ones(n - 1, np.int32), undirected=True)
    p = chl.local_pairs(g, 5000, 3, seed=4).astype(np.int64)
    d = p[:, 1] - p[:, 0]
    assert np.abs(d).max() <= 3
    assert set(np.unique(d).tolist()) == set(range(-3, 4))
    assert np.array_equal(p, chl.local_pairs(g, 5000, 3, seed=4).astype(np.int64))  # seeded
    assert np.array_equal(chl.local_pairs(g, 10, 0, seed=1)[:, 0], chl.local_pairs(g, 10, 0, seed=1)[:, 1])


@pytest.mark.parametrize("name", sorted(k for k, v in golden_manifest().items() if "degree_stdout" in v))
def test_degree_order_matches_reference_tool(chl, name):
    """The host degree order, n and m equal what the reference's degree tool
    (distributed_memory/degree.cpp:54-67) printed and wrote for the same graph file."""
    man = golden_manifest()[name]
    g = chl.Graph.load(os.path.join(GOLD, name + ".gr"))
    assert f"Graph has {g.n} vertices and {g.m} arcs\n" == man["degree_stdout"]
    want = chl.order_load(os.path.join(GOLD, name + ".degree"))
    assert np.array_equal(g.order_degree(reference_ties=True), want)


DEGREE_USAGE_CASES = [[], ["-h"], ["-o"], ["-o", "x"], ["g.gr"], ["-o", "x", "a", "b"], ["-q", "g"],
                      ["-o", "x", "--", "g"], ["--"]]


@pytest.mark.parametrize("args", DEGREE_USAGE_CASES)
def test_degree_tool_usage(args):
    """Command-line errors of plant_amd/bin/degree: the reference tool's usage text
    on stdout and exit status 1 (degree.cpp:32-52), before any device use."""
    if not os.path.exists(DEGREE_TOOL):
        pytest.skip("plant_amd/bin/degree not built")
    r = subprocess.run([DEGREE_TOOL] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert r.stdout == f"Usage: {DEGREE_TOOL} -o ordering graph\n  -o ordering\tFile with the vertex order\n"
    if os.path.exists(REF_DEGREE):
        ref = subprocess.run([REF_DEGREE] + args, capture_output=True, text=True, timeout=60)
        assert ref.returncode == r.returncode
        assert ref.stdout.replace(REF_DEGREE, DEGREE_TOOL) == r.stdout


def test_degree_tool_unreadable_graph(tmp_path):
    if not os.path.exists(DEGREE_TOOL):
        pytest.skip("plant_amd/bin/degree not built")
    missing = str(tmp_path / "none.gr")
    r = subprocess.run([DEGREE_TOOL, "-o", str(tmp_path / "o"), missing], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert r.stdout == ""
    assert r.stderr.endswith(f"Unable to read graph from file {missing}\n")
    assert not (tmp_path / "o").exists()


def test_ctypes_structs_match_header(tmp_path):
    """The binding's ctypes structures have the sizes and field offsets that the C
    compiler gives the structures of include/chl.h."""
    import ctypes as C
    from plant_amd import chl as chlmod
    structs = {"chl_build_opts": chlmod.BuildOpts, "chl_build_stats": chlmod.BuildStats,
               "chl_server_opts": chlmod.ServerOpts, "chl_server_stats": chlmod.ServerStats}
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "chl.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c11", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    got = {}
    for ln in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        a, b, c = ln.split()
        got[(a, b)] = int(c)
    for cname, cls in structs.items():
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, (cname, f)
