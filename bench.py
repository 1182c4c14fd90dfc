"""Benchmark: canonical hub label build on one GPU.

Default workload (road): n = 1,000,000 points uniform in the unit square, each
joined to its 3 nearest neighbours, symmetrised, integer weight
round(1e4 * euclidean) + 1; vertex order = sampled betweenness from 256 seeded
roots; PLaNT trees. Workload rmat / rmat-w16: R-MAT scale 20, edge factor 3,
symmetrised, unit weights / weights in [1, 16], degree order, hybrid PLaNT to
pruned trees. One step = one complete label build (all n roots) with the graph
and order already resident in the library, under learned batching (--static:
static rank-order batching); the value is labels per second (entries of both
sides, as Labeling::get_total counts them). The other batching policy is timed
once beside it and its labels compared.

Extra fields: the dominant kernel's roofline (algorithmic bytes of the tree
kernel over its device time, measured with HIP events on the launch stream), a
CPU baseline (the oracle's GLL on all host cores over a bounded rank prefix of
the same graph), and the query path (merge-join kernel on 4096-pair batches and
the continuous-batching server's p50/p99 latency).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload road|rmat|rmat-w16] [--mode plant|gll|hybrid]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
MODES = {"gll": 0, "plant": 1, "hybrid": 2}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def query_bench(chl, L, n, batches=200, batch=4096, seed=11):
    """Queries/s of back-to-back 4096-pair batches through the server, with p50/p99."""
    lib = chl.lib()
    vp = C.c_void_p

    class SOpts(C.Structure):
        _fields_ = [("streams", C.c_int), ("max_inflight_pairs", C.c_int64), ("max_batch_pairs", C.c_int64),
                    ("reserved", C.c_int * 8)]

    class SStats(C.Structure):
        _fields_ = [("completed", C.c_int64), ("pairs", C.c_int64), ("rejected", C.c_int64),
                    ("p50_us", C.c_double), ("p99_us", C.c_double), ("mean_us", C.c_double),
                    ("max_us", C.c_double), ("qps", C.c_double), ("reserved", C.c_double * 8)]

    lib.chl_server_opts_default.argtypes = [C.POINTER(SOpts)]
    lib.chl_server_create.argtypes = [vp, C.POINTER(SOpts), C.POINTER(vp)]
    lib.chl_server_submit.argtypes = [vp, vp, C.c_int64, vp, C.POINTER(C.c_uint64)]
    lib.chl_server_poll.argtypes = [vp, C.c_uint64, C.c_int]
    lib.chl_server_drain.argtypes = [vp]
    lib.chl_server_stats_get.argtypes = [vp, C.POINTER(SStats)]
    lib.chl_server_stats_reset.argtypes = [vp]
    lib.chl_server_free.argtypes = [vp]
    o = SOpts()
    lib.chl_server_opts_default(C.byref(o))
    o.streams = 8
    srv = vp()
    if lib.chl_server_create(L._h, C.byref(o), C.byref(srv)) != 0:
        raise RuntimeError(lib.chl_last_error().decode())
    rng = np.random.default_rng(seed)
    pairs = [rng.integers(0, n, (batch, 2), dtype=np.uint32) for _ in range(16)]
    outs = [np.zeros(batch, np.int32) for _ in range(batches)]
    t = C.c_uint64()
    # warm-up
    for k in range(16):
        lib.chl_server_submit(srv, pairs[k].ctypes.data, batch, outs[k].ctypes.data, C.byref(t))
    lib.chl_server_drain(srv)
    lib.chl_server_stats_reset(srv)
    tickets = []
    t0 = time.perf_counter()
    for k in range(batches):
        while True:
            rc = lib.chl_server_submit(srv, pairs[k % 16].ctypes.data, batch, outs[k].ctypes.data, C.byref(t))
            if rc == 0:
                break
            if rc != -9:
                raise RuntimeError(lib.chl_last_error().decode())
        tickets.append(t.value)
    lib.chl_server_drain(srv)
    wall = time.perf_counter() - t0
    st = SStats()
    lib.chl_server_stats_get(srv, C.byref(st))
    for tk in tickets:
        lib.chl_server_poll(srv, tk, 1)
    lib.chl_server_free(srv)
    # a direct batched call for the kernel-only rate
    big = rng.integers(0, n, (1 << 20, 2), dtype=np.uint32)
    L.query(big[:4096])
    t1 = time.perf_counter()
    L.query(big)
    direct = time.perf_counter() - t1
    return {"batch_pairs": batch, "batches": batches, "queries_per_s": batches * batch / wall,
            "p50_us": st.p50_us, "p99_us": st.p99_us, "max_us": st.max_us,
            "direct_1M_pairs_queries_per_s": big.shape[0] / direct}


def copy_calibration(nbytes=1 << 30):
    """One device-to-device copy of nbytes (a plain copy kernel of known size, the
    scale for FETCH_SIZE / WRITE_SIZE counters)."""
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {nbytes} bytes device to device")


def cpu_baseline(g, order, seconds):
    """The oracle's GLL (run_paraPLL restated) on all host cores over the rank prefix
    it completes within `seconds` of labeling+sorting+cleaning time."""
    from oracle import oracle as orc
    threads = os.cpu_count() or 1
    src, dst, w = g.arc_list(1)
    og = orc.OGraph.from_arcs(g.n, src, dst, w)
    L, s = orc.gll_bounded(og, order, threads, seconds)
    tsec = float(s[0] + s[1] + s[2])
    labels = L.total()
    return {"value": labels / tsec, "unit": "labels/s", "cores": threads, "kind": "port",
            "sample": f"oracle GLL (shared_memory/akiba.hpp:132-217 restated), {threads} threads, rank prefix "
                      f"[0, {int(s[4])}) of the same graph and order: {labels} labels in {tsec:.2f} s "
                      f"(lt+st+ct, {int(s[3])} supersteps)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="road", choices=["road", "rmat", "rmat-w16"],
                    help="road: 1M-vertex 3-NN road-like graph (default); rmat / rmat-w16: R-MAT scale 20, "
                         "edge factor 3, unit weights / weights in [1, 16], degree order, hybrid mode")
    ap.add_argument("--n", type=int, default=0, help="vertices (road) or 2^scale (rmat); 0 = the workload's own")
    ap.add_argument("--mode", default=None, choices=sorted(MODES), help="default: plant (road), hybrid (rmat)")
    ap.add_argument("--static", action="store_true", help="time static rank-order batching (default: learned)")
    ap.add_argument("--no-compare", action="store_true", help="skip the timed build under the other batching")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--delta", type=int, default=0)
    ap.add_argument("--slots", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-query", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    from plant_amd import chl
    if chl.device_count() < 1:
        print("bench.py needs a gfx950 device", file=sys.stderr)
        sys.exit(1)

    t = time.perf_counter()
    if args.workload == "road":
        mode = args.mode or "plant"
        g = chl.Graph.knn(args.n or 1_000_000, 3, seed=1)
        order, score = g.order_betweenness(256, seed=7, with_score=True)
        desc = (f"road-like 3-NN, n={g.n}, {g.arcs(1) // 2} undirected edges, w=round(1e4*d)+1, "
                f"256-root sampled-betweenness order")
    else:
        mode = args.mode or "hybrid"
        scale = int(np.log2(args.n)) if args.n else 20
        wmax = 16 if args.workload == "rmat-w16" else 1
        g = chl.Graph.rmat(scale, 3 << scale, wmax=wmax, seed=1)
        order = g.order_degree(seed=2)
        _, score = g.order_betweenness(64, seed=7, with_score=True)  # a predictor feature only
        desc = (f"R-MAT scale {scale} (a=.57 b=.19 c=.19 d=.05), edge factor 3, symmetrised, n={g.n}, "
                f"{g.arcs(1) // 2} undirected edges, weights in [1, {wmax}], degree order")
    log(f"graph n={g.n} arcs={g.arcs(1)} + order in {time.perf_counter() - t:.1f} s")
    if args.copy_calibration:
        copy_calibration()
    learned = not args.static
    opts = dict(mode=MODES[mode], batching=1 if learned else 0, batch_roots=args.batch,
                delta=args.delta, slots=args.slots, vertex_score=score.ctypes.data)
    for k in range(args.warmup):
        L = chl.build(g, order, **opts)
        log(f"warmup {k}: {L.stats.total_s:.3f} s")
        del L
    times, stats = [], []
    L = None
    for k in range(args.steps):
        if L is not None:
            del L
        t0 = time.perf_counter()
        L = chl.build(g, order, **opts)
        times.append(time.perf_counter() - t0)
        stats.append(L.stats.as_dict())
        log(f"step {k}: {times[-1]:.3f} s labels={L.stats.labels} avg={L.stats.avg_label:.1f}")
    ms = 1e3 * float(np.mean(times))
    s = stats[-1]
    labels = int(s["labels"])
    relax_s = s["relax_ms"] / 1e3
    achieved = s["relax_bytes"] / relax_s / 1e9 if relax_s > 0 else 0.0
    line = {
        "metric": "label build s + labels/s (learned vs static batching); query/s, p50/p99 us",
        "value": labels / (ms / 1e3), "unit": "labels/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"{desc}, {'learned' if learned else 'static rank-order'} batches",
                   "mode": mode, "build_s": ms / 1e3, "labels": labels, "avg_label": s["avg_label"],
                   "batches": s["batches"], "batch_roots": s["batch_used"], "delta": s["delta_used"],
                   "slots": s["slots_used"], "lt_s": s["lt"], "st_s": s["st"], "ct_s": s["ct"],
                   "mlp_ms": s["mlp_ms"], "pred_mae_learned": s["pred_mae_learned"],
                   "pred_mae_static": s["pred_mae_static"], "pred_roots": s["pred_roots"],
                   "switch_rank": s["switch_rank"], "switch_trials": s["switch_trials"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "tree_kernel",
                     "bytes_per_build": s["relax_bytes"], "kernel_ms_per_build": s["relax_ms"]},
    }
    if not args.no_compare:
        other = dict(opts, batching=0 if learned else 1)
        t0 = time.perf_counter()
        L2 = chl.build(g, order, **other)
        t_other = time.perf_counter() - t0
        sides = (0,) if g.symmetric else (0, 1)  # a symmetric graph stores one side
        same = all(np.array_equal(a, b) for side in sides for a, b in zip(L.copy(side), L2.copy(side)))
        del L2
        t_static, t_learned = (t_other, ms / 1e3) if learned else (ms / 1e3, t_other)
        line["batching"] = {"static_build_s": t_static, "learned_build_s": t_learned,
                            "learned_speedup": t_static / t_learned, "labels_identical": same}
        log(f"other batching: {t_other:.3f} s, labels identical: {same}")
    if not args.no_query:
        line["query"] = query_bench(chl, L, g.n)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(g, order, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
