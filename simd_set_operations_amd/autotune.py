"""Autotuner: batched throughput of every registered 2-set entry and tile shape over a grid of pair
shapes, and, for graph edge intersections, the merge tile shape and the skew threshold that splits
the edges between the search and merge entries.

Grid: larger-set size 2^6 .. 2^14, skew (larger / smaller) 2^0 .. 2^10, selectivity 0.1 / 0.5 /
0.9, density 0.1 (element space = 10 x larger), the planted-shared-then-disjoint data of the
benchmark workloads.  Each cell is a fixed-shape CSR batch of about 2^26 input elements (at most
2^18 pairs); every entry runs it as a device batch (plan, tile kernel, scan, placement, with
outputs) and is timed over replayed runs with HIP events, after a check of its per-pair counts
against the planted counts.  The table (``tuning.json``) feeds :mod:`.tuning`, which prices a
batch per entry and picks the cheapest (``hip_auto``).  It also holds the search entries' probe
rates against one indexed resident set (the skewed workload's shape, per smaller-set size).

Run on the GPU:  python -m simd_set_operations_amd.autotune --out simd_set_operations_amd/tuning.json
"""
import argparse
import json
import sys
import time

import numpy as np

# entries with distinct kernels, and the tile shapes of those that have several ("name@shape":
# merge steps per lane 20/16/24/12, probes per lane 2/1/4; include/setops_hip.h SETOPS_SHAPE)
ALGOS = ("hip_branchless_merge", "hip_branchless_merge@1", "hip_branchless_merge@2", "hip_branchless_merge@3",
         "hip_shuffling", "hip_qfilter", "hip_galloping", "hip_galloping@1", "hip_galloping@2", "hip_binary_search",
         "hip_binary_search@1", "hip_binary_search@2", "hip_simd_galloping")
GRID = {
    "log2_larger": [6, 8, 10, 12, 14],
    "log2_skew": [0, 3, 6, 10],
    "selectivity": [0.1, 0.5, 0.9],
}
CELL_ELEMENTS = 1 << 26
MAX_PAIRS = 1 << 18


def cell_shape(log2_larger, log2_skew):
    larger = 1 << log2_larger
    smaller = max(1, larger >> log2_skew)
    pairs = int(min(MAX_PAIRS, max(64, CELL_ELEMENTS // (larger + smaller))))
    return larger, smaller, pairs


def time_entry(torch, algo, A, a_off, B, b_off, shared, runs):
    from .batch import DeviceBatch

    db = DeviceBatch(algo, A, a_off, B, b_off)
    stream = torch.cuda.current_stream()
    try:
        for _ in range(2):
            db.run(stream)
        torch.cuda.synchronize()
        _, cnt, _ = db.results()
        if not np.all(cnt == shared):
            raise RuntimeError(f"{algo}: per-pair counts differ from the planted {shared}")
        for _ in range(2):  # the replayed launch graph
            db.run(stream)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(runs):
            db.run(stream)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / runs
    finally:
        db.close()


def run(out_path, runs=5, grid=None, algos=ALGOS, log=print):
    import torch

    from . import datagen, intersect

    if not torch.cuda.is_available():
        raise RuntimeError("the autotuner measures on a GPU; none is visible")
    intersect.init(0)
    grid = grid or GRID
    cells = []
    t0 = time.perf_counter()
    for l in grid["log2_larger"]:
        for s in grid["log2_skew"]:
            larger, smaller, pairs = cell_shape(l, s)
            for sel in grid["selectivity"]:
                A, a_off, B, b_off, shared = datagen.fixed_batch(pairs, larger, smaller, sel)
                E = int(A.size + B.size)
                g = {}
                for algo in algos:
                    ms = time_entry(torch, algo, A, a_off, B, b_off, shared, runs)
                    g[algo] = E / (ms / 1e3) / 1e9
                best = max(g, key=g.get)
                cells.append({"log2_larger": l, "log2_skew": s, "selectivity": sel, "larger": larger,
                              "smaller": smaller, "pairs": pairs, "elements": E, "gelems": g, "best": best})
                log(f"[autotune] L=2^{l} skew=2^{s} sel={sel} pairs={pairs}: best {best} "
                    + " ".join(f"{a[4:]}={v:.0f}" for a, v in g.items()))
    graph = graph_split(torch, log=log)
    resident = resident_cells(torch, runs=runs, log=log)
    table = {
        "device": torch.cuda.get_device_name(0),
        "created_unix": int(time.time()),
        "seconds": round(time.perf_counter() - t0, 1),
        "unit": "Gelem/s (input elements, sum |A|+|B|, per second; batch with outputs)",
        "density": 0.1,
        "cell_elements": CELL_ELEMENTS,
        "grid": grid,
        "algorithms": list(algos),
        "cells": cells,
        "graph_skew_split": graph,
        "resident": resident,
    }
    with open(out_path, "w") as f:
        json.dump(table, f, indent=1)
    return table


# resident-pool cells: one 2^24-element set over 2^28 values, indexed, searched by 2^20 probes in
# sets of 2^4 .. 2^10 (each with a tenth planted from the large set) -- the skewed workload's shape
RESIDENT_LOG2_SMALL = (4, 6, 8, 10)
RESIDENT_ALGOS = ("hip_galloping", "hip_galloping@1", "hip_galloping@2", "hip_binary_search", "hip_binary_search@1",
                  "hip_binary_search@2", "hip_simd_galloping")


def resident_cells(torch, runs=5, log=print, log2_small=RESIDENT_LOG2_SMALL, algos=RESIDENT_ALGOS):
    """Probes per second of the search entries against one indexed resident set, per smaller-set
    size: GatherBatch(..., index=True) over the skewed workload's large set, counts checked against
    the planted ones, replayed runs timed with HIP events."""
    from . import datagen
    from .batch import GatherBatch

    stream = torch.cuda.current_stream()
    cells = []
    for ls in log2_small:
        m = 1 << ls
        P = (1 << 20) // m
        shared = max(1, m // 10)
        L, S = datagen.skewed(nsmall=P, small=m, shared=shared)
        pool = S.reshape(-1)
        a_beg = np.arange(P, dtype=np.uint64) * m
        dL = torch.from_numpy(L.view(np.int32)).cuda()
        g = {}
        for algo in algos:
            gb = GatherBatch(algo, pool, a_beg, a_beg + m, dL, np.zeros(P, np.uint64), np.full(P, L.size, np.uint64),
                             index=True)
            try:
                for _ in range(2):
                    gb.run(stream)
                torch.cuda.synchronize()
                _, cnt, _ = gb.results()
                if not np.all(cnt == shared):
                    raise RuntimeError(f"{algo}: resident-pool counts differ from the planted {shared}")
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(runs):
                    gb.run(stream)
                e1.record(stream)
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / runs
            finally:
                gb.close()
            g[algo] = P * m / (ms / 1e3) / 1e9
        best = max(g, key=g.get)
        cells.append({"log2_small": ls, "small": m, "pairs": P, "large": int(L.size), "gprobes": g, "best": best})
        log(f"[autotune] resident 2^24 indexed, small=2^{ls} pairs={P}: best {best} "
            + " ".join(f"{a[4:]}={v:.1f}" for a, v in g.items()))
        del dL
    return cells


GRAPH_SCALE = 20
GRAPH_THRESHOLDS = (8.0, 16.0, 32.0, 64.0, 128.0, 512.0)
GRAPH_MERGE_SHAPES = ("hip_branchless_merge", "hip_branchless_merge@1", "hip_branchless_merge@2",
                      "hip_branchless_merge@3")


def graph_split(torch, scale=GRAPH_SCALE, thresholds=GRAPH_THRESHOLDS, reps=3, log=print):
    """The merge tile shape and the skew threshold for graph edge intersections (count-only, as the
    triangle count runs them).  On an R-MAT graph (0.57, 0.19, 0.19, 0.05) at 2^scale vertices,
    edge factor 16: every hip_branchless_merge tile shape over all edges (the fastest is
    "merge_entry"), then every T, edges whose larger neighbour list is more than T times the
    smaller going to hip_galloping over the indexed CSR and the rest to the merge entry.  Every
    run's sum is checked against the plain merge's; returns the fastest T with the timings."""
    from . import datagen
    from .graph import EdgeIntersections, split_by_skew

    rp, col = datagen.rmat_graph(scale=scale)
    u, v = datagen.graph_edges(rp, col)
    deg = np.diff(rp).astype(np.int64)
    dcol = torch.from_numpy(col.view(np.int32)).cuda()
    stream = torch.cuda.current_stream()

    def timed(batches):
        for b in batches:
            b.run(stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            for b in batches:
                b.run(stream)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps, sum(b.totals()[0] for b in batches)

    plain = EdgeIntersections("hip_branchless_merge", dcol, rp[:-1], rp[1:], u, v)
    ms_plain, total = timed([plain])
    plain.close()
    res = {"scale": scale, "edges": int(u.size), "plain_merge_ms": ms_plain, "ms": {}, "merge_ms": {}}
    # the merge entry's tile shape for gather batches over one CSR (their tiles stage every pair's
    # windows separately, so the pair-shape cells, measured on CSR batches, do not price them)
    for name in GRAPH_MERGE_SHAPES:
        m = EdgeIntersections(name, dcol, rp[:-1], rp[1:], u, v)
        ms, t = timed([m])
        m.close()
        if t != total:
            raise RuntimeError(f"graph merge {name}: sum {t} differs from the plain merge's {total}")
        res["merge_ms"][name] = ms
        log(f"[autotune] graph 2^{scale} merge {name}: {ms:.3f} ms")
    merge = min(res["merge_ms"], key=res["merge_ms"].get)
    res["merge_entry"] = merge
    for thr in thresholds:
        sk = split_by_skew(deg[u], deg[v], thr)
        a = EdgeIntersections("hip_galloping", dcol, rp[:-1], rp[1:], u, v, select=sk, index=True)
        b = EdgeIntersections(merge, dcol, rp[:-1], rp[1:], u, v, select=~sk)
        ms, t = timed([a, b])
        a.close()
        b.close()
        if t != total:
            raise RuntimeError(f"graph split at {thr}: sum {t} differs from the plain merge's {total}")
        res["ms"][str(thr)] = ms
        log(f"[autotune] graph 2^{scale} skew split at {thr:g}: {ms:.3f} ms (plain merge {ms_plain:.3f} ms)")
    res["threshold"] = float(min(res["ms"], key=lambda k: res["ms"][k]))
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default="tuning.json")
    ap.add_argument("--runs", type=int, default=5)
    args = ap.parse_args()
    run(args.out, args.runs, log=lambda m: (print(m), sys.stdout.flush()))


if __name__ == "__main__":
    main()
