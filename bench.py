#!/usr/bin/env python3
"""Benchmark: batched uneven 2-set intersection on one MI355X.

Workload (the batched configuration of this path): 65,536 pairs of sorted unique u32; larger
size log-uniform in [2^6, 2^14]; smaller = larger / r, r log-uniform in [1, 1024]; element space
10 x larger; selectivity ~ U[0,1] with shared elements planted first.  One step = one batch
intersection (plan, split, tile, scan and placement kernels, replayed as one HIP graph) producing
per-pair counts, per-pair offsets and the compacted results, with the inputs already resident in
HBM.

Printed (one JSON line): whole-job throughput in input elements/s (E = sum |A|+|B|, the
reference's element_count convention, scripts/results/process.py:62-77), the main kernel's
roofline figures (HIP events on the launch stream), the plain one-launch-per-pair GPU path on a
sample of pairs, and the CPU baselines (the reference's own QFilter, the branchless_merge
restatement and a restatement of the croaring comparator, one pair per host thread).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
# the kernel whose time is the roofline's (per headline entry)
DOMINANT_KERNEL = {
    "hip_branchless_merge": "setops::merge_tile_kernel<256, 20, 130, false, false, false, 0, false>",
    "hip_shuffling": "setops::merge_tile_kernel<256, 20, 130, false, false, false, 1, false>",
    "hip_qfilter": "setops::merge_tile_kernel<256, 20, 130, false, false, false, 2, false>",
    "hip_galloping": "setops::search_tile_kernel<256, 2, 130, 3072, false, false, 1, false>",
    "hip_binary_search": "setops::search_tile_kernel<256, 2, 130, 3072, false, false, 0, false>",
    "hip_simd_galloping": "setops::search_tile_kernel<256, 2, 130, 3072, false, false, 2, false>",
}


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "merge_tile_kernel_uneven_batch_traffic.txt")
TRACE_FILE = os.path.join(ROOT, "profiles", "uneven_batch_kernel_trace.txt")
# Host threads for the CPU baselines: the host cores this process may run on, at most 16 (the
# host-CPU share a GPU of this pool comes with, whatever the machine's total core count is).
CPU_SHARE = 16


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(CPU_SHARE, n))


def host_info():
    """What the host offers: logical CPUs, the affinity set, and lscpu's model and topology."""
    info = {"cpu_count": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    try:
        import subprocess

        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True, timeout=10).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "NUMA node(s)": "numa_nodes", "CPU max MHz": "max_mhz"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except Exception:
        pass
    return info


def profiled_avg_us(kernel):
    """The kernel's average duration (us) in the committed rocprofv3 kernel-trace summary, or None."""
    try:
        with open(TRACE_FILE) as f:
            for line in f:
                if line.startswith(kernel + " "):
                    return float(line[len(kernel):].split()[1])
    except (OSError, ValueError, IndexError):
        pass
    return None


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary (FETCH_SIZE and
    WRITE_SIZE scaled by a plain copy of known size); None when absent."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                    return int(float(line.split(":")[-1]))
    except OSError:
        pass
    return None


def algorithmic_bytes(E, out_total, P):
    # inputs once (4E), offsets read (2 x 8 B per pair), outputs once (4 per element),
    # per-pair count (4 B) and offset (8 B) written
    return 4 * E + 16 * P + 4 * out_total + 12 * P


def cpu_baselines(A, a_off, B, b_off, E, budget_s):
    """Times the reference's QFilter (compiled from its source), the branchless_merge restatement,
    the reference's widest SIMD block intersection restated with x86 intrinsics (shuffling_avx512,
    or shuffling_avx2 without AVX-512) and the croaring comparator's restatement (roaring bitmaps
    built outside the timed region, as the reference's harness prepares them) over the whole
    batch, one pair per host thread; each repeats until ~budget_s / 2.  Keys starting with
    "countonly_" time a count-only pass and are not candidates for the baseline (the GPU value
    writes every result)."""
    import oracle

    threads = host_threads()
    res = {}
    rbat = oracle.RoaringBatch(A, a_off, B, b_off)
    for key, mat in (("port_croaring_and", True), ("countonly_port_croaring_and_cardinality", False)):
        rbat.run(threads, materialize=mat)  # warm
        reps, el = 0, 0.0
        while True:
            el += rbat.run(threads, materialize=mat)
            reps += 1
            if el > budget_s / 2 or reps >= 200:
                break
        res[key] = {"gelems": E * reps / el / 1e9, "reps": reps, "seconds": el}
    rbat.close()
    if oracle.ref_available():
        rb = oracle.RefQfilterBatch(A, a_off, B, b_off)
        rb.run(threads)  # warm
        reps, t0 = 0, time.perf_counter()
        while True:
            rb.run(threads)
            reps += 1
            el = time.perf_counter() - t0
            if el > budget_s / 2 or reps >= 200:
                break
        res["reference_qfilter_c"] = {"gelems": E * reps / el / 1e9, "reps": reps, "seconds": el}
    reps, t0 = 0, time.perf_counter()
    while True:
        oracle.batch_intersect("branchless_merge", A, a_off, B, b_off, nthreads=threads, with_output=True)
        reps += 1
        el = time.perf_counter() - t0
        if el > budget_s / 2 or reps >= 200:
            break
    res["port_branchless_merge"] = {"gelems": E * reps / el / 1e9, "reps": reps, "seconds": el}
    simd = oracle.fastest_simd_block_algo()
    if simd:
        reps, t0 = 0, time.perf_counter()
        while True:
            oracle.batch_intersect(simd, A, a_off, B, b_off, nthreads=threads, with_output=True)
            reps += 1
            el = time.perf_counter() - t0
            if el > budget_s / 2 or reps >= 200:
                break
        res["port_" + simd] = {"gelems": E * reps / el / 1e9, "reps": reps, "seconds": el}
    return res, threads


def plain_gpu(torch, A, a_off, B, b_off, sample, stream):
    """The plain drop-in: one launch sequence per pair (device-resident inputs), back to back on
    one stream, over an evenly spaced sample of pairs."""
    import ctypes

    from simd_set_operations_amd import _native
    from simd_set_operations_amd.intersect import algo_id

    lib = _native.hip()
    algo = algo_id("hip_branchless_merge")
    P = a_off.size - 1
    idx = np.linspace(0, P - 1, num=min(sample, P)).astype(np.int64)
    # each pair 16-byte aligned in its own pool slot
    sets_a = [A[a_off[p]:a_off[p + 1]] for p in idx]
    sets_b = [B[b_off[p]:b_off[p + 1]] for p in idx]

    def pack(sets):
        lens = np.array([s.size for s in sets], np.int64)
        pad = (lens + 3) // 4 * 4
        start = np.concatenate([[0], np.cumsum(pad)[:-1]])
        buf = np.zeros(int(pad.sum()) + 4, np.uint32)
        for s, st in zip(sets, start):
            buf[st:st + s.size] = s
        return torch.from_numpy(buf.view(np.int32)).cuda(), start, lens

    dA, sa, la = pack(sets_a)
    dB, sb, lb = pack(sets_b)
    cap = int(np.minimum(la, lb).max()) + 4
    dO = torch.empty(cap, dtype=torch.int32, device="cuda")
    dT = torch.zeros(2, dtype=torch.int64, device="cuda")
    pa, pb, po, pt = dA.data_ptr(), dB.data_ptr(), dO.data_ptr(), dT.data_ptr()
    s = stream.cuda_stream

    def one(k):
        _native.check(lib.setops_hip_intersect2_dev(algo, pa + 4 * int(sa[k]), None, int(la[k]), pb + 4 * int(sb[k]),
                                                    None, int(lb[k]), po, None, cap, pt, s))

    for k in range(min(16, len(idx))):
        one(k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(len(idx)):
        one(k)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    E_s = int(la.sum() + lb.sum())
    return {"gelems": E_s / el / 1e9, "pairs": int(len(idx)), "us_per_pair": el / len(idx) * 1e6}


SKEWED_ENTRIES = ("hip_galloping", "hip_binary_search", "hip_simd_galloping", "hip_galloping:index",
                  "hip_binary_search:index", "hip_simd_galloping:index", "hip_auto:index")


def skewed_leg(torch, steps, cpu_threads, with_cpu, entries=SKEWED_ENTRIES):
    """The skewed workload: one resident large set (2^24 over 2^28) against 4,096 sets of 256 with
    26 planted matches each; the large set is read in place by every pair (gather batch).  Reported
    on unique bytes: the large set once, the small sets once, outputs and counts once.  `entries`
    are registered names (``name@shape`` for a tile shape, ``hip_auto`` for the tuning table's
    choice), each with ``:index`` to search through the B-pool index."""
    from simd_set_operations_amd import datagen
    from simd_set_operations_amd.batch import GatherBatch

    L, S = datagen.skewed()
    P, m = S.shape
    pool = S.reshape(-1)
    a_beg = np.arange(P, dtype=np.uint64) * m
    a_end = a_beg + m
    b_beg = np.zeros(P, np.uint64)
    b_end = np.full(P, L.size, np.uint64)
    shared = datagen.SKEWED["shared"]
    E_unique = L.size + pool.size
    out_total = P * shared
    B_unique = 4 * E_unique + 4 * out_total + 4 * P
    res = {"pairs": P, "large": int(L.size), "small": m, "elements_unique": E_unique,
           "elements_streamed": int(P * (L.size + m)), "bytes_unique": B_unique}
    dL = torch.from_numpy(L.view(np.int32)).cuda()
    stream = torch.cuda.current_stream()
    for entry in entries:
        algo, _, how = entry.partition(":")
        indexed = how == "index"
        # the index is built once for the resident set, outside the timed region
        gb = GatherBatch(algo, pool, a_beg, a_end, dL, b_beg, b_end, index=indexed)
        key = (gb.algo if algo == "hip_auto" else algo) + ("_indexed" if indexed else "")
        if algo == "hip_auto":
            key = "auto_" + key
        reps = steps
        for _ in range(2):
            gb.run(stream)
        torch.cuda.synchronize()
        _, cnt, _ = gb.results()
        if not np.all(cnt == shared):
            log(f"skewed {key}: counts differ from the planted {shared}")
            sys.exit(3)
        gb.kernel_time()
        gb.set_timing(True)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(reps):
            gb.run(stream)
        ev1.record(stream)
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / reps
        k_ms, k_runs = gb.kernel_time()
        k_ms /= max(k_runs, 1)
        gb.close()
        res[key] = {"ms_per_batch": ms, "kernel_ms": k_ms, "gelems_unique": E_unique / (ms / 1e3) / 1e9,
                    "hbm_frac_unique": B_unique / (ms / 1e3) / (HBM_PEAK_GBS * 1e9)}
        log(f"skewed {key}: {res[key]}")
    if with_cpu:
        import oracle

        t0 = time.perf_counter()
        cnt, _, _ = oracle.batch_intersect_gather("galloping", pool, a_beg, a_end, L, b_beg, b_end,
                                                  nthreads=cpu_threads, with_output=True)
        el = time.perf_counter() - t0
        if not np.all(cnt == shared):
            log("skewed cpu galloping: counts differ")
            sys.exit(3)
        res["cpu_port_galloping"] = {"ms_per_batch": el * 1e3, "threads": cpu_threads,
                                     "gelems_unique": E_unique / el / 1e9}
        log(f"skewed cpu: {res['cpu_port_galloping']}")
    return res


def graph_leg(torch, cpu_threads, with_cpu, scale):
    """The graph workload: R-MAT (0.57, 0.19, 0.19, 0.05) at 2^scale vertices, edge factor 16,
    symmetrized and deduplicated.  For every edge u < v, |N(u) & N(v)| summed (= 3T), count-only,
    on the CSR in place (gather batches): plain merge (default tile shape, and "merge_tuned" in the
    tile shape the autotuner persisted for graphs, when that differs), BSR merge (popcount sum), the
    oriented lists (= T, the independent check 3 x T), and a skew split -- galloping for edges whose
    larger list exceeds `threshold` x the smaller, the tuned merge entry for the rest -- over a small
    threshold sweep that includes the threshold the autotuner persisted in tuning.json
    ("tuned_split"; "best_split" is the fastest on this graph).  Throughput uses the streamed element count
    sum_{u<v} (d_u + d_v) = sum_v d_v^2."""
    from simd_set_operations_amd import datagen
    from simd_set_operations_amd.graph import EdgeIntersections, csr_to_bsr, split_by_skew

    t0 = time.perf_counter()
    rp, col = datagen.rmat_graph(scale=scale)
    u, v = datagen.graph_edges(rp, col)
    up = datagen.upper_starts(rp, col)
    brp, bases, states = csr_to_bsr(rp, col)
    deg = np.diff(rp).astype(np.int64)
    E = int((deg[u] + deg[v]).sum())
    res = {"scale": scale, "vertices": int(rp.size - 1), "edges_u_lt_v": int(u.size),
           "avg_degree": float(col.size / (rp.size - 1)), "max_degree": int(deg.max()), "elements_streamed": E,
           "bsr_words": int(bases.size), "gen_seconds": time.perf_counter() - t0}
    log(f"graph: {res}")
    dcol = torch.from_numpy(col.view(np.int32)).cuda()
    dbases = torch.from_numpy(bases.view(np.int32)).cuda()
    dstates = torch.from_numpy(states.view(np.int32)).cuda()
    stream = torch.cuda.current_stream()

    def timed(batches, reps=2):
        for b in batches:
            b.run(stream)
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(reps):
            for b in batches:
                b.run(stream)
        ev1.record(stream)
        torch.cuda.synchronize()
        tot = [b.totals() for b in batches]
        return ev0.elapsed_time(ev1) / reps, sum(t[0] for t in tot), sum(t[1] for t in tot)

    e = EdgeIntersections("hip_branchless_merge", dcol, rp[:-1], rp[1:], u, v)
    ms, s3, _ = timed([e])
    e.close()
    res["merge"] = {"ms": ms, "sum": s3, "gelems": E / ms / 1e6}
    from simd_set_operations_amd import tuning

    merge = tuning.graph_merge_entry()  # the tile shape the autotuner measured on a smaller R-MAT graph
    res["tuned_merge_entry"] = merge
    if merge != "hip_branchless_merge":
        e = EdgeIntersections(merge, dcol, rp[:-1], rp[1:], u, v)
        ms, sm, _ = timed([e])
        e.close()
        res["merge_tuned"] = {"ms": ms, "sum": sm, "gelems": E / ms / 1e6}
        if sm != s3:
            log(f"graph tuned merge {merge}: sum {sm} differs from {s3}")
            sys.exit(3)
    e = EdgeIntersections(merge, dcol, up, rp[1:], u, v)
    ms, t, _ = timed([e])
    e.close()
    res["oriented_merge"] = {"ms": ms, "triangles": t}
    e = EdgeIntersections("hip_branchless_merge_bsr", dbases, brp[:-1], brp[1:], u, v, states=dstates)
    ms, _, s3b = timed([e])
    e.close()
    res["merge_bsr"] = {"ms": ms, "sum": s3b, "gelems": E / ms / 1e6}
    best = None
    tuned = tuning.graph_split_threshold()  # measured by the autotuner on a smaller R-MAT graph
    res["tuned_threshold"] = tuned
    sweep = [(32.0, False), (128.0, False), (512.0, False), (32.0, True), (128.0, True)]
    if (tuned, True) not in sweep:
        sweep.append((tuned, True))
    for thr, indexed in sweep:
        sk = split_by_skew(deg[u], deg[v], thr)
        a = EdgeIntersections("hip_galloping", dcol, rp[:-1], rp[1:], u, v, select=sk, index=indexed)
        b = EdgeIntersections(merge, dcol, rp[:-1], rp[1:], u, v, select=~sk)
        ms, sh, _ = timed([a, b])
        a.close()
        b.close()
        key = f"split_{int(thr)}" + ("_indexed" if indexed else "")
        res[key] = {"ms": ms, "sum": sh, "search_edges": int(sk.sum()), "gelems": E / ms / 1e6}
        if sh != s3:
            log(f"graph {key}: sum {sh} differs from {s3}")
            sys.exit(3)
        if best is None or ms < res[best]["ms"]:
            best = key
    res["best_split"] = best
    res["tuned_split"] = f"split_{int(tuned)}_indexed"
    if s3 != 3 * t or s3b != s3:
        log(f"graph: sums disagree (merge {s3}, bsr {s3b}, oriented {t})")
        sys.exit(3)
    res["triangles"] = t
    log(f"graph results: { {k: res[k] for k in res if k.startswith(('merge', 'oriented', 'split'))} }")
    if with_cpu:
        import oracle

        rng = np.random.Generator(np.random.PCG64(7))
        sel = np.sort(rng.choice(u.size, size=max(1, u.size // 200), replace=False))
        su, sv = u[sel], v[sel]
        Es = int((deg[su] + deg[sv]).sum())
        t0 = time.perf_counter()
        oracle.batch_intersect_gather("branchless_merge", col, rp[su], rp[su + 1], col, rp[sv], rp[sv + 1],
                                      nthreads=cpu_threads, with_output=False)
        el = time.perf_counter() - t0
        res["cpu_port_branchless_merge"] = {"sample_edges": int(sel.size), "sample_elements": Es, "seconds": el,
                                            "threads": cpu_threads, "gelems": Es / el / 1e9}
        log(f"graph cpu: {res['cpu_port_branchless_merge']}")
    return res


def _median_us(fn, warmup_s=1.0, min_runs=5, run_s=1.0):
    """Harness::time's shape (benchmark/src/timer/harness.rs:35-62): calls for ~warmup_s untimed,
    then at least min_runs timed calls (more while run_s lasts); the median in microseconds."""
    t_end = time.perf_counter() + warmup_s
    while time.perf_counter() < t_end:
        fn()
    times = []
    t_end = time.perf_counter() + run_s
    while len(times) < min_runs or (time.perf_counter() < t_end and len(times) < 10000):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)) * 1e6, len(times)


def pair_leg(torch, with_cpu):
    """The single plain pair: 2^16 + 2^16 sorted u32 over 2^20 with 19,661 shared.  The CPU
    branchless_merge restatement and the reference's own QFilter (one thread), and the GPU plain
    drop-in through the C ABI -- host pointers (copies included) and device-resident sets."""
    import ctypes

    from simd_set_operations_amd import _native, datagen
    from simd_set_operations_amd.intersect import algo_id

    a, b = datagen.plain_pair()
    E = int(a.size + b.size)
    res = {"elements": E, "results": int(datagen.PLAIN_PAIR["shared"])}
    lib = _native.hip()
    algo = algo_id("hip_branchless_merge")
    out = np.zeros(min(a.size, b.size) + 4, np.uint32)
    n = ctypes.c_size_t()
    vp = ctypes.c_void_p

    def gpu_host():
        _native.check(lib.setops_hip_intersect2(algo, vp(a.ctypes.data), a.size, vp(b.ctypes.data), b.size,
                                                vp(out.ctypes.data), out.size, ctypes.byref(n)))

    gpu_host()
    if n.value != res["results"]:
        log(f"plain pair: {n.value} results, expected {res['results']}")
        sys.exit(3)
    us, runs = _median_us(gpu_host)
    res["gpu_host_pointers"] = {"us": us, "runs": runs, "gelems": E / us / 1e3}
    dA = torch.from_numpy(a.view(np.int32)).cuda()
    dB = torch.from_numpy(b.view(np.int32)).cuda()
    dO = torch.empty(out.size, dtype=torch.int32, device="cuda")
    dT = torch.zeros(2, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream()

    def gpu_dev():
        _native.check(lib.setops_hip_intersect2_dev(algo, dA.data_ptr(), None, a.size, dB.data_ptr(), None, b.size,
                                                    dO.data_ptr(), None, out.size, dT.data_ptr(), stream.cuda_stream))
        stream.synchronize()

    us, runs = _median_us(gpu_dev)
    res["gpu_device_resident"] = {"us": us, "runs": runs, "gelems": E / us / 1e3}
    if with_cpu:
        import oracle

        us, runs = _median_us(lambda: oracle.branchless_merge(a, b))
        res["cpu_port_branchless_merge"] = {"us": us, "runs": runs, "gelems": E / us / 1e3, "threads": 1}
        if oracle.ref_available():
            R = oracle.ref()
            Aa = oracle._aligned(a, a.size + 8)
            Ba = oracle._aligned(b, b.size + 8)
            Oa = oracle._aligned(np.zeros(0, np.uint32), min(a.size, b.size) + 8)
            args = (vp(Aa.ctypes.data), a.size, vp(Ba.ctypes.data), b.size, vp(Oa.ctypes.data))
            if R.intersect_qfilter_uint_b4_v2(*args) != res["results"]:
                log("plain pair: reference qfilter count differs")
                sys.exit(3)
            us, runs = _median_us(lambda: R.intersect_qfilter_uint_b4_v2(*args))
            res["cpu_reference_qfilter_c"] = {"us": us, "runs": runs, "gelems": E / us / 1e3, "threads": 1}
    return res


def bsr_leg(torch, A, a_off, B, b_off, out_total, steps, warmup):
    """The uneven batch in the BSR form (base = v >> 5, 32-bit state; setops/src/bsr.rs:140-164),
    encoded outside the timed region as the reference's harness does (harness.rs:121-122), through
    hip_branchless_merge_bsr with outputs.  Throughput counts the original elements; the result's
    popcount total must equal the plain result count."""
    from simd_set_operations_amd.batch import DeviceBatch
    from simd_set_operations_amd.graph import csr_to_bsr

    a_bo, a_bases, a_states = csr_to_bsr(a_off, A)
    b_bo, b_bases, b_states = csr_to_bsr(b_off, B)
    E = int(A.size + B.size)
    db = DeviceBatch("hip_branchless_merge_bsr", a_bases, a_bo, b_bases, b_bo, A_states=a_states, B_states=b_states)
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        db.run(stream)
    torch.cuda.synchronize()
    words, elems = db.totals()
    if elems != out_total:
        log(f"bsr: popcount total {elems} differs from {out_total}")
        sys.exit(3)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        db.run(stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    db.kernel_time()
    db.set_timing(True)
    for _ in range(steps):
        db.run(stream)
    torch.cuda.synchronize()
    k_ms, runs = db.kernel_time()
    db.close()
    r = {"words_in": int(a_bases.size + b_bases.size), "elements_in": E, "result_words": words,
         "result_elements": elems, "ms_per_step": ms, "kernel_ms": k_ms / max(runs, 1), "gelems": E / ms / 1e6}
    log(f"bsr: {r}")
    return r


def kset_leg(with_cpu):
    """k-set queries through the host-pointer k-set call (the plain drop-in for svs_generic and
    small_adaptive): sets generated as the reference's gen_kset (generators.rs:113-135) with
    max_len 2^18, skew factor 1 (set i has 2^18 / (i+1) elements), density 1%, selectivity 10%.
    Per query: median of 5 calls (upload, launches, download), svs over merge / galloping and
    small_adaptive (one launch per 1 + 8 sets, no intermediate round trips)."""
    from simd_set_operations_amd import experiment, intersect

    res = {}
    for k in (3, 5, 8):
        props = {"set_count": k, "density": 10, "selectivity": 100, "max_len": 18, "skewness_factor": 1000}
        queries = [experiment.gen_kset(props, np.random.Generator(np.random.PCG64([77, k, i]))) for i in range(3)]
        r = {"elements_per_query": int(sum(s.size for s in queries[0]))}
        for algo in ("hip_branchless_merge", "hip_galloping", "hip_small_adaptive"):
            times = []
            for sets in queries:
                exp = intersect.run_kset(sets, algo)  # warm
                ts = []
                for _ in range(5):
                    t0 = time.perf_counter()
                    got = intersect.run_kset(sets, algo)
                    ts.append(time.perf_counter() - t0)
                if got.size != exp.size:
                    log(f"kset {algo}: result changed between calls")
                    sys.exit(3)
                times.append(float(np.median(ts)) * 1e3)
            r[algo] = {"ms_per_query": float(np.mean(times))}
        if with_cpu:
            import oracle

            for name, fn in (("cpu_port_small_adaptive", oracle.small_adaptive),
                             ("cpu_port_svs_branchless_merge", lambda q: oracle.svs("branchless_merge", q))):
                t0 = time.perf_counter()
                for sets in queries:
                    got = fn(sets)
                    if not np.array_equal(got, intersect.run_kset(sets, "hip_small_adaptive")):
                        log(f"kset: {name} differs from hip_small_adaptive")
                        sys.exit(3)
                r[name] = {"ms_per_query": (time.perf_counter() - t0) / len(queries) * 1e3, "threads": 1}
        res[f"k{k}"] = r
        log(f"kset k={k}: {r}")
    return res


def serving_leg(torch, rates, seconds_per_rate, plain_budget_s, mode):
    """The served k-set workload: a device-resident corpus of 2^20 sorted lists (bounded-Zipf(1.1)
    lengths in [2^4, 2^18] over [0, 2^26)); queries of k ~ U{2..6} lists by Zipf(1.0) popularity,
    answered smallest-first, arriving as Poisson streams.  The continuous-batching engine (each
    admitted query answered whole in one batch of k-set groups, two steps in flight on two streams,
    admission against the free HBM less a reserve; its request loop in the library) is compared
    with the same engine driven by its Python loop and with the plain path (one query at a time,
    one launch sequence per SvS round) on the same stream.  Per-query elements follow the
    reference's convention (sum of the k list sizes)."""
    from simd_set_operations_amd import serving

    t0 = time.perf_counter()
    corpus = serving.Corpus(mode=mode)
    res = {"lists": corpus.nlists, "length_mode": mode, "elements": int(corpus.lens.sum()),
           "corpus_gb": corpus.total * 4 / 1e9, "gen_seconds": time.perf_counter() - t0}
    log(f"serving corpus: {res}")
    eng = serving.Engine(corpus, "hip_galloping")
    ids, ks = serving.make_queries(corpus, 4000, seed=11)
    eng.serve(ids, ks, np.zeros(ids.shape[0]))  # warm up
    for i, rate in enumerate(rates):
        nq = int(rate * seconds_per_rate)
        ids, ks = serving.make_queries(corpus, nq, seed=100 + i)
        arr = serving.poisson_arrivals(nq, rate, seed=200 + i)
        lens = np.where(ids >= 0, corpus.lens[np.maximum(ids, 0)], 0)
        E = int(lens.sum())
        wids, wks = serving.make_queries(corpus, 2000, seed=300 + i)  # warm the engine at this rate
        eng.serve(wids, wks, serving.poisson_arrivals(2000, rate, seed=400 + i))
        lat, _, st = eng.serve(ids, ks, arr)
        r = {"queries": nq, "engine": serving.latency_summary(lat), "engine_steps": st["steps"],
             "engine_queries_per_step": st["pairs"] / max(st["steps"], 1), "engine_qps": st["served"] / st["wall_s"],
             "engine_gelems": E / st["wall_s"] / 1e9}
        eng.serve(wids, wks, serving.poisson_arrivals(2000, rate, seed=400 + i), loop="python")
        lat_p, _, st_p = eng.serve(ids, ks, arr, loop="python")
        r["engine_python_loop"] = serving.latency_summary(lat_p)
        r["engine_python_loop_steps"] = st_p["steps"]
        if plain_budget_s > 0:
            plat, _, pst = serving.plain_serve(corpus, ids, ks, arr, timeout_s=plain_budget_s)
            r["plain"] = serving.latency_summary(plat)
            r["plain_qps"] = pst["served"] / pst["wall_s"]
            r["plain_saturated"] = pst["served"] < nq
        res[f"{int(rate)}qps"] = r
        log(f"serving {int(rate)} qps: {r}")
    eng.close()
    del corpus
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--npairs", type=int, default=65536)
    ap.add_argument("--algo", default="hip_branchless_merge", help="registered entry for the headline batch")
    ap.add_argument("--compare", default="hip_qfilter,hip_shuffling,hip_galloping",
                    help="comma-separated entries also timed on the headline batch (extras)")
    ap.add_argument("--cpu-budget", type=float, default=6.0, help="seconds of CPU baseline timing")
    ap.add_argument("--plain-sample", type=int, default=2048)
    ap.add_argument("--no-baselines", action="store_true")
    ap.add_argument("--no-skewed", action="store_true", help="skip the skewed resident-set leg")
    ap.add_argument("--skewed-only", default="", metavar="ENTRIES",
                    help="only run the skewed leg over these comma-separated entries (name[@shape][:index]) and "
                         "print its results as one JSON line: one configuration per process for profiling")
    ap.add_argument("--no-graph", action="store_true", help="skip the R-MAT graph leg")
    ap.add_argument("--graph-scale", type=int, default=22)
    ap.add_argument("--no-serving", action="store_true", help="skip the served k-set leg")
    ap.add_argument("--no-kset", action="store_true", help="skip the host-pointer k-set leg")
    ap.add_argument("--no-bsr", action="store_true", help="skip the BSR form of the uneven batch")
    ap.add_argument("--no-pair", action="store_true", help="skip the single plain pair (2^16 + 2^16)")
    ap.add_argument("--serving-seconds", type=float, default=0.5, help="seconds of arrivals per rate")
    ap.add_argument("--serving-plain-seconds", type=float, default=5.0,
                    help="time limit of the plain one-query-at-a-time path per rate (0: not run)")
    ap.add_argument("--corpus-mode", choices=("iid", "rank"), default="iid")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (3 times) and exit: the reference kernel of "
                         "known traffic that scales the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the library runs on one GPU and has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.arange(n, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}))
        return

    from simd_set_operations_amd import datagen, intersect
    from simd_set_operations_amd.batch import DeviceBatch

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    intersect.init(0)

    if args.skewed_only:
        entries = tuple(e.strip() for e in args.skewed_only.split(",") if e.strip())
        print(json.dumps({"skewed": skewed_leg(torch, args.steps, 1, False, entries)}, default=float))
        return

    t0 = time.perf_counter()
    A, a_off, B, b_off, shared = datagen.uneven_batch(npairs=args.npairs, with_shared=True)
    P = a_off.size - 1
    E = int(A.size + B.size)
    log(f"generated P={P} E={E} in {time.perf_counter() - t0:.1f}s")

    stream = torch.cuda.current_stream()

    def time_batch(algo):
        """Warm up, check, then time `steps` runs with outputs and `steps` count-only runs."""
        db = DeviceBatch(algo, A, a_off, B, b_off)
        for _ in range(args.warmup):
            db.run(stream)
        torch.cuda.synchronize()
        # correctness gate on the bench workload: per-pair counts equal the planted shared counts
        off, cnt, _ = db.results()
        if not np.array_equal(cnt, shared):
            log(f"{algo}: per-pair counts differ from the planted shared counts")
            sys.exit(3)
        out_total = int(off[-1])
        r = {"results": out_total, "info": db.info()}
        for mode in ("out", "count"):
            co = mode == "count"
            # steps: the run as a caller makes it (a repeated run replays the captured launch graph)
            for _ in range(3):
                db.run(stream, count_only=co)
            torch.cuda.synchronize()
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            for _ in range(args.steps):
                db.run(stream, count_only=co)
            ev1.record(stream)
            torch.cuda.synchronize()
            step = ev0.elapsed_time(ev1) / args.steps
            # the tile kernel alone: HIP events recorded around it on the launch stream (eager runs)
            db.kernel_time()  # reset
            db.set_timing(True)
            for _ in range(args.steps):
                db.run(stream, count_only=co)
            torch.cuda.synchronize()
            k_ms, runs = db.kernel_time()
            db.set_timing(False)
            r[mode] = {"ms_per_step": step, "kernel_ms": k_ms / max(runs, 1)}
            r[mode]["gelems"] = E / (r[mode]["ms_per_step"] / 1e3) / 1e9
        words, _ = db.totals()
        if words != out_total:
            log(f"{algo}: count-only total {words} differs from {out_total}")
            sys.exit(3)
        db.close()
        log(f"{algo}: {r}")
        return r

    head = time_batch(args.algo)
    out_total = head["results"]
    step_ms = head["out"]["ms_per_step"]
    kern_ms = head["out"]["kernel_ms"]
    value = E / (step_ms / 1e3) / 1e9
    Bytes = algorithmic_bytes(E, out_total, P)
    achieved = Bytes / (kern_ms / 1e3) / 1e9
    log(f"step {step_ms:.4f} ms, main kernel {kern_ms:.4f} ms, {value:.1f} Gelem/s, {achieved:.0f} GB/s")

    extra = {"count_only": head["count"], "compare": {}}
    try:
        from simd_set_operations_amd import tuning

        pred = tuning.predict(np.diff(a_off), np.diff(b_off))
        extra["autotune"] = {"uneven_batch_choice": min(pred, key=pred.get), "predicted_ms": pred,
                             "skewed_choice": tuning.choose(np.full(4096, 256), np.full(4096, 1 << 24))}
    except FileNotFoundError:
        extra["autotune"] = None
    for name in [x for x in args.compare.split(",") if x and x != args.algo]:
        c = time_batch(name)
        c["roofline_frac"] = Bytes / (c["out"]["kernel_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS
        extra["compare"][name] = c
    cpu = None
    if not args.no_baselines:
        pg = plain_gpu(torch, A, a_off, B, b_off, args.plain_sample, stream)
        extra["plain_gpu_per_pair"] = pg
        extra["speedup_vs_plain_gpu"] = value / pg["gelems"]
        log(f"plain per-pair GPU: {pg}")
        res, threads = cpu_baselines(A, a_off, B, b_off, E, args.cpu_budget)
        best_kind, best = max(((k, v) for k, v in res.items() if not k.startswith("countonly_")),
                              key=lambda kv: kv[1]["gelems"])
        cpu = {
            "value": best["gelems"],
            "unit": "Gelem/s",
            "cores": threads,
            "kind": "reference" if best_kind.startswith("reference") else "port",
            "sample": f"full {P}-pair batch x {best['reps']} repetitions, one pair per host thread "
                      f"({best_kind})",
        }
        import oracle

        dram = oracle.stream_read_gbs(threads)
        cpu["host"] = host_info()
        cpu["dram_read_gbs"] = dram
        extra["cpu_all"] = res
        extra["speedup_vs_cpu"] = value / best["gelems"]
        # the ratio beside the two memory rooflines: HBM peak over the host's measured DRAM read
        extra["cpu_vs_gpu_rooflines"] = {"gpu_hbm_peak_gbs": HBM_PEAK_GBS, "host_dram_read_gbs": dram,
                                         "bandwidth_ratio": HBM_PEAK_GBS / dram,
                                         "speedup_vs_cpu": value / best["gelems"],
                                         "cpu_bytes_per_s_fraction_of_dram": best["gelems"] * 4 / dram}
        log(f"cpu: {res}, host {cpu['host']}, dram read {dram:.1f} GB/s")
    if not args.no_pair:
        extra["plain_pair"] = pair_leg(torch, not args.no_baselines)
        log(f"plain pair: {extra['plain_pair']}")
    if not args.no_bsr:
        extra["bsr"] = bsr_leg(torch, A, a_off, B, b_off, out_total, args.steps, args.warmup)
    if not args.no_kset:
        extra["kset"] = kset_leg(not args.no_baselines)
    if not args.no_serving:
        extra["serving"] = serving_leg(torch, (5e4, 2e5, 8e5), args.serving_seconds, args.serving_plain_seconds,
                                       args.corpus_mode)
    if not args.no_graph:
        extra["graph"] = graph_leg(torch, host_threads(), not args.no_baselines, args.graph_scale)
    if not args.no_skewed:
        extra["skewed"] = skewed_leg(torch, args.steps, host_threads(), not args.no_baselines)

    prof_us = profiled_avg_us(DOMINANT_KERNEL[args.algo])
    rec = {
        "metric": "batched intersect Gelem/s + %HBM roofline; p50/p99 query latency vs plain",
        "value": value,
        "unit": "Gelem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded SplitMix64 generator, planted shared elements)",
        "config": {
            "workload": "uneven_batch",
            "pairs": P,
            "elements": E,
            "results": out_total,
            "larger_log2_range": [6, 14],
            "max_skew": 1024,
            "density": 0.1,
            "algorithm": args.algo,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(DOMINANT_KERNEL[args.algo]),
            "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT),
            "kernel": DOMINANT_KERNEL[args.algo],
            "kernel_ms": kern_ms,
            "algorithmic_bytes": Bytes,
            # the same bytes over the kernel's average in the committed rocprofv3 summary, and over
            # the whole step (plan, split, tile kernel, scan, placement)
            "frac_profile": (Bytes / (prof_us * 1e-6) / 1e9 / HBM_PEAK_GBS) if prof_us else None,
            "profile_avg_us": prof_us,
            "profile_source": os.path.relpath(TRACE_FILE, ROOT),
            "step_frac": Bytes / (step_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
        },
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
