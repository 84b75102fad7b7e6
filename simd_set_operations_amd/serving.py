"""Served k-set queries over a device-resident corpus: continuous batching of k-set groups.

A corpus of sorted u32 lists stays in one device pool.  A query names k lists, smallest first; its
answer is their intersection, what SvS gives (svs_generic, setops/src/intersect/svs.rs:50-73).  The
engine answers every admitted query completely in one step: a step is one batch of k-set groups
(setops_hip_batch_set_further) -- the query's first two lists are a pair of a gather batch and its
other lists are that pair's further sets, in which every element of the pair's intersection is
looked up in turn (small_adaptive's walk, setops/src/intersect/adaptive.rs:107-147).

* Two steps are in flight, on two streams, each with its own batch handle and output buffer: while
  the device runs one step, the host admits the queries that arrived meanwhile and launches the
  next (the device may run both at once); then it waits for the older step, finishes its queries
  (results and offsets come back by copies enqueued behind the step) and frees their memory.
* Admission is in arrival order against a device-memory budget -- the free HBM at start-up less a
  stated reserve -- with each query holding its worst-case cost while in flight, and against the
  pair slots of a step.

Latency is measured from a query's (Poisson) arrival time to the host's view of the end of the step
that answered it.  The plain baseline answers the same query stream one query at a time with one
launch sequence per SvS round (setops_hip_intersect2_dev), each round synchronised to size the
next.
"""
import ctypes
import time
from collections import deque

import numpy as np

from . import _native
from .intersect import algo_id

_c = ctypes

CORPUS = dict(nlists=1 << 20, lmin=16, lmax=1 << 18, zipf_s=1.1, universe=1 << 26, seed=1005)
QUERIES = dict(kmin=2, kmax=6, zipf_s=1.0, seed=1006)


def _uniform(seed, n):
    from .datagen import splitmix64

    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def zipf_lengths(n, lmin, lmax, s, seed, mode="iid"):
    """List lengths: i.i.d. bounded Zipf, P(L) ~ L^-s on [lmin, lmax] ("iid"), or by rank,
    L_r = max(lmin, floor(lmax * r^-s)) for r = 1..n in a seeded order ("rank")."""
    if mode == "iid":
        support = np.arange(lmin, lmax + 1, dtype=np.float64)
        cdf = np.cumsum(support ** -s)
        cdf /= cdf[-1]
        idx = np.searchsorted(cdf, _uniform(seed, n), side="right")
        return (lmin + np.minimum(idx, support.size - 1)).astype(np.uint64)
    r = np.arange(1, n + 1, dtype=np.float64)
    lens = np.maximum(lmin, np.floor(lmax * r ** -s)).astype(np.uint64)
    return lens[np.argsort(_uniform(seed ^ 0x5A5A, n), kind="stable")]


class Corpus:
    """Sorted unique lists generated in device memory, in one pool (each list 16-byte aligned)."""

    def __init__(self, nlists=CORPUS["nlists"], lmin=CORPUS["lmin"], lmax=CORPUS["lmax"], zipf_s=CORPUS["zipf_s"],
                 universe=CORPUS["universe"], seed=CORPUS["seed"], mode="iid"):
        import torch

        self.torch = torch
        self.lens = zipf_lengths(nlists, lmin, lmax, zipf_s, seed, mode)
        if np.any(self.lens > universe):
            raise ValueError("a list is longer than the element space")
        # list i at pool[beg[i] .. beg[i] + lens[i]), starts rounded up to 16 bytes (the plain
        # per-pair path loads 16-byte chunks from a list's first element)
        padded = (self.lens + np.uint64(3)) // np.uint64(4) * np.uint64(4)
        self.beg = np.zeros(nlists, np.uint64)
        np.cumsum(padded[:-1], out=self.beg[1:])
        self.total = int(self.beg[-1] + padded[-1])
        self.nlists = nlists
        self.universe = universe
        self.pool = torch.empty(self.total + 64, dtype=torch.int32, device="cuda")
        d_beg = torch.from_numpy(self.beg.view(np.int64)).cuda()
        d_len = torch.from_numpy(self.lens.view(np.int64)).cuda()
        rc = _native.datagen_hip().setops_gen_sorted_lists_dev(seed, universe, d_beg.data_ptr(), d_len.data_ptr(),
                                                                nlists, self.pool.data_ptr(), None)
        if rc != 0:
            raise RuntimeError(f"corpus generation failed ({rc})")
        torch.cuda.synchronize()

    def list(self, i):
        """Host copy of list i."""
        a = int(self.beg[i])
        return self.pool[a:a + int(self.lens[i])].cpu().numpy().view(np.uint32).copy()


def make_queries(corpus, nq, kmin=QUERIES["kmin"], kmax=QUERIES["kmax"], zipf_s=QUERIES["zipf_s"],
                 seed=QUERIES["seed"]):
    """Queries as an (nq, kmax) array of list ids ordered smallest-first (padding -1) and k per
    query.  k ~ U{kmin..kmax}; ids by Zipf(s) popularity over a seeded permutation of the lists,
    distinct within a query."""
    n = corpus.nlists
    ranks = np.arange(1, n + 1, dtype=np.float64)
    cdf = np.cumsum(ranks ** -zipf_s)
    cdf /= cdf[-1]
    perm = np.argsort(_uniform(seed, n), kind="stable")
    u = _uniform(seed + 1, nq * kmax * 2)
    ks = kmin + (_uniform(seed + 2, nq) * (kmax - kmin + 1)).astype(np.int64)
    ids = np.full((nq, kmax), -1, np.int64)
    pos = 0
    for q in range(nq):
        chosen = []
        while len(chosen) < ks[q]:
            if pos >= u.size:
                u = np.concatenate([u, _uniform(seed + 3 + pos, nq * kmax)])
            x = int(perm[min(np.searchsorted(cdf, u[pos], side="right"), n - 1)])
            pos += 1
            if x not in chosen:
                chosen.append(x)
        chosen.sort(key=lambda i: (int(corpus.lens[i]), i))
        ids[q, :ks[q]] = chosen
    return ids, ks


def poisson_arrivals(nq, rate, seed):
    """Arrival times (seconds from 0) of a Poisson stream at `rate` queries/s."""
    gaps = -np.log(1.0 - _uniform(seed, nq)) / rate
    return np.cumsum(gaps)


class Admission:
    """Admission of arriving queries against a device-memory budget and a pair-slot limit (host
    logic only).  A query reserves ``cost`` words while it is in flight; queries are admitted in
    arrival order, and the first that does not fit stops the admission (no overtaking)."""

    def __init__(self, budget_words, max_pairs):
        self.budget = int(budget_words)
        self.max_pairs = int(max_pairs)
        self.reserved = 0
        self.active = 0

    def admit(self, pending, cost, limit=None):
        """Pops from the deque `pending` the queries that fit now (at most `limit`)."""
        add = []
        room = self.max_pairs if limit is None else min(self.max_pairs, limit)
        while pending and self.active + len(add) < self.max_pairs and len(add) < room:
            q = pending[0]
            c = int(cost[q])
            if self.reserved + c > self.budget:
                break
            pending.popleft()
            self.reserved += c
            add.append(q)
        self.active += len(add)
        return add

    def release(self, words, n):
        self.reserved -= int(words)
        self.active -= int(n)


def query_cost_words(corpus, ids):
    """Worst-case device words a query holds in flight: its result (at most its smallest list,
    min(|L0|, |L1|) with the lists smallest-first), the result's staging and the plan of its
    probe tiles (about two more words per probe), plus a fixed 256 words."""
    need = corpus.lens[np.maximum(ids[:, 0], 0)].astype(np.int64)
    return need, 3 * need + 256


RESERVE_UNITS = 1 << 24      # probes a slot's batch is sized for at start-up
RESERVE_OUT_WORDS = 1 << 22  # result words a slot's buffers hold at start-up


class _Slot:
    """One step in flight: a gather batch handle over the corpus pool with its own stream, output
    buffer and host staging (grow-only)."""

    def __init__(self, eng):
        torch = eng.torch
        self.torch = torch
        lib = eng.lib
        self.stream = torch.cuda.Stream()
        h = _c.c_void_p()
        one = torch.tensor([0, 0, 0, 0], dtype=torch.int64, device="cuda")
        pool = eng.c.pool
        _native.check(lib.setops_hip_batch_create_gather(
            _c.byref(h), algo_id(eng.algo), pool.data_ptr(), None, pool.numel(), one.data_ptr(), one.data_ptr() + 8,
            pool.data_ptr(), None, pool.numel(), one.data_ptr() + 16, one.data_ptr() + 24, 1))
        self.handle = h
        self._one = one
        self.d_idx = self.h_idx = None
        self.out = self.h_res = None
        # sized ahead for a step of max_pairs queries and 2^24 probes with 2^22 result words:
        # growing later frees and reallocates device memory, which waits for the device
        _native.check(lib.setops_hip_batch_reserve(h, eng.max_pairs, RESERVE_UNITS))
        self._grow(13 * eng.max_pairs + 64, RESERVE_OUT_WORDS)
        self.d_off = torch.zeros(eng.max_pairs + 1, dtype=torch.int64, device="cuda")
        self.d_cnt = torch.zeros(eng.max_pairs, dtype=torch.int32, device="cuda")
        self.d_tot = torch.zeros(2, dtype=torch.int64, device="cuda")
        self.h_off = torch.zeros(eng.max_pairs + 1, dtype=torch.int64).pin_memory()
        self.done = torch.cuda.Event()
        self.queries = None

    def _grow(self, words_idx, words_out):
        torch = self.torch
        if self.d_idx is None or self.d_idx.numel() < words_idx:
            n = max(words_idx, 1 << 16)
            self.d_idx = torch.empty(n, dtype=torch.int64, device="cuda")
            self.h_idx = torch.empty(n, dtype=torch.int64).pin_memory()
        if self.out is None or self.out.numel() < words_out:
            n = max(words_out, 1 << 20)
            self.out = torch.empty(n, dtype=torch.int32, device="cuda")
            self.h_res = torch.empty(n, dtype=torch.int32).pin_memory()

    def close(self, lib):
        if self.handle:
            lib.setops_hip_batch_destroy(self.handle)
            self.handle = None


class Engine:
    """Continuous batching of k-set queries over a resident Corpus.

    A step is one batch of k-set groups (setops_hip_batch_set_further): each admitted query is the
    pair (list 0, list 1) with its lists 2 .. k-1 as further sets, so it is answered completely in
    that step (no intermediate results, no host round trip per list).  Two steps are in flight on
    two streams: while the device runs one, the host admits the queries that arrived meanwhile and
    launches the next, which can run beside it; then it waits for the older step and finishes its
    queries.  Admission is by the device memory free at start-up (``budget_bytes``, default: the
    free HBM less ``reserve_bytes``) and by pair slots per step (``max_pairs``).

    ``index``: index the corpus pool once (every 32nd element), so the search entries probe long
    lists and further sets with about one 128-byte line per lookup."""

    def __init__(self, corpus, algo="hip_galloping", max_pairs=16384, index=True, reserve_bytes=16 << 30,
                 budget_bytes=None, nslots=2):
        torch = corpus.torch
        self.c = corpus
        self.torch = torch
        self.algo = algo
        self.max_pairs = int(max_pairs)
        self.lib = _native.hip()
        if budget_bytes is None:
            free, _ = torch.cuda.mem_get_info()
            budget_bytes = free - int(reserve_bytes)
        if budget_bytes <= 0:
            raise ValueError(f"no device memory left for serving (free HBM minus a {reserve_bytes} byte reserve)")
        self.budget_bytes = int(budget_bytes)
        if nslots < 2:
            raise ValueError("the engine needs at least two step slots")
        # the native loop takes any number of slots (steps launched round-robin, the host waiting
        # for the oldest); the Python loop alternates between the first two
        self.slots = [_Slot(self) for _ in range(int(nslots))]
        if index:
            # every pair and every further set reads corpus lists, which never change
            for sl in self.slots:
                _native.check(self.lib.setops_hip_batch_index_b(sl.handle))

    def close(self):
        for sl in self.slots:
            sl.close(self.lib)

    @staticmethod
    def _stream_rows(c, ids, ks, need):
        """Per-query step rows of a whole stream, laid out once before it is served: the pair's
        bounds (A begin, A end, B begin, B end), the sizes, and the further sets' bounds as one
        flat array with an offset per query (queries are admitted in arrival order, so a step's
        queries are a contiguous range and take slices of these)."""
        l0 = ids[:, 0]
        l1 = ids[:, 1]
        b0, e0 = c.beg[l0], c.beg[l0] + c.lens[l0]
        b1, e1 = c.beg[l1], c.beg[l1] + c.lens[l1]
        pairs = np.stack([b0, e0, b1, e1], axis=0).astype(np.uint64)  # (4, nq)
        nx = (ks - 2).astype(np.int64)
        xoff = np.zeros(ids.shape[0] + 1, np.int64)
        np.cumsum(nx, out=xoff[1:])
        X = int(xoff[-1])
        xq = np.repeat(np.arange(ids.shape[0]), nx)
        xj = np.arange(X) - np.repeat(xoff[:-1], nx) + 2
        xl = ids[xq, xj] if X else np.zeros(0, np.int64)
        xb = c.beg[xl].astype(np.uint64)
        xe = (c.beg[xl] + c.lens[xl]).astype(np.uint64)
        need_pfx = np.zeros(ids.shape[0] + 1, np.int64)
        np.cumsum(need, out=need_pfx[1:])
        return {"pairs": pairs, "na": np.ascontiguousarray(c.lens[l0], np.uint64),
                "nb": np.ascontiguousarray(c.lens[l1], np.uint64), "xoff": xoff, "xb": xb, "xe": xe,
                "need_pfx": need_pfx}

    def _launch(self, sl, q0, q1, rows):
        """Enqueue one step for queries q0 .. q1-1 on slot `sl` (nothing waits here)."""
        P = q1 - q0
        x0, x1 = int(rows["xoff"][q0]), int(rows["xoff"][q1])
        X = x1 - x0
        out_words = int(rows["need_pfx"][q1] - rows["need_pfx"][q0]) + 64
        words = 5 * P + 1 + 2 * X
        sl._grow(words, out_words)
        h = sl.h_idx.numpy().view(np.uint64)
        h[0:4 * P].reshape(4, P)[:] = rows["pairs"][:, q0:q1]
        np.subtract(rows["xoff"][q0:q1 + 1], x0, out=h[4 * P:5 * P + 1], casting="unsafe")
        h[5 * P + 1:5 * P + 1 + X] = rows["xb"][x0:x1]
        h[5 * P + 1 + X:5 * P + 1 + 2 * X] = rows["xe"][x0:x1]
        na = rows["na"][q0:q1]
        nb = rows["nb"][q0:q1]
        # upload, rebind, further sets, run and the copies back: one call on the slot's stream
        _native.check(self.lib.setops_hip_batch_step(
            sl.handle, sl.h_idx.data_ptr(), sl.d_idx.data_ptr(), P, X, na.ctypes.data, nb.ctypes.data,
            sl.out.data_ptr(), sl.out.numel(), sl.d_off.data_ptr(), sl.d_cnt.data_ptr(), sl.d_tot.data_ptr(),
            sl.h_off.data_ptr(), sl.h_res.data_ptr(), out_words, sl.stream.cuda_stream))
        sl.done.record(sl.stream)
        sl.queries = np.arange(q0, q1)

    def serve(self, ids, ks, arrivals, keep_results=False, timeout_s=60.0, loop="native"):
        """Serve the query stream; returns (latency seconds per query, results list or None, stats).

        ``loop="native"``: the request loop runs in the library (setops_hip_serve), a few
        microseconds of host work per step; ``loop="python"``: the same loop written here (the
        admission, the step rows and the two slots are the same, so are the results)."""
        if loop == "native":
            return self._serve_native(ids, ks, arrivals, keep_results, timeout_s)
        if loop != "python":
            raise ValueError(f"unknown request loop {loop!r}")
        return self._serve_python(ids, ks, arrivals, keep_results, timeout_s)

    def _slot_desc(self, sl):
        # rows and results: as many words per step as the Python loop's buffers hold
        return _native.ServeSlot(
            sl.handle.value, min(sl.h_idx.numel(), sl.d_idx.numel()), min(sl.out.numel(), sl.h_res.numel()),
            sl.d_tot.data_ptr(), self.max_pairs, sl.stream.cuda_stream)

    def _serve_native(self, ids, ks, arrivals, keep_results, timeout_s):
        import gc

        nq = ids.shape[0]
        need, cost = query_cost_words(self.c, ids)
        budget = self.budget_bytes // 4
        if np.any(cost > budget):
            raise ValueError("a query needs more device memory than the serving budget")
        rows = self._stream_rows(self.c, ids, ks, need)
        if nq:
            # a slot's step buffers hold at least the largest query's step (steps never grow)
            one_rows = int((5 + 2 * np.diff(rows["xoff"])).max()) + 1
            for sl in self.slots:
                sl._grow(one_rows, int(need.max()) + 64)
        self.torch.cuda.synchronize()
        keep = []  # the host arrays the library reads, alive for the call

        def arr(x, dt):
            a = np.ascontiguousarray(x, dtype=dt)
            keep.append(a)
            return a.ctypes.data

        arrivals = np.ascontiguousarray(arrivals, dtype=np.float64)
        q = _native.ServeQueries(nq, arr(rows["pairs"], np.uint64), arr(rows["na"], np.uint64),
                                 arr(rows["nb"], np.uint64), arr(rows["xoff"], np.int64), arr(rows["xb"], np.uint64),
                                 arr(rows["xe"], np.uint64), arr(rows["need_pfx"], np.int64), arr(cost, np.int64),
                                 arr(arrivals, np.float64))
        slots = (_native.ServeSlot * len(self.slots))(*[self._slot_desc(sl) for sl in self.slots])
        done_t = np.full(nq, np.nan)
        st = np.zeros(4, np.uint64)
        res = np.empty(max(int(rows["need_pfx"][-1]), 1), np.uint32) if keep_results else None
        counts = np.zeros(max(nq, 1), np.uint64) if keep_results else None
        gc_was = gc.isenabled()
        gc.collect()
        gc.disable()
        try:
            _native.check(self.lib.setops_hip_serve(
                _c.byref(slots), len(self.slots), _c.byref(q), budget, self.max_pairs, float(timeout_s),
                done_t.ctypes.data, res.ctypes.data if keep_results else None,
                counts.ctypes.data if keep_results else None, st.ctypes.data))
        finally:
            if gc_was:
                gc.enable()
        results = None
        if keep_results:
            pfx = rows["need_pfx"]
            results = [res[int(pfx[i]):int(pfx[i]) + int(counts[i])].copy() if np.isfinite(done_t[i]) else None
                       for i in range(nq)]
        stats = {"steps": int(st[0]), "pairs": int(st[1]), "served": int(st[2]), "wall_s": float(st[3]) / 1e9,
                 "budget_bytes": self.budget_bytes}
        return done_t - arrivals, results, stats

    def _serve_python(self, ids, ks, arrivals, keep_results, timeout_s):
        import gc

        nq = ids.shape[0]
        need, cost = query_cost_words(self.c, ids)
        adm = Admission(self.budget_bytes // 4, self.max_pairs)
        if np.any(cost > adm.budget):
            raise ValueError("a query needs more device memory than the serving budget")
        gc_was = gc.isenabled()
        gc.collect()
        gc.disable()  # no collector pauses inside the timed loop
        rows = self._stream_rows(self.c, ids, ks, need)
        done_t = np.full(nq, np.nan)
        results = [None] * nq if keep_results else None
        pending = deque()
        nxt = 0
        steps = pairs = ndone = 0
        cur = 0
        t0 = time.perf_counter()
        while ndone < nq:
            now = time.perf_counter() - t0
            if now > timeout_s:
                break
            while nxt < nq and arrivals[nxt] <= now:
                pending.append(nxt)
                nxt += 1
            sl, other = self.slots[cur], self.slots[cur ^ 1]
            if pending:  # this slot is free: its last step was finished below
                add = adm.admit(pending, cost)
                if add:
                    if add[-1] - add[0] + 1 != len(add):
                        raise RuntimeError("admission must take queries in arrival order")
                    self._launch(sl, add[0], add[-1] + 1, rows)
                    steps += 1
                    pairs += len(add)
            if other.queries is not None:
                other.done.synchronize()
                t_end = time.perf_counter() - t0
                qs = other.queries
                off = other.h_off[:qs.size + 1].numpy().view(np.uint64)
                done_t[qs] = t_end
                ndone += qs.size
                adm.release(int(cost[qs].sum()), qs.size)
                if keep_results:
                    res = other.h_res.numpy().view(np.uint32)
                    for j, q in enumerate(qs.tolist()):
                        results[q] = res[int(off[j]):int(off[j + 1])].copy()
                other.queries = None
            elif sl.queries is None and not pending and nxt < nq:
                wait = arrivals[nxt] - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 1e-3))
            cur ^= 1
        for sl in self.slots:  # a timed-out stream leaves no step behind
            if sl.queries is not None:
                sl.done.synchronize()
                sl.queries = None
        lat = done_t - arrivals
        stats = {"steps": steps, "pairs": pairs, "served": int(np.isfinite(done_t).sum()),
                 "wall_s": time.perf_counter() - t0, "budget_bytes": self.budget_bytes}
        if gc_was:
            gc.enable()
        return lat, results, stats


def plain_serve(corpus, ids, ks, arrivals, algo="hip_galloping", timeout_s=20.0, keep_results=False):
    """One query at a time, one launch sequence per SvS round (setops_hip_intersect2_dev), in
    arrival order; a query starts when it has arrived and the previous one is done."""
    torch = corpus.torch
    lib = _native.hip()
    a = algo_id(algo)
    nq = ids.shape[0]
    cap = (int(corpus.lens.max()) + 63) // 64 * 64  # both halves 256-byte aligned
    scratch = torch.empty(2 * cap + 128, dtype=torch.int32, device="cuda")
    tot = torch.zeros(2, dtype=torch.int64, device="cuda")
    h_tot = torch.zeros(2, dtype=torch.int64).pin_memory()
    s = torch.cuda.Stream()
    pool = corpus.pool.data_ptr()
    buf = [scratch.data_ptr(), scratch.data_ptr() + 4 * (cap + 64)]
    done_t = np.full(nq, np.nan)
    results = [None] * nq if keep_results else None
    t0 = time.perf_counter()
    for q in range(nq):
        now = time.perf_counter() - t0
        if now > timeout_s:
            break
        if arrivals[q] > now:
            time.sleep(arrivals[q] - now)
        l0 = ids[q, 0]
        ap, an = pool + 4 * int(corpus.beg[l0]), int(corpus.lens[l0])
        k = int(ks[q])
        for r in range(1, k):
            lb = ids[q, r]
            out = buf[r & 1]
            _native.check(lib.setops_hip_intersect2_dev(a, ap, None, an, pool + 4 * int(corpus.beg[lb]), None,
                                                        int(corpus.lens[lb]), out, None, an, tot.data_ptr(),
                                                        s.cuda_stream))
            with torch.cuda.stream(s):
                h_tot.copy_(tot, non_blocking=True)
            s.synchronize()
            ap, an = out, int(h_tot[0])
            if an == 0:
                break
        if keep_results:
            off = (ap - scratch.data_ptr()) // 4
            results[q] = scratch[off:off + an].cpu().numpy().view(np.uint32).copy()
        done_t[q] = time.perf_counter() - t0
    served = int(np.isfinite(done_t).sum())
    return done_t - arrivals, results, {"served": served, "wall_s": time.perf_counter() - t0}


def latency_summary(lat):
    v = lat[np.isfinite(lat)]
    if v.size == 0:
        return {"served": 0}
    return {"served": int(v.size), "p50_ms": float(np.percentile(v, 50) * 1e3), "p99_ms": float(np.percentile(v, 99) * 1e3),
            "mean_ms": float(v.mean() * 1e3), "max_ms": float(v.max() * 1e3)}
