"""Served k-set queries over a device-resident corpus: continuous batching of SvS rounds.

A corpus of sorted u32 lists stays in one device pool.  A query names k lists; it is answered by
SvS smallest-first (svs_generic, setops/src/intersect/svs.rs:50-73): the running result meets the
next list, k-1 rounds.  The engine runs the rounds of *all* in-flight queries as one gather batch
per step:

* new arrivals are admitted while their worst-case intermediate results fit the free-HBM budget
  (an arena of device words, double-buffered) and a pair slot is free;
* every active query contributes one pair to the step -- (list 0, list 1) in its first round, then
  (its last result in the arena, next list) -- so queries at different rounds overlap on the device;
* after the step the host reads the per-pair counts and offsets, finishes the queries whose last
  round this was (or whose result became empty: every later round is empty too), and moves the
  others on.  A finished query's result is copied to the host in the same step.

Latency is measured from a query's (Poisson) arrival time to the end of the step that finished it.
The plain baseline answers the same query stream one query at a time with one launch sequence per
round (setops_hip_intersect2_dev), each round synchronised to size the next.
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
    """Sorted unique lists generated in device memory, followed by a result arena (two halves) in
    the same pool, so one gather batch can read corpus lists and earlier results alike."""

    def __init__(self, nlists=CORPUS["nlists"], lmin=CORPUS["lmin"], lmax=CORPUS["lmax"], zipf_s=CORPUS["zipf_s"],
                 universe=CORPUS["universe"], seed=CORPUS["seed"], mode="iid", arena_words=1 << 26):
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
        self.half = int(arena_words) // 2
        self.arena = [self.total + (-self.total % 64), 0]
        self.arena[1] = self.arena[0] + self.half + (-self.half % 64)
        words = self.arena[1] + self.half + 64
        self.pool = torch.empty(words, dtype=torch.int32, device="cuda")
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


class Engine:
    """Continuous-batching SvS engine over a Corpus (one gather batch handle, rebound every step).

    ``index``: index the corpus pool once (every 32nd element), so search-policy entries probe long
    lists in global memory with about one 128-byte line per lookup."""

    def __init__(self, corpus, algo="hip_galloping", max_pairs=16384, index=True):
        torch = corpus.torch
        self.c = corpus
        self.torch = torch
        self.algo = algo
        self.max_pairs = max_pairs
        lib = _native.hip()
        self.lib = lib
        self.d_idx = torch.zeros(4 * max_pairs, dtype=torch.int64, device="cuda")
        self.h_idx = torch.zeros(4 * max_pairs, dtype=torch.int64).pin_memory()
        self.d_out_off = torch.zeros(max_pairs + 1, dtype=torch.int64, device="cuda")
        self.d_counts = torch.zeros(max_pairs, dtype=torch.int32, device="cuda")
        self.d_tot = torch.zeros(2, dtype=torch.int64, device="cuda")
        self.h_off = torch.zeros(max_pairs + 1, dtype=torch.int64).pin_memory()
        self.stream = torch.cuda.Stream()
        # a gather handle over the whole pool, first bound to one dummy pair
        h = _c.c_void_p()
        one = torch.tensor([0, 0, 0, 0], dtype=torch.int64, device="cuda")
        pool = corpus.pool.data_ptr()
        _native.check(lib.setops_hip_batch_create_gather(
            _c.byref(h), algo_id(algo), pool, None, corpus.pool.numel(), one.data_ptr(), one.data_ptr() + 8, pool, None,
            corpus.pool.numel(), one.data_ptr() + 16, one.data_ptr() + 24, 1))
        self.handle = h
        self._keep = one
        if index:
            # the B side of every round is a corpus list, never the arena, so an index built
            # once over the pool stays valid where it is read
            _native.check(lib.setops_hip_batch_index_b(h))

    def close(self):
        if self.handle:
            self.lib.setops_hip_batch_destroy(self.handle)
            self.handle = None

    def _step(self, a_beg, a_len, b_beg, b_len, half):
        """One batched round for len(a_beg) pairs; returns (out_off[P+1], results on host)."""
        torch = self.torch
        P = a_beg.size
        h = self.h_idx.numpy().view(np.uint64)
        h[0:P] = a_beg
        h[P:2 * P] = a_beg + a_len
        h[2 * P:3 * P] = b_beg
        h[3 * P:4 * P] = b_beg + b_len
        s = self.stream
        with torch.cuda.stream(s):
            self.d_idx[:4 * P].copy_(self.h_idx[:4 * P], non_blocking=True)
            base = self.d_idx.data_ptr()
            na = np.ascontiguousarray(a_len, np.uint64)
            nb = np.ascontiguousarray(b_len, np.uint64)
            _native.check(self.lib.setops_hip_batch_rebind(
                self.handle, base, base + 8 * P, base + 16 * P, base + 24 * P, P,
                na.ctypes.data_as(_c.c_void_p), nb.ctypes.data_as(_c.c_void_p)))
            out_ptr = self.c.pool.data_ptr() + 4 * self.c.arena[half]
            _native.check(self.lib.setops_hip_batch_run(
                self.handle, out_ptr, None, self.c.half, self.d_out_off.data_ptr(), self.d_counts.data_ptr(),
                self.d_tot.data_ptr(), s.cuda_stream))
            self.h_off[:P + 1].copy_(self.d_out_off[:P + 1], non_blocking=True)
        s.synchronize()
        off = self.h_off[:P + 1].numpy().view(np.uint64).copy()
        total = int(off[-1])
        a0 = self.c.arena[half]
        res = self.c.pool[a0:a0 + total].cpu().numpy().view(np.uint32) if total else np.zeros(0, np.uint32)
        return off, res

    def serve(self, ids, ks, arrivals, keep_results=False, timeout_s=60.0):
        """Serve the query stream; returns (latency seconds per query, results list or None, stats)."""
        import gc

        c = self.c
        nq = ids.shape[0]
        lens = c.lens
        gc_was = gc.isenabled()
        gc.collect()
        gc.disable()  # no collector pauses inside the timed loop
        need = lens[np.maximum(ids[:, 0], 0)].astype(np.int64)  # worst-case intermediate words
        if np.any(need > c.half):
            raise ValueError("a query's smallest list exceeds the arena half")
        done_t = np.full(nq, np.nan)
        results = [None] * nq if keep_results else None
        pending = deque()
        nxt = 0
        # active state (numpy, one row per active query)
        act_q = np.zeros(0, np.int64)
        act_r = np.zeros(0, np.int64)  # rounds done
        act_ab = np.zeros(0, np.uint64)
        act_al = np.zeros(0, np.uint64)
        reserved = 0
        half = 0
        steps = 0
        pairs = 0
        ndone = 0
        t0 = time.perf_counter()
        while ndone < nq:
            now = time.perf_counter() - t0
            if now > timeout_s:
                break
            while nxt < nq and arrivals[nxt] <= now:
                pending.append(nxt)
                nxt += 1
            # admit in arrival order within the budget and the pair slots
            add = []
            while pending and act_q.size + len(add) < self.max_pairs and reserved + need[pending[0]] <= c.half:
                q = pending.popleft()
                reserved += int(need[q])
                add.append(q)
            if add:
                aq = np.array(add, np.int64)
                l0 = ids[aq, 0]
                act_q = np.concatenate([act_q, aq])
                act_r = np.concatenate([act_r, np.zeros(aq.size, np.int64)])
                act_ab = np.concatenate([act_ab, c.beg[l0]])
                act_al = np.concatenate([act_al, lens[l0]])
            if act_q.size == 0:
                if nxt < nq:
                    wait = arrivals[nxt] - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 1e-3))
                continue
            lb = ids[act_q, act_r + 1]
            off, res = self._step(act_ab, act_al, c.beg[lb], lens[lb], half)
            steps += 1
            pairs += act_q.size
            t_end = time.perf_counter() - t0
            cnt = np.diff(off)
            act_r += 1
            fin = (act_r >= ks[act_q] - 1) | (cnt == 0)
            if fin.any():
                fq = act_q[fin]
                done_t[fq] = t_end
                ndone += fq.size
                reserved -= int(need[fq].sum())
                if keep_results:
                    for j in np.flatnonzero(fin):
                        results[act_q[j]] = res[int(off[j]):int(off[j + 1])].copy()
            keep = ~fin
            act_q, act_r = act_q[keep], act_r[keep]
            act_ab = (np.uint64(c.arena[half]) + off[:-1])[keep]
            act_al = cnt[keep]
            half ^= 1
        lat = done_t - arrivals
        stats = {"steps": steps, "pairs": pairs, "served": int(np.isfinite(done_t).sum()),
                 "wall_s": time.perf_counter() - t0}
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
