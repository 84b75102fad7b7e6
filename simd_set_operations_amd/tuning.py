"""Algorithm choice from the autotuner's table (``tuning.json``, written by :mod:`.autotune`).

The table holds, for every cell of a (larger-set size, skew, selectivity) grid, the batched
throughput of each registered 2-set entry measured on one MI355X.  A batch is priced per entry
by a cost model over that table: every pair is binned to its nearest (size, skew) cell, and the
predicted time of an entry is the sum over cells of the cell's elements divided by the entry's
throughput there (averaged over the selectivities, which a caller does not know in advance).
``choose`` returns the cheapest entry, in the tile shape the table prices cheapest
(``name@shape``); ``hip_auto`` batches resolve to it at creation.  The table also holds the
graph skew-split threshold (``graph_split_threshold``) and the search entries' rates against an
indexed resident pool (``choose_resident``, for ``hip_auto`` gather batches created with
``index=True``).
"""
import json
import os

import numpy as np

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")

_table = None


def load(path=TABLE_PATH):
    """The tuning table (cached); raises FileNotFoundError when it was never written."""
    global _table
    if path != TABLE_PATH:
        with open(path) as f:
            return json.load(f)
    if _table is None:
        with open(path) as f:
            _table = json.load(f)
    return _table


def _cell_rates(table, algos):
    """{(log2_larger, log2_skew): {algo: Gelem/s averaged over the selectivities}}."""
    acc = {}
    for c in table["cells"]:
        key = (c["log2_larger"], c["log2_skew"])
        d = acc.setdefault(key, {a: [] for a in algos})
        for a in algos:
            if a in c["gelems"]:
                d[a].append(c["gelems"][a])
    return {k: {a: float(np.mean(v)) for a, v in d.items() if v} for k, d in acc.items()}


def predict(a_sizes, b_sizes, table=None, algos=None):
    """Predicted batch time (ms) of each entry for pairs of sizes (a_sizes[i], b_sizes[i])."""
    table = table or load()
    algos = algos or table["algorithms"]
    rates = _cell_rates(table, algos)
    ls = np.array(sorted({k[0] for k in rates}), float)
    ss = np.array(sorted({k[1] for k in rates}), float)
    a = np.asarray(a_sizes, np.float64)
    b = np.asarray(b_sizes, np.float64)
    big = np.maximum(np.maximum(a, b), 1.0)
    small = np.maximum(np.minimum(a, b), 1.0)
    lb = np.log2(big)
    sb = np.log2(big / small)
    il = np.abs(lb[:, None] - ls[None, :]).argmin(axis=1)
    isk = np.abs(sb[:, None] - ss[None, :]).argmin(axis=1)
    elems = a + b
    out = {}
    for alg in algos:
        ms = 0.0
        for i, l in enumerate(ls):
            for j, sk in enumerate(ss):
                m = (il == i) & (isk == j)
                if not m.any():
                    continue
                r = rates.get((int(l), int(sk)), {}).get(alg)
                if not r:
                    ms = float("inf")
                    break
                ms += float(elems[m].sum()) / (r * 1e9) * 1e3
            if ms == float("inf"):
                break
        out[alg] = ms
    return out


def choose(a_sizes, b_sizes, table=None, algos=None):
    """The entry with the smallest predicted batch time."""
    pred = predict(a_sizes, b_sizes, table, algos)
    return min(pred, key=pred.get)


def choose_resident(a_sizes, b_sizes, table=None):
    """For pairs that search one indexed resident pool (GatherBatch(..., index=True)): the search
    entry and tile shape with the smallest predicted time, from the table's resident-pool cells
    (probes per second with a 2^24-element indexed set, per smaller-set size); each pair is priced
    at the cell nearest its smaller size.  None without such cells."""
    table = table or load()
    cells = table.get("resident")
    if not cells:
        return None
    sm = np.array([c["log2_small"] for c in cells], float)
    probes = np.minimum(np.asarray(a_sizes, np.float64), np.asarray(b_sizes, np.float64))
    idx = np.abs(np.log2(np.maximum(probes, 1.0))[:, None] - sm[None, :]).argmin(axis=1)
    algos = sorted(set.intersection(*(set(c["gprobes"]) for c in cells)))
    cost = {a: sum(float(probes[idx == i].sum()) / (cells[i]["gprobes"][a] * 1e9) for i in range(len(cells)))
            for a in algos}
    return min(cost, key=cost.get)


def graph_merge_entry(table=None, default="hip_branchless_merge"):
    """The merge entry (with its tile shape) the autotuner measured fastest for graph edge
    intersections (gather batches over one CSR), or `default` without a measurement."""
    try:
        table = table or load()
    except FileNotFoundError:
        return default
    g = table.get("graph_skew_split")
    return str(g["merge_entry"]) if g and "merge_entry" in g else default


def graph_split_threshold(table=None, default=32.0):
    """The skew threshold the autotuner measured for graph edge intersections (larger list over
    smaller beyond which an edge goes to the search entry), or `default` without a measurement."""
    try:
        table = table or load()
    except FileNotFoundError:
        return default
    g = table.get("graph_skew_split")
    return float(g["threshold"]) if g and "threshold" in g else default
