"""Edge-wise neighbour intersection over a CSR graph (a caller of the batched 2-set path).

For every edge (u, v) with u < v the batch intersects N(u) with N(v) in place in the CSR (a
gather batch: no list is copied) and sums the counts.  Over the full lists the sum is 3T (every
triangle is seen from each of its three edges); over the oriented lists N+(x) = {y in N(x) : y > x}
it is T.  The BSR form runs the same pairs on the BsrVec encoding of each list (base = v >> 5,
state bit v & 31, setops/src/bsr.rs:140-164) and sums popcounts (visitor.rs:399-403).

Skewed edges (a hub against a leaf) can go to a search-policy entry (galloping) and the rest to a
merge-policy entry: ``split_by_skew`` gives the two edge sets for a skew threshold.
"""
import numpy as np

from .batch import GatherBatch


def csr_to_bsr(rowptr, col):
    """BSR encoding of every neighbour list: (bsr_rowptr, bases, states)."""
    n = rowptr.size - 1
    deg = np.diff(rowptr).astype(np.int64)
    if col.size == 0:
        return np.zeros(n + 1, np.uint64), np.zeros(0, np.uint32), np.zeros(0, np.uint32)
    src = np.repeat(np.arange(n, dtype=np.int64), deg)
    base = (col >> np.uint32(5)).astype(np.uint32)
    bit = np.left_shift(np.uint32(1), (col & np.uint32(31)).astype(np.uint32))
    new = np.ones(col.size, bool)
    new[1:] = (base[1:] != base[:-1]) | (src[1:] != src[:-1])
    starts = np.flatnonzero(new)
    states = np.bitwise_or.reduceat(bit, starts).astype(np.uint32)
    bases = base[starts]
    words = np.bincount(src[starts], minlength=n).astype(np.uint64)
    brp = np.zeros(n + 1, np.uint64)
    np.cumsum(words, out=brp[1:])
    return brp, bases, states


def split_by_skew(deg_u, deg_v, threshold):
    """Boolean mask of the edges whose larger list is more than `threshold` times the smaller
    (they go to the search policy)."""
    lo = np.minimum(deg_u, deg_v).astype(np.float64)
    hi = np.maximum(deg_u, deg_v).astype(np.float64)
    return hi > threshold * np.maximum(lo, 1.0)


class EdgeIntersections:
    """Count-only gather batches over the edges (u, v), u < v, of a symmetric CSR already on the
    device (col / bases / states as device tensors)."""

    def __init__(self, algo, pool, beg, end, u, v, states=None, select=None, index=False):
        uu, vv = (u, v) if select is None else (u[select], v[select])
        self.pairs = int(uu.size)
        self.batch = None
        if self.pairs:
            self.batch = GatherBatch(algo, pool, beg[uu], end[uu], pool, beg[vv], end[vv], A_states=states,
                                     B_states=states, outputs=False)
            if index:
                self.batch.index_b()  # search-policy entries: the CSR pool indexed once

    def run(self, stream=None):
        if self.batch is not None:
            self.batch.run(stream, count_only=True)

    def totals(self):
        """(words, elements) summed over the edges of this batch."""
        return self.batch.totals() if self.batch is not None else (0, 0)

    def close(self):
        if self.batch is not None:
            self.batch.close()
            self.batch = None
