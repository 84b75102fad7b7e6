"""The serving engine's admission logic (host only): queries are admitted in arrival order while
their in-flight cost fits the device-memory budget and the step's pair slots; a small budget
refuses admission until memory is released."""
from collections import deque

import numpy as np

from simd_set_operations_amd.serving import Admission, query_cost_words


def test_admits_in_order_within_the_budget():
    adm = Admission(budget_words=1000, max_pairs=8)
    cost = np.array([300, 300, 300, 500, 100])
    pending = deque(range(5))
    assert adm.admit(pending, cost) == [0, 1, 2]
    # query 3 does not fit; query 4 would, but does not overtake it
    assert list(pending) == [3, 4]
    assert adm.admit(pending, cost) == []
    adm.release(300, 1)
    assert adm.admit(pending, cost) == []  # 600 + 500 > 1000
    adm.release(300, 1)
    assert adm.admit(pending, cost) == [3, 4]
    assert adm.reserved == 300 + 500 + 100 and adm.active == 3


def test_small_budget_refuses_admission():
    adm = Admission(budget_words=64, max_pairs=100)
    cost = np.array([65, 10, 10])
    pending = deque(range(3))
    assert adm.admit(pending, cost) == []
    assert adm.reserved == 0 and adm.active == 0 and len(pending) == 3


def test_pair_slots_limit_a_step():
    adm = Admission(budget_words=10 ** 9, max_pairs=4)
    cost = np.ones(10, np.int64)
    pending = deque(range(10))
    assert adm.admit(pending, cost) == [0, 1, 2, 3]
    assert adm.admit(pending, cost) == []  # every slot is in flight
    adm.release(2, 2)
    assert adm.admit(pending, cost, limit=1) == [4]
    assert adm.admit(pending, cost) == [5]


def test_query_cost_follows_the_smallest_list():
    class FakeCorpus:
        lens = np.array([16, 1000, 50, 7], np.uint64)

    ids = np.array([[3, 2, 1, -1], [2, 1, -1, -1]])
    need, cost = query_cost_words(FakeCorpus, ids)
    assert need.tolist() == [7, 50]
    assert cost.tolist() == [3 * 7 + 256, 3 * 50 + 256]


def test_step_rows_follow_the_batch_layout():
    """A step of queries q0 .. q1-1 packs, from the stream's rows, exactly what
    setops_hip_batch_rebind / _set_further read: A begin/end and B begin/end of every pair (the
    query's first two lists), the further-set offsets per pair and the further sets' bounds."""
    from types import SimpleNamespace

    from simd_set_operations_amd.serving import Engine

    rng = np.random.Generator(np.random.PCG64(5))
    lens = rng.integers(1, 1000, size=50).astype(np.uint64)
    beg = np.zeros(50, np.uint64)
    beg[1:] = np.cumsum(lens[:-1])
    c = SimpleNamespace(beg=beg, lens=lens)
    nq, kmax = 40, 6
    ks = rng.integers(2, kmax + 1, size=nq)
    ids = np.full((nq, kmax), -1, np.int64)
    for q in range(nq):
        ids[q, :ks[q]] = rng.choice(50, size=ks[q], replace=False)
    need = lens[ids[:, 0]].astype(np.int64)
    rows = Engine._stream_rows(c, ids, ks, need)
    for q0, q1 in ((0, 1), (3, 17), (17, 40)):
        P = q1 - q0
        qs = np.arange(q0, q1)
        np.testing.assert_array_equal(rows["pairs"][:, q0:q1],
                                      [beg[ids[qs, 0]], beg[ids[qs, 0]] + lens[ids[qs, 0]],
                                       beg[ids[qs, 1]], beg[ids[qs, 1]] + lens[ids[qs, 1]]])
        x0, x1 = rows["xoff"][q0], rows["xoff"][q1]
        xs = [int(i) for q in qs for i in ids[q, 2:ks[q]]]
        assert x1 - x0 == len(xs)
        np.testing.assert_array_equal(rows["xb"][x0:x1], beg[xs])
        np.testing.assert_array_equal(rows["xe"][x0:x1], beg[xs] + lens[xs])
        np.testing.assert_array_equal(rows["xoff"][q0:q1 + 1] - x0, np.concatenate([[0], np.cumsum(ks[qs] - 2)]))
        assert rows["need_pfx"][q1] - rows["need_pfx"][q0] == need[qs].sum()
        np.testing.assert_array_equal(rows["na"][q0:q1], lens[ids[qs, 0]])
        np.testing.assert_array_equal(rows["nb"][q0:q1], lens[ids[qs, 1]])
        assert P == qs.size
