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
