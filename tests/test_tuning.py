"""The autotuner's table and the cost model that picks an entry from it (CPU only)."""
import math

import numpy as np
import pytest

from simd_set_operations_amd import autotune, datagen, tuning


def _table(rates):
    """A two-cell table: rates[(log2_larger, log2_skew)] = {algo: Gelem/s}."""
    cells = []
    for (l, s), g in rates.items():
        for sel in (0.1, 0.9):
            cells.append({"log2_larger": l, "log2_skew": s, "selectivity": sel, "gelems": dict(g)})
    algos = sorted({a for g in rates.values() for a in g})
    return {"algorithms": algos, "cells": cells}


T = _table({(10, 0): {"m": 100.0, "g": 10.0}, (10, 10): {"m": 10.0, "g": 100.0}})


def test_choose_follows_the_cheaper_entry_per_cell():
    assert tuning.choose([1024] * 8, [1024] * 8, table=T) == "m"
    assert tuning.choose([1] * 8, [1024] * 8, table=T) == "g"


def test_predict_sums_elements_over_rates():
    # one balanced pair (2048 elements) and one skewed pair (1025 elements)
    p = tuning.predict([1024, 1], [1024, 1024], table=T)
    assert math.isclose(p["m"], (2048 / 100e9 + 1025 / 10e9) * 1e3)
    assert math.isclose(p["g"], (2048 / 10e9 + 1025 / 100e9) * 1e3)


def test_mixed_batch_weighs_elements():
    # many skewed pairs but most elements in one balanced pair: the balanced cell decides
    a = [1] * 10 + [1024] * 4
    b = [1024] * 14
    pm = tuning.predict(a, b, table=T)
    assert pm["m"] != pm["g"]
    assert tuning.choose(a, b, table=T) == min(pm, key=pm.get)


def test_resident_choice_follows_the_smaller_set_size():
    t = {"resident": [
        {"log2_small": 4, "gprobes": {"a": 10.0, "b": 20.0}},
        {"log2_small": 10, "gprobes": {"a": 30.0, "b": 5.0}},
    ]}
    assert tuning.choose_resident([16] * 100, [1 << 24] * 100, table=t) == "b"
    assert tuning.choose_resident([1024] * 100, [1 << 24] * 100, table=t) == "a"
    # a mix is priced in probes per cell: 16,000 probes at 2^4 and 102,400 at 2^10
    mixed_a, mixed_b = [16] * 1000 + [1024] * 100, [1 << 24] * 1100
    ca = 16000 / 10e9 + 102400 / 30e9
    cb = 16000 / 20e9 + 102400 / 5e9
    assert tuning.choose_resident(mixed_a, mixed_b, table=t) == ("a" if ca < cb else "b")
    assert tuning.choose_resident([16], [1 << 24], table={"resident": []}) is None


def test_committed_table_has_resident_cells():
    t = tuning.load()
    cells = t["resident"]
    assert [c["log2_small"] for c in cells] == list(autotune.RESIDENT_LOG2_SMALL)
    for c in cells:
        assert set(c["gprobes"]) == set(autotune.RESIDENT_ALGOS)
        assert all(v > 0 for v in c["gprobes"].values())
        assert c["best"] == max(c["gprobes"], key=c["gprobes"].get)
        assert c["pairs"] * c["small"] == 1 << 20
    # the skewed workload's shape (4,096 x 256 against one indexed 2^24 set) resolves to a search entry
    assert tuning.choose_resident([256] * 4096, [1 << 24] * 4096).startswith(("hip_binary_search", "hip_galloping",
                                                                           "hip_simd_galloping"))


def test_committed_table_covers_the_grid():
    t = tuning.load()
    grid = t["grid"]
    assert grid == autotune.GRID
    keys = {(c["log2_larger"], c["log2_skew"], c["selectivity"]) for c in t["cells"]}
    want = {(l, s, v) for l in grid["log2_larger"] for s in grid["log2_skew"] for v in grid["selectivity"]}
    assert keys == want
    for c in t["cells"]:
        assert set(c["gelems"]) == set(t["algorithms"])
        assert all(v > 0 for v in c["gelems"].values())
        assert c["best"] == max(c["gelems"], key=c["gelems"].get)
        larger, smaller, pairs = autotune.cell_shape(c["log2_larger"], c["log2_skew"])
        assert (c["larger"], c["smaller"], c["pairs"]) == (larger, smaller, pairs)
        assert c["elements"] == pairs * (larger + smaller)


def test_committed_table_choices():
    # the uneven batch is merge work; a 2^24 resident set against sets of 256 is search work
    A, a_off, B, b_off = datagen.uneven_batch(npairs=4096)
    assert tuning.choose(np.diff(a_off), np.diff(b_off)).partition("@")[0] == "hip_branchless_merge"
    pick = tuning.choose(np.full(4096, 256), np.full(4096, 1 << 24)).partition("@")[0]
    assert pick in ("hip_galloping", "hip_binary_search")


def test_choose_returns_a_tile_shape():
    t = _table({(10, 0): {"m": 100.0, "m@1": 120.0, "g": 10.0}, (10, 10): {"m": 10.0, "g": 100.0, "g@2": 90.0}})
    assert tuning.choose([1024] * 8, [1024] * 8, table=t) == "m@1"
    assert tuning.choose([1] * 8, [1024] * 8, table=t) == "g"


def test_graph_split_threshold_from_the_table():
    assert tuning.graph_split_threshold(table={"cells": []}) == 32.0
    t = {"cells": [], "graph_skew_split": {"threshold": 64.0, "ms": {"32.0": 2.0, "64.0": 1.5}}}
    assert tuning.graph_split_threshold(table=t) == 64.0


def test_graph_merge_entry_from_the_table():
    assert tuning.graph_merge_entry(table={"cells": []}) == "hip_branchless_merge"
    t = {"cells": [], "graph_skew_split": {"threshold": 32.0, "merge_entry": "hip_branchless_merge@2",
                                           "merge_ms": {"hip_branchless_merge": 2.0, "hip_branchless_merge@2": 1.8}}}
    assert tuning.graph_merge_entry(table=t) == "hip_branchless_merge@2"


def test_committed_table_graph_section():
    g = tuning.load()["graph_skew_split"]
    assert set(g["merge_ms"]) == set(autotune.GRAPH_MERGE_SHAPES)
    assert g["merge_entry"] == min(g["merge_ms"], key=g["merge_ms"].get)
    assert str(g["threshold"]) in g["ms"] and g["threshold"] == float(min(g["ms"], key=g["ms"].get))


@pytest.mark.parametrize("larger,smaller,sel", [(64, 64, 0.5), (1024, 8, 0.9), (256, 1, 1.0), (100, 30, 0.0)])
def test_fixed_batch_shapes_and_planted_counts(larger, smaller, sel):
    A, a_off, B, b_off, shared = datagen.fixed_batch(50, larger, smaller, sel)
    assert shared == round(sel * smaller)
    assert A.size == 50 * smaller and B.size == 50 * larger
    for p in range(50):
        a = A[a_off[p]:a_off[p + 1]].astype(np.int64)
        b = B[b_off[p]:b_off[p + 1]].astype(np.int64)
        assert np.all(np.diff(a) > 0) and np.all(np.diff(b) > 0)
        assert b.max() < max(int(np.ceil(larger / 0.1)), larger + smaller)
        assert np.intersect1d(a, b).size == shared
