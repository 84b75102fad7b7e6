"""The experiment-driven harness (schema, x values, generators, dataset layout, results shape),
run on the CPU with oracle timers; the GPU run is in test_gpu_parity.py."""
import json
import os

import numpy as np
import pytest

import oracle
from simd_set_operations_amd import datafile
from simd_set_operations_amd import experiment as ex

DOC = """
[algorithm_sets]
merges = ["branchless_merge", "shuffling_avx512"]

[[experiment]]
name = "sel"
title = "2-set, selectivity"
dataset = "two_sel"
relative_to = "branchless_merge"
algorithms = ["branchless_merge", "galloping", "no_such_algorithm"]

[[experiment]]
name = "kset"
title = "4-set, skew"
dataset = "four_skew"
algorithm_set = "merges"

[[dataset]]
name = "two_sel"
type = "synthetic"
set_count = 2
gen_count = 3
vary = "selectivity"
selectivity = 0
to = 1000
step = 500
skewness_factor = 1000
density = 10
max_len = 10

[[dataset]]
name = "four_skew"
type = "synthetic"
set_count = 4
gen_count = 2
vary = "skew"
skewness_factor = 0
to = 2000
step = 1000
density = 50
selectivity = 300
max_len = 9
"""


@pytest.fixture()
def doc(tmp_path):
    p = tmp_path / "experiment.toml"
    p.write_text(DOC)
    return ex.load(p)


def test_schema_and_x_values(doc):
    two, four = doc["dataset"]
    assert ex.xvalues(two) == [0, 500, 1000]
    assert ex.xvalues(four) == [0, 1000, 2000]
    assert ex.props_at_x(two, 500)["selectivity"] == 500
    assert ex.props_at_x(four, 2000)["skewness_factor"] == 2000
    assert ex.get_algorithms(doc, doc["experiment"][1]) == ["branchless_merge", "shuffling_avx512"]
    assert doc["experiment"][1]["relative_to"] is None


@pytest.mark.parametrize("bad", [
    DOC.replace('type = "synthetic"\nset_count = 4', 'type = "nonsense"\nset_count = 4'),
    DOC.replace('vary = "skew"', 'vary = "colour"'),
    DOC.replace('algorithm_set = "merges"', 'algorithm_set = "merges"\nalgorithms = ["x"]'),
    DOC.replace('dataset = "four_skew"', 'dataset = "missing"'),
    DOC.replace("step = 1000", "step = 0"),
])
def test_schema_rejects(tmp_path, bad):
    p = tmp_path / "bad.toml"
    p.write_text(bad)
    with pytest.raises(ex.ExperimentError):
        ex.load(p)


def test_get_skew_is_k_to_the_f():
    assert [ex.get_skew(k, 0) for k in range(4)] == [1, 1, 1, 1]
    assert [ex.get_skew(k, 1000) for k in range(4)] == [1, 2, 3, 4]
    assert [ex.get_skew(k, 2000) for k in range(4)] == [1, 4, 9, 16]
    assert ex.get_skew(1, 500) == 1  # floor(2^0.5)


@pytest.mark.parametrize("sel,skew,density", [(0, 0, 10), (300, 1000, 10), (1000, 3000, 1), (700, 0, 1000)])
def test_gen_twoset_sizes_domain_and_shared(sel, skew, density):
    props = {"set_count": 2, "density": density, "selectivity": sel, "max_len": 11, "skewness_factor": skew}
    rng = np.random.Generator(np.random.PCG64(5))
    small, large = ex.gen_twoset(props, rng)
    large_len = 1 << 11
    small_len = large_len // ex.get_skew(1, skew)
    max_value = int(min(large_len / (density / 1000), 2 ** 31 - 1))
    want = int(sel / 1000 * small_len)
    if small_len + large_len - want > max_value:  # space too small: shared raised to fit
        want = small_len + large_len - max_value
    assert small.size == small_len and large.size == large_len
    for s in (small, large):
        assert np.all(np.diff(s.astype(np.int64)) > 0) and int(s.max()) < max_value
    assert oracle.branchless_merge(small, large).size == want


def test_gen_kset_sizes_and_shared_core():
    props = {"set_count": 5, "density": 20, "selectivity": 400, "max_len": 10, "skewness_factor": 1500}
    sets = ex.gen_kset(props, np.random.Generator(np.random.PCG64(9)))
    sizes = [1024 // ex.get_skew(i, 1500) for i in reversed(range(5))]
    assert [s.size for s in sets] == sizes and sizes == sorted(sizes)
    core = int(0.4 * sizes[0])
    common = oracle.svs("branchless_merge", sets)
    assert common.size >= core  # every set holds the shared core (other values may coincide)
    for s in sets:
        assert np.all(np.diff(s.astype(np.int64)) > 0)


def test_generate_layout_and_skip(doc, tmp_path):
    d = tmp_path / "datasets"
    logs = []
    ex.generate(doc, str(d), log=logs.append)
    assert logs == ["Building two_sel", "Building four_skew"]
    assert sorted(os.listdir(d / "two_sel")) == ["0", "1000", "500"]
    assert sorted(os.listdir(d / "two_sel" / "500")) == ["0", "1", "2"]
    with open(d / "four_skew.json") as f:
        assert json.load(f) == doc["dataset"][1]
    sets = datafile.read(d / "four_skew" / "2000" / "1")
    assert [s.size for s in sets] == [512 // ex.get_skew(i, 2000) for i in reversed(range(4))]
    a, b = datafile.read(d / "two_sel" / "1000" / "0")
    assert oracle.branchless_merge(a, b).size == a.size  # selectivity 1: small inside large
    logs.clear()
    ex.generate(doc, str(d), log=logs.append)
    assert logs == ["Skipping two_sel", "Skipping four_skew"]


class _OracleTimer:
    def __init__(self, f):
        self.f = f
        self.calls = 0

    def prepare(self, sets):
        return sets

    def run(self, sets):
        self.calls += 1
        return oracle.svs("branchless_merge", sets).size if len(sets) > 2 else self.f(sets[0], sets[1]).size


def test_run_results_shape(doc, tmp_path):
    d = tmp_path / "datasets"
    ex.generate(doc, str(d), log=lambda m: None)
    timers = {"branchless_merge": _OracleTimer(oracle.branchless_merge),
              "galloping": _OracleTimer(lambda a, b: oracle.intersect2("galloping", a, b)),
              "shuffling_avx512": _OracleTimer(oracle.branchless_merge)}
    logs = []
    res = ex.run(doc, str(d), resolver=timers.get, warmup_s=0.0, log=logs.append)
    assert set(res) == {"experiments", "datasets", "algorithm_sets"}
    assert [e["name"] for e in res["experiments"]] == ["sel", "kset"]
    two = res["datasets"]["two_sel"]
    assert two["info"] == doc["dataset"][0]
    assert set(two["algos"]) == {"branchless_merge", "galloping", "no_such_algorithm"}
    assert two["algos"]["no_such_algorithm"] == []
    assert "  unknown algorithm no_such_algorithm" in logs
    runs = two["algos"]["galloping"]
    assert [r["x"] for r in runs] == [0, 500, 1000]
    for r in runs:
        assert len(r["times"]) == 3 and all(t > 0 for t in r["times"])
        assert r["l1d"] == {"rd_access": None, "rd_miss": None, "wr_access": None, "wr_miss": None}
        assert r["branches"] is None
    assert set(res["datasets"]["four_skew"]["algos"]) == {"branchless_merge", "shuffling_avx512"}
    json.dumps(res)  # serialisable as the reference's results file
    # one experiment selected: only its dataset runs
    res = ex.run(doc, str(d), experiments=["kset"], resolver=timers.get, warmup_s=0.0, log=lambda m: None)
    assert list(res["datasets"]) == ["four_skew"] and [e["name"] for e in res["experiments"]] == ["kset"]


def test_gpu_entry_names():
    reg = {"hip_branchless_merge", "hip_shuffling", "hip_shuffling_bsr", "hip_galloping"}
    assert ex.gpu_entry_name("branchless_merge", reg) == "hip_branchless_merge"
    assert ex.gpu_entry_name("shuffling_avx512", reg) == "hip_shuffling"
    assert ex.gpu_entry_name("shuffling_sse_bsr", reg) == "hip_shuffling_bsr"
    assert ex.gpu_entry_name("hip_galloping", reg) == "hip_galloping"
    assert ex.gpu_entry_name("shuffling_avx2_bsr_branch", reg) == "hip_shuffling_bsr"
    assert ex.gpu_entry_name("shuffling_sse_branch", reg) == "hip_shuffling"
    reg2 = reg | {"hip_simd_galloping", "hip_simd_galloping_bsr", "hip_broadcast_bsr"}
    assert ex.gpu_entry_name("galloping_avx512", reg2) == "hip_simd_galloping"
    assert ex.gpu_entry_name("galloping_sse_bsr", reg2) == "hip_simd_galloping_bsr"
    assert ex.gpu_entry_name("galloping", reg2) == "hip_galloping"
    assert ex.gpu_entry_name("broadcast_avx2_bsr_branch", reg2) == "hip_broadcast_bsr"
    assert ex.gpu_entry_name("croaring", reg) is None
