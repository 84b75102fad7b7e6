"""The reference's experiment-driven harness, driving the GPU entries.

Mirrors the reference's benchmark crate:

* the experiment file (benchmark/src/schema.rs:14-86): ``[[experiment]]`` entries (name, title,
  dataset, ``algorithms`` or ``algorithm_set``, optional ``relative_to``), ``[[dataset]]`` entries
  (``type = "synthetic"`` with vary / to / step / gen_count and the intersection properties
  set_count, density, selectivity, max_len, skewness_factor; percentages in thousandths) and
  ``[algorithm_sets]``;
* x values and the properties at x (lib.rs:27-58);
* synthetic generation (generators.rs:30-259): 2-set with planted shared elements, k-set with a
  shared core, sizes max_len / k^f;
* dataset directories datasets/<name>/<x>/<i> in the reference's datafile format, with a
  <name>.json info file that makes regeneration a no-op when unchanged (bin/generate.rs:68-221);
* the benchmark run (bin/benchmark.rs:60-243): each dataset is run once for the union of the
  algorithms its experiments name; per x value and algorithm every datafile is timed once after a
  warmup of 1 s split across the files (timer/harness.rs:35-62); results are written in the
  reference's Results JSON shape (x, times in ns; the CPU perf-counter fields are null).

Algorithm names resolve to GPU entries: a registered name is used as is, otherwise ``hip_`` + the
name, with the reference's SIMD-width and ``_branch`` suffixes dropped (``shuffling_avx512_branch``
-> ``hip_shuffling``: vector width and branching are CPU notions).  ``galloping_sse/avx2/avx512``
are the reference's SIMD galloping (simd_galloping.rs) and resolve to ``hip_simd_galloping``.  A
name with no GPU entry is reported and skipped, as the reference skips an unknown algorithm.  Generation uses a seeded PCG64 stream per datafile instead
of the reference's thread-local RNG, so datasets are reproducible.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

PERCENT = 1000
PARAMETERS = {"density": "density", "selectivity": "selectivity", "size": "max_len", "skew": "skewness_factor",
              "set_count": "set_count"}
INTERSECTION_FIELDS = ("set_count", "density", "selectivity", "max_len", "skewness_factor")
SIMD_SUFFIXES = ("_sse", "_avx2", "_avx512")


class ExperimentError(ValueError):
    pass


# ---- experiment file ------------------------------------------------------------------------------

def _u32(d, k, where):
    v = d.get(k)
    if not isinstance(v, int) or v < 0 or v > 0xFFFFFFFF:
        raise ExperimentError(f"{where}: '{k}' must be a u32")
    return v


def parse(doc):
    """Validates a parsed experiment document (schema.rs:14-86); returns it with defaults filled."""
    for key in ("experiment", "dataset", "algorithm_sets"):
        if key not in doc:
            raise ExperimentError(f"missing '{key}'")
    sets = doc["algorithm_sets"]
    if not isinstance(sets, dict) or not all(isinstance(v, list) for v in sets.values()):
        raise ExperimentError("algorithm_sets must map names to lists of algorithms")
    for d in doc["dataset"]:
        where = f"dataset {d.get('name', '?')}"
        if not isinstance(d.get("name"), str):
            raise ExperimentError(f"{where}: 'name' must be a string")
        if d.get("type") == "synthetic":
            if d.get("vary") not in PARAMETERS:
                raise ExperimentError(f"{where}: unknown vary '{d.get('vary')}'")
            for k in ("to", "step") + INTERSECTION_FIELDS:
                _u32(d, k, where)
            if not isinstance(d.get("gen_count"), int) or d["gen_count"] < 0:
                raise ExperimentError(f"{where}: 'gen_count' must be a count")
            if d["step"] == 0:
                raise ExperimentError(f"{where}: 'step' must be positive")
        elif d.get("type") == "real":
            for k in ("set_count_start", "set_count_end"):
                _u32(d, k, where)
        else:
            raise ExperimentError(f"{where}: unknown type '{d.get('type')}'")
    names = {d["name"] for d in doc["dataset"]}
    for e in doc["experiment"]:
        where = f"experiment {e.get('name', '?')}"
        for k in ("name", "title", "dataset"):
            if not isinstance(e.get(k), str):
                raise ExperimentError(f"{where}: '{k}' must be a string")
        if ("algorithms" in e) == ("algorithm_set" in e):
            raise ExperimentError(f"{where}: exactly one of 'algorithms' and 'algorithm_set'")
        if e["dataset"] not in names:
            raise ExperimentError(f"{where}: unknown dataset '{e['dataset']}'")
        e.setdefault("relative_to", None)
    return doc


def load(path):
    """Reads and validates an experiment TOML file."""
    try:
        import tomllib
    except ImportError:  # Python < 3.11
        import tomli as tomllib
    with open(path, "rb") as f:
        return parse(tomllib.load(f))


def get_algorithms(doc, entry):
    """lib.rs:60-69: the entry's list, or the named algorithm set."""
    if "algorithms" in entry:
        return list(entry["algorithms"])
    name = entry["algorithm_set"]
    if name not in doc["algorithm_sets"]:
        raise ExperimentError(f"algorithm set {name} not found")
    return list(doc["algorithm_sets"][name])


def xvalues(info):
    """lib.rs:27-44: from the varied property's value to ``to`` by ``step`` (inclusive)."""
    if info["type"] == "real":
        return list(range(info["set_count_start"], info["set_count_end"] + 1))
    begin = info[PARAMETERS[info["vary"]]]
    return list(range(begin, info["to"] + 1, info["step"]))


def props_at_x(info, x):
    """lib.rs:46-58: the intersection properties with the varied one set to x."""
    props = {k: info[k] for k in INTERSECTION_FIELDS}
    props[PARAMETERS[info["vary"]]] = x
    return props


# ---- synthetic generation (generators.rs) ------------------------------------------------------------

def get_skew(set_index, skew_factor):
    """generators.rs:224-229: size ratio (k+1)^(f/1000) of the set at 0-based index k."""
    return int(float(set_index + 1) ** (skew_factor / PERCENT))


def _max_value(max_len, density):
    # (len as f64 / density) as i32: saturating, so density 0 gives i32::MAX
    v = max_len / density if density > 0 else float("inf")
    return int(min(v, 2 ** 31 - 1))


def shuffled_set(n, max_value, rng):
    """generators.rs:84-111: n distinct values of [0, max_value) in random order."""
    if n * 2 < max_value:
        items = np.zeros(0, np.int64)
        while items.size < n:
            need = n - items.size
            items = np.unique(np.concatenate([items, rng.integers(0, max_value, size=need * 2, dtype=np.int64)]))
        rng.shuffle(items)
        return items[:n]
    return rng.permutation(max_value)[:n].astype(np.int64)


def sorted_set_containing(include, n, max_value, rng):
    """generators.rs:138-215: a sorted set of n values of [0, max_value) holding every value of
    ``include``.  (The reference's dense branch keeps n values besides ``include``; this keeps
    n in total, as its sparse branch and its callers' sizes do.)"""
    need = n - include.size
    if need < 0:
        raise ExperimentError("set shorter than its shared core")
    if n * 2 < max_value:
        inc = np.asarray(include, np.int64)
        others = np.zeros(0, np.int64)
        while others.size < need:
            draw = rng.integers(0, max_value, size=(n - others.size) * 2, dtype=np.int64)
            others = np.unique(np.concatenate([others, draw[~np.isin(draw, inc)]]))
        rng.shuffle(others)
        others = others[:need]
    else:
        mask = np.ones(max_value, bool)
        mask[np.asarray(include, np.int64)] = False
        others = np.flatnonzero(mask)
        rng.shuffle(others)
        others = others[:need]
    return np.sort(np.concatenate([others, np.asarray(include, np.int64)]))


def gen_twoset(props, rng):
    """generators.rs:30-78: (small, large) sorted sets; large = 2^max_len elements, small =
    large / 2^(f/1000); floor(selectivity x small) shared; element space large / density (shared
    count raised when the space is too small for the rest)."""
    large_len = 1 << props["max_len"]
    small_len = large_len // get_skew(1, props["skewness_factor"])
    max_value = _max_value(large_len, props["density"] / PERCENT)
    shared = int(props["selectivity"] / PERCENT * small_len)
    gen_count = shared + (small_len + large_len - 2 * shared)
    if gen_count > max_value:
        shared = small_len + large_len - max_value
        gen_count = max_value
        if shared > small_len:
            raise ExperimentError(f"no {large_len}-element set fits an element space of {max_value}")
    values = shuffled_set(gen_count, max_value, rng)
    both, rest = values[:shared], values[shared:]
    small = np.sort(np.concatenate([both, rest[:small_len - shared]]))
    large = np.sort(np.concatenate([both, rest[small_len - shared:]]))
    return small.astype(np.uint32), large.astype(np.uint32)


def gen_kset(props, rng):
    """generators.rs:113-135: set_count sorted sets, smallest first; set k (0-based from the
    largest) has 2^max_len / (k+1)^(f/1000) elements; every set holds the same
    floor(selectivity x smallest) shared values."""
    max_len = 1 << props["max_len"]
    k = props["set_count"]
    max_value = _max_value(max_len, props["density"] / PERCENT)
    min_len = max_len // get_skew(k - 1, props["skewness_factor"])
    shared = shuffled_set(int(props["selectivity"] / PERCENT * min_len), max_value, rng)
    return [sorted_set_containing(shared, max_len // get_skew(i, props["skewness_factor"]), max_value,
                                  rng).astype(np.uint32) for i in reversed(range(k))]


def generate_intersection(props, rng):
    """bin/generate.rs:209-221: 2-set or k-set by the set count."""
    if props["set_count"] == 2:
        return list(gen_twoset(props, rng))
    return gen_kset(props, rng)


def _rng(name, x, i, seed):
    h = np.frombuffer(name.encode(), np.uint8).astype(np.uint64).sum() if name else 0
    return np.random.Generator(np.random.PCG64([seed, int(h), len(name), x, i]))


def generate_dataset(info, datasets, seed=1234, log=print):
    """bin/generate.rs:68-207: writes datasets/<name>/<x>/<i> and datasets/<name>.json; skips a
    dataset whose info file already matches."""
    from . import datafile

    if info["type"] != "synthetic":
        raise ExperimentError(f"dataset {info['name']}: only synthetic datasets are generated here")
    info_path = os.path.join(datasets, info["name"] + ".json")
    if os.path.exists(info_path):
        with open(info_path) as f:
            if json.load(f) == info:
                log(f"Skipping {info['name']}")
                return False
        log(f"Rebuilding {info['name']}")
    else:
        log(f"Building {info['name']}")
    root = os.path.join(datasets, info["name"])
    if os.path.isdir(root):
        import shutil

        shutil.rmtree(root)
    for x in xvalues(info):
        xdir = os.path.join(root, str(x))
        os.makedirs(xdir, exist_ok=True)
        props = props_at_x(info, x)
        for i in range(info["gen_count"]):
            datafile.write(os.path.join(xdir, str(i)), generate_intersection(props, _rng(info["name"], x, i, seed)))
    os.makedirs(datasets, exist_ok=True)
    with open(info_path, "w") as f:
        json.dump(info, f)
    return True


def generate(doc, datasets, names=None, seed=1234, log=print):
    for d in doc["dataset"]:
        if names and d["name"] not in names:
            continue
        generate_dataset(d, datasets, seed, log)


# ---- benchmark run (bin/benchmark.rs) ----------------------------------------------------------------

def gpu_entry_name(name, registered):
    """The GPU entry for a reference algorithm name, or None."""
    if name in registered:
        return name
    base = name
    if base.endswith("_branch"):  # branch vs branch-free is a CPU notion too
        base = base[:-len("_branch")]
    bsr = base.endswith("_bsr")
    if bsr:
        base = base[:-4]
    for suffix in SIMD_SUFFIXES:
        if base.endswith(suffix):
            base = base[:-len(suffix)]
            if base == "galloping":  # galloping_sse/avx2/avx512 name simd_galloping.rs's entries
                base = "simd_galloping"
    cand = "hip_" + base + ("_bsr" if bsr else "")
    return cand if cand in registered else None


class GpuTimer:
    """Times one GPU entry on a datafile's sets like Timer::run (timer.rs:52-72): 2-set entries on
    two sets, the k-set call (svs) on more; BSR entries get their sets encoded outside the timed
    region (harness.rs:121-122).  ``resident``: plain 2-set inputs are uploaded once and each run
    is the device-pointer call plus a stream synchronisation."""

    def __init__(self, entry, count_only=False, resident=False):
        from . import _native

        self.entry = entry
        self.count_only = count_only
        self.resident = resident
        self.bsr = bool(_native.hip().setops_hip_algo_flags(_algo_id(entry)) & _native.ALGO_BSR)

    def prepare(self, sets):
        from .bsr import BsrVec

        if len(sets) > 2 and self.bsr:
            raise ExperimentError("k-set intersection not supported")
        if self.bsr:
            return [BsrVec.from_sorted(s) for s in sets]
        if self.resident and len(sets) == 2:
            return _Resident(self.entry, sets[0], sets[1], self.count_only)
        return sets

    def run(self, data):
        from . import intersect
        from .visitor import Counter

        if isinstance(data, _Resident):
            return data.run()
        if self.bsr:
            _, n = intersect.intersect2_bsr(self.entry, data[0], data[1], count_only=self.count_only)
            return n
        if len(data) == 2:
            if self.count_only:
                c = Counter()
                intersect.lookup(self.entry)(data[0], data[1], c)
                return c.count()
            return intersect.run_2set_c(data[0], data[1], self.entry).size
        return intersect.run_kset(data, self.entry).size


def _algo_id(name):
    from . import intersect

    return intersect.algo_id(name)


class _Resident:
    """Device copies of one pair and the device-pointer call on them."""

    def __init__(self, entry, a, b, count_only):
        import ctypes

        import torch

        from . import _native

        self.torch = torch
        self.lib = _native.hip()
        self.algo = _algo_id(entry)
        self.a = torch.from_numpy(np.ascontiguousarray(a, np.uint32).view(np.int32)).cuda() if a.size else \
            torch.zeros(4, dtype=torch.int32, device="cuda")
        self.b = torch.from_numpy(np.ascontiguousarray(b, np.uint32).view(np.int32)).cuda() if b.size else \
            torch.zeros(4, dtype=torch.int32, device="cuda")
        self.na, self.nb = int(a.size), int(b.size)
        self.cap = max(min(self.na, self.nb), 1)
        self.out = None if count_only else torch.empty(self.cap, dtype=torch.int32, device="cuda")
        self.tot = torch.zeros(2, dtype=torch.int64, device="cuda")
        self.stream = torch.cuda.current_stream()
        self._c = ctypes

    def run(self):
        from . import _native

        _native.check(self.lib.setops_hip_intersect2_dev(
            self.algo, self.a.data_ptr(), None, self.na, self.b.data_ptr(), None, self.nb,
            self.out.data_ptr() if self.out is not None else None, None, self.cap if self.out is not None else 0,
            self.tot.data_ptr(), self.stream.cuda_stream))
        self.stream.synchronize()
        return None


def gpu_resolver(count_only=False, resident=False):
    from . import intersect

    registered = intersect.registered()

    def resolve(name):
        entry = gpu_entry_name(name, registered)
        return GpuTimer(entry, count_only, resident) if entry else None

    return resolve


def time_datafile(timer, sets, warmup_s):
    """harness.rs:35-62: run for the warmup time, then time one run (ns)."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < warmup_s:
        timer.run(timer.prepare(sets))
    data = timer.prepare(sets)
    s = time.perf_counter_ns()
    timer.run(data)
    return time.perf_counter_ns() - s


def _result_run(x):
    cache = {"rd_access": None, "rd_miss": None, "wr_access": None, "wr_miss": None}
    return {"x": x, "times": [], "l1d": dict(cache), "l1i": dict(cache), "ll": dict(cache), "branches": None,
            "branch_misses": None, "cpu_stalled_front": None, "cpu_stalled_back": None, "instructions": None,
            "cpu_cycles": None, "cpu_cycles_ref": None}


def run(doc, datasets, experiments=None, resolver=None, warmup_s=1.0, log=print):
    """bin/benchmark.rs:60-243: the Results document for the selected experiments (all if none)."""
    from . import datafile

    resolver = resolver or gpu_resolver()
    chosen = [e for e in doc["experiment"] if not experiments or e["name"] in experiments]
    dataset_algos = {}
    for e in chosen:
        dataset_algos.setdefault(e["dataset"], set()).update(get_algorithms(doc, e))
    if not dataset_algos:
        raise ExperimentError("no algorithm matches found")
    results = {}
    for info in doc["dataset"]:
        algos = dataset_algos.get(info["name"])
        if not algos:
            continue
        log(info["name"])
        algo_results = {a: [] for a in sorted(algos)}
        for x in xvalues(info):
            log(f"[x: {x:4}]")
            xdir = os.path.join(datasets, info["name"], str(x))
            if not os.path.isdir(xdir):
                raise ExperimentError(f"unable to open {xdir}: dataset not generated")
            files = sorted((f for f in os.listdir(xdir) if f.isdigit()), key=int)
            for name, runs in algo_results.items():
                log(f"  {name}")
                timer = resolver(name)
                if timer is None:
                    log(f"  unknown algorithm {name}")
                    continue
                r = _result_run(x)
                warm = warmup_s / max(len(files), 1)
                for f in files:
                    sets = datafile.read(os.path.join(xdir, f))
                    try:
                        r["times"].append(time_datafile(timer, sets, warm))
                    except ExperimentError as err:
                        log(f"warn: {err}")
                        break
                runs.append(r)
        results[info["name"]] = {"info": info, "algos": algo_results}
    return {"experiments": chosen, "datasets": results, "algorithm_sets": doc["algorithm_sets"]}


def main(argv=None):
    ap = argparse.ArgumentParser(description="experiment-driven harness over the GPU entries")
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate", help="write the synthetic datasets of an experiment file")
    g.add_argument("-e", "--experiment", required=True)
    g.add_argument("-d", "--datasets", required=True)
    g.add_argument("names", nargs="*")
    r = sub.add_parser("run", help="time the experiments' algorithms on their datasets")
    r.add_argument("-e", "--experiment", required=True)
    r.add_argument("-d", "--datasets", required=True)
    r.add_argument("-o", "--out", required=True)
    r.add_argument("--count-only", action="store_true")
    r.add_argument("--resident", action="store_true", help="2-set inputs already in device memory")
    r.add_argument("--warmup", type=float, default=1.0, help="seconds of warmup per x value and algorithm")
    r.add_argument("experiments", nargs="*")
    args = ap.parse_args(argv)
    doc = load(args.experiment)
    if args.cmd == "generate":
        generate(doc, args.datasets, args.names)
        return
    from . import intersect

    intersect.init(0)
    res = run(doc, args.datasets, args.experiments, gpu_resolver(args.count_only, args.resident), args.warmup,
              log=lambda m: (print(m), sys.stdout.flush()))
    with open(args.out, "w") as f:
        json.dump(res, f)


if __name__ == "__main__":
    main()
