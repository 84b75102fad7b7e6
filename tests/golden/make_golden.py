"""Writes the golden vectors of tests/golden/ from the reference's own compiled code.

Inputs are seeded (numpy PCG64); expected outputs come from the reference's
setops/ffi/qfilter/qfilter.cpp (intersect_qfilter_uint_b4_v2 and intersect_qfilter_bsr_b4_v2),
compiled from its source into oracle/_ref/ by oracle/Makefile -- not from the oracle, so the
fixtures pin the oracle instead of restating it.  The reference's known-answer cases
(setops/tests/unit_tests.rs:6-59) are included as written there.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def sorted_unique(rng, n, universe):
    return np.unique(rng.integers(0, universe, size=n, dtype=np.int64)).astype(np.uint32)


def pair_cases(rng):
    cases = []
    # (len_a, len_b, universe): similar sizes, skewed, tiny, dense, empty
    shapes = [(0, 0, 10), (0, 17, 100), (5, 0, 100), (1, 1, 2), (3, 3, 4), (7, 9, 16), (15, 16, 40),
              (16, 16, 40), (17, 33, 64), (100, 100, 150), (100, 100, 1000), (1000, 1000, 1500),
              (1000, 1000, 100000), (37, 4000, 8000), (4000, 37, 8000), (1, 5000, 10000), (4096, 4096, 8192),
              (20000, 20000, 60000), (300, 20000, 40000), (8192, 8192, 1 << 17)]
    for la, lb, u in shapes:
        a = sorted_unique(rng, la, u)
        b = sorted_unique(rng, lb, u)
        # plant some shared elements so selectivity is not only chance
        if a.size and b.size:
            k = min(a.size, b.size) // 3
            if k:
                b = np.unique(np.concatenate([b, rng.choice(a, size=k, replace=False)])).astype(np.uint32)
        cases.append((a, b))
    return cases


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libqfilter_ref.so is not built (make -C oracle with the reference present)")
    rng = np.random.Generator(np.random.PCG64(20240601))
    arrays = {}
    for i, (a, b) in enumerate(pair_cases(rng)):
        arrays[f"pair{i}_a"] = a
        arrays[f"pair{i}_b"] = b
        arrays[f"pair{i}_out"] = oracle.ref_qfilter(a, b, v2=True)
        ba, sa = oracle.bsr_from_sorted(a)
        bb, sb = oracle.bsr_from_sorted(b)
        ob, os_ = oracle.ref_qfilter_bsr(ba, sa, bb, sb, v2=True)
        arrays[f"pair{i}_bsr_bases"] = ob
        arrays[f"pair{i}_bsr_states"] = os_
    # setops/tests/unit_tests.rs:6-45 (known answers) and :47-59 (galloping case; note that the
    # reference's `1<<12 + 1` parses as 1 << 13)
    kat = [
        ([1, 2, 3, 4], [1, 2, 3, 4, 5], [1, 2, 3, 4]),
        ([0, 4, 5, 8], [1, 2, 3, 6], []),
        ([1, 4, 5], [1, 4, 5], [1, 4, 5]),
        ([10, 42], [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 22, 25, 28, 39, 42, 43, 47, 49], [10, 42]),
        ([1, 3, 5, 8, 9, 10, 14, 15, 18, 20], [1, 2, 3, 4, 9, 10, 11, 15, 20, 21], [1, 3, 9, 10, 15, 20]),
        ([1 << 13], list(range(12345)), [1 << 13]),
    ]
    for i, (a, b, o) in enumerate(kat):
        arrays[f"kat{i}_a"] = np.array(a, np.uint32)
        arrays[f"kat{i}_b"] = np.array(b, np.uint32)
        arrays[f"kat{i}_out"] = np.array(o, np.uint32)
    np.savez_compressed(os.path.join(OUT, "pairs.npz"), **arrays)
    n_pairs = len(pair_cases(np.random.Generator(np.random.PCG64(1))))
    print(f"wrote {len(arrays)} arrays ({n_pairs} pairs, {len(kat)} known answers) to tests/golden/pairs.npz")


if __name__ == "__main__":
    main()
