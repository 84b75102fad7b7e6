"""The C++ API (simd_set_operations_amd/cpp/setops.hpp) and its test program tests/cpp/
property_tests.cpp, built next to the library by the package Makefile."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "simd_set_operations_amd", "cpp", "build", "property_tests")


def test_cpp_program_is_built_and_lists_the_registry():
    assert os.path.exists(BIN), "run __graft_entry__.build() (make -C simd_set_operations_amd/csrc)"
    r = subprocess.run([BIN, "--registry"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    names = {line.split()[0] for line in r.stdout.splitlines()}
    for n in ("hip_branchless_merge", "hip_galloping", "hip_binary_search", "hip_shuffling", "hip_qfilter",
              "hip_branchless_merge_bsr", "hip_galloping_bsr", "hip_shuffling_bsr", "hip_qfilter_bsr",
              "hip_simd_galloping", "hip_simd_galloping_bsr", "hip_broadcast", "hip_broadcast_bsr"):
        assert n in names


@pytest.mark.gpu
def test_cpp_property_tests_pass_on_the_gpu():
    sys.stdout.write("[gpu-case] C++ property tests\n")
    sys.stdout.flush()
    r = subprocess.run([BIN], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    sys.stdout.write(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-4000:]
    assert "0 failures" in r.stdout
