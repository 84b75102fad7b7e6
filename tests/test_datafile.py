"""The reference's datafile format (benchmark/src/datafile.rs): the round trips of its own tests
(datafile.rs:184-226) and the reader's error kinds."""
import os

import numpy as np
import pytest

from simd_set_operations_amd import datafile

# the sets of datafile.rs test_pair / test_kset / test_many_empty_sets / test_large_sets
REF_CASES = {
    "pair": [[0, 4, 10, 20, 21, 26, 99], [0, 5, 6]],
    "kset": [
        [122, 120, 161, 155, 97, 86, 36, 32, 80, 9, 149, 140, 200, 82, 143, 30, 71, 33],
        [19, 6, 141, 90, 194, 167, 119, 125, 156, 197, 79, 98, 160, 28, 42, 111, 124, 18],
        [74, 196, 139, 159, 78, 48, 168, 199, 169, 26, 15, 55, 182, 59, 10, 49, 165, 75],
        [99, 190, 11, 107, 193, 69, 38, 137, 65, 44, 127, 62, 146, 135, 94, 50, 105, 123],
        [41, 72, 68, 198, 195, 31, 134, 46, 103, 7, 61, 185, 188, 47, 152, 70, 189, 130],
        [126, 117, 148, 8, 178, 93, 166, 138, 171, 158, 131, 183, 157, 176, 16, 40, 34],
        [52, 54, 67, 122, 106, 12, 187, 150, 22, 172, 64, 163, 101, 37, 110, 170, 144],
        [113, 181, 87, 24, 60, 88, 164, 51, 63, 179, 43, 92, 186, 84, 25, 95, 115, 133],
        [85, 96, 132, 1, 56, 53, 89, 136, 21, 45, 27, 29, 58, 116, 2, 118, 39, 23, 20, 3],
        [100, 154, 4, 17, 129, 174, 57, 73, 145, 112, 14, 177, 184, 76, 91, 104, 83, 151],
        [173, 162, 81, 13, 114, 77, 66, 5, 191, 108, 180, 147, 175, 109, 35, 153, 128],
        [142, 192, 102],
    ],
    "many_empty": [[], [], [], [], []],
    "large": [list(range((1 << 16) - 2)), list(range((1 << 17) + 5)), list(range((1 << 14) // 3))],
}


@pytest.mark.parametrize("case", sorted(REF_CASES))
def test_write_read_round_trip(tmp_path, case):
    p = tmp_path / "f"
    datafile.write(p, REF_CASES[case])
    got = datafile.read(p)
    assert len(got) == len(REF_CASES[case])
    for g, e in zip(got, REF_CASES[case]):
        assert np.array_equal(g, np.array(e, np.uint32))


def test_layout_bytes(tmp_path):
    p = tmp_path / "f"
    datafile.write(p, [[1, 2], [3]])
    raw = open(p, "rb").read()
    assert raw[:3] == bytes([0xE9, 0xAA, 0x05]) and raw[3] & 1 == 1
    assert np.frombuffer(raw[4:8], np.uint32)[0] == 2
    assert list(np.frombuffer(raw[8:16], np.uint32)) == [2, 1]
    assert list(np.frombuffer(raw[16:], np.int32)) == [1, 2, 3]


def test_reader_errors(tmp_path):
    p = tmp_path / "f"
    datafile.write(p, [[1, 2], [3]])
    raw = bytearray(open(p, "rb").read())
    bad = tmp_path / "b"
    for mutate, kind in ((lambda r: r.__setitem__(0, 0), "bad magic"),
                         (lambda r: r.__setitem__(3, r[3] ^ 1), "bad endianness"),
                         (lambda r: r.__setitem__(slice(4, 8), np.uint32(1).tobytes()), "bad set count")):
        r = bytearray(raw)
        mutate(r)
        open(bad, "wb").write(bytes(r))
        with pytest.raises(datafile.DatafileError) as e:
            datafile.read(bad)
        assert e.value.kind == kind
    open(bad, "wb").write(bytes(raw[:-2]))
    with pytest.raises(datafile.DatafileError) as e:
        datafile.read(bad)
    assert e.value.kind == "io"
    with pytest.raises(datafile.DatafileError):
        datafile.write(bad, [[1]])


def test_dataset_directory(tmp_path):
    for x in ("0", "10", "5"):
        os.makedirs(tmp_path / x)
        for i in range(3):
            datafile.write(tmp_path / x / str(i), [[int(x), i], [i]])
    ds = datafile.read_dataset(tmp_path)
    assert list(ds) == ["0", "5", "10"]
    assert [list(s[1]) for s in ds["10"]] == [[0], [1], [2]]
