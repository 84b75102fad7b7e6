"""The reference's binary set files (benchmark/src/datafile.rs:5-166), read and written.

Layout: 3-byte magic E9 AA 05, one flag byte (bit 0: written little-endian), u32 set count
(>= 2), the u32 length of every set, then every set's elements as i32, in order.  Sets of the
path are sorted unique u32 with values below 2^31, so the i32 payload reads as u32 unchanged.

``read`` / ``write`` keep the reference's checks and error kinds (BadMagic, BadEndianness,
BadSetCount, and I/O errors for truncated files).  ``read_dataset`` walks a dataset directory
laid out as the reference's generator writes it (datasets/<name>/<x>/<i>, benchmark/src/
bin/generate.rs:152-221) and returns the sets of every datafile per x value.
"""
import os
import sys

import numpy as np

MAGIC = bytes([0xE9, 0xAA, 0x05])
LITTLE_ENDIAN_BIT = 1
MIN_SET_COUNT = 2


class DatafileError(Exception):
    """A datafile the reference's reader would reject (kind: 'bad magic', 'bad endianness',
    'bad set count', 'io')."""

    def __init__(self, kind, detail=""):
        super().__init__(f"{kind}{': ' + detail if detail else ''}")
        self.kind = kind


def _little():
    return sys.byteorder == "little"


def write(path, sets):
    """Writes sets (sequences of ints below 2^31) as a datafile (datafile.rs:117-160)."""
    if len(sets) < MIN_SET_COUNT or len(sets) > 0xFFFFFFFF:
        raise DatafileError("bad set count", str(len(sets)))
    arrs = [np.ascontiguousarray(s, dtype=np.int32) for s in sets]
    with open(path, "wb") as f:
        f.write(MAGIC + bytes([LITTLE_ENDIAN_BIT if _little() else 0]))
        f.write(np.uint32(len(arrs)).tobytes())
        f.write(np.array([a.size for a in arrs], dtype=np.uint32).tobytes())
        for a in arrs:
            f.write(a.tobytes())


def read(path):
    """Reads a datafile (datafile.rs:66-115); returns a list of uint32 arrays."""
    with open(path, "rb") as f:
        header = f.read(8)
        if len(header) < 8:
            raise DatafileError("io", "truncated header")
        if header[:3] != MAGIC:
            raise DatafileError("bad magic")
        if bool(header[3] & LITTLE_ENDIAN_BIT) != _little():
            raise DatafileError("bad endianness", "system is " + ("little endian" if _little() else "big endian"))
        count = int(np.frombuffer(header[4:8], dtype=np.uint32)[0])
        if count < MIN_SET_COUNT:
            raise DatafileError("bad set count", str(count))
        raw = f.read(4 * count)
        if len(raw) < 4 * count:
            raise DatafileError("io", "truncated lengths")
        lengths = np.frombuffer(raw, dtype=np.uint32)
        sets = []
        for n in lengths:
            data = f.read(4 * int(n))
            if len(data) < 4 * int(n):
                raise DatafileError("io", "truncated set")
            sets.append(np.frombuffer(data, dtype=np.int32).view(np.uint32).copy())
    return sets


def read_dataset(root):
    """{x: [sets of datafile 0, sets of datafile 1, ...]} for a directory root/<x>/<i>."""
    out = {}
    for x in sorted(os.listdir(root), key=lambda s: (len(s), s)):
        d = os.path.join(root, x)
        if not os.path.isdir(d):
            continue
        files = sorted((f for f in os.listdir(d) if f.isdigit()), key=int)
        out[x] = [read(os.path.join(d, f)) for f in files]
    return out
