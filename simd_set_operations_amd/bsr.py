"""Base-and-State Representation (mirrors setops/src/bsr.rs).

A sorted u32 set is stored as parallel arrays: ``bases`` (value >> 5, strictly increasing) and
``states`` (32-bit mask of value & 31, never zero).  Encoding and decoding are vectorised numpy
and run on the host, as the reference encodes outside its timed region
(benchmark/src/timer/harness.rs:121-122).
"""
import numpy as np

BSR_WIDTH = 32  # bsr.rs:136
BSR_SHIFT = 5  # bsr.rs:137
BSR_MASK = BSR_WIDTH - 1  # bsr.rs:138


class BsrVec:
    """BsrVec (bsr.rs:48-110)."""

    __slots__ = ("bases", "states")

    def __init__(self, bases=None, states=None):
        self.bases = np.zeros(0, np.uint32) if bases is None else np.ascontiguousarray(bases, dtype=np.uint32)
        self.states = np.zeros(0, np.uint32) if states is None else np.ascontiguousarray(states, dtype=np.uint32)
        if self.bases.shape != self.states.shape:
            raise ValueError("bases and states differ in length")

    @classmethod
    def from_sorted(cls, sorted_values):
        """bsr.rs:140-164: one word per distinct base, bit (v & 31) set for every value."""
        v = np.asarray(sorted_values, dtype=np.uint32)
        if v.size == 0:
            return cls()
        base = v >> np.uint32(BSR_SHIFT)
        bit = np.left_shift(np.uint32(1), v & np.uint32(BSR_MASK), dtype=np.uint32)
        starts = np.flatnonzero(np.concatenate(([True], base[1:] != base[:-1])))
        states = np.bitwise_or.reduceat(bit, starts).astype(np.uint32)
        return cls(base[starts], states)

    def to_sorted_set(self):
        """bsr.rs:77-88: expand every state's bits in ascending order."""
        if self.bases.size == 0:
            return np.zeros(0, np.uint32)
        bits = np.unpackbits(self.states.view(np.uint8).reshape(-1, 4), axis=1, bitorder="little")
        w, b = np.nonzero(bits)
        return ((self.bases[w].astype(np.uint64) << BSR_SHIFT) | b.astype(np.uint64)).astype(np.uint32)

    def append(self, base, state):
        """bsr.rs:69-75 (with the debug assertions made unconditional)."""
        if state == 0:
            raise ValueError("BSR state must be non-zero")
        if self.bases.size and int(self.bases[-1]) >= int(base):
            raise ValueError("BSR bases must be strictly increasing")
        self.bases = np.append(self.bases, np.uint32(base))
        self.states = np.append(self.states, np.uint32(state))

    def element_count(self):
        return int(np.unpackbits(self.states.view(np.uint8)).sum())

    def bsr_ref(self):
        return self

    def __len__(self):
        return int(self.bases.size)

    def __eq__(self, other):
        return (
            isinstance(other, BsrVec)
            and np.array_equal(self.bases, other.bases)
            and np.array_equal(self.states, other.states)
        )

    def __repr__(self):
        return f"BsrVec(len={len(self)})"
