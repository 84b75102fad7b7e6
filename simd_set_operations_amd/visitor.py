"""Output sinks ("visitors") of the slice-in / visitor-out contract.

Mirrors setops/src/visitor.rs of the reference: a 2-set function visits every result element
exactly once, in ascending order.  The GPU entries produce a whole result buffer per call, so
besides the per-element ``visit`` every sink has a bulk ``visit_many`` that is equivalent to
calling ``visit`` on each element in order.
"""
import numpy as np


class Visitor:
    """Trait Visitor<T> (visitor.rs:15-17)."""

    def visit(self, value):
        raise NotImplementedError

    def visit_many(self, values):
        for v in np.asarray(values).tolist():
            self.visit(v)


class Clearable:
    """Trait Clearable (visitor.rs:19-21)."""

    def clear(self):
        raise NotImplementedError


class Counter(Visitor, Clearable):
    """Counts visits without storing them (visitor.rs:24-48).  For BSR visits it counts the
    elements, i.e. popcount(state) per visited word (visitor.rs:399-403)."""

    def __init__(self):
        self._count = 0

    def visit(self, value):
        self._count += 1

    def visit_many(self, values):
        self._count += int(np.asarray(values).size)

    def visit_bsr(self, base, state):
        self._count += bin(int(state) & 0xFFFFFFFF).count("1")

    def visit_bsr_many(self, bases, states):
        s = np.asarray(states, dtype=np.uint32)
        self._count += int(np.unpackbits(s.view(np.uint8)).sum())

    def add(self, n):
        self._count += int(n)

    def count(self):
        return self._count

    def clear(self):
        self._count = 0


class VecWriter(Visitor, Clearable):
    """Appends visits to a growable vector (visitor.rs:51-97).  Like VecWriter<T>, it holds the
    element type it is given: int32 visits (the int32 mode) stay int32, anything else is u32."""

    def __init__(self, capacity=0, dtype=None):
        self._parts = []
        self._n = 0
        self._dtype = dtype

    def _as(self, values):
        v = np.asarray(values)
        if self._dtype is None:
            self._dtype = np.int32 if v.dtype == np.int32 else np.uint32
        return v.astype(self._dtype, copy=False)

    def visit(self, value):
        self._parts.append(self._as(np.asarray([value])))
        self._n += 1

    def visit_many(self, values):
        v = self._as(values)
        if v.size:
            self._parts.append(v.copy())
            self._n += v.size

    def clear(self):
        self._parts = []
        self._n = 0

    def as_array(self):
        if not self._parts:
            return np.zeros(0, dtype=self._dtype or np.uint32)
        if len(self._parts) > 1:
            self._parts = [np.concatenate(self._parts)]
        return self._parts[0]

    def __len__(self):
        return self._n


class SliceWriter(Visitor, Clearable):
    """Writes visits into a caller-provided buffer (visitor.rs:100-131).  Overrunning the buffer
    raises IndexError, as indexing past the slice panics in the reference."""

    def __init__(self, data):
        self.data = data
        self._pos = 0

    def visit(self, value):
        self.data[self._pos] = value
        self._pos += 1

    def visit_many(self, values):
        v = np.asarray(values)
        end = self._pos + v.size
        if end > len(self.data):
            raise IndexError(f"SliceWriter overflow: {end} > {len(self.data)}")
        self.data[self._pos:end] = v
        self._pos = end

    def position(self):
        return self._pos

    def clear(self):
        self._pos = 0

    def as_array(self):
        return np.asarray(self.data[: self._pos])


class BsrWriter(Clearable):
    """BSR sink collecting (base, state) words (the BsrVec visitor, setops/src/bsr.rs:48-75):
    bases must arrive strictly increasing with non-zero states."""

    def __init__(self):
        self.bases = []
        self.states = []

    def visit_bsr(self, base, state):
        self.bases.append(np.asarray([base], dtype=np.uint32))
        self.states.append(np.asarray([state], dtype=np.uint32))

    def visit_bsr_many(self, bases, states):
        b = np.asarray(bases, dtype=np.uint32)
        s = np.asarray(states, dtype=np.uint32)
        if b.size:
            self.bases.append(b.copy())
            self.states.append(s.copy())

    def clear(self):
        self.bases = []
        self.states = []

    def into_bsr(self):
        from .bsr import BsrVec

        if not self.bases:
            return BsrVec(np.zeros(0, np.uint32), np.zeros(0, np.uint32))
        return BsrVec(np.concatenate(self.bases), np.concatenate(self.states))
