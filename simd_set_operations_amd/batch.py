"""Device-resident batches of set pairs (the batched form of Intersect2 / Intersect2Bsr).

Device memory comes from PyTorch (plumbing only); the library receives raw pointers through the
C ABI (``setops_hip_batch_*``).  A :class:`DeviceBatch` uploads a CSR batch once, plans once
(sizes only), and can be run many times on a stream.  A :class:`GatherBatch` takes pairs as
(begin, end) index ranges into two pools, so pairs can share sets (graph edges over one CSR,
queries over one corpus).
"""
import ctypes

import numpy as np

from . import _native

_c = ctypes


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise _native.SetopsError(_native.SETOPS_ERR_NO_DEVICE, "no GPU visible to torch")
    return torch


def _is_tensor(x):
    return type(x).__module__.startswith("torch")


def to_device_u32(arr, torch):
    """Upload a uint32 numpy array (as int32 bits); returns a tensor whose storage is 16-byte
    aligned (torch allocations are 256-byte aligned).  A device tensor is used as it is."""
    if _is_tensor(arr):
        return arr
    a = np.ascontiguousarray(arr)
    a = a.view(np.uint32) if a.dtype == np.int32 else np.ascontiguousarray(a, dtype=np.uint32)
    t = torch.from_numpy(a.view(np.int32)) if a.size else torch.zeros(4, dtype=torch.int32)
    return t.to("cuda")


def to_device_u64(arr, torch):
    if _is_tensor(arr):
        return arr
    a = np.ascontiguousarray(arr, dtype=np.uint64)
    return torch.from_numpy(a.view(np.int64)).to("cuda")


def resolve_algo(algo, a_sizes, b_sizes, bsr, indexed=False):
    """``hip_auto``: the entry and tile shape (``name@shape``) the tuning table prices cheapest for
    these pair sizes (its BSR twin, in the default shape, for BSR sets; for plain sets searched
    through an indexed resident pool, the cheapest search entry against one); any other name is
    returned as it is."""
    if algo != "hip_auto":
        return algo
    from . import tuning

    best = tuning.choose_resident(a_sizes, b_sizes) if indexed and not bsr else None
    best = best or tuning.choose(a_sizes, b_sizes)
    if bsr:
        from .intersect import registered

        best = best.partition("@")[0]
        twin = best + "_bsr"
        return twin if twin in registered() else "hip_galloping_bsr" if best == "hip_binary_search" else \
            "hip_branchless_merge_bsr"
    return best


class _Batch:
    def _outputs(self, capacity):
        torch = self.torch
        self.capacity = int(capacity)
        cap = max(self.capacity, 4)
        self.d_out = torch.empty(cap, dtype=torch.int32, device="cuda")
        self.d_out_s = torch.empty(cap, dtype=torch.int32, device="cuda") if self.bsr else None
        self.d_out_off = torch.empty(self.P + 1, dtype=torch.int64, device="cuda")
        self.d_counts = torch.empty(self.P, dtype=torch.int32, device="cuda")
        self.d_totals = torch.zeros(2, dtype=torch.int64, device="cuda")


def _signed(signed, A, B):
    """int32 mode (SETOPS_SIGNED_I32): as asked, else when a numpy pool is int32."""
    if signed is not None:
        return bool(signed)
    return any(not _is_tensor(x) and np.asarray(x).dtype == np.int32 for x in (A, B))


class DeviceBatch(_Batch):
    """CSR batch on the GPU: pair p = (A[a_off[p]:a_off[p+1]], B[b_off[p]:b_off[p+1]]).

    int32 pools (or ``signed=True``) select the int32 mode: sets sorted as signed integers."""

    def __init__(self, algo, A, a_off, B, b_off, A_states=None, B_states=None, signed=None):
        torch = _torch()
        from .intersect import algo_id

        lib = _native.hip()
        self.torch = torch
        self.algo = algo
        a_off = np.ascontiguousarray(a_off, dtype=np.uint64)
        b_off = np.ascontiguousarray(b_off, dtype=np.uint64)
        if a_off.size != b_off.size or a_off.size < 2:
            raise ValueError("a_off and b_off must both have P+1 >= 2 entries")
        if a_off[0] != 0 or b_off[0] != 0:
            raise ValueError("offsets must start at 0")
        if np.any(np.diff(a_off.astype(np.int64)) < 0) or np.any(np.diff(b_off.astype(np.int64)) < 0):
            raise ValueError("offsets must be non-decreasing")
        self.P = a_off.size - 1
        self.a_len = int(a_off[-1])
        self.b_len = int(b_off[-1])
        if len(A) != self.a_len or len(B) != self.b_len:
            raise ValueError("pool lengths must equal the last offsets")
        self.a_sizes = np.diff(a_off)
        self.b_sizes = np.diff(b_off)
        self.requested = algo
        algo = self.algo = resolve_algo(algo, self.a_sizes, self.b_sizes, A_states is not None)
        self.dA = to_device_u32(A, torch)
        self.dB = to_device_u32(B, torch)
        self.dA_off = to_device_u64(a_off, torch)
        self.dB_off = to_device_u64(b_off, torch)
        self.bsr = A_states is not None
        self.signed = _signed(signed, A, B)
        self.dAs = to_device_u32(A_states, torch) if self.bsr else None
        self.dBs = to_device_u32(B_states, torch) if self.bsr else None
        h = _c.c_void_p()
        _native.check(
            lib.setops_hip_batch_create(
                _c.byref(h), algo_id(algo) | (_native.SIGNED_I32 if self.signed else 0), self.dA.data_ptr(), self.dAs.data_ptr() if self.bsr else None,
                self.a_len, self.dA_off.data_ptr(), self.dB.data_ptr(), self.dBs.data_ptr() if self.bsr else None,
                self.b_len, self.dB_off.data_ptr(), self.P,
            )
        )
        self.handle = h
        self._outputs(np.minimum(self.a_sizes, self.b_sizes).sum())

    # -- running ----------------------------------------------------------------------------
    def run(self, stream=None, count_only=False):
        """Launch one batch intersection on `stream` (torch stream or None = current stream)."""
        lib = _native.hip()
        s = stream if stream is not None else self.torch.cuda.current_stream()
        _native.check(
            lib.setops_hip_batch_run(
                self.handle,
                None if count_only else self.d_out.data_ptr(),
                None if (count_only or not self.bsr) else self.d_out_s.data_ptr(),
                0 if count_only else self.d_out.numel(),
                self.d_out_off.data_ptr(),
                self.d_counts.data_ptr(),
                self.d_totals.data_ptr(),
                s.cuda_stream,
            )
        )

    def index_b(self):
        """Index the B pool (every 32nd element) for the search-policy entries: larger sets that
        are searched in global memory then cost about one 128-byte line per probe.  For resident
        B pools shared by many pairs; the pool must not change afterwards where pairs read it."""
        _native.check(_native.hip().setops_hip_batch_index_b(self.handle))
        return self

    def check(self):
        _native.check(_native.hip().setops_hip_batch_check(self.handle))

    def set_timing(self, on=True):
        _native.check(_native.hip().setops_hip_batch_set_timing(self.handle, 1 if on else 0))

    def kernel_time(self):
        """(summed main-kernel ms, runs) since the last call (HIP events on the run stream)."""
        ms = _c.c_double(0)
        runs = _c.c_uint64(0)
        _native.check(_native.hip().setops_hip_batch_kernel_time(self.handle, _c.byref(ms), _c.byref(runs)))
        return ms.value, runs.value

    def info(self):
        nt, g, u = _c.c_uint64(), _c.c_uint64(), _c.c_uint64()
        _native.check(_native.hip().setops_hip_batch_info(self.handle, _c.byref(nt), _c.byref(g), _c.byref(u)))
        return {"tiles": nt.value, "grid": g.value, "units": u.value}

    # -- results ------------------------------------------------------------------------------
    def results(self):
        """Host copies: (out_off[P+1], counts[P], data[, states]) of the last run."""
        self.torch.cuda.synchronize()
        self.check()
        off = self.d_out_off.cpu().numpy().view(np.uint64).copy()
        cnt = self.d_counts.cpu().numpy().view(np.uint32).copy()
        total = int(off[-1])
        data = self.d_out[:total].cpu().numpy().view(np.int32 if self.signed else np.uint32).copy()
        if self.bsr:
            st = self.d_out_s[:total].cpu().numpy().view(np.uint32).copy()
            return off, cnt, data, st
        return off, cnt, data

    def totals(self):
        self.torch.cuda.synchronize()
        t = self.d_totals.cpu().numpy().view(np.uint64)
        return int(t[0]), int(t[1])

    def close(self):
        if self.handle:
            _native.hip().setops_hip_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GatherBatch(DeviceBatch):
    """Gather batch on the GPU: pair p = (A[a_beg[p]:a_end[p]], B[b_beg[p]:b_end[p]]).

    Pools and index arrays may be numpy arrays (uploaded) or device tensors (used in place).
    ``capacity`` sizes the output (default: sum of min(|A_p|, |B_p|), computed on the host when
    the index arrays are numpy).  ``outputs=False`` allocates no result buffer (count-only use).
    ``index=True`` indexes the B pool at creation (``index_b``)."""

    def __init__(self, algo, A, a_beg, a_end, B, b_beg, b_end, A_states=None, B_states=None, capacity=None,
                 outputs=True, signed=None, index=False):
        torch = _torch()
        from .intersect import algo_id

        lib = _native.hip()
        self.torch = torch
        self.algo = algo
        self.P = int(len(a_beg))
        if not (len(a_end) == len(b_beg) == len(b_end) == self.P) or self.P == 0:
            raise ValueError("a_beg, a_end, b_beg, b_end must have the same non-zero length")
        la = lb = None
        if not _is_tensor(a_beg):
            la = np.asarray(a_end, np.int64) - np.asarray(a_beg, np.int64)
            lb = np.asarray(b_end, np.int64) - np.asarray(b_beg, np.int64)
            if np.any(la < 0) or np.any(lb < 0):
                raise ValueError("end must not precede begin")
        if capacity is None:
            if la is None:
                capacity = int(torch.minimum(a_end - a_beg, b_end - b_beg).sum().item())
            else:
                capacity = int(np.minimum(la, lb).sum())
        self.requested = algo
        if algo == "hip_auto" and la is None:
            la = (a_end - a_beg).cpu().numpy()
            lb = (b_end - b_beg).cpu().numpy()
        algo = self.algo = resolve_algo(algo, la, lb, A_states is not None, indexed=index)
        self.dA = to_device_u32(A, torch)
        self.dB = to_device_u32(B, torch)
        self.a_len = int(self.dA.numel()) if len(A) else 0
        self.b_len = int(self.dB.numel()) if len(B) else 0
        self.dA_beg = to_device_u64(a_beg, torch)
        self.dA_end = to_device_u64(a_end, torch)
        self.dB_beg = to_device_u64(b_beg, torch)
        self.dB_end = to_device_u64(b_end, torch)
        self.bsr = A_states is not None
        self.signed = _signed(signed, A, B)
        self.dAs = to_device_u32(A_states, torch) if self.bsr else None
        self.dBs = to_device_u32(B_states, torch) if self.bsr else None
        h = _c.c_void_p()
        _native.check(
            lib.setops_hip_batch_create_gather(
                _c.byref(h), algo_id(algo) | (_native.SIGNED_I32 if self.signed else 0), self.dA.data_ptr(), self.dAs.data_ptr() if self.bsr else None,
                self.a_len, self.dA_beg.data_ptr(), self.dA_end.data_ptr(), self.dB.data_ptr(),
                self.dBs.data_ptr() if self.bsr else None, self.b_len, self.dB_beg.data_ptr(),
                self.dB_end.data_ptr(), self.P,
            )
        )
        self.handle = h
        self._outputs(capacity if outputs else 0)
        if index:
            self.index_b()

    # -- k-set groups --------------------------------------------------------------------------
    def set_further(self, x_off, x_beg=None, x_end=None):
        """k-set groups: pair p's results must also be in every further set B[x_beg[j]:x_end[j]],
        j in [x_off[p], x_off[p+1]) (setops_hip_batch_set_further; search entries, plain sets).
        ``x_off=None`` detaches the further sets."""
        lib = _native.hip()
        if x_off is None:
            _native.check(lib.setops_hip_batch_set_further(self.handle, None, None, None))
            self._further = None
            return self
        torch = self.torch

        def dev(a):
            if _is_tensor(a):
                return a if a.numel() else torch.zeros(1, dtype=torch.int64, device="cuda")
            a = np.ascontiguousarray(a, dtype=np.uint64)
            return to_device_u64(a if a.size else np.zeros(1, np.uint64), torch)

        self._further = (dev(x_off), dev(x_beg), dev(x_end))
        _native.check(lib.setops_hip_batch_set_further(self.handle, *(t.data_ptr() for t in self._further)))
        return self


class KsetBatch(GatherBatch):
    """k-set groups over one pool, all answered in one batch: group g is the sets
    pool[beg[i]:end[i]] for i in [group_off[g], group_off[g+1]) (k >= 2 each), in the caller's order
    (smallest first, as svs_generic takes them, setops/src/intersect/svs.rs:50-73).  The first two
    sets of a group are its pair; the others are its further sets (set_further).  Results: one
    ascending k-way intersection per group, in group order (``results()``)."""

    def __init__(self, algo, pool, beg, end, group_off, index=False, capacity=None, signed=None):
        beg = np.ascontiguousarray(beg, np.uint64)
        end = np.ascontiguousarray(end, np.uint64)
        group_off = np.ascontiguousarray(group_off, np.int64)
        G = group_off.size - 1
        k = np.diff(group_off)
        if G < 1 or np.any(k < 2):
            raise ValueError("every group needs at least two sets")
        first = group_off[:-1]
        further = np.ones(beg.size, bool)
        further[first] = False
        further[first + 1] = False
        x_off = (group_off - 2 * np.arange(G + 1)).astype(np.uint64)
        super().__init__(algo, pool, beg[first], end[first], pool, beg[first + 1], end[first + 1], capacity=capacity,
                         signed=signed, index=index)
        self.set_further(x_off, beg[further], end[further])
