"""superintervals_amd -- MI355X superset-index build and batched overlap queries.

Host-side Python mirror of the reference's Python binding
(src/superintervals/intervalmap.pyx of kcleal/superintervals v1.0.0) over the
C-ABI of libsuperintervals_hip.so.  Every build and query runs on the GPU; the
package has no CPU fallback and fails to import when the library is missing.

    DeviceIndex   -- the device-resident index (si_device.h) for numpy host
                     arrays or torch device tensors; batch entry points.
    IntervalMap   -- the reference's Python IntervalMap API (add / build / count /
                     search_idxs / search_values / *_batch / from_arrays ...).
    workload_intervals / workload_queries -- deterministic synthetic inputs.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from ._native import SI_ORDER_DESC, SI_ORDER_VECTOR, SI_INDEX_NONE, SIError

__all__ = ["DeviceIndex", "LearnedIndex", "SetTable", "IntervalMap", "SIError", "SI_ORDER_DESC", "SI_ORDER_VECTOR",
           "SI_INDEX_NONE", "workload_intervals", "workload_queries", "add_flags", "lib_path"]

lib_path = N.LIB_PATH
_ORDERS = {"desc": SI_ORDER_DESC, "vector": SI_ORDER_VECTOR,
           SI_ORDER_DESC: SI_ORDER_DESC, SI_ORDER_VECTOR: SI_ORDER_VECTOR}


def _ptr(a) -> int:
    """Address of a numpy array or torch tensor (None -> 0)."""
    if a is None:
        return 0
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()   # torch tensor


def _is_device(a) -> bool:
    return not isinstance(a, np.ndarray) and getattr(a, "is_cuda", False)


def _stream_of(stream) -> int:
    if stream is None:
        return 0
    return int(getattr(stream, "cuda_stream", stream))


def add_flags(starts: np.ndarray, ends: np.ndarray):
    """(start_sorted, end_sorted) as add() leaves them after appending the arrays in
    order (superintervals.hpp:95-105, c_superintervals.h:408-413)."""
    starts = np.asarray(starts)
    ends = np.asarray(ends)
    if starts.size < 2:
        return True, True
    ds = np.diff(starts.astype(np.int64))
    de = np.diff(ends.astype(np.int64))
    bad = np.flatnonzero(ds < 0)
    k = bad[0] if bad.size else ds.size
    start_sorted = bad.size == 0
    end_sorted = not bool(np.any((ds[:k] == 0) & (de[:k] > 0)))
    return start_sorted, end_sorted


def workload_intervals(config: int, n: int, seed: int = 1, coord_dtype=None, sorted: bool = True):
    """n intervals of workload `config` (1-4) as (starts, ends) numpy arrays."""
    dt = np.dtype(coord_dtype or (np.int64 if config == 3 else np.int32))
    s = np.empty(n, dt)
    e = np.empty(n, dt)
    N.check(N.lib.siWorkloadIntervals(config, seed, n, dt.itemsize, int(sorted), _ptr(s), _ptr(e)),
            "siWorkloadIntervals")
    return s, e


def workload_queries(config: int, n: int, seed: int = 2, coord_dtype=None):
    """n queries of workload `config` (1-4) as (qs, qe) numpy arrays."""
    dt = np.dtype(coord_dtype or (np.int64 if config == 3 else np.int32))
    s = np.empty(n, dt)
    e = np.empty(n, dt)
    N.check(N.lib.siWorkloadQueries(config, seed, n, dt.itemsize, _ptr(s), _ptr(e)),
            "siWorkloadQueries")
    return s, e


class DeviceIndex:
    """A built superset index resident in HBM (si_device.h siIndex)."""

    def __init__(self, coord_dtype=np.int32):
        self.dtype = np.dtype(coord_dtype)
        if self.dtype not in (np.dtype(np.int32), np.dtype(np.int64)):
            raise TypeError("coordinates must be int32 or int64")
        h = N.lib.siIndexCreate(self.dtype.itemsize)
        if not h:
            raise SIError(N.SI_ERR_HIP, "siIndexCreate")
        self._h = h
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            N.lib.siIndexDestroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter exit the module's globals may already be cleared; the process
        # releases its device memory then anyway
        if N is not None and getattr(N, "lib", None) is not None:
            self.close()

    @property
    def handle(self) -> int:
        return self._h

    def __len__(self):
        return self.n

    # ---- build ------------------------------------------------------------
    def build(self, starts, ends, values=None, start_sorted=None, end_sorted=None, stream=None):
        """Build from host numpy arrays or device tensors (same coordinate dtype).
        The sortedness flags default to what add() would have recorded."""
        on_dev = _is_device(starts)
        if not on_dev:
            starts = np.ascontiguousarray(starts, dtype=self.dtype)
            ends = np.ascontiguousarray(ends, dtype=self.dtype)
            if values is not None:
                values = np.ascontiguousarray(values, dtype=np.int32)
        if len(starts) != len(ends):
            raise ValueError("starts and ends must have the same length")
        if start_sorted is None or end_sorted is None:
            if on_dev:
                ss, es = add_flags(starts.cpu().numpy(), ends.cpu().numpy())
            else:
                ss, es = add_flags(starts, ends)
            start_sorted = ss if start_sorted is None else start_sorted
            end_sorted = es if end_sorted is None else end_sorted
        n = len(starts)
        N.check(N.lib.siIndexBuild(self._h, _ptr(starts), _ptr(ends), _ptr(values), n,
                                   int(bool(start_sorted)), int(bool(end_sorted)), int(on_dev),
                                   _stream_of(stream)), "siIndexBuild")
        self.n = n
        return self

    @property
    def last_build_ms(self) -> float:
        return N.lib.siIndexLastBuildMs(self._h)

    def download(self, values: bool = False):
        """Sorted starts, ends, perm (uint32) and branch (int64, -1 = none)."""
        n = self.n
        s = np.empty(n, self.dtype)
        e = np.empty(n, self.dtype)
        p = np.empty(n, np.uint32)
        b = np.empty(n, np.uint64)
        v = np.empty(n, np.int32) if values else None
        N.check(N.lib.siIndexDownload(self._h, _ptr(s), _ptr(e), _ptr(v), _ptr(p), _ptr(b)),
                "siIndexDownload")
        out = {"starts": s, "ends": e, "perm": p, "branch": b.view(np.int64)}
        if values:
            out["values"] = v
        return out

    def device_arrays(self):
        return {"starts": N.lib.siIndexDeviceStarts(self._h), "ends": N.lib.siIndexDeviceEnds(self._h),
                "branch": N.lib.siIndexDeviceBranch(self._h), "perm": N.lib.siIndexDevicePerm(self._h)}

    # ---- host-buffer queries ----------------------------------------------
    def _q(self, qs, qe):
        qs = np.ascontiguousarray(qs, dtype=self.dtype)
        qe = np.ascontiguousarray(qe, dtype=self.dtype)
        if qs.shape != qe.shape:
            raise ValueError("starts and ends must have the same length")
        return qs, qe

    def upper_bound(self, xs):
        xs = np.ascontiguousarray(xs, dtype=self.dtype)
        out = np.empty(xs.size, np.uint32)
        N.check(N.lib.siUpperBoundHost(self._h, _ptr(xs), xs.size, _ptr(out)), "siUpperBoundHost")
        return out

    def count(self, qs, qe):
        qs, qe = self._q(qs, qe)
        out = np.empty(qs.size, np.uint32)
        N.check(N.lib.siCountHost(self._h, _ptr(qs), _ptr(qe), qs.size, _ptr(out)), "siCountHost")
        return out

    def has_overlaps(self, qs, qe):
        qs, qe = self._q(qs, qe)
        out = np.empty(qs.size, np.uint8)
        N.check(N.lib.siHasOverlapsHost(self._h, _ptr(qs), _ptr(qe), qs.size, _ptr(out)),
                "siHasOverlapsHost")
        return out.astype(bool)

    def coverage(self, qs, qe):
        qs, qe = self._q(qs, qe)
        c = np.empty(qs.size, np.uint32)
        v = np.empty(qs.size, np.int32)
        N.check(N.lib.siCoverageHost(self._h, _ptr(qs), _ptr(qe), qs.size, _ptr(c), _ptr(v)),
                "siCoverageHost")
        return c, v

    def search(self, qs, qe, order="desc", values=False):
        """CSR result (offsets[nq+1] uint64, idx uint32 or values int32)."""
        qs, qe = self._q(qs, qe)
        nq = qs.size
        off = np.empty(nq + 1, np.uint64)
        tot = C.c_uint64(0)
        rc = N.lib.siSearchHost(self._h, _ptr(qs), _ptr(qe), nq, _ORDERS[order], _ptr(off), None,
                                None, 0, C.byref(tot))
        if rc not in (N.SI_OK, N.SI_ERR_CAPACITY):
            N.check(rc, "siSearchHost")
        total = int(tot.value)
        out = np.empty(total, np.int32 if values else np.uint32)
        if total:
            N.check(N.lib.siSearchHost(self._h, _ptr(qs), _ptr(qe), nq, _ORDERS[order], _ptr(off),
                                       None if values else _ptr(out), _ptr(out) if values else None,
                                       total, C.byref(tot)), "siSearchHost")
        return off, out

    def search_keys(self, qs, qe, items=False):
        """Batched search_keys (items=False) or search_items (items=True), superintervals.hpp
        :913-971: offsets[nq+1] and, per hit in descending walk order, its start and end
        (and payload).  Runs on the device: plan, offsets, descending fill, key gather."""
        import torch
        qs, qe = self._q(qs, qe)
        nq = qs.size
        dev = torch.device("cuda:0")
        tqs, tqe = torch.from_numpy(qs).to(dev), torch.from_numpy(qe).to(dev)
        off = torch.empty(nq + 1, dtype=torch.int64, device=dev)
        plan = torch.empty(max(self.plan_bytes(nq), 256), dtype=torch.uint8, device=dev)
        if nq:
            self.plan_device(tqs, tqe, plan)
            self.offsets_planned(plan, nq, off)
        else:
            off.zero_()
        total = int(off[-1].item())
        tdt = torch.int32 if self.dtype == np.int32 else torch.int64
        idx = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
        ks = torch.empty(max(total, 1), dtype=tdt, device=dev)
        ke = torch.empty(max(total, 1), dtype=tdt, device=dev)
        kv = torch.empty(max(total, 1), dtype=torch.int32, device=dev) if items else None
        if total:
            self.fill_planned(plan, nq, off, idx, order="desc")
            N.check(N.lib.siGatherKeys(self._h, _ptr(idx), total, _ptr(ks), _ptr(ke), _ptr(kv), None),
                    "siGatherKeys")
        torch.cuda.synchronize()
        out = [off.cpu().numpy().astype(np.uint64), ks[:total].cpu().numpy(), ke[:total].cpu().numpy()]
        if items:
            out.append(kv[:total].cpu().numpy())
        return tuple(out)

    def _set_op_inputs(self, a_s, a_e, order):
        import torch
        a_s, a_e = self._q(a_s, a_e)
        off, idx = self.search(a_s, a_e, order=order)
        dev = torch.device("cuda:0")
        t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
        return a_s, a_e, off, idx, t, dev

    def intersection_pieces(self, a_s, a_e, a_data=None):
        """Intersection of the intervals (a_s[k], a_e[k]) with this index, superintervals.hpp
        :1158-1174: for each k in order and each hit h in the C++ vector order, the piece
        [max(a_s, starts[h]), min(a_e, ends[h])].  Returns (starts, ends, k, h): the source
        interval k of the first set and the sorted position h in this index of each piece.
        A piece has start > end when either interval does; callers drop those, as the
        reference does."""
        import torch
        a_s, a_e, off, idx, t, dev = self._set_op_inputs(a_s, a_e, "vector")
        na, total = a_s.size, int(off[-1])
        tdt = torch.int32 if self.dtype == np.int32 else torch.int64
        os_, oe = (torch.empty(max(total, 1), dtype=tdt, device=dev) for _ in range(2))
        ok = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
        if total:
            ts, te, toff, tidx = t(a_s), t(a_e), t(off.view(np.int64)), t(idx)
            tk = torch.arange(na, dtype=torch.int32, device=dev)
            N.check(N.lib.siIntersectPieces(self._h, _ptr(ts), _ptr(te), _ptr(tk), na, _ptr(toff), _ptr(tidx),
                                            _ptr(os_), _ptr(oe), _ptr(ok), None, None), "siIntersectPieces")
            torch.cuda.synchronize()
        return (os_[:total].cpu().numpy(), oe[:total].cpu().numpy(), ok[:total].cpu().numpy(),
                idx[:total].astype(np.int64))

    def difference_pieces(self, a_s, a_e):
        """Difference of the intervals (a_s[k], a_e[k]) and this index, superintervals.hpp
        :1188-1211: the parts of each interval k that no interval here covers, in the
        reference's order.  Returns (starts, ends, k)."""
        import torch
        a_s, a_e, off, idx, t, dev = self._set_op_inputs(a_s, a_e, "desc")
        na = a_s.size
        tdt = torch.int32 if self.dtype == np.int32 else torch.int64
        ts, te, toff = t(a_s), t(a_e), t(off.view(np.int64))
        tidx = t(idx) if idx.size else torch.empty(1, dtype=torch.int32, device=dev)
        poff = torch.empty(na + 1, dtype=torch.int64, device=dev)
        ws = torch.empty(max(self.workspace_bytes(na), 256), dtype=torch.uint8, device=dev)
        N.check(N.lib.siDifferenceOffsets(self._h, _ptr(ts), _ptr(te), na, _ptr(toff), _ptr(tidx), _ptr(poff),
                                          _ptr(ws), ws.numel(), None), "siDifferenceOffsets")
        torch.cuda.synchronize()
        total = int(poff[-1].item())
        os_, oe = (torch.empty(max(total, 1), dtype=tdt, device=dev) for _ in range(2))
        ok = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
        if total:
            tk = torch.arange(na, dtype=torch.int32, device=dev)
            N.check(N.lib.siDifferenceFill(self._h, _ptr(ts), _ptr(te), _ptr(tk), na, _ptr(toff), _ptr(tidx),
                                           _ptr(poff), _ptr(os_), _ptr(oe), _ptr(ok), None), "siDifferenceFill")
            torch.cuda.synchronize()
        return os_[:total].cpu().numpy(), oe[:total].cpu().numpy(), ok[:total].cpu().numpy()

    def walk_stats_device(self, qs_t, qe_t, stream=None):
        """(hits, ends reads, branch reads, found) totals of the reference walk."""
        import torch
        tot = torch.zeros(4, dtype=torch.int64, device=qs_t.device)
        N.check(N.lib.siWalkStats(self._h, _ptr(qs_t), _ptr(qe_t), qs_t.numel(), _ptr(tot),
                                  _stream_of(stream)), "siWalkStats")
        return [int(x) for x in tot.cpu().tolist()]

    # ---- device-buffer queries (torch tensors, stream-ordered) --------------
    def count_device(self, qs_t, qe_t, counts_t, stream=None):
        N.check(N.lib.siCount(self._h, _ptr(qs_t), _ptr(qe_t), qs_t.numel(), _ptr(counts_t),
                              _stream_of(stream)), "siCount")

    def offsets_device(self, qs_t, qe_t, offsets_t, counts_t=None, workspace_t=None, stream=None):
        ws_bytes = workspace_t.numel() * workspace_t.element_size() if workspace_t is not None else 0
        N.check(N.lib.siSearchOffsets(self._h, _ptr(qs_t), _ptr(qe_t), qs_t.numel(), _ptr(counts_t),
                                      _ptr(offsets_t), _ptr(workspace_t), ws_bytes,
                                      _stream_of(stream)), "siSearchOffsets")

    def fill_device(self, qs_t, qe_t, offsets_t, idx_t, order="desc", stream=None):
        N.check(N.lib.siSearchFill(self._h, _ptr(qs_t), _ptr(qe_t), qs_t.numel(), _ORDERS[order],
                                   _ptr(offsets_t), _ptr(idx_t), _stream_of(stream)), "siSearchFill")

    def values_fill_device(self, qs_t, qe_t, offsets_t, vals_t, stream=None):
        N.check(N.lib.siSearchValuesFill(self._h, _ptr(qs_t), _ptr(qe_t), qs_t.numel(),
                                         _ptr(offsets_t), _ptr(vals_t), _stream_of(stream)),
                "siSearchValuesFill")

    @staticmethod
    def workspace_bytes(nq: int) -> int:
        return int(N.lib.siSearchWorkspaceBytes(nq))

    # ---- planned batches (torch tensors) ------------------------------------
    def plan_bytes(self, nq: int) -> int:
        return int(N.lib.siPlanBytes(self._h, nq))

    def plan_device(self, qs_t, qe_t, plan_t, stream=None):
        """Sort the batch by end and compute upper_bound positions into plan_t (a device
        tensor of at least plan_bytes(nq) bytes)."""
        N.check(N.lib.siPlanQueries(self._h, _ptr(qs_t), _ptr(qe_t), qs_t.numel(), _ptr(plan_t),
                                    plan_t.numel() * plan_t.element_size(), _stream_of(stream)),
                "siPlanQueries")

    def count_planned(self, plan_t, nq, counts_t, stream=None):
        N.check(N.lib.siCountPlanned(self._h, _ptr(plan_t), nq, _ptr(counts_t), _stream_of(stream)),
                "siCountPlanned")

    def offsets_planned(self, plan_t, nq, offsets_t, counts_t=None, stream=None):
        N.check(N.lib.siSearchOffsetsPlanned(self._h, _ptr(plan_t), nq, _ptr(counts_t), _ptr(offsets_t),
                                             _stream_of(stream)), "siSearchOffsetsPlanned")

    def fill_planned(self, plan_t, nq, offsets_t, idx_t, order="desc", stream=None, capacity=None):
        """Fill pass; with capacity, only output slots [0, capacity) are written."""
        if capacity is None:
            N.check(N.lib.siSearchFillPlanned(self._h, _ptr(plan_t), nq, _ORDERS[order], _ptr(offsets_t),
                                              _ptr(idx_t), _stream_of(stream)), "siSearchFillPlanned")
        else:
            N.check(N.lib.siSearchFillPlannedBounded(self._h, _ptr(plan_t), nq, _ORDERS[order], _ptr(offsets_t),
                                                     _ptr(idx_t), int(capacity), _stream_of(stream)),
                    "siSearchFillPlannedBounded")

    def values_fill_planned(self, plan_t, nq, offsets_t, vals_t, stream=None):
        N.check(N.lib.siSearchValuesFillPlanned(self._h, _ptr(plan_t), nq, _ptr(offsets_t), _ptr(vals_t),
                                                _stream_of(stream)), "siSearchValuesFillPlanned")


class SetTable:
    """Many resident int32 indexes answered in one batch (si_device.h siSetTable): slot ->
    DeviceIndex; every query names its slot.  Results equal each slot's own DeviceIndex."""

    def __init__(self, max_sets: int):
        h = N.lib.siSetTableCreate(4, max_sets)
        if not h:
            raise SIError(N.SI_ERR_ARG, "siSetTableCreate")
        self._h = h
        self.max_sets = max_sets
        self.slots = [None] * max_sets       # keeps each DeviceIndex alive while it is in a slot

    def close(self):
        if getattr(self, "_h", None):
            N.lib.siSetTableDestroy(self._h)
            self._h = None
        self.slots = []

    def __del__(self):
        if N is not None and getattr(N, "lib", None) is not None:   # (cleared at interpreter exit)
            self.close()

    def put(self, slot: int, index):
        """Put a built int32 DeviceIndex into slot (None empties it)."""
        if index is not None and index.dtype != np.dtype(np.int32):
            raise TypeError("set tables hold int32 indexes")
        N.check(N.lib.siSetTablePut(self._h, slot, index.handle if index is not None else None), "siSetTablePut")
        self.slots[slot] = index

    def plan_bytes(self, nq: int) -> int:
        return int(N.lib.siSetTablePlanBytes(self._h, nq))

    def offsets_device(self, set_ids_t, qs_t, qe_t, plan_t, offsets_t, counts_t=None, stream=None):
        N.check(N.lib.siSetTableSearchOffsets(self._h, _ptr(set_ids_t), _ptr(qs_t), _ptr(qe_t), qs_t.numel(),
                                              _ptr(counts_t), _ptr(offsets_t), _ptr(plan_t),
                                              plan_t.numel() * plan_t.element_size(), _stream_of(stream)),
                "siSetTableSearchOffsets")

    def fill_device(self, plan_t, nq, offsets_t, idx_t, order="desc", capacity=None, stream=None):
        cap = idx_t.numel() if capacity is None else int(capacity)
        N.check(N.lib.siSetTableSearchFill(self._h, _ptr(plan_t), nq, _ORDERS[order], _ptr(offsets_t), _ptr(idx_t),
                                           cap, _stream_of(stream)), "siSetTableSearchFill")

    def search(self, set_ids, qs, qe, order="desc"):
        """Host-buffer convenience: CSR (offsets uint64, positions uint32) of the batch."""
        import torch
        dev = torch.device("cuda:0")
        nq = int(np.asarray(qs).size)
        sid = torch.from_numpy(np.ascontiguousarray(set_ids, np.uint32).view(np.int32)).to(dev)
        tqs = torch.from_numpy(np.ascontiguousarray(qs, np.int32)).to(dev)
        tqe = torch.from_numpy(np.ascontiguousarray(qe, np.int32)).to(dev)
        plan = torch.empty(max(self.plan_bytes(nq), 256), dtype=torch.uint8, device=dev)
        off = torch.empty(nq + 1, dtype=torch.int64, device=dev)
        self.offsets_device(sid, tqs, tqe, plan, off)
        torch.cuda.synchronize()
        total = int(off[-1].item())
        idx = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
        if total:
            self.fill_device(plan, nq, off, idx, order=order, capacity=total)
        torch.cuda.synchronize()
        return off.cpu().numpy().astype(np.uint64), idx[:total].cpu().numpy().view(np.uint32)


class LearnedIndex:
    """Learned position model attached to a DeviceIndex (si_device.h siLearned)."""

    def __init__(self, index: DeviceIndex, leaves: int = 4096):
        h = N.lib.siLearnedCreate(index.handle, leaves)
        if not h:
            raise SIError(N.SI_ERR_ARG, "siLearnedCreate")
        self._h = h
        self.index = index        # keeps the index alive
        self.leaves = leaves

    def close(self):
        if getattr(self, "_h", None):
            N.lib.siLearnedDestroy(self._h)
            self._h = None

    def __del__(self):
        if N is not None and getattr(N, "lib", None) is not None:   # (cleared at interpreter exit)
            self.close()

    def train(self, steps: int = 400, batch: int = 65536, lr: float = 3e-3, stream=None):
        N.check(N.lib.siLearnedTrain(self._h, steps, batch, lr, _stream_of(stream)), "siLearnedTrain")
        return self.info()

    @staticmethod
    def param_count() -> int:
        return int(N.lib.siLearnedParamCount())

    def params(self) -> np.ndarray:
        p = np.empty(self.param_count(), np.float32)
        N.check(N.lib.siLearnedGetParams(self._h, _ptr(p)), "siLearnedGetParams")
        return p

    def norm(self):
        kmin = C.c_int64(0)
        inv = C.c_float(0)
        N.check(N.lib.siLearnedNorm(self._h, C.byref(kmin), C.byref(inv)), "siLearnedNorm")
        return int(kmin.value), float(inv.value)

    def info(self) -> dict:
        e = C.c_int64(0)
        loss = C.c_float(0)
        used = C.c_uint32(0)
        N.check(N.lib.siLearnedInfo(self._h, C.byref(e), C.byref(loss), C.byref(used)), "siLearnedInfo")
        return {"max_abs_err": int(e.value), "loss": float(loss.value), "used_leaves": int(used.value)}

    def gradient(self, positions, params=None) -> np.ndarray:
        pos = np.ascontiguousarray(positions, np.uint32)
        out = np.empty(self.param_count() + 1, np.float32)
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        N.check(N.lib.siLearnedGradient(self._h, _ptr(p), _ptr(pos), pos.size, _ptr(out)), "siLearnedGradient")
        return out

    def upper_bound_device(self, xs_t, out_t, stats_t=None, stream=None):
        N.check(N.lib.siUpperBoundLearned(self._h, _ptr(xs_t), xs_t.numel(), _ptr(out_t), _ptr(stats_t),
                                          _stream_of(stream)), "siUpperBoundLearned")


class IntervalMap:
    """The reference's Python IntervalMap (intervalmap.pyx:17-819), query path only.

    Intervals are end-inclusive int32 pairs with an arbitrary Python payload.  build()
    sorts and indexes on the GPU; queries run on the GPU.  As in the reference,
    search_idxs / search_idxs_batch return the C++ vector order (leading contiguous run
    ascending, then descending, superintervals.hpp:879-905) and search_values the
    strictly descending order.  The batch calls return numpy arrays (count_batch) and
    lists of numpy arrays instead of Python lists.  intersection and difference follow
    superintervals.hpp:1158-1211 (the C++ set operations; the first set need not be built).
    """

    def __init__(self):
        self._starts: list = []
        self._ends: list = []
        self._vals: list = []
        self._dirty = True
        self._ix = None
        self._s = self._e = None
        self._perm = None
        self._flags = [True, True]

    # ---- storage (intervalmap.pyx:38-240) ----------------------------------
    def __len__(self):
        return self.size()

    def __getitem__(self, index):
        return self.at(index)

    def size(self):
        return len(self._starts) if self._dirty else int(self._s.size)

    def add(self, start: int, end: int, value=None):
        self._materialize_lists()
        ss, es = self._flags
        if ss and self._starts:
            ss = not (start < self._starts[-1])
            if ss and start == self._starts[-1] and end > self._ends[-1]:
                es = False
        self._flags = [ss, es]
        self._starts.append(int(start))
        self._ends.append(int(end))
        self._vals.append(value)
        self._dirty = True

    def reserve(self, n: int):
        pass

    def clear(self):
        self.__init__()

    @classmethod
    def from_arrays(cls, starts, ends, values=None):
        starts = np.asarray(starts, dtype=np.int32)
        ends = np.asarray(ends, dtype=np.int32)
        if starts.shape[0] != ends.shape[0]:
            raise ValueError("starts and ends must have the same length")
        if values is not None and len(values) != starts.shape[0]:
            raise ValueError("values length must match starts/ends length")
        m = cls()
        m._starts = starts.tolist()
        m._ends = ends.tolist()
        m._vals = list(values) if values is not None else [None] * starts.shape[0]
        m._flags = list(add_flags(starts, ends))
        m.build()
        return m

    def _materialize_lists(self):
        if not self._dirty and self._s is not None:
            self._starts = self._s.tolist()
            self._ends = self._e.tolist()
            self._vals = [self._vals_sorted[k] for k in range(self._s.size)]
            self._flags = [True, True]
            self._dirty = True

    def build(self):
        if not self._dirty:
            return
        s = np.asarray(self._starts, dtype=np.int32)
        e = np.asarray(self._ends, dtype=np.int32)
        if self._ix is None:
            self._ix = DeviceIndex(np.int32)
        self._ix.build(s, e, None, self._flags[0], self._flags[1])
        arr = self._ix.download()
        self._s, self._e, self._perm = arr["starts"], arr["ends"], arr["perm"]
        self._vals_sorted = [self._vals[p] for p in self._perm.tolist()]
        self._dirty = False
        self._flags = [True, True]

    def _need_built(self):
        if self._dirty:
            raise RuntimeError("build() must be called after add() and before queries")

    def at(self, index: int):
        if self.size() == 0 or index < 0 or index >= self.size():
            raise IndexError("Index out of range")
        if self._dirty:
            return self._starts[index], self._ends[index], self._vals[index]
        return int(self._s[index]), int(self._e[index]), self._vals_sorted[index]

    def starts_at(self, index: int):
        return self.at(index)[0]

    def ends_at(self, index: int):
        return self.at(index)[1]

    def data_at(self, index: int):
        return self.at(index)[2]

    # ---- single queries (intervalmap.pyx:241-361) ------------------------
    def has_overlaps(self, start: int, end: int) -> bool:
        if self.size() == 0:
            return False
        self._need_built()
        return bool(self._ix.has_overlaps([start], [end])[0])

    def count(self, start: int, end: int) -> int:
        if self.size() == 0:
            return 0
        self._need_built()
        return int(self._ix.count([start], [end])[0])

    def search_idxs(self, start: int, end: int) -> list:
        if self.size() == 0:
            return []
        self._need_built()
        _, idx = self._ix.search([start], [end], order="vector")
        return idx.astype(np.int64).tolist()

    def search_values(self, start: int, end: int) -> list:
        if self.size() == 0:
            return []
        self._need_built()
        _, idx = self._ix.search([start], [end], order="desc")
        return [self._vals_sorted[k] for k in idx.tolist()]

    def search_keys(self, start: int, end: int) -> list:
        if self.size() == 0:
            return []
        self._need_built()
        _, idx = self._ix.search([start], [end], order="desc")
        return [(int(self._s[k]), int(self._e[k])) for k in idx.tolist()]

    def search_items(self, start: int, end: int) -> list:
        if self.size() == 0:
            return []
        self._need_built()
        _, idx = self._ix.search([start], [end], order="desc")
        return [(int(self._s[k]), int(self._e[k]), self._vals_sorted[k]) for k in idx.tolist()]

    def coverage(self, start: int, end: int):
        if self.size() == 0:
            return 0, 0
        self._need_built()
        c, v = self._ix.coverage([start], [end])
        return int(c[0]), int(v[0])

    # ---- batch queries (intervalmap.pyx:363-494) ---------------------------
    def _batch_args(self, starts, ends):
        starts = np.ascontiguousarray(starts, dtype=np.int32)
        ends = np.ascontiguousarray(ends, dtype=np.int32)
        if starts.shape[0] != ends.shape[0]:
            raise ValueError("starts and ends must have the same length")
        return starts, ends

    def count_batch(self, starts, ends) -> np.ndarray:
        starts, ends = self._batch_args(starts, ends)
        if self.size() == 0:
            return np.zeros(starts.shape[0], np.uint32)
        self._need_built()
        return self._ix.count(starts, ends)

    def search_idxs_batch(self, starts, ends) -> list:
        starts, ends = self._batch_args(starts, ends)
        if self.size() == 0:
            return [np.zeros(0, np.uint32) for _ in range(starts.shape[0])]
        self._need_built()
        off, idx = self._ix.search(starts, ends, order="vector")
        return np.split(idx, off[1:-1].astype(np.int64))

    def search_values_batch(self, starts, ends) -> list:
        starts, ends = self._batch_args(starts, ends)
        if self.size() == 0:
            return [[] for _ in range(starts.shape[0])]
        self._need_built()
        off, idx = self._ix.search(starts, ends, order="desc")
        vals = self._vals_sorted
        o = off.tolist()
        il = idx.tolist()
        return [[vals[k] for k in il[o[q]:o[q + 1]]] for q in range(starts.shape[0])]

    # ---- set operations (superintervals.hpp:1158-1211) ------------------------
    def _set_arrays(self):
        """This set's intervals in its current order (as added when not built, else sorted)."""
        if self._dirty:
            return (np.asarray(self._starts, np.int32), np.asarray(self._ends, np.int32), self._vals)
        return self._s, self._e, self._vals_sorted

    def intersection(self, other):
        """Regions covered by both sets (superintervals.hpp:1158-1174): a new, unindexed
        IntervalMap with every overlapping piece of this set's interval k and other's
        interval h, for k in this set's order and h in search_idxs order, carrying k's
        value.  other must be built."""
        other._need_built()
        s, e, vals = self._set_arrays()
        out = IntervalMap()
        if s.size == 0 or other.size() == 0:
            return out
        ps, pe, pk, _ = other._ix.intersection_pieces(s, e)
        for a, b, k in zip(ps.tolist(), pe.tolist(), pk.tolist()):
            if a <= b:   # superintervals.hpp:1167: empty pieces (an interval with start > end) are dropped
                out.add(a, b, vals[k])
        return out

    def difference(self, other):
        """Regions of this set that other does not cover (superintervals.hpp:1188-1211): a
        new, unindexed IntervalMap of the uncovered pieces of each interval, in order, with
        its value.  other must be built."""
        other._need_built()
        s, e, vals = self._set_arrays()
        out = IntervalMap()
        if s.size == 0:
            return out
        if other.size() == 0:   # nothing covered: every interval with start <= end, whole
            keep = np.flatnonzero(s <= e)
            ps, pe, pk = s[keep], e[keep], keep
        else:
            ps, pe, pk = other._ix.difference_pieces(s, e)
        for a, b, k in zip(ps.tolist(), pe.tolist(), pk.tolist()):
            out.add(a, b, vals[k])
        return out
