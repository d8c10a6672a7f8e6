"""Continuous-batching query server over many resident indexes.

No counterpart in the reference.  Many built interval sets stay resident in HBM and a
stream of query requests (each aimed at one of them) is answered with the same results
each request would get alone (search_idxs positions in descending order, or counts).

Design
 * Residency: the interval sets are packed into one device index with 64-bit coordinates,
   set k shifted by k * 2^33.  Shifted sets occupy disjoint coordinate ranges, so no
   overlap and no branch pointer crosses sets: a query of set k finds exactly set k's hits,
   at positions base_k + (its position in set k alone), in the same order.  One kernel
   launch therefore serves queries of any mix of sets (cross-index batching).
 * Requests are admitted in arrival order while their HBM demand fits the budget.  Stage 1
   (queries on the device, plan, offsets) needs ResidentSet.stage1_bytes(n) for n queries
   (about 73 bytes per query plus a fixed 13 KB); stage 2 (fill) needs 4 bytes per
   result, known only after stage 1.  A batch joins all admitted requests waiting for the
   same stage; several batches are in flight on separate streams, and finished batches
   release their memory to the next admissions (continuous batching).
 * Results go back to the host as numpy arrays; latency = completion time - arrival time.
"""
from __future__ import annotations

import collections
import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import DeviceIndex

SHIFT = 1 << 33


@dataclass
class Request:
    rid: int
    index_id: int
    qs: np.ndarray
    qe: np.ndarray
    kind: str = "search"          # "search" (descending positions) or "count"
    arrival: float = 0.0
    done: float = 0.0
    offsets: np.ndarray = None
    idx: np.ndarray = None            # positions in the packed index (int32 view of the result buffer)
    base: int = 0                     # packed position of the target set's first interval
    counts: np.ndarray = None
    event: threading.Event = field(default_factory=threading.Event)

    def positions(self) -> np.ndarray:
        """Result positions in the target set alone (uint32), in descending order per query."""
        return (self.idx.astype(np.int64) - self.base).astype(np.uint32)

    @property
    def nq(self):
        return int(self.qs.size)

    @property
    def latency(self):
        return self.done - self.arrival


class ResidentSet:
    """Interval sets packed into one device index (see the module docstring)."""

    def __init__(self, sets):
        starts, ends, base = [], [], []
        pos = 0
        for k, (s, e) in enumerate(sets):
            s = np.asarray(s, np.int64)
            e = np.asarray(e, np.int64)
            if s.size and (s.min() < -(1 << 31) or e.max() >= (1 << 31)):
                raise ValueError("interval sets must have int32 coordinates")
            starts.append(s + k * SHIFT)
            ends.append(e + k * SHIFT)
            base.append(pos)
            pos += s.size
        self.base = np.array(base + [pos], np.int64)
        self.sizes = np.diff(self.base)
        self.n = pos
        self.index = DeviceIndex(np.int64)
        t0 = time.perf_counter()
        self.index.build(np.concatenate(starts), np.concatenate(ends))
        self.build_s = time.perf_counter() - t0

    def stage1_bytes(self, nq: int) -> int:
        """Device bytes of stage 1 for nq queries: int64 starts and ends, the plan, offsets."""
        return 16 * nq + self.index.plan_bytes(nq) + 8 * (nq + 1)

    def shift(self, index_id, qs, qe):
        off = np.int64(index_id) * SHIFT
        return np.asarray(qs, np.int64) + off, np.asarray(qe, np.int64) + off


class Server:
    """Continuous batching over a ResidentSet.  Call submit() from any thread; run() (or a
    background thread started by start()) drives the GPU."""

    def __init__(self, resident: ResidentSet, hbm_budget_bytes: int = 8 << 30, streams: int = 3,
                 max_batch_queries: int = 1 << 22, keep_results: bool = True):
        import torch
        torch.cuda.init()          # torch's device state before its first stream is created
        self.torch = torch
        self.res = resident
        self.budget = int(hbm_budget_bytes)
        self.in_use = 0
        self.max_batch = int(max_batch_queries)
        self.keep_results = keep_results       # False: record latency only (load generators)
        self.dev = torch.device("cuda:0")
        self.streams = [torch.cuda.Stream() for _ in range(streams)]
        self.pending = collections.deque()      # waiting for stage 1
        self.ready = collections.deque()        # stage 1 done, waiting for stage 2 memory
        self.inflight = []                      # (stage, requests, tensors, event, bytes, stream)
        self.lock = threading.Lock()
        self.next_id = 0
        self.completed = []
        self._stop = False
        self._thread = None
        self.stats = {"batches1": 0, "batches2": 0, "max_in_use": 0, "rejected": 0}

    # ---- client side --------------------------------------------------------
    def submit(self, index_id, qs, qe, kind="search", arrival=None) -> Request:
        if not (0 <= index_id < len(self.res.sizes)):
            raise ValueError("unknown index id %d" % index_id)
        qs = np.ascontiguousarray(qs, np.int32)
        qe = np.ascontiguousarray(qe, np.int32)
        if qs.shape != qe.shape:
            raise ValueError("starts and ends must have the same length")
        with self.lock:
            r = Request(self.next_id, int(index_id), qs, qe, kind,
                        arrival if arrival is not None else time.perf_counter())
            self.next_id += 1
            if self.res.stage1_bytes(r.nq) > self.budget:
                self.stats["rejected"] += 1
                raise MemoryError("request of %d queries exceeds the HBM budget" % r.nq)
            self.pending.append(r)
        return r

    # ---- scheduler ----------------------------------------------------------
    def _free_stream(self):
        busy = {id(b[5]) for b in self.inflight}
        for s in self.streams:
            if id(s) not in busy:
                return s
        return None

    def _launch_stage1(self):
        torch = self.torch
        if self.ready:
            return False                     # batches waiting for fill memory go first
        with self.lock:
            batch, nq, need = [], 0, 0
            while self.pending:
                r = self.pending[0]
                b = self.res.stage1_bytes(r.nq)
                if batch and (nq + r.nq > self.max_batch):
                    break
                if self.in_use + need + b > self.budget:
                    break
                batch.append(self.pending.popleft())
                nq += r.nq
                need += b
        if not batch:
            return False
        st = self._free_stream()
        if st is None:
            with self.lock:
                self.pending.extendleft(reversed(batch))
            return False
        self.in_use += need
        self.stats["max_in_use"] = max(self.stats["max_in_use"], self.in_use)
        ix = self.res.index
        shifted = [self.res.shift(r.index_id, r.qs, r.qe) for r in batch]
        qs = np.concatenate([a for a, _ in shifted]) if nq else np.zeros(0, np.int64)
        qe = np.concatenate([b for _, b in shifted]) if nq else np.zeros(0, np.int64)
        with torch.cuda.stream(st):
            tqs = torch.from_numpy(qs).to(self.dev, non_blocking=False)
            tqe = torch.from_numpy(qe).to(self.dev, non_blocking=False)
            plan = torch.empty(max(ix.plan_bytes(nq), 256), dtype=torch.uint8, device=self.dev)
            off = torch.empty(nq + 1, dtype=torch.int64, device=self.dev)
            if nq:
                ix.plan_device(tqs, tqe, plan, stream=st)
                ix.offsets_planned(plan, nq, off, stream=st)
            else:
                off.zero_()
            off_h = torch.empty(nq + 1, dtype=torch.int64, pin_memory=True)
            off_h.copy_(off, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(st)
        self.stats["batches1"] += 1
        self.inflight.append((1, batch, {"qs": tqs, "qe": tqe, "plan": plan, "off": off, "off_h": off_h},
                              ev, need, st))
        return True

    def _launch_stage2(self):
        torch = self.torch
        if not self.ready:
            return False
        group = self.ready[0]
        batch, t, need1 = group
        total = int(t["off_h"][-1])
        need2 = 4 * total
        # Wait for memory only while a fill batch is in flight: that batch gives memory
        # back when it ends.  Otherwise the holders are this queue's own waiting batches,
        # which cannot release before the head runs, so the head goes over the budget.
        if self.in_use + need2 > self.budget and any(item[0] == 2 for item in self.inflight):
            return False
        st = self._free_stream()
        if st is None:
            return False
        self.ready.popleft()
        self.in_use += need2
        self.stats["max_in_use"] = max(self.stats["max_in_use"], self.in_use)
        nq = t["qs"].numel()
        with torch.cuda.stream(st):
            idx = torch.empty(max(total, 1), dtype=torch.int32, device=self.dev)
            if total:
                self.res.index.fill_planned(t["plan"], nq, t["off"], idx, order="desc", stream=st)
            idx_h = torch.empty(max(total, 1), dtype=torch.int32, pin_memory=True)
            idx_h.copy_(idx, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(st)
        t["idx"] = idx
        t["idx_h"] = idx_h
        self.stats["batches2"] += 1
        self.inflight.append((2, batch, t, ev, need1 + need2, st))
        return True

    def _finish(self, stage, batch, t, need):
        now = time.perf_counter()
        off = t["off_h"].numpy()
        q0 = 0
        if stage == 1:
            needs_fill = [r for r in batch if r.kind == "search"]
            for r in batch:
                o = off[q0:q0 + r.nq + 1]
                q0 += r.nq
                if r.kind == "count":
                    r.counts = np.diff(o).astype(np.uint32)
                    r.done = now
                    r.event.set()
                    self.completed.append(r)
            if needs_fill:
                self.ready.append((batch, t, need))
            else:
                self.in_use -= need
            return
        idx = t["idx_h"].numpy()
        for r in batch:
            o = off[q0:q0 + r.nq + 1]
            q0 += r.nq
            if r.kind != "search":
                continue
            r.base = int(self.res.base[r.index_id])
            if self.keep_results:
                r.offsets = (o - o[0]).astype(np.uint64)
                r.idx = idx[o[0]:o[-1]]
            r.done = now
            r.event.set()
            self.completed.append(r)
        self.in_use -= need

    def poll(self) -> bool:
        """One scheduling round; returns True if anything is still queued or in flight."""
        still = []
        for item in self.inflight:
            stage, batch, t, ev, need, st = item
            if ev.query():
                self._finish(stage, batch, t, need)
            else:
                still.append(item)
        self.inflight = still
        progressed = True
        while progressed:
            progressed = self._launch_stage2() or self._launch_stage1()
        return bool(self.pending or self.ready or self.inflight)

    def run(self, until_idle=True, sleep=0.0002):
        while True:
            busy = self.poll()
            if until_idle and not busy:
                return
            if self._stop and not busy:
                return
            time.sleep(sleep)

    def start(self):
        self._stop = False
        self._thread = threading.Thread(target=self.run, kwargs={"until_idle": False}, daemon=True)
        self._thread.start()

    def stop(self):
        self._stop = True
        if self._thread:
            self._thread.join()
            self._thread = None


def poisson_arrivals(rate_per_s, count, seed):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, count))


def log_uniform_int(rng, lo, hi, size=None):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size)).astype(np.int64)


def serve_stream(server: Server, requests, arrivals, timeout_s=120.0):
    """Submit requests at their arrival offsets (seconds) and serve until all complete.
    Returns (latencies in seconds, wall seconds)."""
    t0 = time.perf_counter()
    k = 0
    deadline = t0 + timeout_s
    while k < len(requests) or server.poll():
        now = time.perf_counter() - t0
        while k < len(requests) and arrivals[k] <= now:
            idx_id, qs, qe, kind = requests[k]
            server.submit(idx_id, qs, qe, kind, arrival=t0 + arrivals[k])
            k += 1
        server.poll()
        if time.perf_counter() > deadline:
            raise TimeoutError("serving stream did not finish")
        time.sleep(0.0001)
    wall = time.perf_counter() - t0
    lat = np.array([r.latency for r in server.completed])
    return lat, wall
