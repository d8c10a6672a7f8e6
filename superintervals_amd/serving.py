"""Continuous-batching query server over many resident indexes.

No counterpart in the reference (its only multi-index structure is one IntervalMap per
sequence name, examples/bed-intersect-si.rs:100-123).  Many built interval sets stay
resident in HBM and a stream of query requests, each aimed at one of them, is answered with
exactly the results each request would get alone (search_idxs positions in descending
order, or counts).

Design
 * Residency: every set is its own int32 index (DeviceIndex) in a slot of a SetTable
   (si_device.h siSetTable).  Sets are added or evicted at any time without touching the
   others; a batch names each query's slot, so one plan and one fill launch serve queries of
   any mix of sets (cross-index batching), and positions are each set's own.
 * Shapes: a batch of n queries is padded to a bucket of B = 2^k >= n queries (the padding
   queries name the table's empty slot and find nothing).  Each bucket has cached device
   buffers and two captured graphs replayed for every batch of that shape: stage 1 (plan +
   offsets) on the B queries, stage 2 (fill) into an output of C = 2^j >= the batch's
   total positions (graph per capacity).  Stage 2 is sized from stage 1's total, read back.
 * Uploads and results go through pinned host buffers with asynchronous copies on the
   batch's stream; several batches are in flight on separate streams.
 * Admission by HBM: a batch is admitted while the device memory of the batches in flight
   plus the cached buffers of idle shapes fits the budget; idle cached shapes are dropped
   first when it does not.  A fill that cannot get memory waits only while another batch is
   in flight (the only holder that frees memory), otherwise it runs over the budget, so the
   queue cannot deadlock.  A request whose smallest batch exceeds the budget is rejected.
 * Latency = completion time (results in host memory) - arrival time.
"""
from __future__ import annotations

import collections
import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import DeviceIndex, SetTable

MIN_BUCKET = 1024          # smallest padded batch (queries)
MIN_CAPACITY = 1 << 14     # smallest fill output (positions)


def bucket(n: int, lo: int) -> int:
    """Smallest power of two >= max(n, lo)."""
    return max(lo, 1 << (max(int(n), 1) - 1).bit_length())


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
    idx: np.ndarray = None        # positions in the target set (uint32)
    counts: np.ndarray = None
    event: threading.Event = field(default_factory=threading.Event)

    def positions(self) -> np.ndarray:
        """Result positions in the target set, descending per query."""
        return self.idx

    @property
    def nq(self):
        return int(self.qs.size)

    @property
    def latency(self):
        return self.done - self.arrival


class ResidentSets:
    """Interval sets resident in HBM, one int32 index per slot of a set table."""

    def __init__(self, sets=(), max_sets: int = 64):
        self.table = SetTable(max_sets)
        self.max_sets = max_sets
        self.sizes = {}
        self.build_s = 0.0
        for k, (s, e) in enumerate(sets):
            self.add(k, s, e)

    def add(self, slot: int, starts, ends):
        """Build a set and put it into slot (replacing what was there)."""
        s = np.asarray(starts)
        e = np.asarray(ends)
        if s.size and (s.min() < -(1 << 31) or e.max() >= (1 << 31)):
            raise ValueError("interval sets must have int32 coordinates")
        t0 = time.perf_counter()
        ix = DeviceIndex(np.int32).build(s.astype(np.int32), e.astype(np.int32))
        self.build_s += time.perf_counter() - t0
        old = self.table.slots[slot]
        self.table.put(slot, ix)
        self.sizes[slot] = int(s.size)
        if old is not None and old not in self.table.slots:
            old.close()

    def evict(self, slot: int):
        old = self.table.slots[slot]
        self.table.put(slot, None)
        self.sizes.pop(slot, None)
        if old is not None and old not in self.table.slots:
            old.close()

    def index(self, slot: int) -> DeviceIndex:
        return self.table.slots[slot]

    @property
    def n(self) -> int:
        return int(sum(self.sizes.values()))

    def stage1_bytes(self, nq: int) -> int:
        """Device bytes of the stage-1 buffers of the bucket of nq queries: slots, starts,
        ends, the plan, offsets."""
        b = bucket(nq, MIN_BUCKET)
        return 12 * b + self.table.plan_bytes(b) + 8 * (b + 1)


class _Shape:
    """Cached buffers and captured graphs of one bucket of B queries."""

    def __init__(self, torch, res: ResidentSets, B: int, stream):
        self.torch = torch
        self.B = B
        dev = torch.device("cuda:0")
        self.sid = torch.full((B,), res.max_sets, dtype=torch.int32, device=dev)
        self.qs = torch.zeros(B, dtype=torch.int32, device=dev)
        self.qe = torch.zeros(B, dtype=torch.int32, device=dev)
        self.plan = torch.empty(res.table.plan_bytes(B), dtype=torch.uint8, device=dev)
        self.off = torch.zeros(B + 1, dtype=torch.int64, device=dev)
        self.h_sid = torch.full((B,), res.max_sets, dtype=torch.int32, pin_memory=True)
        self.h_qs = torch.zeros(B, dtype=torch.int32, pin_memory=True)
        self.h_qe = torch.zeros(B, dtype=torch.int32, pin_memory=True)
        self.h_off = torch.zeros(B + 1, dtype=torch.int64, pin_memory=True)
        self.bytes = res.stage1_bytes(B)
        self.res = res
        self.fills = {}                      # capacity -> (idx, graph)
        self.fill_bytes = 0
        with torch.cuda.stream(stream):      # one eager run first (library state), then capture
            res.table.offsets_device(self.sid, self.qs, self.qe, self.plan, self.off, stream=stream)
        stream.synchronize()
        self.g1 = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.g1, stream=stream):
            res.table.offsets_device(self.sid, self.qs, self.qe, self.plan, self.off, stream=stream)

    def fill(self, C: int, stream):
        if C not in self.fills:
            torch = self.torch
            idx = torch.empty(C, dtype=torch.int32, device="cuda")
            with torch.cuda.stream(stream):   # one eager run first (library state), then capture
                self.res.table.fill_device(self.plan, self.B, self.off, idx, order="desc", capacity=C, stream=stream)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=stream):
                self.res.table.fill_device(self.plan, self.B, self.off, idx, order="desc", capacity=C, stream=stream)
            self.fills[C] = (idx, g)
            self.fill_bytes += 4 * C
        return self.fills[C]


class Server:
    """Continuous batching over ResidentSets.  Call submit() from any thread; run() (or a
    background thread started by start()) drives the GPU."""

    def __init__(self, resident: ResidentSets, hbm_budget_bytes: int = 8 << 30, streams: int = 3,
                 max_batch_queries: int = 1 << 22, keep_results: bool = True):
        import torch
        torch.cuda.init()          # torch's device state before its first stream is created
        self.torch = torch
        self.res = resident
        self.budget = int(hbm_budget_bytes)
        self.max_batch = int(max_batch_queries)
        self.keep_results = keep_results       # False: record latency only (load generators)
        self.streams = [torch.cuda.Stream() for _ in range(streams)]
        # one pinned result buffer per stream, grown on demand: a batch's results are copied
        # out of it before its stream takes the next batch
        self.h_res = {id(s): torch.empty(0, dtype=torch.int32) for s in self.streams}
        self.pending = collections.deque()      # waiting for stage 1
        self.ready = collections.deque()        # stage 1 done, waiting for stage 2
        self.inflight = []                      # (stage, requests, shape, nq, event, stream, result buffer)
        self.idle = collections.OrderedDict()   # id -> idle cached shape (least recently used first)
        self.cached_bytes = 0                   # device bytes of every cached shape (idle or busy)
        self.lock = threading.Lock()
        self.next_id = 0
        self.completed = []
        self._stop = False
        self._thread = None
        self.stats = {"batches1": 0, "batches2": 0, "max_in_use": 0, "rejected": 0, "shapes": 0, "graphs": 0}

    @property
    def in_use(self) -> int:
        """Device bytes held by batches in flight or waiting for their fill."""
        busy = [it[2] for it in self.inflight] + [r[1] for r in self.ready]
        return int(sum(s.bytes + s.fill_bytes for s in busy))

    # ---- client side --------------------------------------------------------
    def submit(self, index_id, qs, qe, kind="search", arrival=None) -> Request:
        if not (0 <= index_id < self.res.max_sets) or self.res.index(index_id) is None:
            raise ValueError("no set in slot %d" % index_id)
        qs = np.ascontiguousarray(qs, np.int32)
        qe = np.ascontiguousarray(qe, np.int32)
        if qs.shape != qe.shape:
            raise ValueError("starts and ends must have the same length")
        with self.lock:
            r = Request(self.next_id, int(index_id), qs, qe, kind,
                        arrival if arrival is not None else time.perf_counter())
            self.next_id += 1
            if self.res.stage1_bytes(r.nq) > self.budget or r.nq > self.max_batch:
                self.stats["rejected"] += 1
                raise MemoryError("request of %d queries exceeds the HBM budget" % r.nq)
            self.pending.append(r)
        return r

    def add_set(self, slot, starts, ends):
        """Add or replace a resident set; in-flight batches finish first."""
        self.drain()
        self.res.add(slot, starts, ends)

    def evict_set(self, slot):
        self.drain()
        self.res.evict(slot)

    def drain(self):
        while self.inflight or self.ready:
            self.poll(admit=False)
            time.sleep(0.0002)

    # ---- memory ---------------------------------------------------------------
    def _reserve(self, need: int) -> bool:
        """Room for need more device bytes, dropping idle cached shapes if necessary."""
        while self.cached_bytes + need > self.budget and self.idle:
            _, sh = self.idle.popitem(last=False)
            self.cached_bytes -= sh.bytes + sh.fill_bytes
            del sh
        return self.cached_bytes + need <= self.budget

    def _shape(self, B: int, stream):
        for key, sh in self.idle.items():
            if sh.B == B:
                del self.idle[key]
                return sh
        if not self._reserve(self.res.stage1_bytes(B)) and self.inflight:
            return None
        sh = _Shape(self.torch, self.res, B, stream)
        self.cached_bytes += sh.bytes
        self.stats["shapes"] += 1
        self.stats["graphs"] += 1
        return sh

    # ---- scheduler ----------------------------------------------------------
    def _free_stream(self):
        busy = {id(it[5]) for it in self.inflight}
        for s in self.streams:
            if id(s) not in busy:
                return s
        return None

    def _launch_stage1(self):
        torch = self.torch
        if self.ready:
            return False                     # batches waiting for fill memory go first
        st = self._free_stream()
        if st is None:
            return False
        # join waiting requests while the padded batch's stage-1 memory fits what the
        # batches in flight leave of the budget (the first alone may wait for them)
        avail = self.budget - self.in_use
        with self.lock:
            batch, nq = [], 0
            while self.pending:
                m = nq + self.pending[0].nq
                if m > self.max_batch or self.res.stage1_bytes(m) > avail:
                    if batch or self.inflight:
                        break
                batch.append(self.pending.popleft())
                nq = m
        if not batch:
            return False
        B = bucket(nq, MIN_BUCKET)
        sh = self._shape(B, st)
        if sh is None:                       # no memory yet: wait for a batch in flight
            with self.lock:
                self.pending.extendleft(reversed(batch))
            return False
        # pinned staging: slots, starts, ends; the padding names the empty slot
        h_sid, h_qs, h_qe = sh.h_sid.numpy(), sh.h_qs.numpy(), sh.h_qe.numpy()
        q0 = 0
        for r in batch:
            h_sid[q0:q0 + r.nq] = r.index_id
            h_qs[q0:q0 + r.nq] = r.qs
            h_qe[q0:q0 + r.nq] = r.qe
            q0 += r.nq
        h_sid[nq:] = self.res.max_sets
        with torch.cuda.stream(st):
            sh.sid.copy_(sh.h_sid, non_blocking=True)
            sh.qs.copy_(sh.h_qs, non_blocking=True)
            sh.qe.copy_(sh.h_qe, non_blocking=True)
            sh.g1.replay()
            sh.h_off.copy_(sh.off, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(st)
        self.stats["batches1"] += 1
        self.inflight.append((1, batch, sh, nq, ev, st, None))
        self.stats["max_in_use"] = max(self.stats["max_in_use"], self.in_use)
        return True

    def _launch_stage2(self):
        torch = self.torch
        if not self.ready:
            return False
        batch, sh, nq = self.ready[0]
        total = int(sh.h_off[nq])
        C = bucket(total, MIN_CAPACITY)
        st = self._free_stream()
        if st is None:
            return False
        if C not in sh.fills:
            if not self._reserve(4 * C) and self.inflight:
                return False                 # wait for memory another batch will free
            self.cached_bytes += 4 * C
            self.stats["graphs"] += 1
        idx, g2 = sh.fill(C, st)
        self.ready.popleft()
        h_idx = self.h_res[id(st)]
        if h_idx.numel() < total:
            h_idx = self.h_res[id(st)] = torch.empty(max(total, 2 * h_idx.numel()), dtype=torch.int32,
                                                     pin_memory=True)
        with torch.cuda.stream(st):
            g2.replay()
            if total:
                h_idx[:total].copy_(idx[:total], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(st)
        self.stats["batches2"] += 1
        self.inflight.append((2, batch, sh, nq, ev, st, h_idx))
        self.stats["max_in_use"] = max(self.stats["max_in_use"], self.in_use)
        return True

    def _release(self, sh):
        self.idle[id(sh)] = sh

    def _finish(self, stage, batch, sh, nq, h_idx):
        now = time.perf_counter()
        off = sh.h_off.numpy()
        if stage == 1:
            q0 = 0
            for r in batch:
                if r.kind == "count":
                    r.counts = np.diff(off[q0:q0 + r.nq + 1]).astype(np.uint32)
                    r.done = now
                    r.event.set()
                    self.completed.append(r)
                q0 += r.nq
            if any(r.kind == "search" for r in batch):
                self.ready.append((batch, sh, nq))
            else:
                self._release(sh)
            return
        idx = h_idx.numpy()
        q0 = 0
        for r in batch:
            o = off[q0:q0 + r.nq + 1]
            q0 += r.nq
            if r.kind != "search":
                continue
            if self.keep_results:
                r.offsets = (o - o[0]).astype(np.uint64)
                r.idx = idx[o[0]:o[-1]].view(np.uint32).copy()
            r.done = now
            r.event.set()
            self.completed.append(r)
        self._release(sh)

    def poll(self, admit: bool = True) -> bool:
        """One scheduling round; returns True if anything is still queued or in flight."""
        still = []
        for item in self.inflight:
            stage, batch, sh, nq, ev, st, h_idx = item
            if ev.query():
                self._finish(stage, batch, sh, nq, h_idx)
            else:
                still.append(item)
        self.inflight = still
        progressed = True
        while progressed:
            progressed = self._launch_stage2() or (admit and self._launch_stage1())
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
