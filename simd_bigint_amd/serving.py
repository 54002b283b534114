"""Continuous-batching request loop for multiply / square requests.

Requests arrive over time. Each names device-resident operands (slices of
an operand pool that stays in HBM, so nothing is copied per request) and a
result buffer the caller allocated beforehand. The server owns ONE device
arena of `budget_bytes`, allocated when it is built; every launch takes its
transform workspace (sba_batch_ws_bytes of its jobs) from that arena by
first fit and returns it when the launch's event has fired, so the loop
itself allocates nothing (sba_batch_launch_ws allocates nothing either).

A waiting request is admitted when its share of workspace fits in the free
arena; every admitted request that is not yet running leaves as ONE batch
(all its size classes in one launch sequence) on the first idle stream.
Requests that arrive while earlier batches run are admitted as workspace
frees up and run concurrently on the other streams: continuous batching,
not a queue in front of one multiply at a time.

The loop never spins: with requests still to arrive it sleeps until the next
arrival (or a short quantum while launches are in flight); with only
completions left it blocks on the oldest launch's event (blocking-sync
events, so the host thread waits in the driver).

Latency of a request = completion time - arrival time, both on the host
clock; completion is when the loop observes its launch's event.
"""
from __future__ import annotations

import bisect
import ctypes
import time
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import SbaJob, SBA_OK, lib

_ALIGN = 256
POLL_S = 20e-6                      # sleep quantum while launches are in flight


@dataclass
class Request:
    rid: int
    arrival: float                  # seconds from the start of the run
    a: object                       # torch device tensor (view into the pool)
    b: Optional[object]             # None: square of a
    r: object = None                # result tensor (caller memory, >= limbs)
    admitted: float = -1.0
    launched: float = -1.0
    done: float = -1.0
    need: int = 0                   # workspace bytes while in flight

    @property
    def limbs(self) -> int:
        return self.a.numel() + (self.b.numel() if self.b is not None else self.a.numel())


def _ws_bytes(jobs_arr, n) -> int:
    return int(lib().sba_batch_ws_bytes(ctypes.cast(jobs_arr, ctypes.c_void_p), n))


def request_need(req: Request) -> int:
    """Workspace of the request run alone (its size class's transform
    buffers plus one job-table slot): an upper bound of its share in any
    batch."""
    arr = (SbaJob * 1)()
    arr[0] = SbaJob(None, None, req.a.numel(), None if req.b is None else 1,
                    0 if req.b is None else req.b.numel())
    return _ws_bytes(arr, 1)


class Arena:
    """First-fit allocator over one preallocated device buffer (offsets
    only; the buffer is allocated once by the Server)."""

    def __init__(self, size: int):
        self.size = size
        self.free: List[Tuple[int, int]] = [(0, size)]     # sorted (offset, length)
        self.used = 0

    def alloc(self, n: int) -> Optional[int]:
        n = (n + _ALIGN - 1) // _ALIGN * _ALIGN
        for i, (off, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (off + n, ln - n)
                self.used += n
                return off
        return None

    def release(self, off: int, n: int) -> None:
        n = (n + _ALIGN - 1) // _ALIGN * _ALIGN
        self.used -= n
        i = bisect.bisect_left(self.free, (off, 0))
        self.free.insert(i, (off, n))
        # coalesce with the neighbours
        if i + 1 < len(self.free) and self.free[i][0] + self.free[i][1] == self.free[i + 1][0]:
            self.free[i] = (self.free[i][0], self.free[i][1] + self.free[i + 1][1])
            self.free.pop(i + 1)
        if i > 0 and self.free[i - 1][0] + self.free[i - 1][1] == self.free[i][0]:
            self.free[i - 1] = (self.free[i - 1][0], self.free[i - 1][1] + self.free[i][1])
            self.free.pop(i)


@dataclass
class ServeStats:
    latencies_ms: List[float] = field(default_factory=list)
    batches: int = 0
    max_batch: int = 0
    batch_sizes: List[int] = field(default_factory=list)
    wall_s: float = 0.0
    peak_ws_bytes: int = 0

    def summary(self) -> dict:
        lat = np.array(self.latencies_ms)
        return {"requests": len(lat), "p50_ms": round(float(np.percentile(lat, 50)), 4),
                "p99_ms": round(float(np.percentile(lat, 99)), 4),
                "mean_ms": round(float(lat.mean()), 4),
                "muls_per_s": round(len(lat) / self.wall_s, 1),
                "batches": self.batches, "max_batch": self.max_batch,
                "mean_batch": round(float(np.mean(self.batch_sizes)), 2) if self.batch_sizes
                else 0.0,
                "peak_ws_bytes": int(self.peak_ws_bytes)}


class Server:
    """Continuous batching over `nstreams` HIP streams with all workspace
    in one preallocated arena of `budget_bytes` device bytes."""

    def __init__(self, budget_bytes: int, nstreams: int = 3, max_batch: int = 256):
        import torch
        self.torch = torch
        self.budget = int(budget_bytes)
        self.max_batch = max_batch
        self.dev = torch.device("cuda:0")
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(nstreams)]
        # the only device allocation: the workspace arena and the job tables' host side
        self.arena_buf = torch.empty(self.budget, dtype=torch.uint8, device=self.dev)
        self.arena = Arena(self.budget)
        self.events = [torch.cuda.Event(blocking=True) for _ in range(nstreams)]
        self.jobs = (SbaJob * max_batch)()
        self.busy = [None] * nstreams           # ([requests], arena offset, bytes)

    def _launch(self, k: int, reqs: List[Request], now: float) -> bool:
        arr = self.jobs
        for i, q in enumerate(reqs):
            arr[i] = SbaJob(q.r.data_ptr(), q.a.data_ptr(), q.a.numel(),
                            None if q.b is None else q.b.data_ptr(),
                            0 if q.b is None else q.b.numel())
        ws_bytes = _ws_bytes(arr, len(reqs))
        off = self.arena.alloc(ws_bytes)
        if off is None:                           # fragmented: try again after a completion
            return False
        st = self.streams[k]
        rc = lib().sba_batch_launch_ws(ctypes.cast(arr, ctypes.c_void_p), len(reqs),
                                       ctypes.c_void_p(self.arena_buf.data_ptr() + off),
                                       ws_bytes, ctypes.c_void_p(int(st.cuda_stream)))
        if rc != SBA_OK:
            self.arena.release(off, ws_bytes)
            raise RuntimeError("sba_batch_launch_ws returned %d" % rc)
        self.events[k].record(st)
        for q in reqs:
            q.launched = now
        self.busy[k] = (reqs, off, ws_bytes)
        return True

    def _complete(self, k: int, t0: float, stats: ServeStats) -> int:
        reqs, off, n = self.busy[k]
        fin = time.perf_counter() - t0
        for q in reqs:
            q.done = fin
            stats.latencies_ms.append((fin - q.arrival) * 1e3)
        self.arena.release(off, n)
        self.busy[k] = None
        return len(reqs)

    def run(self, requests: List[Request]) -> ServeStats:
        """Serve `requests` (sorted by arrival, result buffers allocated) in
        real time."""
        torch = self.torch
        stats = ServeStats()
        for q in requests:
            if q.r is None or q.r.numel() < q.limbs or not q.r.is_contiguous():
                raise ValueError("request %d needs a contiguous result buffer of %d limbs"
                                 % (q.rid, q.limbs))
        waiting: List[Request] = []
        admitted: List[Request] = []
        reserved = 0                              # workspace promised to admitted requests
        nxt = 0
        done = 0
        order: List[int] = []                     # launch order of busy streams (oldest first)
        t0 = time.perf_counter()
        while done < len(requests):
            now = time.perf_counter() - t0
            while nxt < len(requests) and requests[nxt].arrival <= now:
                q = requests[nxt]
                q.need = request_need(q)
                if q.need > self.budget:
                    raise RuntimeError("request %d needs more than the whole arena" % q.rid)
                waiting.append(q)
                nxt += 1
            # completions free workspace
            for k in list(order):
                if self.events[k].query():
                    done += self._complete(k, t0, stats)
                    order.remove(k)
            inflight = sum(s[2] for s in self.busy if s is not None)
            # admission by workspace, FIFO
            while waiting and inflight + reserved + waiting[0].need <= self.budget:
                q = waiting.pop(0)
                q.admitted = now
                reserved += q.need
                admitted.append(q)
            # continuous batching: every admitted request goes out on an idle stream
            launched = False
            if admitted:
                for k, slot in enumerate(self.busy):
                    if slot is None:
                        batch = admitted[: self.max_batch]
                        if self._launch(k, batch, now):
                            admitted = admitted[self.max_batch:]
                            reserved -= sum(q.need for q in batch)
                            order.append(k)
                            stats.batches += 1
                            stats.max_batch = max(stats.max_batch, len(batch))
                            stats.batch_sizes.append(len(batch))
                            launched = True
                        break
            stats.peak_ws_bytes = max(stats.peak_ws_bytes, self.arena.used)
            if launched:
                continue
            # nothing to launch now: wait without spinning
            if nxt < len(requests):
                dt = requests[nxt].arrival - (time.perf_counter() - t0)
                if order:
                    dt = min(dt, POLL_S)
                if dt > 0:
                    time.sleep(dt)
            elif order:
                self.events[order[0]].synchronize()    # only completions left: block
        torch.cuda.synchronize()
        stats.wall_s = time.perf_counter() - t0
        return stats


def poisson_requests(n: int, rate_per_s: float, pool, seed: int = 5, lo: int = 256,
                     hi: int = 1 << 20, sqr_frac: float = 0.2,
                     results: bool = True) -> List[Request]:
    """n requests with exponential inter-arrival times at `rate_per_s`; 80%
    multiplies / 20% squares; operand sizes log-uniform in [lo, hi] limbs,
    taken as slices of the device-resident random pool (a 1-D tensor of at
    least 2*hi limbs). results: allocate each request's result buffer now
    (caller memory, before any serving starts)."""
    import torch
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    span = np.log2(hi / lo)
    P = pool.numel()
    for i in range(n):
        t += rng.exponential(1.0 / rate_per_s)
        an = int(lo * 2 ** (rng.random() * span))
        oa = int(rng.integers(0, P - an + 1))
        a = pool[oa:oa + an]
        b = None
        if rng.random() >= sqr_frac:
            bn = int(lo * 2 ** (rng.random() * span))
            ob = int(rng.integers(0, P - bn + 1))
            b = pool[ob:ob + bn]
        q = Request(i, t, a, b)
        if results:
            q.r = torch.empty(q.limbs, dtype=torch.int64, device=pool.device)
        out.append(q)
    return out
