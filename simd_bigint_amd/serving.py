"""Continuous-batching request loop for multiply / square requests.

Requests arrive over time (each names device-resident operands: a slice of
an operand pool that stays in HBM, so nothing is copied per request). The
server admits a waiting request when its HBM need -- result limbs plus the
transform workspace of its size class -- fits in the free budget, and
launches every admitted request that is not yet running as ONE batch
(sba_batch_launch_ws: all size classes of the batch in one launch sequence)
on the first idle stream. Requests that arrive while earlier batches run
are admitted as budget frees up and run concurrently on the other streams:
continuous batching, not a queue in front of one multiply at a time.

Latency of a request = completion time - arrival time, both on the host
clock; completion is observed by polling the batch's stream event.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import SbaJob, SBA_OK, lib


@dataclass
class Request:
    rid: int
    arrival: float                  # seconds from the start of the run
    a: object                       # torch device tensor (view into the pool)
    b: Optional[object]             # None: square of a
    r: object = None                # result tensor (allocated on admission)
    admitted: float = -1.0
    launched: float = -1.0
    done: float = -1.0
    need: int = 0                   # HBM bytes held while in flight

    @property
    def limbs(self) -> int:
        return self.a.numel() + (self.b.numel() if self.b is not None else self.a.numel())


def _ws_bytes(jobs_arr, n) -> int:
    return int(lib().sba_batch_ws_bytes(ctypes.cast(jobs_arr, ctypes.c_void_p), n))


def request_need(req: Request) -> int:
    arr = (SbaJob * 1)()
    arr[0] = SbaJob(None, None, req.a.numel(), None if req.b is None else 1,
                    0 if req.b is None else req.b.numel())
    return _ws_bytes(arr, 1) + 8 * req.limbs


@dataclass
class ServeStats:
    latencies_ms: List[float] = field(default_factory=list)
    batches: int = 0
    max_batch: int = 0
    wall_s: float = 0.0
    peak_hbm_bytes: int = 0

    def summary(self) -> dict:
        lat = np.array(self.latencies_ms)
        return {"requests": len(lat), "p50_ms": round(float(np.percentile(lat, 50)), 4),
                "p99_ms": round(float(np.percentile(lat, 99)), 4),
                "mean_ms": round(float(lat.mean()), 4),
                "muls_per_s": round(len(lat) / self.wall_s, 1),
                "batches": self.batches, "max_batch": self.max_batch,
                "peak_hbm_bytes": int(self.peak_hbm_bytes)}


class Server:
    """HBM-budgeted continuous batching over `nstreams` HIP streams."""

    def __init__(self, budget_bytes: int, nstreams: int = 3, max_batch: int = 256):
        import torch
        self.torch = torch
        self.budget = int(budget_bytes)
        self.max_batch = max_batch
        self.dev = torch.device("cuda:0")
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(nstreams)]
        self.busy = [None] * nstreams           # (event, [requests], workspace)
        self.used = 0

    def _launch(self, k: int, reqs: List[Request], now: float) -> None:
        torch = self.torch
        arr = (SbaJob * len(reqs))()
        for i, q in enumerate(reqs):
            arr[i] = SbaJob(q.r.data_ptr(), q.a.data_ptr(), q.a.numel(),
                            None if q.b is None else q.b.data_ptr(),
                            0 if q.b is None else q.b.numel())
        ws_bytes = _ws_bytes(arr, len(reqs))
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=self.dev)
        st = self.streams[k]
        ws.record_stream(st)
        rc = lib().sba_batch_launch_ws(ctypes.cast(arr, ctypes.c_void_p), len(reqs),
                                       ctypes.c_void_p(ws.data_ptr()), ws_bytes,
                                       ctypes.c_void_p(int(st.cuda_stream)))
        if rc != SBA_OK:
            raise RuntimeError("sba_batch_launch_ws returned %d" % rc)
        ev = torch.cuda.Event()
        ev.record(st)
        for q in reqs:
            q.launched = now
        self.busy[k] = (ev, reqs, ws)

    def run(self, requests: List[Request]) -> ServeStats:
        """Serve `requests` (sorted by arrival) in real time."""
        torch = self.torch
        stats = ServeStats()
        pending = list(requests)
        waiting: List[Request] = []
        admitted: List[Request] = []
        nxt = 0
        t0 = time.perf_counter()
        done = 0
        while done < len(requests):
            now = time.perf_counter() - t0
            while nxt < len(pending) and pending[nxt].arrival <= now:
                q = pending[nxt]
                q.need = request_need(q)
                if q.need > self.budget:
                    raise RuntimeError("request %d needs more than the whole budget" % q.rid)
                waiting.append(q)
                nxt += 1
            # completions free budget
            for k, slot in enumerate(self.busy):
                if slot is not None and slot[0].query():
                    fin = time.perf_counter() - t0
                    for q in slot[1]:
                        q.done = fin
                        stats.latencies_ms.append((fin - q.arrival) * 1e3)
                        self.used -= q.need
                        done += 1
                    self.busy[k] = None
            # admission by HBM budget, FIFO
            while waiting and self.used + waiting[0].need <= self.budget:
                q = waiting.pop(0)
                q.r = torch.empty(q.limbs, dtype=torch.int64, device=self.dev)
                q.admitted = now
                self.used += q.need
                admitted.append(q)
            stats.peak_hbm_bytes = max(stats.peak_hbm_bytes, self.used)
            # continuous batching: every admitted request goes out on an idle stream
            if admitted:
                for k, slot in enumerate(self.busy):
                    if slot is None:
                        batch = admitted[: self.max_batch]
                        admitted = admitted[self.max_batch:]
                        self._launch(k, batch, now)
                        stats.batches += 1
                        stats.max_batch = max(stats.max_batch, len(batch))
                        break
            if nxt >= len(pending) and not waiting and not admitted:
                time.sleep(0)                    # only completions left to observe
        torch.cuda.synchronize()
        stats.wall_s = time.perf_counter() - t0
        return stats


def poisson_requests(n: int, rate_per_s: float, pool, seed: int = 5, lo: int = 256,
                     hi: int = 1 << 20, sqr_frac: float = 0.2) -> List[Request]:
    """n requests with exponential inter-arrival times at `rate_per_s`; 80%
    multiplies / 20% squares; operand sizes log-uniform in [lo, hi] limbs,
    taken as slices of the device-resident random pool (a 1-D tensor of at
    least 2*hi limbs)."""
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
        out.append(Request(i, t, a, b))
    return out
