"""Continuous-batching request loop for forward / adjoint CBS solves.

Requests (a phantom RI, illumination fields, forward or adjoint) arrive over
time.  The scheduler admits a waiting request as soon as a worker slot is free
and its HBM footprint fits in the remaining budget; admitted requests run
concurrently, each on its own solver handle and HIP stream, so their kernels
overlap on the device.  Solver handles - with their potential and 2D
projection kernels resident in HBM - are cached per grid configuration and
reused by later requests of the same shape.  Per-request latency (arrival to
completion), queueing delay and throughput are recorded.

An adjoint request is the forward solve on conj(flip_z(RI)) with the conjugated
fields, flipped and conjugated back, as BACKWARD_SOLVER_MULTI forms it
(Codes/BACKWARD_SOLVER_MULTI.m:48-50).

The admission/scheduling logic is independent of the GPU: `executor` is any
callable (request) -> result, so the policy is testable on the CPU.
"""
from __future__ import annotations

import heapq
import itertools
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Any, Callable, Optional

import numpy as np

GB = 1 << 30


def estimate_footprint(sim_size, ri_size, conv_xy, n_rhs, io=True):
    """Device bytes of one solve: V, three state arrays per rhs, the conv-grid
    work buffer and three kernels of the 2D projections, staging of u_in/u_out
    (mirrors the library's cbs_footprint plus the resident kernels)."""
    nvox = int(np.prod(sim_size))
    nri = int(np.prod(ri_size))
    b = 8 * nvox                                   # V
    b += 3 * 8 * nvox * n_rhs                      # Field, f, spectrum
    b += 2 * 8 * nri * n_rhs                       # u_in / u_out staging
    if io:
        n2 = int(conv_xy[0]) * int(conv_xy[1])
        b += 8 * n2 * ri_size[2] * n_rhs + 2 * 8 * n2 * n_rhs   # conv work + 2D spectra
        b += 3 * 8 * n2 * ri_size[2]                             # resident kernels
        b += 8 * nri * n_rhs                                      # u_in on device
    return b


@dataclass(order=True)
class Request:
    arrival: float
    rid: int = field(compare=False)
    kind: str = field(compare=False, default='forward')       # 'forward' | 'adjoint'
    params: dict = field(compare=False, default_factory=dict)
    RI: Any = field(compare=False, default=None)
    fields: Any = field(compare=False, default=None)          # (x, y, 1, A)
    footprint: int = field(compare=False, default=0)
    # filled by the scheduler
    t_admit: float = field(compare=False, default=float('nan'))
    t_done: float = field(compare=False, default=float('nan'))
    result: Any = field(compare=False, default=None)
    error: Optional[str] = field(compare=False, default=None)
    cancelled: bool = field(compare=False, default=False)

    @property
    def latency(self):
        return self.t_done - self.arrival

    @property
    def queue_delay(self):
        return self.t_admit - self.arrival


class Scheduler:
    """Admission by slots and HBM footprint; continuous (not batch-synchronous)."""

    def __init__(self, executor: Callable[[Request], Any], hbm_budget: int, slots: int = 4,
                 timeout_s: Optional[float] = None, clock=time.perf_counter):
        self.executor = executor
        self.hbm_budget = int(hbm_budget)
        self.slots = int(slots)
        self.timeout_s = timeout_s
        self.clock = clock
        self._lock = threading.Condition()
        self._waiting: list = []
        self._in_use = 0
        self._running = 0
        self._seq = itertools.count()
        self.done: list = []
        self.rejected: list = []

    def submit(self, req: Request):
        """Queue a request; requests larger than the whole budget are rejected."""
        with self._lock:
            if req.footprint > self.hbm_budget:
                req.error = 'rejected: footprint %.2f GB exceeds the HBM budget %.2f GB' % (
                    req.footprint / GB, self.hbm_budget / GB)
                req.t_done = self.clock()
                self.rejected.append(req)
                return False
            heapq.heappush(self._waiting, (req.arrival, next(self._seq), req))
            self._lock.notify_all()
            return True

    def cancel(self, rid):
        """Cancel a request that is still waiting; running requests are not interrupted."""
        with self._lock:
            for _, _, r in self._waiting:
                if r.rid == rid and not r.cancelled:
                    r.cancelled = True
                    r.error = 'cancelled'
                    return True
        return False

    def _pick(self):
        """First waiting request (arrival order) that fits; smaller later ones may pass a
        large head-of-line request only when it cannot fit even after others finish."""
        now = self.clock()
        keep = []
        chosen = None
        while self._waiting:
            item = heapq.heappop(self._waiting)
            r = item[2]
            if r.cancelled:
                r.t_done = now
                self.done.append(r)
                continue
            if self.timeout_s is not None and now - r.arrival > self.timeout_s:
                r.error = 'timed out in queue after %.3f s' % (now - r.arrival)
                r.t_done = now
                self.done.append(r)
                continue
            if chosen is None and self._in_use + r.footprint <= self.hbm_budget:
                chosen = r
                continue
            keep.append(item)
        for it in keep:
            heapq.heappush(self._waiting, it)
        return chosen

    def _run_one(self, r: Request):
        try:
            r.result = self.executor(r)
        except Exception as e:   # noqa: BLE001 - a failing request must not stop the loop
            r.error = '%s: %s' % (type(e).__name__, e)
        finally:
            with self._lock:
                r.t_done = self.clock()
                self._in_use -= r.footprint
                self._running -= 1
                self.done.append(r)
                self._lock.notify_all()

    def run(self, requests, realtime=True):
        """Feed requests at their arrival times (seconds from now when realtime)
        and serve until all have finished.  Returns the finished requests."""
        t0 = self.clock()
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            r.arrival = t0 + r.arrival if realtime else t0
        pending = list(reqs)
        pool = ThreadPoolExecutor(max_workers=self.slots)
        try:
            while True:
                with self._lock:
                    now = self.clock()
                    while pending and pending[0].arrival <= now:
                        self.submit(pending.pop(0))
                    while self._running < self.slots:
                        r = self._pick()
                        if r is None:
                            break
                        r.t_admit = self.clock()
                        self._in_use += r.footprint
                        self._running += 1
                        pool.submit(self._run_one, r)
                    if not pending and not self._waiting and self._running == 0:
                        break
                    wait = 0.05
                    if pending:
                        wait = max(0.0, min(wait, pending[0].arrival - now))
                    self._lock.wait(timeout=wait)
        finally:
            pool.shutdown(wait=True)
        return self.done + self.rejected


def summarize(done, voxel_iters_of):
    ok = [r for r in done if r.error is None]
    lat = np.array([r.latency for r in ok]) * 1e3
    q = np.array([r.queue_delay for r in ok]) * 1e3
    if not ok:
        return {'completed': 0, 'failed': len(done)}
    span = max(r.t_done for r in ok) - min(r.arrival for r in ok)
    vi = float(sum(voxel_iters_of(r) for r in ok))
    return {'completed': len(ok), 'failed': len(done) - len(ok),
            'latency_ms': {'p50': float(np.percentile(lat, 50)), 'p99': float(np.percentile(lat, 99)),
                           'mean': float(lat.mean())},
            'queue_ms': {'p50': float(np.percentile(q, 50)), 'p99': float(np.percentile(q, 99))},
            'span_s': span, 'gvoxel_iter_per_s': vi / span / 1e9}


class SolverCache:
    """GPU executor: one solver per (grid configuration, worker slot), with its 2D
    projection kernels computed once per configuration and kept resident."""

    def __init__(self, device=0, options=None):
        self.device = device
        self.options = options or {}
        self._lock = threading.Lock()
        self._free = {}        # key -> list of idle solvers
        self._kernels = {}     # key -> host kernel arrays (shared by solvers of that key)

    @staticmethod
    def key(p):
        return (tuple(p['size']), tuple(p['RI_xy_size']) if 'RI_xy_size' in p else None,
                tuple(np.round(p['boundary_thickness'], 6)), p['iterations_number'], p['wavelength'], p['NA'],
                p['RI_bg'], tuple(p['resolution']), bool(p['acyclic']))

    def _get(self, p):
        from .forward_solver import FORWARD_SOLVER_CONVERGENT_BORN
        from ._lib import check, lib, CBS_MEM_HOST
        from .forward_solver import _fptr
        k = self.key(p)
        with self._lock:
            lst = self._free.setdefault(k, [])
            if lst:
                return k, lst.pop()
            kern = self._kernels.get(k)
        s = FORWARD_SOLVER_CONVERGENT_BORN(p, device=self.device, options=self.options)
        if kern is None:
            s._ensure_io()
            with self._lock:
                self._kernels[k] = s._io_arrays
        else:
            s._io_arrays = kern
            check(lib.cbs_set_io_kernels(s._h, *[_fptr(a) for a in kern], CBS_MEM_HOST))
            s._io_ready = True
        return k, s

    def prewarm(self, params_list, log=None):
        """Build the solver and its 2D projection kernels for each configuration ahead
        of the request stream (one idle solver per configuration is kept)."""
        for p in params_list:
            t0 = time.perf_counter()
            k, s = self._get(p)
            self._put(k, s)
            if log:
                log('prewarmed %s in %.1f s' % (list(p['size']), time.perf_counter() - t0))

    def _put(self, k, s):
        with self._lock:
            self._free[k].append(s)

    def __call__(self, r: Request):
        k, s = self._get(r.params)
        try:
            if r.kind == 'forward':
                s.set_RI(r.RI)
                return s.solve(r.fields)
            if r.kind == 'adjoint':
                # the adjoint potential conj(flip_z(RI)) (BACKWARD_SOLVER_MULTI.m:48), formed on the device
                s.set_RI(r.RI, flip_conj=True)
                t, rf, u = s.solve(np.conj(r.fields))
                return t, rf, (None if u is None else np.flip(np.conj(u), axis=2))
            raise ValueError('unknown request kind %r' % r.kind)
        finally:
            self._put(k, s)

    def close(self):
        with self._lock:
            for lst in self._free.values():
                for s in lst:
                    s.close()
            self._free.clear()
