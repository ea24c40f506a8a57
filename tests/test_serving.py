"""Request loop: admission by slots and HBM footprint, continuous admission,
rejection, cancellation, queue time-out and failure isolation (CPU, fake
executor); on the GPU, each result in a concurrent mixed stream equals the same
request solved alone."""
import threading
import time

import numpy as np
import pytest

from inverse_solver_amd.serving import Scheduler, Request, estimate_footprint, summarize, GB


class FakeExec:
    def __init__(self, dur=0.05, fail_ids=()):
        self.dur = dur
        self.fail_ids = set(fail_ids)
        self.lock = threading.Lock()
        self.running = 0
        self.max_running = 0
        self.mem = 0
        self.max_mem = 0
        self.order = []

    def __call__(self, r):
        with self.lock:
            self.running += 1
            self.mem += r.footprint
            self.max_running = max(self.max_running, self.running)
            self.max_mem = max(self.max_mem, self.mem)
            self.order.append(r.rid)
        try:
            time.sleep(self.dur)
            if r.rid in self.fail_ids:
                raise RuntimeError('injected failure')
            return r.rid
        finally:
            with self.lock:
                self.running -= 1
                self.mem -= r.footprint


def _reqs(n, fp=1 * GB, gap=0.0):
    return [Request(arrival=i * gap, rid=i, footprint=fp) for i in range(n)]


def test_slots_and_budget_are_respected():
    ex = FakeExec()
    s = Scheduler(ex, hbm_budget=3 * GB, slots=4)
    done = s.run(_reqs(10, fp=1 * GB), realtime=False)
    assert len(done) == 10 and all(r.error is None for r in done)
    assert ex.max_running <= 3 and ex.max_mem <= 3 * GB          # memory binds before slots
    s2 = Scheduler(FakeExec(), hbm_budget=100 * GB, slots=2)
    ex2 = s2.executor
    s2.run(_reqs(6), realtime=False)
    assert ex2.max_running == 2


def test_oversize_request_rejected_and_others_served():
    ex = FakeExec(dur=0.01)
    s = Scheduler(ex, hbm_budget=4 * GB, slots=2)
    reqs = _reqs(3)
    reqs.append(Request(arrival=0.0, rid=99, footprint=5 * GB))
    done = s.run(reqs, realtime=False)
    rej = [r for r in done if r.rid == 99][0]
    assert rej.error.startswith('rejected')
    assert sorted(r.rid for r in done if r.error is None) == [0, 1, 2]


def test_continuous_admission_of_late_arrivals():
    """A request arriving while others run starts as soon as a slot frees,
    not after the whole first group has finished."""
    ex = FakeExec(dur=0.2)
    s = Scheduler(ex, hbm_budget=100 * GB, slots=2)
    reqs = [Request(arrival=0.0, rid=0, footprint=GB), Request(arrival=0.0, rid=1, footprint=GB),
            Request(arrival=0.05, rid=2, footprint=GB)]
    reqs[1].params = {}
    ex.dur = 0.2
    done = {r.rid: r for r in s.run(reqs, realtime=True)}
    # rid 2 waits for the first free slot (~0.2 s after start), not for both
    assert done[2].t_admit - done[0].arrival < 0.35
    assert done[2].t_admit >= min(done[0].t_done, done[1].t_done) - 1e-3


def test_cancel_and_timeout_and_failure_isolation():
    ex = FakeExec(dur=0.1, fail_ids={1})
    s = Scheduler(ex, hbm_budget=100 * GB, slots=1, timeout_s=0.25)
    reqs = _reqs(5)
    th = threading.Timer(0.02, lambda: s.cancel(4))
    th.start()
    done = {r.rid: r for r in s.run(reqs, realtime=False)}
    th.join()
    assert done[0].error is None
    assert 'injected failure' in done[1].error
    assert done[4].error == 'cancelled'
    assert any(done[i].error and 'timed out' in done[i].error for i in (2, 3))


def test_footprint_estimate_scales():
    a = estimate_footprint([160] * 3, [128] * 3, [256, 256], 1)
    b = estimate_footprint([160] * 3, [128] * 3, [256, 256], 49)
    assert b > 40 * (a - 3 * 8 * 256 * 256 * 128 - 8 * 160 ** 3) * 0.5
    assert estimate_footprint([576, 576, 320], [512, 512, 256], [1024, 1024], 30) < 288 * GB


def test_summary():
    ex = FakeExec(dur=0.01)
    s = Scheduler(ex, hbm_budget=100 * GB, slots=2)
    done = s.run(_reqs(4), realtime=False)
    out = summarize(done, lambda r: 1e9)
    assert out['completed'] == 4 and out['latency_ms']['p99'] >= out['latency_ms']['p50']


@pytest.mark.gpu
def test_gpu_mixed_stream_equals_alone():
    from helpers import params, sphere_ri
    from inverse_solver_amd.serving import SolverCache
    from inverse_solver_amd.phantoms import circle_illumination
    rng = np.random.default_rng(0)
    cfgs = [(16, 16, 17), (24, 24, 17)]
    reqs = []
    for i in range(6):
        ri = cfgs[i % 2]
        p = params(ri, iterations=10)
        RI = sphere_ri(ri, 0.3 + 0.05 * (i % 3), n_s=1.36 + 0.01 * i)
        A = 1 + i % 3
        u = circle_illumination(ri[:2], A, 1.2, 0.532, [0.1] * 3)
        reqs.append(Request(arrival=0.01 * i, rid=i, kind='adjoint' if i % 2 else 'forward', params=p, RI=RI,
                            fields=u, footprint=estimate_footprint([n + 8 for n in ri], ri, [2 * ri[0], 2 * ri[1]],
                                                                   A)))
    cache = SolverCache()
    s = Scheduler(cache, hbm_budget=8 * GB, slots=3)
    done = {r.rid: r for r in s.run(reqs, realtime=True)}
    alone = SolverCache()
    for i, r in enumerate(reqs):
        assert done[i].error is None, done[i].error
        ref = alone(r)
        for a, b in zip(done[i].result, ref):
            if b is not None:
                np.testing.assert_allclose(a, b, rtol=0, atol=1e-6 * np.abs(b).max())
    cache.close()
    alone.close()
