"""The serving loop's scheduling (horses3d_amd/serve.py) with stand-in engines and a stand-in
clock: every request completes with the result it has alone, no case is admitted before it
arrives, the number of running cases and the memory budget bound admission, and the latency
figures are those of the recorded times.  Host only; the GPU form is in test_engine_gpu.py."""
import time

import numpy as np
import pytest

from horses3d_amd import serve


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1e-4   # every look at the clock costs 0.1 ms
        return self.t

    def sleep(self, s):
        self.t += s


class FakeEngine:
    """Work of n steps takes n * 1 ms of the clock; the state counts the steps taken."""
    live = 0
    most = 0

    def __init__(self, r, clock):
        self.clock, self.busy_until, self.steps, self.r = clock, 0.0, 0, r
        FakeEngine.live += 1
        FakeEngine.most = max(FakeEngine.most, FakeEngine.live)

    def set_state(self, Q):
        self.Q = np.array(Q, dtype=float)

    def device_bytes(self):
        return 1000 * self.r.nodes

    def done(self):
        return self.clock.t >= self.busy_until

    def advance_async(self, n, cfl, dcfl):
        assert self.done()
        self.busy_until = self.clock.t + 1e-3 * n
        self.steps += n
        self.Q = self.Q + n * cfl

    def advance(self, n, cfl, dcfl, t=0.0):
        self.advance_async(n, cfl, dcfl)
        self.clock.t = self.busy_until
        return 0.1 * self.steps

    def time(self):
        return 0.1 * self.steps

    def get_state(self):
        return self.Q.copy()

    def close(self):
        FakeEngine.live -= 1


def requests(n=12, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nodes = int(rng.integers(100, 1000))
        case = {"off": np.array([0, nodes])}
        out.append(serve.Request(case, {}, np.full((nodes, 5), float(k)), steps=int(rng.integers(5, 60)),
                                 cfl=0.5, arrival=float(rng.exponential(0.01) * k)))
    return out


def run(reqs, **kw):
    clock = Clock()
    FakeEngine.live = FakeEngine.most = 0
    return serve.serve(reqs, engine_factory=lambda r: FakeEngine(r, clock), clock=clock, sleep=clock.sleep,
                       poll_s=0.0, workers=0, **kw)


def test_every_request_completes_as_alone():
    reqs = requests()
    res, st = run(reqs, chunk=7, max_active=4, device_memory=lambda: 1 << 40)
    clock = Clock()
    alone, _ = serve.one_at_a_time(reqs, engine_factory=lambda r: FakeEngine(r, clock), clock=clock,
                                   sleep=clock.sleep)
    assert st["cases"] == len(reqs) and FakeEngine.live == 0
    for r, a, q in zip(res, alone, reqs):
        assert np.array_equal(r.Q, a.Q) and r.t == a.t
        assert np.array_equal(r.Q, q.Q0 + q.steps * 0.5)
        assert r.admitted >= q.arrival and r.finished >= r.admitted
    lat = np.array([r.latency for r in res])
    assert st["p50_latency_s"] == np.percentile(lat, 50) and st["p99_latency_s"] == np.percentile(lat, 99)
    assert st["dof_updates"] == sum(q.nodes * q.steps for q in reqs)


def test_admission_bounds():
    reqs = requests(16, seed=1)
    for q in reqs:
        q.arrival = 0.0
    run(reqs, chunk=5, max_active=3, device_memory=lambda: 1 << 40)
    assert FakeEngine.most == 3
    # a memory budget of 2 MB beside a 100-byte reserve: the live cases never hold more
    free = {"b": 2 * 1000 * 1000 + 100, "low": 1 << 60}
    clock = Clock()

    def factory(r):
        e = FakeEngine(r, clock)
        free["b"] -= 1000 * r.nodes
        free["low"] = min(free["low"], free["b"])
        close = e.close

        def closing():
            free["b"] += 1000 * r.nodes
            close()
        e.close = closing
        return e

    FakeEngine.live = FakeEngine.most = 0
    res, st = serve.serve(reqs, engine_factory=factory, clock=clock, sleep=clock.sleep, poll_s=0.0, chunk=5,
                          max_active=16, device_memory=lambda: free["b"], hbm_reserve=100, workers=0)
    assert st["cases"] == len(reqs) and st["max_concurrent"] == FakeEngine.most > 1
    assert free["low"] >= 100 and free["b"] == 2 * 1000 * 1000 + 100


class WallClock:
    """Stand-in engine clock on wall time (work of n steps takes n ms)."""

    @property
    def t(self):
        return time.perf_counter()


def test_threaded_setup_and_retirement():
    """With set-up and retirement on worker threads (real clock), results are unchanged."""
    reqs = requests(10, seed=3)
    for q in reqs:
        q.arrival *= 0.1
    wall = WallClock()
    res, st = serve.serve(reqs, engine_factory=lambda r: FakeEngine(r, wall), poll_s=1e-4, chunk=9,
                          max_active=4, device_memory=lambda: 1 << 40, workers=2)
    assert st["cases"] == len(reqs) and FakeEngine.live == 0
    for r, q in zip(res, reqs):
        assert np.array_equal(r.Q, q.Q0 + q.steps * 0.5)
        assert r.admitted >= q.arrival


def test_failed_setup_retires_that_case_alone():
    """A case whose set-up fails is retired alone with its error; the other cases complete with
    the results they have alone, and every engine is closed."""
    reqs = requests(6, seed=5)
    for q in reqs:
        q.arrival = 0.0
    clock = Clock()
    made = []

    def factory(r):
        if r is reqs[3]:
            raise RuntimeError("no memory for this case")
        e = FakeEngine(r, clock)
        made.append(e)
        return e

    FakeEngine.live = 0
    res, st = serve.serve(reqs, engine_factory=factory, clock=clock, sleep=clock.sleep, poll_s=0.0, chunk=5,
                          max_active=16, device_memory=lambda: 1 << 40, workers=0)
    assert made and FakeEngine.live == 0
    assert res[3].error and "no memory" in res[3].error and res[3].Q is None
    assert st["cases"] == 5 and st["failed"] == 1
    for k, (r, q) in enumerate(zip(res, reqs)):
        if k != 3:
            assert r.error is None and np.array_equal(r.Q, q.Q0 + q.steps * 0.5)


class DivergingEngine(FakeEngine):
    """Diverges (raises the engine's NaN error on polling) once it has taken 7 steps."""

    def done(self):
        d = super().done()
        if d and self.steps >= 7:
            from horses3d_amd.engine import HX3DDiverged
            raise HX3DDiverged("numerical divergence", 5)
        return d


def test_diverging_case_retired_alone():
    """A case whose steps diverge stops there with its error; the others are unaffected."""
    reqs = requests(8, seed=9)
    for q in reqs:
        q.steps = max(q.steps, 12)
    clock = Clock()
    FakeEngine.live = 0
    factory = lambda r: (DivergingEngine if r is reqs[2] else FakeEngine)(r, clock)
    res, st = serve.serve(reqs, engine_factory=factory, clock=clock, sleep=clock.sleep, poll_s=0.0, chunk=5,
                          max_active=16, device_memory=lambda: 1 << 40, workers=0)
    assert FakeEngine.live == 0
    assert res[2].error and "divergence" in res[2].error and res[2].Q is not None
    assert st["failed"] == 1 and st["cases"] == len(reqs) - 1
    for k, (r, q) in enumerate(zip(res, reqs)):
        if k != 2:
            assert r.error is None and np.array_equal(r.Q, q.Q0 + q.steps * 0.5)


class FakeMulti:
    """A stand-in several-case engine: n steps take n * 1 ms of the clock for the whole group;
    each active case's state counts its own steps."""
    live = 0

    def __init__(self, rs, clock, diverge=None):
        self.rs, self.clock, self.busy_until = rs, clock, 0.0
        self.active = [True] * len(rs)
        self.steps = [0] * len(rs)
        self.diverge = diverge or {}
        FakeMulti.live += 1

    def set_states(self, Q):
        self.Q = [np.array(q, dtype=float) for q in Q]

    def device_bytes(self):
        return 1000 * sum(r.nodes for r in self.rs)

    def done(self):
        return self.clock.t >= self.busy_until

    def advance(self, n, cfl, dcfl):
        assert self.done()
        self.busy_until = self.clock.t + 1e-3 * n
        for k in range(len(self.rs)):
            if self.active[k]:
                self.steps[k] += n
                self.Q[k] = self.Q[k] + n * cfl

    def set_active(self, k, on):
        self.active[k] = on

    def status(self):
        bad = np.array([k in self.diverge and self.steps[k] >= self.diverge[k] for k in range(len(self.rs))])
        return 0.1 * np.array(self.steps, dtype=float), bad

    def get_case_state(self, k):
        return self.Q[k].copy()

    def close(self):
        FakeMulti.live -= 1


def test_batched_serving_completes_every_case_as_alone():
    """Several-case engines: every request ends with its own step count and the state it has
    alone, cases are grouped only by their batch key, and a case that diverges is retired alone."""
    reqs = [serve.Request({"off": q.case["off"], "N": np.full((2, 3), 3 + k % 2)}, q.physics, q.Q0, q.steps,
                          cfl=q.cfl, arrival=q.arrival) for k, q in enumerate(requests(20, seed=13))]
    clock = Clock()
    FakeMulti.live = 0
    made = []

    def factory(rs):
        keys = {serve.batch_key(r) for r in rs}
        assert len(keys) == 1
        m = FakeMulti(rs, clock, diverge={k: 3 for k, r in enumerate(rs) if r is reqs[7]})
        made.append(m)
        return m

    res, st = serve.serve_batched(reqs, batch_factory=factory, clock=clock, sleep=clock.sleep, poll_s=0.0, chunk=7,
                                  min_batch=3, max_wait_s=0.01, device_memory=lambda: 1 << 40, workers=0)
    assert FakeMulti.live == 0 and len(made) >= 2
    assert res[7].error and "divergence" in res[7].error
    assert st["failed"] == 1 and st["cases"] == len(reqs) - 1
    for k, (r, q) in enumerate(zip(res, reqs)):
        assert r.admitted >= q.arrival
        if k != 7:
            assert r.error is None and np.array_equal(r.Q, q.Q0 + q.steps * 0.5), k
            assert r.t == pytest.approx(0.1 * q.steps)


def test_built_requests_are_made_at_set_up_and_dropped_after_upload():
    """A request given a builder makes its arrays when its case is set up and drops them once
    they are on the device, in every loop; the results are those of requests given the arrays."""
    base = requests(8, seed=5)
    made = []

    def lazy(q):
        def build():
            made.append(q.nodes)
            return {"off": q.case["off"]}, q.Q0.copy()
        return serve.Request(None, {}, None, q.steps, cfl=q.cfl, arrival=q.arrival, build=build, nodes=q.nodes)

    ref, _ = run(base, device_memory=lambda: 1 << 40)
    for loop in ("serve", "one_at_a_time"):
        made.clear()
        reqs = [lazy(q) for q in base]
        clock = Clock()
        f = (lambda r: FakeEngine(r, clock))
        if loop == "serve":
            res, _ = serve.serve(reqs, engine_factory=f, clock=clock, sleep=clock.sleep, poll_s=0.0, workers=0,
                                 device_memory=lambda: 1 << 40)
        else:
            res, _ = serve.one_at_a_time(reqs, engine_factory=f, clock=clock, sleep=clock.sleep)
        assert sorted(made) == sorted(q.nodes for q in base), loop
        assert all(q._case is None and q._Q0 is None for q in reqs), loop
        for a, b in zip(res, ref):
            assert np.array_equal(a.Q, b.Q) and a.t == b.t, loop
    digests, _ = run([lazy(q) for q in base], keep_states=False, device_memory=lambda: 1 << 40)
    assert [d.Q for d in digests] == [serve._digest(b.Q) for b in ref]
