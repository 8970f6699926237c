"""Serving: independent simulations admitted as they arrive and stepped together on one GPU.

Each admitted case is an hx3d engine with its own HIP stream and device-resident state.  The
loop admits requests whose arrival time has passed while free HBM (less a reserve) holds the
case and fewer than `max_active` cases are resident; it feeds up to `max_running` of them at a
time, least remaining work first, with chunks of `chunk` steps queued without blocking
(hx3d_advance with no host time, polled with hx3d_query), so the kernels of different cases
overlap on the device; and it retires a case
when its last chunk has finished: state and time are read back and the engine's memory is
freed for the next arrival.  A case's result does not depend on what runs beside it: each
engine computes only from its own arrays, in its own stream order, so the state is the one
the case reaches alone, bit for bit.

A case that fails is retired alone and the others keep running: a case whose steps diverge (a NaN
appears in its state; the engine's RK update flags it, as the reference's checkForNan does after
every step, ExplicitMethods.f90:786) ends with its state at that point and `Result.error` set,
and so does a case whose engine cannot be set up.

`serve` returns per-case results and the aggregate figures (DOF-updates/s, p50/p99 latency
from arrival to completion); `one_at_a_time` runs the same requests in arrival order, each
alone on the plain engine, for comparison."""
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import engine as E

BYTES_PER_NODE_GUESS = 1200.0   # device bytes per node before the first engine is measured


class Request:
    """One simulation: a case dict, engine keyword arguments, initial state, step count and
    CFL numbers, arriving `arrival` seconds after the loop starts.

    A request may instead carry `build`, a function returning (case, Q0), with its node count
    and its order N (one order everywhere, or None for mixed orders): the arrays are then made
    when the case is set up and dropped once they are on the device (release), so a long
    stream never holds every case on the host at once."""

    def __init__(self, case, physics, Q0, steps, cfl=0.4, dcfl=0.4, arrival=0.0, name=None,
                 build=None, nodes=None, order=None):
        self._case, self.physics, self._Q0, self.build = case, physics, Q0, build
        self.steps, self.cfl, self.dcfl, self.arrival = int(steps), cfl, dcfl, float(arrival)
        self.name = name
        if case is not None:
            self.nodes = int(case["off"][-1]) if "off" in case else len(Q0)
            N = np.asarray(case.get("N", [-1]))
            self.order = int(N.flat[0]) if N.flat[0] >= 0 and (N == N.flat[0]).all() else None
        else:
            self.nodes, self.order = int(nodes), order

    def _make(self):
        if self._case is None:
            self._case, self._Q0 = self.build()

    @property
    def case(self):
        self._make()
        return self._case

    @property
    def Q0(self):
        self._make()
        return self._Q0

    def release(self):
        """Drops built arrays (a request given its arrays keeps them)."""
        if self.build is not None:
            self._case, self._Q0 = None, None


def _digest(Q):
    import hashlib
    return None if Q is None else hashlib.sha1(np.ascontiguousarray(Q).tobytes()).hexdigest()


class Result:
    """The end of one request: final state Q and time t (the state where it stopped for a failed
    case, or None if no engine was built), admission and completion times, and error (None, or
    the message of the failure that retired the case).  With keep_states=False the loops keep
    only the SHA-1 digest of each final state's bytes (Q is then that hex string)."""

    def __init__(self, req, Q, t, admitted, finished, error=None):
        self.request, self.Q, self.t = req, Q, t
        self.admitted, self.finished = admitted, finished
        self.latency = finished - req.arrival
        self.error = error


def _stats(results, wall):
    ok = [r for r in results if r.error is None]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
    work = sum(r.request.nodes * r.request.steps for r in ok)
    return dict(cases=len(ok), failed=len(results) - len(ok), wall_s=wall, dof_updates=work,
                dof_updates_per_s=work / wall if wall > 0 else 0.0,
                p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)))


def serve(requests, chunk=20, max_active=16, max_running=2, hbm_reserve=8 << 30, poll_s=2e-4,
          engine_factory=None, device_memory=None, clock=time.perf_counter, sleep=time.sleep, workers=2,
          keep_states=True):
    """Runs the requests with continuous batching; returns (results in request order, stats).

    Up to max_active cases are resident; of these, up to max_running have steps queued at a
    time, chosen by least remaining work (steps x nodes), so short cases finish early instead
    of sharing the device evenly with long ones (two measured best on MI355X: the same
    throughput as four to sixteen, a sixth of their median latency).

    Engine set-up (host tables, uploads) and retirement (read-back, free) run on `workers`
    host threads (the C calls release the interpreter lock), so that the loop keeps the
    running engines fed while a new case is being set up; admission counts the memory of
    cases still being set up."""
    make = engine_factory or (lambda r: E.Engine(r.case, **r.physics))
    freemem = device_memory or (lambda: E.device_memory()[0])
    order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
    pending = list(order)
    preparing = {}   # request index -> (future, reserved bytes)
    active = []      # [index, engine, steps queued so far, admitted time]
    retiring = {}    # request index -> (future, admitted time)
    results = [None] * len(requests)
    per_node = [None]
    most = 0

    def prepare(r):
        try:
            eng = make(r)
            eng.set_state(r.Q0)
        finally:
            r.release()
        eng.t = 0.0
        b = eng.device_bytes()
        if b > 0:
            per_node[0] = max(per_node[0] or 0.0, b / r.nodes)
        return eng

    def finish(eng):
        try:
            t = eng.time()
            Q = eng.get_state()
        finally:
            eng.close()
        return t, (Q if keep_states else _digest(Q))

    pool = ThreadPoolExecutor(max_workers=max(1, workers)) if workers else None
    submit = pool.submit if pool else _run_now
    t0 = clock()
    try:
        while pending or preparing or active or retiring:
            now = clock() - t0
            # admission: arrived requests, in arrival order, while memory and slots allow
            while pending and requests[pending[0]].arrival <= now and len(active) + len(preparing) < max_active:
                i = pending[0]
                r = requests[i]
                need = (per_node[0] or BYTES_PER_NODE_GUESS) * r.nodes
                held = sum(b for _, b in preparing.values())
                if (active or preparing or retiring) and need + held > freemem() - hbm_reserve:
                    break   # wait for running cases to finish and free their memory
                pending.pop(0)
                preparing[i] = (submit(prepare, r), need)
            for i in [k for k, (f, _) in preparing.items() if f.done()]:
                f, _ = preparing.pop(i)
                if f.exception() is not None:
                    # the case cannot be set up: retired alone with the error
                    results[i] = Result(requests[i], None, None, clock() - t0, clock() - t0, error=str(f.exception()))
                    continue
                active.append([i, f.result(), 0, clock() - t0])
            most = max(most, len(active))
            busy = bool(preparing or retiring)
            # idle engines: retire the finished ones; feed the others, least remaining work
            # first, while fewer than max_running have work queued
            running, idle = 0, []
            for slot in list(active):
                i, eng, queued, admitted = slot
                try:
                    done = eng.done()
                except E.HX3DDiverged as exc:
                    # a NaN appeared in this case's state: it stops there, the others go on
                    active.remove(slot)
                    retiring[i] = (submit(finish, eng), admitted, str(exc))
                    continue
                if not done:
                    running += 1
                elif queued < requests[i].steps:
                    idle.append(slot)
                else:
                    active.remove(slot)
                    retiring[i] = (submit(finish, eng), admitted, None)
            idle.sort(key=lambda sl: (requests[sl[0]].steps - sl[2]) * requests[sl[0]].nodes)
            for slot in idle[:max(0, max_running - running)]:
                i, eng, queued, _ = slot
                r = requests[i]
                n = min(chunk, r.steps - queued)
                eng.advance_async(n, r.cfl, r.dcfl)
                slot[2] = queued + n
            busy = busy or running > 0 or bool(idle)
            for i in [k for k, v in retiring.items() if v[0].done()]:
                f, admitted, err = retiring.pop(i)
                t, Q = f.result()
                results[i] = Result(requests[i], Q, t, admitted, clock() - t0, error=err)
            if not busy and pending and not active:
                # nothing running: sleep until the next arrival
                wait = requests[pending[0]].arrival - (clock() - t0)
                if wait > 0:
                    sleep(min(wait, 0.05))
            elif busy:
                sleep(poll_s)
    except BaseException:
        # release every engine still held before passing the error on
        if pool:
            pool.shutdown(wait=True)
            pool = None
        held = [sl[1] for sl in active]
        held += [f.result() for f, _ in preparing.values() if f.done() and f.exception() is None]
        for eng in held:
            eng.close()
        raise
    finally:
        if pool:
            pool.shutdown(wait=True)
    wall = clock() - t0
    st = _stats([r for r in results if r is not None], wall)
    st["max_concurrent"] = most
    return results, st


class _Done:
    """A finished result or error in the form of a future (set-up and retirement on the calling
    thread)."""

    def __init__(self, v, exc=None):
        self.v, self.exc = v, exc

    def done(self):
        return True

    def result(self):
        if self.exc is not None:
            raise self.exc
        return self.v

    def exception(self):
        return self.exc


def _run_now(fn, *a):
    try:
        return _Done(fn(*a))
    except Exception as exc:   # kept for the caller, as a pool's future keeps it
        return _Done(None, exc)


def one_at_a_time(requests, engine_factory=None, clock=time.perf_counter, sleep=time.sleep, keep_states=True,
                  progress=None):
    """The same requests in arrival order, each alone on the plain engine (create, step to the
    end, read back, free) and waiting for its arrival; returns (results, stats)."""
    make = engine_factory or (lambda r: E.Engine(r.case, **r.physics))
    order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
    results = [None] * len(requests)
    stepping = 0.0   # time inside the step loops (the rest is set-up, transfers and waiting)
    t0 = clock()
    for i in order:
        r = requests[i]
        wait = r.arrival - (clock() - t0)
        if wait > 0:
            sleep(wait)
        admitted = clock() - t0
        try:
            eng = make(r)
        except E.HX3DError as exc:
            r.release()
            results[i] = Result(r, None, None, admitted, clock() - t0, error=str(exc))
            continue
        err = None
        try:
            try:
                eng.set_state(r.Q0)
            finally:
                r.release()
            ts = clock()
            try:
                t = eng.advance(r.steps, r.cfl, r.dcfl, t=0.0)
            except E.HX3DDiverged as exc:
                t, err = eng.time(), str(exc)
            stepping += clock() - ts
            Q = eng.get_state()
        finally:
            eng.close()
        results[i] = Result(r, Q if keep_states else _digest(Q), t, admitted, clock() - t0, error=err)
        done = sum(x is not None for x in results)
        if progress and done % progress == 0:
            import sys
            print("[serve] one at a time: %d of %d cases done at %.1f s" % (done, len(requests), clock() - t0),
                  file=sys.stderr, flush=True)
    wall = clock() - t0
    st = _stats(results, wall)
    st["stepping_s"] = stepping
    return results, st


def batch_key(r):
    """Requests that can share one engine: one order N everywhere and the same discretisation."""
    if r.order is None:
        return None   # mixed orders: served alone
    ph = r.physics
    return (r.order, ph.get("riemann", "roe"), ph.get("inviscid", "standard"), ph.get("averaging", "standard"),
            ph.get("les"), ph.get("lambda_stab", 1.0), float(r.cfl), float(r.dcfl))


class _Batch:
    """A running several-case engine: the requests it holds, the steps each has left, and which
    are still active."""

    def __init__(self, idx, eng, admitted, steps):
        self.idx, self.eng, self.admitted = idx, eng, admitted
        self.left = list(steps)
        self.live = [True] * len(idx)


def serve_batched(requests, chunk=50, max_batch=64, max_batches=3, min_batch=8, max_wait_s=0.05,
                  hbm_reserve=8 << 30, poll_s=2e-4, batch_factory=None, device_memory=None,
                  clock=time.perf_counter, sleep=time.sleep, workers=2, progress=None, keep_states=True):
    """Continuous batching with shared launches: arrived requests that can share an engine (one
    order, one discretisation, see batch_key) are gathered into several-case engines
    (engine.MultiEngine), each case with its own physics numbers, CFL dt, time and step count.
    A group is formed when min_batch such requests wait, when the oldest has waited max_wait_s,
    or when the device has nothing else to do, up to max_batch cases and the free HBM (less a
    reserve); at most max_batches engines run at once (their streams overlap on the device).
    An engine takes chunks of at most `chunk` steps, cut so that no case overshoots its step
    count; a case that has taken its steps, or whose state holds a NaN, is switched off and
    its state read back while the others go on.  Returns (results in request order, stats)."""
    make = batch_factory or (lambda rs: E.MultiEngine([r.case for r in rs], [r.physics for r in rs]))
    freemem = device_memory or (lambda: E.device_memory()[0])
    order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
    pending = list(order)
    waiting = {}     # key -> [(request index, arrival)]
    preparing = []   # (future, idx list, reserved bytes)
    running = []     # _Batch
    closing = []     # futures of engines being freed
    results = [None] * len(requests)
    per_node = [None]
    done_count = [0]

    def prepare(rs):
        try:
            eng = make(rs)
            eng.set_states([r.Q0 for r in rs])
        finally:
            for r in rs:
                r.release()
        b = eng.device_bytes()
        if b > 0:
            per_node[0] = max(per_node[0] or 0.0, b / sum(r.nodes for r in rs))
        return eng

    def read_case(eng, pos):
        Q = eng.get_case_state(pos)
        return Q if keep_states else _digest(Q)

    def close(eng):
        eng.close()

    pool = ThreadPoolExecutor(max_workers=max(1, workers)) if workers else None
    submit = pool.submit if pool else _run_now
    t0 = clock()
    try:
        while pending or waiting or preparing or running or closing:
            now = clock() - t0
            while pending and requests[pending[0]].arrival <= now:
                i = pending.pop(0)
                key = batch_key(requests[i])
                waiting.setdefault(key if key is not None else ("alone", i), []).append(i)
            # form groups
            idle_device = not running and not preparing
            for key in list(waiting):
                lst = waiting[key]
                if len(running) + len(preparing) >= max_batches:
                    break
                oldest = min(requests[i].arrival for i in lst)
                if not (len(lst) >= min_batch or now - oldest >= max_wait_s or idle_device):
                    continue
                held = sum(b for _, _, b in preparing)
                room = freemem() - hbm_reserve - held
                take, need = [], 0.0
                for i in lst[:max_batch]:
                    nb_ = (per_node[0] or BYTES_PER_NODE_GUESS) * requests[i].nodes
                    if take and need + nb_ > room:
                        break
                    take.append(i)
                    need += nb_
                if not take or (need > room and (running or preparing)):
                    continue   # wait for memory
                waiting[key] = lst[len(take):]
                if not waiting[key]:
                    del waiting[key]
                preparing.append((submit(prepare, [requests[i] for i in take]), take, need))
                idle_device = False
            for item in [p for p in preparing if p[0].done()]:
                preparing.remove(item)
                f, take, _ = item
                if f.exception() is not None:
                    for i in take:
                        results[i] = Result(requests[i], None, None, clock() - t0, clock() - t0, error=str(f.exception()))
                        done_count[0] += 1
                    continue
                running.append(_Batch(take, f.result(), clock() - t0, [requests[i].steps for i in take]))
            busy = bool(preparing)
            for b in list(running):
                if not b.eng.done():
                    busy = True
                    continue
                # the engine is idle: retire what finished (its state read back here, the others
                # are untouched by the read), then queue the next chunk
                times, bad = b.eng.status()
                for pos, i in enumerate(b.idx):
                    if b.live[pos] and (bad[pos] or b.left[pos] == 0):
                        b.live[pos] = False
                        b.eng.set_active(pos, False)
                        err = "numerical divergence: a NaN appeared in the state" if bad[pos] else None
                        results[i] = Result(requests[i], read_case(b.eng, pos), float(times[pos]), b.admitted,
                                            clock() - t0, error=err)
                        done_count[0] += 1
                        if progress and done_count[0] % progress == 0:
                            import sys
                            print("[serve] %d of %d cases done at %.1f s" % (done_count[0], len(requests), clock() - t0),
                                  file=sys.stderr, flush=True)
                live = [p for p in range(len(b.idx)) if b.live[p]]
                if live:
                    n = min(chunk, min(b.left[p] for p in live))
                    b.eng.advance(n, requests[b.idx[live[0]]].cfl, requests[b.idx[live[0]]].dcfl)
                    for p in live:
                        b.left[p] -= n
                    busy = True
                else:
                    running.remove(b)
                    closing.append(submit(close, b.eng))
            closing = [f for f in closing if not f.done()]
            if not busy and pending and not running and not waiting:
                wait = requests[pending[0]].arrival - (clock() - t0)
                if wait > 0:
                    sleep(min(wait, 0.05))
            elif busy or closing:
                sleep(poll_s)
            elif waiting:
                sleep(poll_s)
    except BaseException:
        if pool:
            pool.shutdown(wait=True)
            pool = None
        for b in running:
            b.eng.close()
        for f, _, _ in preparing:
            if f.done() and f.exception() is None:
                f.result().close()
        raise
    finally:
        if pool:
            pool.shutdown(wait=True)
    wall = clock() - t0
    st = _stats([r for r in results if r is not None], wall)
    return results, st
