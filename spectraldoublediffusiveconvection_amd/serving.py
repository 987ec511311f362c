"""Serving engine: continuous batching of time-integration, Newton and
continuation jobs of mixed sizes on one GPU.

Time integrations.  The jobs of one size class share a plan and a fixed
array of slots on the device (states, Ra_T, Ra_S).  For slot counts
1, 4, 16, 32, 64, ... up to the capacity the class holds a HIP graph of
``chunk`` in-place steps of its first slots (captured once, when the class
is created).  Each engine iteration launches, for every class with fewer
than two chunks in flight, one chunk for its active slots: the graph of the
smallest bucket that holds them (eager steps for a shorter last chunk).
Jobs that finish at the end of that chunk are copied to pinned host buffers
right behind it on the class's stream, the remaining slots are compacted and
waiting jobs are copied in from pinned buffers -- all stream-ordered, so the
host never waits for the device: it records an event per chunk and hands a
finished job back when its event has completed.  Jobs that arrive while
others run join the running batch at the next chunk.  A class is created
when its first job is admitted, if its HBM footprint fits the budget; an
idle class is released when the budget is needed for another.  Every class
steps on its own stream, so classes overlap on the device.

Newton and continuation jobs run on worker threads, each with its own stream
(ctypes releases the GIL during library calls), so their GPU work overlaps
the batched steps.  Workers allocate nothing: the engine thread creates, per
worker and size class, the Krylov object, the device vectors and the pinned
staging buffer before it hands the first job of that class to the worker,
and the workers make only kernel launches, copies between device and pinned
memory, and waits on their own stream -- all inside the shared side of the
engine's Gate, whose exclusive side the engine thread holds while it
captures a class's graphs, so no worker touches the GPU during a capture.
Newton jobs of one class are solved together (sddc_newton_batched, one
batched J v per Arnoldi step).  A continuation converges the steady state at
its Ra_T and then either follows a given path of Ra_T values by Newton
solves, each started from the previous solution (natural parameter,
_NewtonC, Main.py:717-739), or takes ``arc_points`` pseudo-arclength steps
(_ContinC, Main.py:742-950) with step size control, which also passes folds.

``Engine.shutdown()`` stops the workers between jobs and between the points
of a continuation, marks the jobs still queued as cancelled, and joins the
(non-daemon) worker threads, so no thread is inside a library call when the
interpreter exits.

The engine records arrival, admission and completion times of every job;
``Engine.stats()`` gives p50/p99 latency (completion - arrival) and
throughput.
"""
from __future__ import annotations

import collections
import contextlib
import dataclasses
import queue
import threading
import time
from typing import Any, Callable, List, Optional

import numpy as np

from . import operators as _ops


@dataclasses.dataclass
class Job:
    kind: str                      # "time" | "newton" | "continuation"
    N_fm: int
    N_r: int
    Ra: float
    Ra_s: float
    X0: np.ndarray
    d: float = 0.35
    dt: float = 5e-3
    Pr: float = 1.0
    Tau: float = 1.0 / 15.0
    symmetric: bool = False
    nsteps: int = 500              # time integration
    ra_path: Optional[List[float]] = None   # natural continuation: Ra_T values after Ra
    arc_points: int = 0            # pseudo-arclength continuation: points after Ra
    ds: float = 10.0               # its initial arclength step
    sign: float = 1.0              # its initial direction in Ra_T
    tol_newton: float = 1e-8
    tol_gmres: float = 1e-4
    restart: int = 100
    arrival: float = 0.0           # seconds after the stream starts
    id: int = -1
    # filled by the engine
    t_admit: Optional[float] = None
    t_done: Optional[float] = None
    result: Any = None
    info: Any = None
    cancelled: bool = False

    def key(self):
        return (self.N_fm, self.N_r, self.d, self.dt, self.Pr, self.Tau, self.symmetric)

    def newton_key(self):
        return (self.N_fm, self.N_r, self.d, self.Pr, self.Tau, self.symmetric, self.restart)


def class_bytes(N_fm, N_r, capacity, precision):
    """HBM footprint estimate of a time-integration class: operators and the
    segment maps, plus per-slot state and workspace."""
    nr = N_r - 1
    N = N_fm * nr
    rs = 4 if precision == "mixed" else 8
    ops = 3 * N_fm * nr * nr * 8 * 4
    per_slot = N * (8 * (2 + 12 + 3 + 7) + rs * 11)
    return ops + capacity * per_slot


def slot_buckets(capacity):
    """Slot counts with a captured graph: 1, 4, 16, 32, 64, then steps of
    64, below the capacity, and the capacity itself."""
    out = [b for b in (1, 4) if b < capacity]
    b = 16
    while b < capacity:
        out.append(b)
        b = b * 2 if b < 64 else b + 64
    out.append(capacity)
    return out


class Gate:
    """Keeps every other thread off the GPU while the engine thread captures
    graphs.  HIP invalidates a capture (thread-local mode included) when
    another thread works on the device meanwhile -- kernel launches and waits
    on its own non-blocking stream included, as measured on MI355X -- so the
    workers hold the gate shared around their device work and the engine
    holds it exclusively while it creates (captures) or releases a class.
    Writers go first: a waiting capture stops new shared holders."""

    def __init__(self):
        self._cv = threading.Condition()
        self._readers = 0
        self._writer = False
        self._waiting = 0

    @contextlib.contextmanager
    def shared(self):
        with self._cv:
            while self._writer or self._waiting:
                self._cv.wait()
            self._readers += 1
        try:
            yield
        finally:
            with self._cv:
                self._readers -= 1
                if self._readers == 0:
                    self._cv.notify_all()

    @contextlib.contextmanager
    def exclusive(self):
        with self._cv:
            self._waiting += 1
            while self._writer or self._readers:
                self._cv.wait()
            self._waiting -= 1
            self._writer = True
        try:
            yield
        finally:
            with self._cv:
                self._writer = False
                self._cv.notify_all()


class _Pool:
    """Free list of row indices of a pinned host buffer."""

    def __init__(self, n):
        self.free = list(range(n - 1, -1, -1))

    def get(self):
        return self.free.pop() if self.free else None

    def put(self, i):
        self.free.append(i)


class TimeClass:
    """Slots of the time-integration jobs of one size class, with the graphs
    of its slot buckets and the pinned staging of states in and out."""

    def __init__(self, key, capacity, precision, tuning=None, chunk=50):
        import torch

        from . import device as _dev
        N_fm, N_r, d, dt, Pr, Tau, sym = key
        self.key = key
        self.capacity = capacity
        self.chunk = chunk
        self.plan = _dev.Plan(_ops.build(N_fm, N_r, d, dt, Pr, Tau), symmetric=sym,
                              precision=precision, max_batch=capacity, tuning=tuning)
        self.stream = torch.cuda.Stream()
        n3 = 3 * N_fm * (N_r - 1)
        self.X = torch.zeros((capacity, n3), dtype=torch.float64, device="cuda")
        self.Ra = torch.zeros(capacity, dtype=torch.float64, device="cuda")
        self.Ra_s = torch.zeros(capacity, dtype=torch.float64, device="cuda")
        self.pin_in = torch.empty((capacity, n3), dtype=torch.float64, pin_memory=True)
        self.pin_out = torch.empty((2 * capacity, n3), dtype=torch.float64, pin_memory=True)
        self.pool_in, self.pool_out = _Pool(capacity), _Pool(2 * capacity)
        torch.cuda.synchronize()
        self.graphs = {b: self.plan.graph(self.X[:b], self.Ra[:b], self.Ra_s[:b], nsteps=chunk)
                       for b in slot_buckets(capacity)}
        self.jobs: List[Optional[Job]] = [None] * capacity
        self.remaining = np.zeros(capacity, dtype=np.int64)
        self.n = 0
        self.inflight = collections.deque()   # (event, [(job, pin_out row)], [pin_in rows])
        self.held_in: List[int] = []          # pin_in rows read by copies not yet fenced
        self.bytes = class_bytes(N_fm, N_r, capacity, precision)
        self.instance_steps = 0

    def free(self):
        return self.capacity - self.n

    def idle(self):
        return self.n == 0 and not self.inflight

    def admit(self, job, now):
        """Copy a job's state into the next slot (stream-ordered); False when
        no staging row is free until the next chunk completes."""
        import torch
        r = self.pool_in.get()
        if r is None:
            return False
        i = self.n
        self.pin_in[r].numpy()[:] = np.asarray(job.X0, dtype=np.float64)
        with torch.cuda.stream(self.stream):
            self.X[i].copy_(self.pin_in[r], non_blocking=True)
            self.Ra[i].fill_(float(job.Ra))
            self.Ra_s[i].fill_(float(job.Ra_s))
        self.held_in.append(r)
        self.jobs[i] = job
        self.remaining[i] = int(job.nsteps)
        self.n += 1
        job.t_admit = now
        return True

    def launch(self):
        """Launch one chunk for the active slots, then the copies out of the
        jobs it finishes and the slot compaction; returns instance-steps."""
        import torch
        n = self.n
        if n == 0 or len(self.inflight) >= 2:
            return 0
        k = int(min(self.chunk, self.remaining[:n].min()))
        done = [i for i in range(n) if self.remaining[i] == k]
        if len(self.pool_out.free) < len(done):
            return 0   # every finished state needs a free pinned row
        with torch.cuda.stream(self.stream):
            if k == self.chunk:
                b = next(b for b in self.graphs if b >= n)
                self.graphs[b].launch()
                launched = b * k
            else:
                self.plan.step(self.X[:n], self.Ra[:n], self.Ra_s[:n], nsteps=k, out=self.X[:n])
                launched = n * k
            self.remaining[:n] -= k
            out = []
            for i in done:
                r = self.pool_out.get()
                self.pin_out[r].copy_(self.X[i], non_blocking=True)
                out.append((self.jobs[i], r))
            for i in sorted(done, reverse=True):
                last = self.n - 1
                if i != last:
                    self.X[i].copy_(self.X[last])
                    self.Ra[i].copy_(self.Ra[last])
                    self.Ra_s[i].copy_(self.Ra_s[last])
                    self.jobs[i] = self.jobs[last]
                    self.remaining[i] = self.remaining[last]
                self.X[last].zero_()
                self.jobs[last] = None
                self.remaining[last] = 0
                self.n -= 1
            ev = torch.cuda.Event()
            ev.record(self.stream)
        self.inflight.append((ev, out, self.held_in))
        self.held_in = []
        self.instance_steps += n * k
        return launched

    def poll(self, now):
        """Jobs whose chunks have completed on the device."""
        finished = []
        while self.inflight and self.inflight[0][0].query():
            _, out, rows_in = self.inflight.popleft()
            for job, r in out:
                job.result = self.pin_out[r].numpy().copy()
                job.t_done = now
                self.pool_out.put(r)
                finished.append(job)
            for r in rows_in:
                self.pool_in.put(r)
        return finished

    def close(self):
        self.stream.synchronize()
        for g in self.graphs.values():
            g.close()
        self.graphs = {}
        self.plan.close()


class NewtonContext:
    """Device state of one worker for one Newton size class: the batched
    Newton solver and the Krylov object of the continuation steps (on one
    plan at dt = 1), the batch's states and parameters, the continuation
    vector Y = (X, Ra_T) and its tangent, and pinned staging.  Created by the
    engine thread; the worker only uses it."""

    def __init__(self, job, batch):
        import torch

        from . import device as _dev
        from .newton import Krylov, NewtonBatch
        plan = _dev.Plan(_ops.build(job.N_fm, job.N_r, job.d, 1.0, job.Pr, job.Tau),
                         symmetric=job.symmetric, max_batch=batch)
        self.nb = NewtonBatch(plan, restart=job.restart, max_batch=batch)
        self.kr = Krylov(plan, restart=job.restart)
        self.batch = batch
        self.n3 = 3 * job.N_fm * (job.N_r - 1)
        self.X = torch.zeros((batch, self.n3), dtype=torch.float64, device="cuda")
        self.ra = torch.zeros(batch, dtype=torch.float64, device="cuda")
        self.ras = torch.zeros(batch, dtype=torch.float64, device="cuda")
        self.Y = torch.zeros(self.n3 + 1, dtype=torch.float64, device="cuda")
        self.Yd = torch.zeros_like(self.Y)
        self.pin = torch.empty((batch, self.n3 + 1), dtype=torch.float64, pin_memory=True)

    def close(self):
        self.nb.close()
        self.kr.close()
        self.kr.plan.close()


class NewtonWorker(threading.Thread):
    """Runs Newton and continuation jobs on its own stream until stopped.
    Newton jobs of one size class waiting together in its queue are solved as
    one batch (up to the context's batch size).  ``solve(worker, jobs)``
    -> [(result, info)] does the work; the default runs the jobs on the device
    through the worker's NewtonContext of their class."""

    def __init__(self, engine, stream=None, solve: Optional[Callable] = None, batch=8):
        super().__init__(daemon=False)
        self.engine = engine
        self.stream = stream
        self.batch = batch
        self.q: "queue.Queue[Optional[Job]]" = queue.Queue()
        self.backlog: "collections.deque[Job]" = collections.deque()
        self.contexts = {}
        self.stop = engine.stop
        self.busy = False
        self.error = None
        self._solve = solve or NewtonWorker.device_solve

    def load(self):
        return self.q.qsize() + len(self.backlog) + (1 if self.busy else 0)

    # ---------------------------------------------------------- device work
    def _rows_to_host(self, ctx, src, rows, n):
        """Copy `rows` vectors of length n from src (device) through pinned memory."""
        ctx.pin[:rows, :n].copy_(src.reshape(rows, -1)[:, :n], non_blocking=True)
        self.stream.synchronize()
        return [ctx.pin[i, :n].numpy().copy() for i in range(rows)]

    def device_solve(self, jobs):
        with self.engine.gate.shared():
            return self._device_solve(jobs)

    def _device_solve(self, jobs):
        job = jobs[0]
        ctx = self.contexts[job.newton_key()]
        n3 = ctx.n3
        opts = dict(tol_newton=job.tol_newton, tol_gmres=job.tol_gmres)
        if job.kind == "newton":
            B = len(jobs)
            for i, j in enumerate(jobs):
                ctx.pin[i, :n3].numpy()[:] = np.asarray(j.X0, dtype=np.float64)
            ctx.X[:B].copy_(ctx.pin[:B, :n3], non_blocking=True)
            for i, j in enumerate(jobs):
                ctx.ra[i].fill_(float(j.Ra))
                ctx.ras[i].fill_(float(j.Ra_s))
            infos = ctx.nb.solve_(ctx.X[:B], ctx.ra[:B], ctx.ras[:B], **opts)
            return list(zip(self._rows_to_host(ctx, ctx.X[:B], B, n3), infos))
        # continuation: Newton solves as a batch of one, then arclength steps
        X1, ra1, ras1 = ctx.X[:1], ctx.ra[:1], ctx.ras[:1]
        ctx.pin[0, :n3].numpy()[:] = np.asarray(job.X0, dtype=np.float64)
        X1.copy_(ctx.pin[:1, :n3], non_blocking=True)
        ras1.fill_(float(job.Ra_s))
        points, infos = [], []
        for ra in [job.Ra] + list(job.ra_path or []):
            if self.stop.is_set():
                infos.append({"cancelled": True})
                return [(points, infos)]
            ra1.fill_(float(ra))
            info = ctx.nb.solve_(X1, ra1, ras1, **opts)[0]
            infos.append(info)
            points.append((ra, self._rows_to_host(ctx, X1, 1, n3)[0]))
            if not info["converged"]:
                return [(points, infos)]
        if job.arc_points > 0:
            ctx.Y[:n3].copy_(X1[0])
            ctx.Y[n3].fill_(float(job.Ra))
            ctx.Yd.zero_()
            ds = job.ds
            for _ in range(job.arc_points):
                if self.stop.is_set():
                    infos.append({"cancelled": True})
                    break
                ds, info = ctx.kr.arclength_step_(ctx.Y, ctx.Yd, ras1, job.sign, ds, **opts)
                infos.append(info)
                if info["status"] != 0:
                    break
                y = self._rows_to_host(ctx, ctx.Y, 1, n3 + 1)[0]
                points.append((float(y[-1]), y[:-1]))
        return [(points, infos)]

    # ------------------------------------------------------------- the loop
    def _next(self):
        """The next unit of work: one continuation job, or up to `batch`
        Newton jobs of one class that are waiting now; None to stop."""
        while True:                         # everything queued now, without waiting
            try:
                self.backlog.append(self.q.get_nowait())
            except queue.Empty:
                break
        if not self.backlog:
            self.backlog.append(self.q.get())
        job = self.backlog.popleft()
        if job is None:
            return None
        if job.kind != "newton":
            return [job]
        batch, rest = [job], collections.deque()
        while self.backlog:
            j = self.backlog.popleft()
            if (j is not None and j.kind == "newton" and j.newton_key() == job.newton_key()
                    and len(batch) < self.batch):
                batch.append(j)
            else:
                rest.append(j)
        self.backlog = rest
        return batch

    def _serve(self):
        while True:
            jobs = self._next()
            if jobs is None:
                return
            if self.stop.is_set():
                for j in jobs:
                    self.engine.cancel(j)
                continue
            self.busy = True
            try:
                results = self._solve(self, jobs)
            except Exception as e:   # reported with the jobs, the engine keeps serving
                results = [(None, {"exception": repr(e)})] * len(jobs)
                self.error = e
            t = self.engine.now()
            for j, (res, info) in zip(jobs, results):
                j.result, j.info, j.t_done = res, info, t
                self.engine.completed(j)
            self.busy = False

    def run(self):
        if self.stream is None:
            self._serve()
            return
        import torch
        with torch.cuda.stream(self.stream):
            self._serve()


class Engine:
    def __init__(self, precision="mixed", capacity=64, hbm_budget=None, chunk=50,
                 newton_workers=2, tuning_table=None, solve=None, newton_batch=8):
        self.precision = precision
        self.capacity = capacity
        self.chunk = chunk
        self.tuning_table = tuning_table
        self._device = solve is None
        if hbm_budget is None:
            hbm_budget = 1 << 62
            if self._device:
                import torch
                free, _ = torch.cuda.mem_get_info()
                hbm_budget = int(0.8 * free)
        self.hbm_budget = hbm_budget
        self.classes = collections.OrderedDict()
        self.stop = threading.Event()
        self.gate = Gate()
        streams = [None] * newton_workers
        if self._device:
            import torch
            streams = [torch.cuda.Stream() for _ in range(newton_workers)]
        self.newton_batch = newton_batch
        self.workers = [NewtonWorker(self, s, solve, newton_batch) for s in streams]
        self._home = {}
        self.done: List[Job] = []
        self.cancelled: List[Job] = []
        self._lock = threading.Lock()
        self.t0 = time.perf_counter()
        self.elapsed = 0.0

    def now(self):
        return time.perf_counter() - self.t0

    def completed(self, job):
        with self._lock:
            self.done.append(job)

    def cancel(self, job):
        job.cancelled = True
        job.info = {"cancelled": True}
        with self._lock:
            self.cancelled.append(job)

    @property
    def instance_steps(self):
        return sum(c.instance_steps for c in self.classes.values()) + self._released_steps

    _released_steps = 0

    # ------------------------------------------------------------ resources
    def _class_for(self, job):
        key = job.key()
        c = self.classes.get(key)
        if c is not None:
            return c
        need = class_bytes(job.N_fm, job.N_r, self.capacity, self.precision)
        used = sum(c.bytes for c in self.classes.values())
        # release idle classes while the new one does not fit
        for k in list(self.classes):
            if used + need <= self.hbm_budget:
                break
            if self.classes[k].idle():
                used -= self.classes[k].bytes
                old = self.classes.pop(k)
                self._released_steps += old.instance_steps
                with self.gate.exclusive():
                    old.close()
        if used + need > self.hbm_budget:
            return None
        tuning = (self.tuning_table.lookup(key, self.capacity, self.precision)
                  if self.tuning_table else None)
        with self.gate.exclusive():   # no worker on the GPU while the graphs are captured
            c = TimeClass(key, self.capacity, self.precision, tuning, self.chunk)
        self.classes[key] = c
        return c

    def _context(self, w, job):
        k = job.newton_key()
        if self._device and k not in w.contexts:
            w.contexts[k] = NewtonContext(job, self.newton_batch)

    def _dispatch(self, job):
        """Hand a Newton or continuation job to a worker, creating that
        worker's device context of the job's class here, on the engine thread.
        Newton jobs of one class go to one worker (its home, assigned in turn
        as classes appear), so that they batch; continuation jobs go to the
        least loaded worker."""
        if job.kind == "newton":
            k = job.newton_key()
            if k not in self._home:
                self._home[k] = self.workers[len(self._home) % len(self.workers)]
            w = self._home[k]
        else:   # the least loaded worker, one that is no Newton class's home if any is free
            homes = set(map(id, self._home.values()))
            others = [x for x in self.workers if id(x) not in homes] or self.workers
            w = min(others, key=lambda x: x.load())
        self._context(w, job)
        w.q.put(job)

    def prepare(self, keys=(), newton_jobs=()):
        """Create the time classes of `keys` and every worker's Newton contexts
        for the classes of `newton_jobs` ahead of a run (graph captures and
        allocations before any job is timed)."""
        for key in keys:
            if key not in self.classes:
                self._class_for(Job("time", key[0], key[1], 0.0, 0.0, None, d=key[2],
                                    dt=key[3], Pr=key[4], Tau=key[5], symmetric=key[6]))
        for job in newton_jobs:
            for w in self.workers:
                self._context(w, job)

    # ------------------------------------------------------------------ run
    def run(self, jobs, realtime=True, max_seconds=None):
        """Serve jobs (arrival times in seconds from now).  With realtime=False
        the clock is virtual: every engine iteration that launches work
        advances it by one unit, and an idle engine jumps to the next arrival;
        arrival times are then in iterations."""
        for i, j in enumerate(jobs):
            if j.id < 0:
                j.id = i
        pending = collections.deque(sorted(jobs, key=lambda j: j.arrival))
        waiting = collections.deque()
        for w in self.workers:
            if not w.is_alive() and not self.stop.is_set():
                w.start()
        self.t0 = time.perf_counter()
        virtual = 0.0
        total = len(jobs)
        base_done = len(self.done)
        while True:
            now = self.now() if realtime else virtual
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            for c in list(self.classes.values()):
                for job in c.poll(now):
                    self.completed(job)
            # admission, FIFO within each kind
            still = collections.deque()
            for job in waiting:
                if job.kind == "time":
                    c = self._class_for(job)
                    if c is None or c.free() == 0 or not c.admit(job, now):
                        still.append(job)
                else:
                    job.t_admit = now
                    self._dispatch(job)
            waiting = still
            launched = 0
            for c in self.classes.values():
                launched += c.launch()
            with self._lock:
                ndone = len(self.done) - base_done
            if ndone >= total:
                break
            if max_seconds is not None and self.now() > max_seconds:
                break
            if not realtime:
                if launched:
                    virtual += 1.0
                elif pending and not waiting and not any(c.n or c.inflight
                                                          for c in self.classes.values()):
                    virtual = max(virtual + 1.0, pending[0].arrival)
                else:
                    time.sleep(0.0002)
            elif launched == 0:
                if pending and not waiting:
                    time.sleep(max(0.0, min(0.002, pending[0].arrival - self.now())))
                else:
                    time.sleep(0.0002)
        if self._device:
            import torch
            torch.cuda.synchronize()
        self.elapsed = self.now()
        return self.done

    def shutdown(self):
        """Stop the workers: jobs still queued are cancelled, a running
        continuation stops at its next point, and every worker thread is
        joined (each library call it may be in is bounded)."""
        self.stop.set()
        for w in self.workers:
            while True:
                try:
                    job = w.q.get_nowait()
                except queue.Empty:
                    break
                if job is not None:
                    self.cancel(job)
            w.q.put(None)
        for w in self.workers:
            if w.is_alive():
                w.join()
        for c in list(self.classes.values()):
            self._released_steps += c.instance_steps
            c.close()
        self.classes.clear()
        for w in self.workers:
            for ctx in w.contexts.values():
                ctx.close()
            w.contexts.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.shutdown()
        return False

    def stats(self):
        done = [j for j in self.done if j.t_done is not None]
        lat = np.array([j.t_done - j.arrival for j in done])
        by_kind = collections.defaultdict(list)
        for j in done:
            by_kind[j.kind].append(j.t_done - j.arrival)
        steps = self.instance_steps
        out = {
            "jobs": len(done),
            "cancelled": len(self.cancelled),
            "elapsed_s": self.elapsed,
            "jobs_per_s": len(done) / self.elapsed if self.elapsed > 0 else 0.0,
            "instance_steps_per_s": steps / self.elapsed if self.elapsed > 0 else 0.0,
            "latency_p50_ms": float(np.percentile(lat, 50) * 1e3) if lat.size else None,
            "latency_p99_ms": float(np.percentile(lat, 99) * 1e3) if lat.size else None,
        }
        for k, v in by_kind.items():
            v = np.array(v)
            out["latency_p50_ms_" + k] = float(np.percentile(v, 50) * 1e3)
            out["latency_p99_ms_" + k] = float(np.percentile(v, 99) * 1e3)
        return out


def poisson_stream(rate, duration, seed=0, mix=(0.6, 0.3, 0.1), sizes=None, nsteps=500,
                   newton_case=None, ra=(2500.0, 3000.0), ra_s=(300.0, 500.0), cont_points=20,
                   arclength=False):
    """Jobs arriving as a Poisson process of `rate` per second over `duration`
    seconds.  Time integrations draw their size from `sizes` (N_r, N_fm);
    Newton and continuation jobs start from `newton_case` (a dict with the
    parameters and a state near a steady solution), or from one of a list of
    such dicts drawn uniformly; their Ra_T is the case's times
    1 + ra_spread * U(-1, 1) (ra_spread 0.02 unless the case sets it) and
    their start the case's state times 1 + perturb * cos(i + phase) (perturb
    0 unless set).  Continuations have `cont_points` points in
    all: natural steps of newton_case["dRa"] in Ra_T, or with `arclength`
    pseudo-arclength steps starting at ds = dRa."""
    rng = np.random.default_rng(seed)
    sizes = sizes or [(r, k) for r in (20, 30, 40) for k in (64, 128, 256, 512)]
    cases = newton_case if isinstance(newton_case, (list, tuple)) else (
        [newton_case] if newton_case is not None else [])
    t, jobs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        u = rng.random()
        if u < mix[0] or not cases:
            nr_, k_ = sizes[rng.integers(len(sizes))]
            x = rng.standard_normal(3 * k_ * (nr_ - 1))
            jobs.append(Job("time", k_, nr_, float(rng.uniform(*ra)), float(rng.choice(ra_s)),
                            1e-3 * x / np.linalg.norm(x), nsteps=nsteps, arrival=t))
        else:
            c = cases[rng.integers(len(cases))] if len(cases) > 1 else cases[0]
            kind = "newton" if u < mix[0] + mix[1] else "continuation"
            ra0 = c["Ra"] * (1.0 + c.get("ra_spread", 0.02) * rng.uniform(-1, 1))
            X0 = c["X"]
            if c.get("perturb", 0.0) > 0.0:   # a smooth relative perturbation of the start
                X0 = X0 * (1.0 + c["perturb"] * np.cos(np.arange(X0.size) + rng.uniform(0, 6.3)))
            path, arc = None, 0
            if kind == "continuation":
                if arclength:
                    arc = cont_points - 1
                else:
                    path = list(ra0 + c.get("dRa", 20.0) * np.arange(1, cont_points))
            jobs.append(Job(kind, c["N_fm"], c["N_r"], ra0, c["Ra_s"], X0, d=c["d"],
                            Pr=c["Pr"], Tau=c["Tau"], symmetric=c.get("symmetric", False),
                            ra_path=path, arc_points=arc, ds=c.get("dRa", 20.0),
                            restart=c.get("restart", 100), arrival=t))
    return jobs
