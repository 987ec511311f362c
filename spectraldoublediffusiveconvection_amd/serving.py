"""Serving engine: continuous batching of time-integration, Newton and
continuation jobs of mixed sizes on one GPU.

Time-integration jobs of one size class share a plan and a fixed array of
slots on the device.  Each engine iteration advances every active slot of
every class by one chunk of steps (the chunk ends at the nearest job's last
step), finished jobs are copied out and their slots compacted, and waiting
jobs are admitted into free slots -- so jobs that arrive while others run
join the running batch at the next iteration.  A class is created when its
first job is admitted, if its HBM footprint fits the budget; a class with no
work is released when the budget is needed for another.  Each class steps on
its own stream, so classes overlap on the device.

Newton jobs (steady states, device Newton-Krylov) and continuation jobs run
on worker threads with their own streams; their GPU work overlaps with the
batched steps (ctypes releases the GIL during library calls).  A
continuation job converges the steady state at its Ra_T and then either
follows a given path of Ra_T values by Newton solves, each started from the
previous solution (natural parameter, _NewtonC, Main.py:717-739), or takes
`arc_points` pseudo-arclength steps (_ContinC, Main.py:742-950) with step
size control, which also passes folds of the branch.

The engine records arrival, admission and completion times of every job;
``Engine.stats()`` gives p50/p99 latency (completion - arrival) and
throughput.
"""
from __future__ import annotations

import collections
import dataclasses
import queue
import threading
import time
from typing import Any, List, Optional

import numpy as np
import torch

from . import device as _dev
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

    def key(self):
        return (self.N_fm, self.N_r, self.d, self.dt, self.Pr, self.Tau, self.symmetric)


def class_bytes(N_fm, N_r, capacity, precision):
    """HBM footprint estimate of a time-integration class: operators and the
    segment maps, plus per-slot state and workspace."""
    nr = N_r - 1
    N = N_fm * nr
    rs = 4 if precision == "mixed" else 8
    ops = 3 * N_fm * nr * nr * 8 * 4
    per_slot = N * (8 * (2 + 12 + 3 + 7) + rs * 11)
    return ops + capacity * per_slot


class TimeClass:
    """Slots of the time-integration jobs of one size class."""

    def __init__(self, key, capacity, precision, tuning=None):
        N_fm, N_r, d, dt, Pr, Tau, sym = key
        self.key = key
        self.capacity = capacity
        self.plan = _dev.Plan(_ops.build(N_fm, N_r, d, dt, Pr, Tau), symmetric=sym,
                              precision=precision, max_batch=capacity, tuning=tuning)
        self.stream = torch.cuda.Stream()
        n3 = 3 * N_fm * (N_r - 1)
        self.X = torch.zeros((capacity, n3), dtype=torch.float64, device="cuda")
        self.Ra = torch.zeros(capacity, dtype=torch.float64, device="cuda")
        self.Ra_s = torch.zeros(capacity, dtype=torch.float64, device="cuda")
        self.jobs: List[Optional[Job]] = [None] * capacity
        self.remaining = np.zeros(capacity, dtype=np.int64)
        self.n = 0
        self.bytes = class_bytes(N_fm, N_r, capacity, precision)

    def free(self):
        return self.capacity - self.n

    def admit(self, job, now):
        i = self.n
        with torch.cuda.stream(self.stream):
            self.X[i].copy_(torch.as_tensor(np.asarray(job.X0, dtype=np.float64)), non_blocking=False)
            self.Ra[i] = float(job.Ra)
            self.Ra_s[i] = float(job.Ra_s)
        self.jobs[i] = job
        self.remaining[i] = int(job.nsteps)
        self.n += 1
        job.t_admit = now

    def advance(self, max_chunk):
        """Launch one chunk for all active slots; returns instance-steps launched."""
        n = self.n
        if n == 0:
            return 0
        k = int(min(max_chunk, self.remaining[:n].min()))
        with torch.cuda.stream(self.stream):
            self.plan.step(self.X[:n], self.Ra[:n], self.Ra_s[:n], nsteps=k, out=self.X[:n])
        self.remaining[:n] -= k
        return k * n

    def harvest(self, now):
        """Finish jobs whose steps are done; compact the slots."""
        n = self.n
        done = [i for i in range(n) if self.remaining[i] == 0]
        if not done:
            return []
        self.stream.synchronize()
        finished = []
        for i in sorted(done, reverse=True):
            job = self.jobs[i]
            job.result = self.X[i].cpu().numpy()
            job.t_done = now
            finished.append(job)
            last = self.n - 1
            if i != last:
                with torch.cuda.stream(self.stream):
                    self.X[i].copy_(self.X[last])
                    self.Ra[i] = self.Ra[last]
                    self.Ra_s[i] = self.Ra_s[last]
                self.jobs[i] = self.jobs[last]
                self.remaining[i] = self.remaining[last]
            self.jobs[last] = None
            self.remaining[last] = 0
            self.n -= 1
        return finished


class NewtonWorker(threading.Thread):
    """Runs Newton and continuation jobs on its own stream."""

    def __init__(self, engine):
        super().__init__(daemon=True)
        self.engine = engine
        self.q: "queue.Queue[Optional[Job]]" = queue.Queue()
        self.busy = False
        self.plans = collections.OrderedDict()
        self.error = None

    def _krylov(self, job):
        from .newton import Krylov
        key = (job.N_fm, job.N_r, job.d, job.Pr, job.Tau, job.symmetric, job.restart)
        kr = self.plans.get(key)
        if kr is None:
            plan = _dev.Plan(_ops.build(job.N_fm, job.N_r, job.d, 1.0, job.Pr, job.Tau),
                             symmetric=job.symmetric)
            kr = Krylov(plan, restart=job.restart)
            self.plans[key] = kr
            while len(self.plans) > 4:
                _, old = self.plans.popitem(last=False)
                old.close()
        return kr

    def solve(self, job):
        kr = self._krylov(job)
        X = torch.as_tensor(np.asarray(job.X0, dtype=np.float64), device="cuda")
        if job.kind == "newton":
            Xs, info = kr.newton(X, job.Ra, job.Ra_s, tol_newton=job.tol_newton,
                                 tol_gmres=job.tol_gmres)
            return Xs.cpu().numpy(), info
        points, infos = [], []
        for ra in [job.Ra] + list(job.ra_path or []):
            X, info = kr.newton(X, ra, job.Ra_s, tol_newton=job.tol_newton,
                                tol_gmres=job.tol_gmres)
            infos.append(info)
            points.append((ra, X.cpu().numpy()))
            if not info["converged"]:
                return points, infos
        if job.arc_points > 0:
            Y = torch.cat((X.reshape(-1), torch.tensor([float(job.Ra)], dtype=torch.float64,
                                                        device=X.device)))
            Y_dot = torch.zeros_like(Y)
            ds = job.ds
            for _ in range(job.arc_points):
                Y, Y_dot, ds, info = kr.arclength_step(Y, Y_dot, job.Ra_s, job.sign, ds,
                                                       tol_newton=job.tol_newton,
                                                       tol_gmres=job.tol_gmres)
                infos.append(info)
                if info["status"] != 0:
                    break
                y = Y.cpu().numpy()
                points.append((float(y[-1]), y[:-1]))
        return points, infos

    def run(self):
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            while True:
                job = self.q.get()
                if job is None:
                    break
                self.busy = True
                try:
                    job.result, job.info = self.solve(job)
                except Exception as e:   # reported with the job, the engine keeps serving
                    job.result, job.info = None, {"exception": repr(e)}
                    self.error = e
                job.t_done = self.engine.now()
                self.engine.completed(job)
                self.busy = False


class Engine:
    def __init__(self, precision="mixed", capacity=64, hbm_budget=None, chunk=25,
                 newton_workers=2, tuning_table=None):
        self.precision = precision
        self.capacity = capacity
        self.chunk = chunk
        self.tuning_table = tuning_table
        if hbm_budget is None:
            free, total = torch.cuda.mem_get_info()
            hbm_budget = int(0.8 * free)
        self.hbm_budget = hbm_budget
        self.classes = collections.OrderedDict()
        self.workers = [NewtonWorker(self) for _ in range(newton_workers)]
        self._rr = 0
        self.done: List[Job] = []
        self._lock = threading.Lock()
        self.instance_steps = 0
        self.t0 = None

    def now(self):
        return time.perf_counter() - self.t0

    def completed(self, job):
        with self._lock:
            self.done.append(job)

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
            if self.classes[k].n == 0:
                used -= self.classes[k].bytes
                del self.classes[k]
        if used + need > self.hbm_budget:
            return None
        tuning = (self.tuning_table.lookup(key, self.capacity, self.precision)
                  if self.tuning_table else None)
        c = TimeClass(key, self.capacity, self.precision, tuning)
        self.classes[key] = c
        return c

    def run(self, jobs, realtime=True, max_seconds=None):
        """Serve jobs (arrival times in seconds from now).  With realtime=False
        the clock is virtual: every engine iteration (one chunk of steps for
        every class) advances it by one unit, and an idle engine jumps to the
        next arrival; arrival times are then in iterations."""
        for i, j in enumerate(jobs):
            if j.id < 0:
                j.id = i
        pending = collections.deque(sorted(jobs, key=lambda j: j.arrival))
        waiting = collections.deque()
        for w in self.workers:
            if not w.is_alive():
                w.start()
        self.t0 = time.perf_counter()
        virtual = 0.0
        total = len(jobs)
        while True:
            now = self.now() if realtime else virtual
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # admission, FIFO within each kind
            still = collections.deque()
            for job in waiting:
                if job.kind == "time":
                    c = self._class_for(job)
                    if c is not None and c.free() > 0:
                        c.admit(job, now)
                    else:
                        still.append(job)
                else:
                    job.t_admit = now
                    w = self.workers[self._rr % len(self.workers)]
                    self._rr += 1
                    w.q.put(job)
            waiting = still
            launched = 0
            for c in self.classes.values():
                launched += c.advance(self.chunk)
            self.instance_steps += launched
            now_done = self.now() if realtime else virtual
            for c in self.classes.values():
                for job in c.harvest(now_done):
                    self.completed(job)
            with self._lock:
                ndone = len(self.done)
            if ndone >= total:
                break
            if max_seconds is not None and self.now() > max_seconds:
                break
            if not realtime:
                if launched == 0 and pending and not waiting:
                    virtual = max(virtual + 1.0, pending[0].arrival)
                else:
                    virtual += 1.0
                if launched == 0 and not pending:
                    time.sleep(0.001)   # only worker jobs left
            elif launched == 0:
                if pending and not waiting:
                    time.sleep(max(0.0, min(0.005, pending[0].arrival - self.now())))
                else:
                    time.sleep(0.001)
        torch.cuda.synchronize()
        self.elapsed = self.now()
        return self.done

    def shutdown(self):
        for w in self.workers:
            if w.is_alive():
                w.q.put(None)
        for w in self.workers:
            if w.is_alive():
                w.join(timeout=30)

    def stats(self):
        lat = np.array([j.t_done - j.arrival for j in self.done if j.t_done is not None])
        by_kind = collections.defaultdict(list)
        for j in self.done:
            by_kind[j.kind].append(j.t_done - j.arrival)
        out = {
            "jobs": len(self.done),
            "elapsed_s": self.elapsed,
            "jobs_per_s": len(self.done) / self.elapsed if self.elapsed > 0 else 0.0,
            "instance_steps_per_s": self.instance_steps / self.elapsed if self.elapsed > 0 else 0.0,
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
    parameters and a state near a steady solution).  Continuations have
    `cont_points` points in all: natural steps of newton_case["dRa"] in Ra_T,
    or with `arclength` pseudo-arclength steps starting at ds = dRa."""
    rng = np.random.default_rng(seed)
    sizes = sizes or [(r, k) for r in (20, 30, 40) for k in (64, 128, 256, 512)]
    t, jobs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        u = rng.random()
        if u < mix[0] or newton_case is None:
            nr_, k_ = sizes[rng.integers(len(sizes))]
            x = rng.standard_normal(3 * k_ * (nr_ - 1))
            jobs.append(Job("time", k_, nr_, float(rng.uniform(*ra)), float(rng.choice(ra_s)),
                            1e-3 * x / np.linalg.norm(x), nsteps=nsteps, arrival=t))
        else:
            c = newton_case
            kind = "newton" if u < mix[0] + mix[1] else "continuation"
            ra0 = c["Ra"] * (1.0 + 0.02 * rng.uniform(-1, 1))
            path, arc = None, 0
            if kind == "continuation":
                if arclength:
                    arc = cont_points - 1
                else:
                    path = list(ra0 + c.get("dRa", 20.0) * np.arange(1, cont_points))
            jobs.append(Job(kind, c["N_fm"], c["N_r"], ra0, c["Ra_s"], c["X"], d=c["d"],
                            Pr=c["Pr"], Tau=c["Tau"], symmetric=c.get("symmetric", False),
                            ra_path=path, arc_points=arc, ds=c.get("dRa", 20.0),
                            restart=c.get("restart", 100), arrival=t))
    return jobs
