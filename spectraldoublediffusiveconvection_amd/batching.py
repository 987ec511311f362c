"""Parameter sweeps over mixed resolutions: many instances, grouped by size
class (N_fm, N_r, d, dt, Pr, Tau), advanced by one HIP graph.

Each size class owns one device plan (operators resident once) and a
contiguous batch of its instances' states; Ra_T and Ra_S are per instance.
``Sweep.graph(nsteps)`` captures nsteps of every class into a single graph in
which the classes run on forked streams (sddc_graph_create_multi), so one
launch advances the whole sweep.  Launch shapes of each class come from the
autotuner's table (autotune.py) when it has an entry.
"""
from __future__ import annotations

import ctypes
from collections import OrderedDict

import numpy as np
import torch

from . import _lib
from . import device as _dev
from . import operators as _ops


def class_key(inst):
    return (int(inst["N_fm"]), int(inst["N_r"]), float(inst["d"]), float(inst["dt"]),
            float(inst["Pr"]), float(inst["Tau"]))


class SizeClass:
    """One plan and the batch of instances of one size class."""

    def __init__(self, key, X0, Ra, Ra_s, precision="mixed", tuning=None, symmetric=False):
        N_fm, N_r, d, dt, Pr, Tau = key
        self.key = key
        self.plan = _dev.Plan(_ops.build(N_fm, N_r, d, dt, Pr, Tau), symmetric=symmetric,
                              precision=precision, max_batch=max(len(X0), 1), tuning=tuning)
        self.X = torch.as_tensor(np.ascontiguousarray(np.asarray(X0, dtype=np.float64)),
                                 device="cuda").reshape(len(X0), -1).contiguous()
        self.Ra = torch.as_tensor(np.asarray(Ra, dtype=np.float64), device="cuda")
        self.Ra_s = torch.as_tensor(np.asarray(Ra_s, dtype=np.float64), device="cuda")

    @property
    def B(self):
        return self.X.shape[0]

    def step(self, nsteps=1):
        self.plan.step(self.X, self.Ra, self.Ra_s, nsteps=nsteps, out=self.X)


class MultiGraph:
    """nsteps in-place steps of several size classes in one HIP graph."""

    def __init__(self, classes, nsteps):
        self.classes = list(classes)
        n = len(self.classes)
        V = ctypes.c_void_p
        plans = (V * n)(*[c.plan._h.value for c in self.classes])
        B = (ctypes.c_int * n)(*[c.B for c in self.classes])
        ra = (V * n)(*[c.Ra.data_ptr() for c in self.classes])
        ras = (V * n)(*[c.Ra_s.data_ptr() for c in self.classes])
        X = (V * n)(*[c.X.data_ptr() for c in self.classes])
        self.stream = torch.cuda.Stream()
        self.stream.wait_stream(torch.cuda.current_stream())
        h = V()
        _lib.check(_lib.lib().sddc_graph_create_multi(n, plans, B, ra, ras, X, int(nsteps),
                                                      V(self.stream.cuda_stream), ctypes.byref(h)))
        self._h = h
        self.nsteps = int(nsteps)

    def launch(self):
        self.stream.wait_stream(torch.cuda.current_stream())
        _lib.check(_lib.lib().sddc_graph_launch(self._h, ctypes.c_void_p(self.stream.cuda_stream)))
        torch.cuda.current_stream().wait_stream(self.stream)

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().sddc_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Sweep:
    """Instances of mixed sizes and parameters, grouped into size classes.

    instances: sequence of dicts with N_fm, N_r, d, dt, Pr, Tau, Ra, Ra_s, X0.
    """

    def __init__(self, instances, precision="mixed", tuning_table=None, symmetric=False):
        groups = OrderedDict()
        self.where = []
        for idx, inst in enumerate(instances):
            key = class_key(inst)
            g = groups.setdefault(key, [])
            self.where.append((key, len(g)))
            g.append(inst)
        self.classes = OrderedDict()
        for key, insts in groups.items():
            tuning = tuning_table.lookup(key, len(insts), precision) if tuning_table else None
            self.classes[key] = SizeClass(
                key, [i["X0"] for i in insts], [i["Ra"] for i in insts],
                [i["Ra_s"] for i in insts], precision=precision, tuning=tuning,
                symmetric=symmetric)

    @property
    def n_instances(self):
        return len(self.where)

    def step(self, nsteps=1):
        for c in self.classes.values():
            c.step(nsteps)

    def graph(self, nsteps=1):
        return MultiGraph(self.classes.values(), nsteps)

    def states(self):
        """Current states, in the order the instances were given (numpy)."""
        torch.cuda.synchronize()
        host = {k: c.X.cpu().numpy() for k, c in self.classes.items()}
        return [host[k][row] for k, row in self.where]


def random_sweep(n, seed=0, N_r=(20, 30, 40), N_fm=(64, 128, 256, 512), d=0.35, dt=5e-3,
                 Pr=1.0, Tau=1.0 / 15.0, Ra=(2500.0, 3000.0), Ra_s=(300.0, 500.0), amp=1e-3):
    """n instances with sizes drawn uniformly from N_r x N_fm, Ra_T uniform in
    Ra, Ra_S drawn from Ra_s, states seeded normal with ||X||_2 = amp."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nr_ = int(rng.choice(N_r))
        k_ = int(rng.choice(N_fm))
        n3 = 3 * k_ * (nr_ - 1)
        x = rng.standard_normal(n3)
        out.append(dict(N_fm=k_, N_r=nr_, d=d, dt=dt, Pr=Pr, Tau=Tau,
                        Ra=float(rng.uniform(*Ra)), Ra_s=float(rng.choice(Ra_s)),
                        X0=amp * x / np.linalg.norm(x)))
    return out
