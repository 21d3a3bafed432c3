"""BatchPlacer: many independent placements advanced together (include/eplace_batch.h).

Each instance runs the same ePlace flow as Placer.run (the reference's
Placer::run, AIEplace.cpp:44-76); the batch shares one device context, one
captured hipGraph per group of Nesterov trials and a device-side optimiser
control, so small instances stop paying per-launch host overhead.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import api
from .netlist import Netlist, generate


class BatchPlacer:
    def __init__(self, netlists: Sequence[Netlist], bins: Sequence[int],
                 params: Optional[api.EpParams] = None, **kw):
        if len(netlists) != len(bins):
            raise ValueError("one bin grid per netlist")
        self.netlists = list(netlists)
        self.bins = [int(b) for b in bins]
        self.params = params if params is not None else api.default_params(**kw)
        n = len(self.netlists)
        structs = [nl.c_struct() for nl in self.netlists]  # keep alive during create
        arr = (C.POINTER(api.EpNetlist) * n)(*[C.pointer(s) for s in structs])
        b = np.asarray(self.bins, np.int32)
        self._b = C.c_void_p()
        api.check(api.lib.ep_batch_create(arr, api.iptr(b), n, C.byref(self.params),
                                          C.byref(self._b)))
        self.n = n

    def close(self):
        if self._b:
            api.check(api.lib.ep_batch_destroy(self._b))
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def reset(self):
        api.check(api.lib.ep_batch_reset(self._b))

    def init(self):
        api.check(api.lib.ep_batch_init(self._b))

    def run(self, max_launches: int = 1 << 20) -> dict:
        info = api.EpBatchRunInfo()
        api.check(api.lib.ep_batch_run(self._b, max_launches, C.byref(info)))
        return {k: getattr(info, k) for k, _ in api.EpBatchRunInfo._fields_}

    def results(self) -> dict:
        it = np.zeros(self.n, np.int32)
        ev = np.zeros(self.n, np.int64)
        hp = np.zeros(self.n, np.float64)
        ov = np.zeros(self.n, np.float64)
        api.check(api.lib.ep_batch_results(
            self._b, api.iptr(it), ev.ctypes.data_as(C.POINTER(C.c_int64)),
            hp.ctypes.data_as(C.POINTER(C.c_double)), ov.ctypes.data_as(C.POINTER(C.c_double))))
        return {"iterations": it, "evals": ev, "hpwl": hp, "overflow": ov}

    def sizes(self, i: int):
        s = [C.c_int32() for _ in range(3)]
        api.check(api.lib.ep_batch_sizes(self._b, i, *[C.byref(v) for v in s]))
        return tuple(v.value for v in s)

    def _pairs(self, fn, i: int):
        n_mov, n_fill, _ = self.sizes(i)
        x = np.zeros(n_mov + n_fill, np.float32)
        y = np.zeros(n_mov + n_fill, np.float32)
        api.check(fn(self._b, i, api.fptr(x), api.fptr(y)))
        return x, y

    def positions(self, i: int):
        return self._pairs(api.lib.ep_batch_positions, i)

    def gradient(self, i: int):
        return self._pairs(api.lib.ep_batch_gradient, i)

    def state(self, i: int) -> dict:
        g, l, a = C.c_double(), C.c_double(), C.c_double()
        api.check(api.lib.ep_batch_state(self._b, i, C.byref(g), C.byref(l), C.byref(a)))
        return {"gamma": g.value, "lambda": l.value, "alpha": a.value}


def batch_cell_counts(n: int = 64, lo: int = 5_000, hi: int = 150_000, seed: int = 1) -> list:
    """Cell counts log-uniform in [lo, hi] (the batched workload's instance sizes)."""
    rng = np.random.default_rng(seed)
    return [int(round(v)) for v in np.exp(rng.uniform(np.log(lo), np.log(hi), size=n))]


def batch_workload(n: int = 64, lo: int = 5_000, hi: int = 150_000, seed: int = 1):
    """n standard-cell instances (seeds 1..n) with bin grids matched to size."""
    from .configs import bins_for_cells

    counts = batch_cell_counts(n, lo, hi, seed)
    nls = [generate("stdcell", seed=k + 1, n_movable=c) for k, c in enumerate(counts)]
    return nls, [bins_for_cells(c) for c in counts]
