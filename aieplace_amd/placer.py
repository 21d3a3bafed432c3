"""Placer: the reference's per-iteration placer API over the HIP C ABI.

Mirrors class Placer of vck5000/host/src/include/AIEplace.h:21-99 (method names
kept): computeAllPartials -> WA gradient kernel, computeOverlaps -> density
scatter, computeElectricFields -> MFMA transform field solve, nudgeAllNodes ->
one Nesterov step with backtracking, performIteration/run/printIterationResults
as in AIEplace.cpp:9-76, 1048-1070.  Errors raise EplaceError (the reference
logs and exits, AIEplace.cpp:95-98).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import api
from .netlist import Netlist


def hpwl_at_overflow(overflow, hpwl, target: float = 0.10, window: int = 12) -> float:
    """HPWL of a run at overflow exactly `target`, from its per-iteration log:
    log HPWL fitted linearly in the overflow over the last `window` iterations
    (the run stops at the first iteration with overflow <= target, so they
    straddle it).  Near the stop HPWL grows ~0.2% per iteration while the
    overflow falls by ~0.0005, so the final HPWL depends on which iteration
    crosses the target first; the value at matched overflow does not."""
    overflow = np.asarray(overflow, np.float64)
    hpwl = np.asarray(hpwl, np.float64)
    a = max(0, len(hpwl) - window)
    c = np.polyfit(overflow[a:], np.log(hpwl[a:]), 1)
    return float(np.exp(np.polyval(c, target)))


class Placer:
    def __init__(self, netlist: Netlist, params: Optional[api.EpParams] = None, **kw):
        self.netlist = netlist
        self.params = params if params is not None else api.default_params(**kw)
        self._ctx = C.c_void_p()
        api.check(api.lib.ep_create(C.byref(netlist.c_struct()), C.byref(self.params),
                                    C.byref(self._ctx)))
        sizes = [C.c_int32() for _ in range(5)]
        api.check(api.lib.ep_get_sizes(self._ctx, *[C.byref(s) for s in sizes]))
        self.n_mov, self.n_fill, self.n_fix, self.n_nets, self.n_pins = (s.value for s in sizes)
        self.n_place = self.n_mov + self.n_fill
        self.bins = (self.params.bins_x, self.params.bins_y)
        self.iteration = 0
        self.history = []

    # ------------------------------------------------------------ lifecycle
    def close(self):
        if self._ctx:
            api.check(api.lib.ep_destroy(self._ctx))
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ state
    def cells(self):
        n_all = self.n_place + self.n_fix
        w = np.zeros(n_all, np.float32)
        h = np.zeros(n_all, np.float32)
        fx = np.zeros(max(self.n_fix, 1), np.float32)
        fy = np.zeros(max(self.n_fix, 1), np.float32)
        api.check(api.lib.ep_get_cells(self._ctx, api.fptr(w), api.fptr(h), api.fptr(fx), api.fptr(fy)))
        return w, h, fx[:self.n_fix], fy[:self.n_fix]

    def nets(self):
        ns = np.zeros(self.n_nets + 1, np.int32)
        pc = np.zeros(max(self.n_pins, 1), np.int32)
        dx = np.zeros(max(self.n_pins, 1), np.float32)
        dy = np.zeros(max(self.n_pins, 1), np.float32)
        perm = np.zeros(max(self.n_nets, 1), np.int32)
        api.check(api.lib.ep_get_nets(self._ctx, api.iptr(ns), api.iptr(pc), api.fptr(dx),
                                      api.fptr(dy), api.iptr(perm)))
        return ns, pc[:self.n_pins], dx[:self.n_pins], dy[:self.n_pins], perm[:self.n_nets]

    def get_positions(self):
        x = np.zeros(self.n_place, np.float32)
        y = np.zeros(self.n_place, np.float32)
        api.check(api.lib.ep_get_positions(self._ctx, api.fptr(x), api.fptr(y)))
        return x, y

    def set_positions(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        assert len(x) == self.n_place and len(y) == self.n_place
        api.check(api.lib.ep_set_positions(self._ctx, api.fptr(x), api.fptr(y)))

    def set_precision(self, precision: int):
        api.check(api.lib.ep_set_precision(self._ctx, int(precision)))

    # ------------------------------------------------------------ per-iteration calls
    def computeAllPartials(self, gamma: float):
        """WA gradient per cell (AIEplace.cpp:256-381 / 574-673). Returns (gx, gy, wa, hpwl)."""
        wa, hp = C.c_double(), C.c_double()
        api.check(api.lib.ep_wa_grad(self._ctx, float(gamma), C.byref(wa), C.byref(hp)))
        gx = np.zeros(self.n_place, np.float32)
        gy = np.zeros(self.n_place, np.float32)
        api.check(api.lib.ep_get_wa_grad(self._ctx, api.fptr(gx), api.fptr(gy)))
        return gx, gy, wa.value, hp.value

    def computeOverlaps(self):
        """Density maps (AIEplace.cpp:937-967). Returns (field_map, mov_map, fix_map, overflow)."""
        ov = C.c_double()
        api.check(api.lib.ep_density(self._ctx, C.byref(ov)))
        nb = self.bins[0] * self.bins[1]
        maps = [np.zeros(nb, np.float32) for _ in range(3)]
        api.check(api.lib.ep_get_density(self._ctx, *[api.fptr(m) for m in maps]))
        return tuple(m.reshape(self.bins) for m in maps) + (ov.value,)

    def computeElectricFields(self, precision: Optional[int] = None):
        """Field of the current density map (AIEplace.cpp:388-561 / 679-934). Returns (ex, ey)."""
        prec = self.params.precision if precision is None else precision
        api.check(api.lib.ep_field(self._ctx, int(prec)))
        nb = self.bins[0] * self.bins[1]
        ex = np.zeros(nb, np.float32)
        ey = np.zeros(nb, np.float32)
        api.check(api.lib.ep_get_field(self._ctx, api.fptr(ex), api.fptr(ey)))
        return ex.reshape(self.bins), ey.reshape(self.bins)

    def field_error(self, precision: int):
        """(err_ex, err_ey): max-norm error of the field in `precision` vs an FP64 device solve."""
        ex, ey = C.c_double(), C.c_double()
        api.check(api.lib.ep_field_error(self._ctx, int(precision), C.byref(ex), C.byref(ey)))
        return ex.value, ey.value

    def gradient(self, gamma: float, lam: float):
        n = self.n_place
        out = {k: np.zeros(n, np.float32) for k in ("gx", "gy", "gwx", "gwy", "fx", "fy")}
        hp, ov = C.c_double(), C.c_double()
        api.check(api.lib.ep_gradient(self._ctx, float(gamma), float(lam),
                                      *[api.fptr(out[k]) for k in ("gx", "gy", "gwx", "gwy", "fx", "fy")],
                                      C.byref(hp), C.byref(ov)))
        out["hpwl"], out["tau"] = hp.value, ov.value
        return out

    def cell_bins(self, stretched: bool = True):
        out = np.zeros(4 * self.n_place, np.int32)
        api.check(api.lib.ep_cell_bins(self._ctx, int(stretched), api.iptr(out)))
        return out.reshape(-1, 4)

    def hpwl(self) -> float:
        v = C.c_double()
        api.check(api.lib.ep_hpwl(self._ctx, C.byref(v)))
        return v.value

    def overflow(self) -> float:
        v = C.c_double()
        api.check(api.lib.ep_overflow(self._ctx, C.byref(v)))
        return v.value

    # ------------------------------------------------------------ learned step
    def set_step_model(self, weights):
        w = np.ascontiguousarray(weights, np.float32)
        api.check(api.lib.ep_set_step_model(self._ctx, api.fptr(w), len(w)))

    def take_step_records(self):
        """(features[n,16], log10 accepted step[n]) recorded since the last call."""
        n = C.c_int32()
        api.check(api.lib.ep_take_step_records(self._ctx, None, None, 0, C.byref(n)))
        f = np.zeros((max(n.value, 1), api.MLP_IN), np.float32)
        t = np.zeros(max(n.value, 1), np.float32)
        api.check(api.lib.ep_take_step_records(self._ctx, api.fptr(f), api.fptr(t), n.value, C.byref(n)))
        return f[:n.value], t[:n.value]

    def profile(self, reps: int = 20):
        stats = (api.EpKernelStat * 16)()
        n = C.c_int32()
        api.check(api.lib.ep_profile(self._ctx, int(reps), stats, 16, C.byref(n)))
        return [dict(name=s.name.decode(), launches=s.launches, avg_us=s.avg_us, bytes=s.bytes,
                     flops=s.flops) for s in stats[:n.value]]

    # ------------------------------------------------------------ optimiser
    def initializePlacement(self):
        api.check(api.lib.ep_init_optimizer(self._ctx))

    def nudgeAllNodes(self) -> dict:
        si = api.EpStepInfo()
        api.check(api.lib.ep_step(self._ctx, C.byref(si)))
        rec = dict(iteration=si.iteration, evals=si.evals, hpwl=si.hpwl, overflow=si.overflow,
                   lam=si.lambda_, gamma=si.gamma, alpha=si.alpha, wa=si.wa,
                   learned_used=si.learned_used, learned_accepted=si.learned_accepted)
        self.history.append(rec)
        self.iteration = si.iteration + 1
        return rec

    def performIteration(self) -> dict:
        return self.nudgeAllNodes()

    def printIterationResults(self, rec: Optional[dict] = None):
        rec = rec or (self.history[-1] if self.history else None)
        if rec:
            print(f"Iteration {rec['iteration']:5d}  HPWL {rec['hpwl']:.6e}  Overflow {rec['overflow']:.4f}")

    def run(self, log_cap: int = 0) -> dict:
        ri = api.EpRunInfo()
        logs = (api.EpStepInfo * max(log_cap, 1))()
        api.check(api.lib.ep_run(self._ctx, C.byref(ri), logs, int(log_cap)))
        out = dict(iterations=ri.iterations, evals=ri.evals, hpwl=ri.hpwl, overflow=ri.overflow,
                   seconds=ri.seconds, learned_proposals=ri.learned_proposals,
                   learned_accepted=ri.learned_accepted)
        out["log"] = [dict(iteration=l.iteration, evals=l.evals, hpwl=l.hpwl, overflow=l.overflow,
                           lam=l.lambda_, gamma=l.gamma, alpha=l.alpha)
                      for l in logs[:min(log_cap, ri.iterations)]]
        return out

    def run_with_report(self, report, design=None, max_iters: Optional[int] = None) -> dict:
        """Placer::run with its operator output (AIEplace.cpp:44-76, 1048-1124): one
        host-controlled Nesterov step per iteration (ep_step), each reported to
        `report` (an aieplace_amd.report.RunReport: table row, snapshot every
        report.snapshot_every iterations), then statistics.md, the final snapshot
        and, for a LEF/DEF design, its DEF."""
        import time
        limit = self.params.max_iters if max_iters is None else max_iters
        t0 = time.perf_counter()
        self.initializePlacement()
        rec, trials = None, 0
        while self.iteration < limit:
            rec = self.nudgeAllNodes()
            trials += rec["evals"]
            report.iteration(self, rec)
            if rec["overflow"] <= self.params.stop_overflow:
                break
        # evals: the Nesterov trials' gradient evaluations (the set-up's are not counted)
        res = dict(iterations=self.iteration, evals=trials,
                   hpwl=rec["hpwl"] if rec else self.hpwl(), overflow=rec["overflow"] if rec else self.overflow(),
                   seconds=time.perf_counter() - t0)
        res["statistics"] = report.final(self, res, self.params, design)
        return res

    def printFinalResults(self, res: dict):
        print(f"Iterations {res['iterations']}  Final HPWL {res['hpwl']:.6e}  "
              f"Final Overflow {res['overflow']:.4f}  ({res['seconds']:.3f} s)")
