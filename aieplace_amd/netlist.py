"""Netlists as numpy structure-of-arrays, and the seeded synthetic generators.

Replaces the reference's DataBase netlist load (vck5000/host/src/DataBase.cpp:7-96)
for synthetic workloads; the generator itself is native code (generator.cpp)
reached through the C ABI (include/eplace_gen.h).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import api


@dataclass
class Netlist:
    """Cells are [movable | fixed]; pins reference cells by index."""

    w: np.ndarray
    h: np.ndarray
    x: np.ndarray
    y: np.ndarray
    n_mov: int
    net_start: np.ndarray
    pin_cell: np.ndarray
    pin_dx: np.ndarray
    pin_dy: np.ndarray
    die: tuple

    def __post_init__(self):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        i = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        self.w, self.h, self.x, self.y = f(self.w), f(self.h), f(self.x), f(self.y)
        self.pin_dx, self.pin_dy = f(self.pin_dx), f(self.pin_dy)
        self.net_start, self.pin_cell = i(self.net_start), i(self.pin_cell)
        self.die = tuple(float(v) for v in self.die)

    @property
    def n_fix(self) -> int:
        return len(self.w) - self.n_mov

    @property
    def n_nets(self) -> int:
        return len(self.net_start) - 1

    @property
    def n_pins(self) -> int:
        return len(self.pin_cell)

    def c_struct(self) -> api.EpNetlist:
        nl = api.EpNetlist()
        nl.n_mov, nl.n_fix = self.n_mov, self.n_fix
        nl.w, nl.h = api.fptr(self.w), api.fptr(self.h)
        nl.x, nl.y = api.fptr(self.x), api.fptr(self.y)
        nl.n_nets, nl.n_pins = self.n_nets, self.n_pins
        nl.net_start, nl.pin_cell = api.iptr(self.net_start), api.iptr(self.pin_cell)
        nl.pin_dx, nl.pin_dy = api.fptr(self.pin_dx), api.fptr(self.pin_dy)
        nl.xl, nl.yl, nl.xh, nl.yh = self.die
        self._keep = nl  # arrays stay referenced by self
        return nl


def plan_fillers(nl: Netlist, params: api.EpParams):
    """Fillers ep_create will add: (n_fill, fw, fh, x, y).  Host only."""
    n = C.c_int32()
    fw, fh = C.c_float(), C.c_float()
    cs = nl.c_struct()
    api.check(api.lib.ep_plan_fillers(C.byref(cs), C.byref(params), C.byref(n), C.byref(fw),
                                      C.byref(fh), None, None, 0))
    x = np.zeros(max(n.value, 1), np.float32)
    y = np.zeros(max(n.value, 1), np.float32)
    api.check(api.lib.ep_plan_fillers(C.byref(cs), C.byref(params), C.byref(n), C.byref(fw),
                                      C.byref(fh), api.fptr(x), api.fptr(y), n.value))
    return n.value, fw.value, fh.value, x[:n.value], y[:n.value]


def generate(kind: str = "stdcell", seed: int = 1, **overrides) -> Netlist:
    """Seeded synthetic instance: kind 'mesh' (24x24 lattice) or 'stdcell'."""
    k = api.EP_GEN_MESH if kind == "mesh" else api.EP_GEN_STDCELL
    gp = api.EpGenParams()
    api.lib.ep_gen_default_params(C.byref(gp), k)
    gp.seed = seed
    for key, v in overrides.items():
        if not hasattr(gp, key):
            raise KeyError(key)
        setattr(gp, key, v)
    dp = C.POINTER(api.EpDesign)()
    api.check(api.lib.ep_gen_create(C.byref(gp), C.byref(dp)))
    try:
        nl = dp.contents.nl
        n_all = nl.n_mov + nl.n_fix
        cp = lambda p, n, t: np.ctypeslib.as_array(p, shape=(max(n, 0),)).astype(t, copy=True) if n > 0 else np.zeros(0, t)
        out = Netlist(
            w=cp(nl.w, n_all, np.float32), h=cp(nl.h, n_all, np.float32),
            x=cp(nl.x, n_all, np.float32), y=cp(nl.y, n_all, np.float32), n_mov=nl.n_mov,
            net_start=cp(nl.net_start, nl.n_nets + 1, np.int32),
            pin_cell=cp(nl.pin_cell, nl.n_pins, np.int32),
            pin_dx=cp(nl.pin_dx, nl.n_pins, np.float32), pin_dy=cp(nl.pin_dy, nl.n_pins, np.float32),
            die=(nl.xl, nl.yl, nl.xh, nl.yh))
    finally:
        api.lib.ep_gen_free(dp)
    return out
