"""LEF/DEF designs in and DEF out (include/eplace_io.h).

Mirrors the reference's DataBase load and store: `DataBase(input_dir)` reads every
.lef of a directory and its floorplan.def (vck5000/host/src/DataBase.cpp:7-96), and
`writeDEF(output_path)` writes `<output_path>/<design>.def` after placement
(DataBase.cpp:538-654, called from AIEplace.cpp:1123).  The parser and writer are
the library's own native code; this module only wraps them.

    d = LefDef.from_dir("designs/simple")          # or LefDef(lef_files, def_file)
    nl = d.netlist()                                # movable | fixed components | IO pins
    with Placer(nl, params) as pl:
        pl.run()
        x, y = pl.get_positions()
    d.writeDEF("out", x, y)                         # out/<design>.def
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Sequence

import numpy as np

from . import api
from .netlist import Netlist


class LefDef:
    """A design read from LEF and DEF files; owns the native handle."""

    def __init__(self, lef_paths: Sequence[str], def_path: str, _handle=None):
        self._h = C.c_void_p()
        if _handle is not None:
            self._h = _handle
        else:
            paths = [os.fsencode(p) for p in lef_paths]
            arr = (C.c_char_p * max(1, len(paths)))(*paths)
            api.check(api.lib.ep_lefdef_read(arr, len(paths), os.fsencode(def_path), C.byref(self._h)))
        info = api.EpLefdefInfo()
        name = C.create_string_buffer(256)
        api.check(api.lib.ep_lefdef_get_info(self._h, C.byref(info), name))
        self.name = name.value.decode()
        self.info = {k: getattr(info, k) for k, _ in api.EpLefdefInfo._fields_ if k != "die"}
        self.info["die"] = tuple(info.die)

    @classmethod
    def from_dir(cls, path: str) -> "LefDef":
        h = C.c_void_p()
        api.check(api.lib.ep_lefdef_read_dir(os.fsencode(path), C.byref(h)))
        return cls([], "", _handle=h)

    def netlist(self) -> Netlist:
        """The placer's netlist view (arrays copied out of the handle)."""
        v = api.EpNetlist()
        api.check(api.lib.ep_lefdef_netlist(self._h, C.byref(v)))
        n_all = v.n_mov + v.n_fix

        def take(ptr, n, dtype):
            if n == 0:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)

        return Netlist(w=take(v.w, n_all, np.float32), h=take(v.h, n_all, np.float32),
                       x=take(v.x, n_all, np.float32), y=take(v.y, n_all, np.float32), n_mov=v.n_mov,
                       net_start=take(v.net_start, v.n_nets + 1, np.int32),
                       pin_cell=take(v.pin_cell, v.n_pins, np.int32),
                       pin_dx=take(v.pin_dx, v.n_pins, np.float32), pin_dy=take(v.pin_dy, v.n_pins, np.float32),
                       die=(v.xl, v.yl, v.xh, v.yh))

    def writeDEF(self, output_path: str, x: np.ndarray, y: np.ndarray) -> str:
        """Writes `<output_path>/<design>.def` with the movable components at x, y
        (the first n_mov entries: ep_get_positions order, fillers after them)."""
        n = self.info["n_mov"]
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float32)[:n])
        y = np.ascontiguousarray(np.asarray(y, dtype=np.float32)[:n])
        if len(x) < n or len(y) < n:
            raise api.EplaceError(api.EP_EINVAL, f"writeDEF: need {n} positions, got {len(x)}")
        os.makedirs(output_path, exist_ok=True)
        api.check(api.lib.ep_lefdef_write_def(self._h, api.fptr(x), api.fptr(y), os.fsencode(output_path)))
        return os.path.join(output_path, self.name + ".def")

    def close(self):
        if self._h:
            api.check(api.lib.ep_lefdef_free(self._h))
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
