"""Row legalisation after global placement (include/eplace_legal.h, native Abacus).

    x, y, info = legalize(netlist, x, y)            # movable lower-left corners in, legal ones out

Row height defaults to the movable cells' most common height, the site width
to the largest width that divides every movable width (for integer-site cells:
the smallest common site).  No counterpart in the reference, whose flow ends at
global placement (vck5000/host/src/AIEplace.cpp:44-76).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import api
from .netlist import Netlist


def default_grid(nl: Netlist):
    """(row height, site width) for the movable cells of nl."""
    h = nl.h[:nl.n_mov].astype(np.float64)
    w = nl.w[:nl.n_mov].astype(np.float64)
    vals, counts = np.unique(np.round(h, 6), return_counts=True)
    row = float(vals[np.argmax(counts)])
    site = float(w.min())
    # largest divisor of every width on a 1e-3 grid
    g = 0
    for v in np.unique(np.round(w * 1000).astype(np.int64)):
        g = int(np.gcd(g, int(v)))
    if g > 0:
        site = g / 1000.0
    return row, site


def legalize(nl: Netlist, x, y, row_height: Optional[float] = None, site_width: Optional[float] = None,
             snap_to_sites: bool = True):
    """Legal movable positions from global ones (the first n_mov entries of x, y are read)."""
    r, s = default_grid(nl)
    row_height = r if row_height is None else row_height
    site_width = s if site_width is None else site_width
    n = nl.n_mov
    xi = np.ascontiguousarray(np.asarray(x, np.float32)[:n])
    yi = np.ascontiguousarray(np.asarray(y, np.float32)[:n])
    if len(xi) < n or len(yi) < n:
        raise api.EplaceError(api.EP_EINVAL, f"legalize: need {n} positions")
    xo = np.zeros(n, np.float32)
    yo = np.zeros(n, np.float32)
    info = api.EpLegalInfo()
    api.check(api.lib.ep_legalize(C.byref(nl.c_struct()), api.fptr(xi), api.fptr(yi), float(row_height),
                                  float(site_width), int(bool(snap_to_sites)), api.fptr(xo), api.fptr(yo),
                                  C.byref(info)))
    return xo, yo, {k: getattr(info, k) for k, _ in api.EpLegalInfo._fields_}
