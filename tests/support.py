"""Shared helpers: oracle designs built from the same netlist the device gets."""
from __future__ import annotations

import numpy as np

from aieplace_amd import api
from aieplace_amd.netlist import Netlist, plan_fillers
from aieplace_amd.placer import hpwl_at_overflow  # noqa: F401  (shared with bench.py)


def oracle_design(O, nl: Netlist, params: api.EpParams):
    """oracle.Design with the fillers ep_create adds; returns (design, x0, y0)."""
    n_fill, fw, fh, fx, fy = plan_fillers(nl, params)
    n_mov, n_fix = nl.n_mov, nl.n_fix
    w = np.concatenate([nl.w[:n_mov], np.full(n_fill, fw, np.float32), nl.w[n_mov:]])
    h = np.concatenate([nl.h[:n_mov], np.full(n_fill, fh, np.float32), nl.h[n_mov:]])
    pc = nl.pin_cell.copy()
    pc[pc >= n_mov] += n_fill
    d = O.Design(w=w.astype(np.float64), h=h.astype(np.float64), n_mov=n_mov, n_fill=n_fill,
                 fix_x=nl.x[n_mov:], fix_y=nl.y[n_mov:], net_start=nl.net_start, pin_cell=pc,
                 pin_dx=nl.pin_dx, pin_dy=nl.pin_dy, die=nl.die,
                 bins=(params.bins_x, params.bins_y))
    x0 = np.concatenate([nl.x[:n_mov], fx]).astype(np.float32)
    y0 = np.concatenate([nl.y[:n_mov], fy]).astype(np.float32)
    # same clamp into the die as ep_create
    x0 = np.minimum(np.maximum(x0, np.float32(nl.die[0])), np.float32(nl.die[2]) - w[:n_mov + n_fill])
    y0 = np.minimum(np.maximum(y0, np.float32(nl.die[1])), np.float32(nl.die[3]) - h[:n_mov + n_fill])
    return d, x0.astype(np.float32), y0.astype(np.float32)


def oracle_params(params: api.EpParams, bw: float, n_nets: int) -> dict:
    # ep_create's default when hpwl_ref <= 0: 0.05 bin widths per net
    href = params.hpwl_ref if params.hpwl_ref > 0 else 0.05 * max(1, n_nets) * np.float32(bw)
    return dict(target_density=params.target_density, gamma_bins=params.gamma_bins,
                gamma_schedule=params.gamma_schedule, lambda_eta=params.lambda_eta,
                hpwl_ref=float(href), mu_max=params.mu_max, mu_min=params.mu_min,
                max_backtrack=params.max_backtrack, stop_overflow=params.stop_overflow,
                max_iters=params.max_iters, init_disp_bins=params.init_disp_bins)


def nudged_start(x, member: int):
    """Start `member` of an ensemble: member 0 is x itself; member m > 0 moves a
    seeded half of the coordinates (rng seed m) to the next FP32 value up."""
    x = np.array(x, dtype=np.float32)
    if member > 0:
        m = np.random.default_rng(member).random(len(x)) < 0.5
        x[m] = np.nextafter(x[m], np.float32(np.inf))
    return x


def netlist_lefdef(nl, units=1000, name="gen"):
    """LEF (one macro per cell) and DEF text for a netlist; cells [movable | fixed]."""
    lef, comps, nets = [], [], []
    q = lambda v: int(round(float(v) * units))
    n_all = len(nl.w)
    pin_names = {}
    for n in range(nl.n_nets):
        for p in range(nl.net_start[n], nl.net_start[n + 1]):
            pin_names.setdefault(int(nl.pin_cell[p]), []).append(p)
    for k in range(n_all):
        lef.append(f"MACRO M{k}\n  SIZE {float(nl.w[k])!r} BY {float(nl.h[k])!r} ;")
        for p in pin_names.get(k, []):
            dx, dy = float(nl.pin_dx[p]), float(nl.pin_dy[p])
            lef.append(f"  PIN P{p} PORT LAYER m1 ; RECT {dx - 0.01!r} {dy - 0.01!r} {dx + 0.01!r} {dy + 0.01!r} ; "
                       f"END END P{p}")
        lef.append(f"END M{k}")
        st = "PLACED" if k < nl.n_mov else "FIXED"
        comps.append(f"- c{k} M{k} + {st} ( {q(nl.x[k])} {q(nl.y[k])} ) N ;")
    for n in range(nl.n_nets):
        conn = " ".join(f"( c{nl.pin_cell[p]} P{p} )" for p in range(nl.net_start[n], nl.net_start[n + 1]))
        nets.append(f"- net{n} {conn} ;")
    die = " ".join(str(q(v)) for v in nl.die)
    dx0, dy0, dx1, dy1 = die.split()
    d = (f"DESIGN {name} ;\nUNITS DISTANCE MICRONS {units} ;\nDIEAREA ( {dx0} {dy0} ) ( {dx1} {dy1} ) ;\n"
         f"COMPONENTS {n_all} ;\n" + "\n".join(comps) + "\nEND COMPONENTS\n"
         f"NETS {nl.n_nets} ;\n" + "\n".join(nets) + "\nEND NETS\nEND DESIGN\n")
    return "\n".join(lef) + "\nEND LIBRARY\n", d


def round_half_away(v):
    """Rounding to the nearest integer, ties away from zero (C llround): how
    positions are converted to DEF database units."""
    return np.sign(v) * np.floor(np.abs(v) + 0.5)


def check_legal(nl, x, y, row, site, tol=1e-4):
    n = nl.n_mov
    xl, yl, xh, yh = nl.die
    w, h = nl.w[:n].astype(np.float64), nl.h[:n].astype(np.float64)
    x, y = np.asarray(x, np.float64), np.asarray(y, np.float64)
    r = (y - yl) / row
    s = (x - xl) / site
    assert np.abs(r - np.rint(r)).max() < tol, "off the rows"
    assert np.abs(s - np.rint(s)).max() < tol, "off the sites"
    assert x.min() >= xl - tol and (x + w).max() <= xh + tol and y.min() >= yl - tol and (y + h).max() <= yh + tol
    rows = np.rint(r).astype(int)
    for rr in np.unique(rows):
        idx = np.where(rows == rr)[0]
        idx = idx[np.argsort(x[idx])]
        assert (x[idx[1:]] >= x[idx[:-1]] + w[idx[:-1]] - tol).all(), f"overlap in row {rr}"
    fx, fy = nl.x[n:].astype(np.float64), nl.y[n:].astype(np.float64)
    fw, fh = nl.w[n:].astype(np.float64), nl.h[n:].astype(np.float64)
    for k in range(len(fx)):
        ox = np.minimum(x + w, fx[k] + fw[k]) - np.maximum(x, fx[k])
        oy = np.minimum(y + h, fy[k] + fh[k]) - np.maximum(y, fy[k])
        assert not ((ox > tol) & (oy > tol)).any(), f"overlap with fixed cell {k}"
