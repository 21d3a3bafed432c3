"""Row legalisation (include/eplace_legal.h, Abacus): host code, no GPU.

Pins: legality by exact geometric checks (rows, sites, die, no overlaps, no
fixed-cell overlap); a legal input is a fixed point; within one row the
cluster solution equals the minimum of the summed squared displacement for the
given order, found here by a general constrained solver (scipy SLSQP).
"""
import numpy as np
import pytest
from scipy.optimize import minimize

import aieplace_amd as A
from aieplace_amd import api
from aieplace_amd.legalize import default_grid, legalize
from aieplace_amd.netlist import Netlist
from support import check_legal


def spread(nl, seed, frac=0.25):
    """Global-placement-like positions: uniform over the die, cells overlapping."""
    rng = np.random.default_rng(seed)
    xl, yl, xh, yh = nl.die
    n = nl.n_mov
    x = rng.uniform(xl, xh - nl.w[:n]).astype(np.float32)
    y = rng.uniform(yl, yh - nl.h[:n]).astype(np.float32)
    return x, y


def tetris(nl, x, y):
    """Greedy baseline without fixed cells: cells in x order, each to the row
    whose frontier (first free site) gives the smallest squared displacement."""
    n = nl.n_mov
    W = int(nl.die[2] - nl.die[0])
    rows = int(nl.die[3] - nl.die[1])
    front = np.zeros(rows)
    r = np.arange(rows)
    sq = 0.0
    for i in np.argsort(x, kind="stable"):
        w = float(nl.w[i])
        pos = np.maximum(front, np.rint(x[i] - nl.die[0]))
        pos = np.where(pos + w <= W, pos, np.inf)
        cost = (pos - (x[i] - nl.die[0])) ** 2 + (r - (y[i] - nl.die[1])) ** 2
        k = int(np.argmin(cost))
        sq += cost[k]
        front[k] = pos[k] + w
    return sq


def test_generated_design_is_legalised():
    nl = A.generate("stdcell", seed=2, n_movable=3000)
    row, site = default_grid(nl)
    assert (row, site) == (1.0, 1.0)
    x, y = spread(nl, 1)
    lx, ly, info = legalize(nl, x, y)
    check_legal(nl, lx, ly, row, site)
    assert info["rows"] == int((nl.die[3] - nl.die[1]) / row) and info["segments"] >= info["rows"]
    # without the fixed block: Abacus moves cells less than the greedy baseline
    free = Netlist(w=nl.w[:nl.n_mov], h=nl.h[:nl.n_mov], x=nl.x[:nl.n_mov], y=nl.y[:nl.n_mov], n_mov=nl.n_mov,
                   net_start=np.array([0]), pin_cell=np.zeros(0), pin_dx=np.zeros(0), pin_dy=np.zeros(0),
                   die=nl.die)
    fx, fy, finfo = legalize(free, x, y)
    check_legal(free, fx, fy, row, site)
    base = tetris(free, x, y)
    print(f"sum sq displacement: abacus {finfo['sum_sq_disp']:.4g}, greedy {base:.4g}")
    assert finfo["sum_sq_disp"] < base
    # a legal placement is a fixed point
    lx2, ly2, info2 = legalize(nl, lx, ly)
    np.testing.assert_array_equal(lx2, lx)
    np.testing.assert_array_equal(ly2, ly)
    assert info2["total_disp"] == 0.0


def one_row(widths, x0, sites):
    n = len(widths)
    w = np.asarray(widths, np.float32)
    return Netlist(w=w, h=np.ones(n, np.float32), x=np.asarray(x0, np.float32), y=np.zeros(n, np.float32),
                   n_mov=n, net_start=np.array([0]), pin_cell=np.zeros(0), pin_dx=np.zeros(0),
                   pin_dy=np.zeros(0), die=(0.0, 0.0, float(sites), 1.0))


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_single_row_is_the_least_squares_optimum(seed):
    rng = np.random.default_rng(seed)
    n, sites = 12, 40
    widths = rng.integers(1, 4, n)
    x0 = np.sort(rng.uniform(-3.0, sites, n)).astype(np.float32)  # some cells start left of the die
    nl = one_row(widths, x0, sites)
    lx, ly, _ = legalize(nl, x0, np.zeros(n, np.float32), 1.0, 1.0, snap_to_sites=False)
    assert (ly == 0).all()
    order = np.argsort(x0, kind="stable")
    cost = float(((lx.astype(np.float64) - x0) ** 2).sum())
    w = widths[order].astype(np.float64)
    cons = [{"type": "ineq", "fun": (lambda v, i=i: v[i + 1] - v[i] - w[i])} for i in range(n - 1)]
    cons += [{"type": "ineq", "fun": lambda v: v[0]}, {"type": "ineq", "fun": lambda v: sites - v[-1] - w[-1]}]
    x00 = x0[order].astype(np.float64)
    start = np.concatenate([[0.0], np.cumsum(w[:-1])])
    res = minimize(lambda v: ((v - x00) ** 2).sum(), start, jac=lambda v: 2 * (v - x00), constraints=cons,
                   method="SLSQP", options=dict(ftol=1e-12, maxiter=500))
    assert res.success, res.message
    assert cost == pytest.approx(res.fun, rel=1e-6, abs=1e-8)
    check_legal(nl, np.rint(lx), ly, 1.0, 1.0)  # after snapping the same order is legal


def test_fixed_cells_block_rows():
    nl0 = A.generate("stdcell", seed=4, n_movable=1500)
    # a fixed block in the middle of the die on top of the generator's macros
    xl, yl, xh, yh = nl0.die
    bx, by = 0.4 * (xh - xl), 0.3 * (yh - yl)
    nl = Netlist(w=np.append(nl0.w, 12.0), h=np.append(nl0.h, 9.0), x=np.append(nl0.x, bx),
                 y=np.append(nl0.y, by), n_mov=nl0.n_mov, net_start=nl0.net_start, pin_cell=nl0.pin_cell,
                 pin_dx=nl0.pin_dx, pin_dy=nl0.pin_dy, die=nl0.die)
    x, y = spread(nl, 3)
    x[:200] = bx + 3.0  # many cells start inside the block
    y[:200] = by + 4.0
    lx, ly, _ = legalize(nl, x, y)
    check_legal(nl, lx, ly, 1.0, 1.0)


def test_errors():
    nl = one_row([1, 2], [0.0, 1.0], 10)
    tall = Netlist(w=nl.w, h=np.float32([1, 2]), x=nl.x, y=nl.y, n_mov=2, net_start=nl.net_start,
                   pin_cell=nl.pin_cell, pin_dx=nl.pin_dx, pin_dy=nl.pin_dy, die=nl.die)
    with pytest.raises(api.EplaceError) as e:
        legalize(tall, nl.x, nl.y, 1.0, 1.0)
    assert e.value.code == api.EP_ENOTSUP
    full = one_row([4, 4, 4], [0.0, 1.0, 2.0], 10)
    with pytest.raises(api.EplaceError) as e:
        legalize(full, full.x, full.y, 1.0, 1.0)
    assert e.value.code == api.EP_EINVAL
