"""A LEF/DEF design through the placer (include/eplace_io.h + include/eplace.h).

A generated design is written as LEF/DEF text and read by ep_lefdef_read; the
placer's gradient on that netlist matches the FP64 oracle on the same netlist
(the bar of the other gradient tests), the placement reaches overflow 0.10, and
ep_lefdef_write_def writes the placed positions back: re-read, they equal the
placer's positions rounded to database units, fixed cells and nets unchanged,
and the HPWL of the re-read design equals the placer's to that rounding.
"""
import sys

import numpy as np
import pytest

import aieplace_amd as A
from aieplace_amd.lefdef import LefDef
from support import netlist_lefdef, oracle_design, round_half_away

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def hpwl(nl):
    px = nl.x[nl.pin_cell].astype(np.float64) + nl.pin_dx
    py = nl.y[nl.pin_cell].astype(np.float64) + nl.pin_dy
    s = 0.0
    for n in range(nl.n_nets):
        a, b = nl.net_start[n], nl.net_start[n + 1]
        s += px[a:b].max() - px[a:b].min() + py[a:b].max() - py[a:b].min()
    return s


def test_lefdef_design_places_and_writes_back(tmp_path, oracle):
    say("case lefdef stdcell n=20000 bins=128")
    gen = A.generate("stdcell", seed=5, n_movable=20_000)
    lef, df = netlist_lefdef(gen, name="gen20k")
    (tmp_path / "cells.lef").write_text(lef)
    (tmp_path / "gen20k.def").write_text(df)
    p = A.placer_params(A.CONFIGS["stdcell_200k"], bins_x=128, bins_y=128)
    with LefDef([str(tmp_path / "cells.lef")], str(tmp_path / "gen20k.def")) as d:
        nl = d.netlist()
        o, _, _ = oracle_design(oracle, nl, p)
        with A.Placer(nl, p) as pl:
            xs, ys = pl.get_positions()
            bw = (nl.die[2] - nl.die[0]) / 128
            g = pl.gradient(2.0 * bw, 3e-3)
            og = o.gradient(xs.astype(np.float64), ys.astype(np.float64), p.target_density, 2.0 * bw, 3e-3)
            for k in ("gwx", "gwy", "fx", "fy", "gx", "gy"):
                err = np.abs(g[k] - og[k]).max() / np.abs(og[k]).max()
                say(f"  {k}: rel err {err:.2e}")
                assert err <= 2e-4, k
            r = pl.run()
            x, y = pl.get_positions()
            h_dev = pl.hpwl()  # FP64 HPWL at the returned positions
        say(f"  placed: it={r['iterations']} hpwl={r['hpwl']:.6e} tau={r['overflow']:.4f}")
        assert r["overflow"] <= p.stop_overflow + 1e-9
        out = d.writeDEF(str(tmp_path / "out"), x, y)
    with LefDef([str(tmp_path / "cells.lef")], out) as d2:
        back = d2.netlist()
    n = nl.n_mov
    dbu = lambda v: (round_half_away(np.asarray(v, np.float64) * 1000) / 1000).astype(np.float32)
    np.testing.assert_array_equal(back.x[:n], dbu(x[:n]))
    np.testing.assert_array_equal(back.y[:n], dbu(y[:n]))
    np.testing.assert_array_equal(back.x[n:], nl.x[n:])
    np.testing.assert_array_equal(back.pin_cell, nl.pin_cell)
    placed = nl
    placed.x = np.concatenate([x[:n], nl.x[n:]]).astype(np.float32)
    placed.y = np.concatenate([y[:n], nl.y[n:]]).astype(np.float32)
    h_placed, h_back = hpwl(placed), hpwl(back)
    say(f"  hpwl placer {h_dev:.6e}, host {h_placed:.6e}, re-read {h_back:.6e}")
    assert h_placed == pytest.approx(h_dev, rel=1e-6)
    # each pin moves by at most half a database unit per axis
    assert abs(h_back - h_placed) <= 2 * 0.0005 * nl.n_pins
