"""A LEF/DEF design through the placer (include/eplace_io.h + include/eplace.h).

A generated design is written as LEF/DEF text and read by ep_lefdef_read; the
placer's gradient on that netlist matches the FP64 oracle on the same netlist
(the bar of the other gradient tests), the placement reaches overflow 0.10, and
ep_lefdef_write_def writes the placed positions back: re-read, they equal the
placer's positions rounded to database units, fixed cells and nets unchanged,
and the HPWL of the re-read design equals the placer's to that rounding.
"""
import os
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


def test_run_with_report_writes_the_run_outputs(tmp_path):
    """Placer.run_with_report on a LEF/DEF design: the operator output of the
    reference's run (statistics.md, placement_<it>.png every 10 iterations and
    at the end, <design>.def), and the same run as the device-controlled ep_run
    (host- and device-side control take the same decisions)."""
    from aieplace_amd.report import RunReport
    say("case report mesh 576 cells")
    gen = A.generate("mesh", seed=1)
    lef, df = netlist_lefdef(gen, name="mesh576")
    (tmp_path / "cells.lef").write_text(lef)
    (tmp_path / "floorplan.def").write_text(df)
    p = A.placer_params(A.CONFIGS["mesh_c1"])
    with LefDef.from_dir(str(tmp_path)) as d:
        nl = d.netlist()
        with A.Placer(nl, p) as pl:
            rep = RunReport("mesh576", out_dir=str(tmp_path / "run"), snapshot_every=10, pixels=256, quiet=True)
            res = pl.run_with_report(rep, design=d)
        with A.Placer(nl, p) as pl2:
            dev = pl2.run(log_cap=10_000)
    say(f"  reported: it={res['iterations']} hpwl={res['hpwl']:.6e} tau={res['overflow']:.4f}; "
        f"ep_run: it={dev['iterations']} hpwl={dev['hpwl']:.6e}")
    assert res["overflow"] <= p.stop_overflow
    assert res["iterations"] == dev["iterations"]
    assert res["evals"] == sum(l["evals"] for l in dev["log"])
    assert res["hpwl"] == pytest.approx(dev["hpwl"], rel=1e-9)
    files = set(os.listdir(tmp_path / "run"))
    want = {f"placement_{k}.png" for k in range(0, res["iterations"], 10)} | {
        f"placement_{res['iterations']}.png", "statistics.md", "mesh576.def"}
    assert want <= files, sorted(want - files)
    stats = (tmp_path / "run" / "statistics.md").read_text()
    assert f"| Iterations | {res['iterations']} |" in stats
    assert f"| Final HPWL | {res['hpwl']:.6e} |" in stats


def test_global_placement_then_legalisation():
    """The flow after global placement: the placed 20k-cell design legalised on
    the host (Abacus) is legal, and its HPWL grows by a few percent."""
    from aieplace_amd.legalize import legalize
    from support import check_legal
    say("case legalise stdcell n=20000 bins=128")
    nl = A.generate("stdcell", seed=5, n_movable=20_000)
    p = A.placer_params(A.CONFIGS["stdcell_200k"], bins_x=128, bins_y=128)
    with A.Placer(nl, p) as pl:
        r = pl.run()
        x, y = pl.get_positions()
        h_global = pl.hpwl()
    lx, ly, info = legalize(nl, x, y)
    check_legal(nl, lx, ly, 1.0, 1.0)
    placed = A.Netlist(w=nl.w, h=nl.h, x=np.concatenate([lx, nl.x[nl.n_mov:]]),
                       y=np.concatenate([ly, nl.y[nl.n_mov:]]), n_mov=nl.n_mov, net_start=nl.net_start,
                       pin_cell=nl.pin_cell, pin_dx=nl.pin_dx, pin_dy=nl.pin_dy, die=nl.die)
    h_legal = hpwl(placed)
    say(f"  global: tau={r['overflow']:.4f} hpwl={h_global:.6e}; legal: hpwl={h_legal:.6e} "
        f"(+{100 * (h_legal / h_global - 1):.2f}%), mean disp {info['total_disp'] / nl.n_mov:.3f}, "
        f"max {info['max_disp']:.2f}, {info['seconds'] * 1e3:.1f} ms")
    assert h_legal < 1.2 * h_global


def test_command_line_run(tmp_path):
    """eplace_run on a LEF/DEF directory: places, legalises and writes the DEF,
    which reads back legal."""
    import subprocess
    from support import check_legal
    say("case eplace_run mesh 576 cells")
    gen = A.generate("mesh", seed=1)
    lef, df = netlist_lefdef(gen, name="mesh576")
    (tmp_path / "cells.lef").write_text(lef)
    (tmp_path / "floorplan.def").write_text(df)
    runner = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "aieplace_amd", "eplace_run")
    r = subprocess.run([runner, str(tmp_path), "--bins", "32", "--density", "1.0", "--legalize",
                        "--out", str(tmp_path / "out")], capture_output=True, text=True, timeout=120)
    say(r.stdout[-600:])
    assert r.returncode == 0, r.stderr
    assert "final HPWL" in r.stdout and "legalised" in r.stdout
    with LefDef([str(tmp_path / "cells.lef")], str(tmp_path / "out" / "mesh576.def")) as d:
        back = d.netlist()
    check_legal(back, back.x[:back.n_mov], back.y[:back.n_mov], 1.0, 1.0)
