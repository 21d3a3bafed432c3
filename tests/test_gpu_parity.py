"""Parity of the HIP product path (through the C ABI) against the FP64 oracle.

Bars: bin indices bit-exact; density maps / overflow / HPWL to FP32 rounding;
WA gradient to 1e-4 of its scale; field within 1e-5 relative (max-norm) of the
FP64 solve in the plain FP32 form; final HPWL within 0.5% of the FP64
trajectory at overflow <= 0.10.
"""
import sys

import numpy as np
import pytest

import aieplace_amd as A
from aieplace_amd import api
import os

from support import hpwl_at_overflow, nudged_start, oracle_design, oracle_params

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(kind, n=20_000, bins=128, seed=1, **kw):
    if kind == "mesh":
        nl = A.generate("mesh", seed=seed)
        p = A.placer_params(A.CONFIGS["mesh_c1"], **kw)
    else:
        nl = A.generate("stdcell", seed=seed, n_movable=n)
        p = A.placer_params(A.CONFIGS["stdcell_200k"], bins_x=bins, bins_y=bins, **kw)
    return nl, p


def spread_positions(d, rng, frac=1.0):
    """Random positions over (a fraction of) the die for cells [mov|fill]."""
    xl, yl, xh, yh = d.die
    n = d.n_place
    cx, cy = (xl + xh) / 2, (yl + yh) / 2
    x = cx + (rng.random(n) - 0.5) * frac * (xh - xl) - d.w[:n] / 2
    y = cy + (rng.random(n) - 0.5) * frac * (yh - yl) - d.h[:n] / 2
    x = np.clip(x, xl, xh - d.w[:n]).astype(np.float32)
    y = np.clip(y, yl, yh - d.h[:n]).astype(np.float32)
    return x, y


CASES = [("mesh", 0, 32), ("stdcell", 20_000, 128), ("stdcell", 20_000, 256)]


@pytest.mark.parametrize("kind,n,bins", CASES)
def test_cell_bins_bit_exact(oracle, kind, n, bins):
    say(f"case cell_bins {kind} n={n} bins={bins}")
    nl, p = make(kind, n, bins)
    d, x0, y0 = oracle_design(oracle, nl, p)
    with A.Placer(nl, p) as pl:
        rng = np.random.default_rng(11)
        x, y = spread_positions(d, rng, 0.6)
        # put some cells exactly on bin boundaries and on the die edges
        bw = np.float32((nl.die[2] - nl.die[0]) / bins)
        k = np.arange(0, d.n_place, 7)
        x[k] = (np.floor(x[k] / bw) * bw).astype(np.float32)
        x[::13] = np.float32(nl.die[0])
        pl.set_positions(x, y)
        xs, ys = pl.get_positions()
        w, h = d.w[:d.n_place].astype(np.float32), d.h[:d.n_place].astype(np.float32)
        bh = np.float32((nl.die[3] - nl.die[1]) / bins)
        for stretched in (True, False):
            got = pl.cell_bins(stretched)
            exp = oracle.cell_bins_f32(xs, ys, w, h, nl.die[0], nl.die[1], bw, bh, bins, bins, stretched)
            assert np.array_equal(got, exp)


@pytest.mark.parametrize("kind,n,bins", CASES)
def test_density_and_overflow(oracle, kind, n, bins):
    say(f"case density {kind} n={n} bins={bins}")
    nl, p = make(kind, n, bins)
    d, _, _ = oracle_design(oracle, nl, p)
    with A.Placer(nl, p) as pl:
        for frac in (0.05, 0.5, 1.0):
            x, y = spread_positions(d, np.random.default_rng(int(frac * 100)), frac)
            pl.set_positions(x, y)
            xs, ys = pl.get_positions()
            fmap, mmap, fix, tau = pl.computeOverlaps()
            xd, yd = xs.astype(np.float64), ys.astype(np.float64)
            ofix = d.density_fixed()
            ofield = d.density(xd, yd, "stretched", 0, d.n_place) + p.target_density * ofix
            omov = d.density(xd, yd, "exact", 0, d.n_mov)
            scale = max(1.0, np.abs(ofield).max())
            assert np.abs(fmap - ofield).max() <= 2e-5 * scale * max(1, np.sqrt(d.n_place / 1000))
            assert np.abs(mmap - omov).max() <= 2e-5 * max(1.0, np.abs(omov).max()) * max(1, np.sqrt(d.n_mov / 1000))
            assert np.abs(fix - ofix).max() <= 1e-5 * max(1.0, ofix.max())
            otau = d.overflow(omov, ofix, p.target_density)
            assert tau == pytest.approx(otau, rel=1e-4, abs=1e-6)
            # area conservation: every smoothed cell lies inside the die when spread
            if frac <= 0.5:
                area = (d.w[:d.n_place] * d.h[:d.n_place]).sum() + p.target_density * ofix.sum()
                assert fmap.sum(dtype=np.float64) == pytest.approx(area, rel=1e-5)


@pytest.mark.parametrize("kind,n,bins", CASES)
def test_wa_gradient_and_hpwl(oracle, kind, n, bins):
    say(f"case wa {kind} n={n} bins={bins}")
    nl, p = make(kind, n, bins)
    d, _, _ = oracle_design(oracle, nl, p)
    with A.Placer(nl, p) as pl:
        for frac, gamma in ((0.05, 0.7), (1.0, 3.0), (1.0, 25.0)):
            x, y = spread_positions(d, np.random.default_rng(3), frac)
            pl.set_positions(x, y)
            xs, ys = pl.get_positions()
            gx, gy, wa, hp = pl.computeAllPartials(gamma)
            ogx, ogy, owa = d.wa_grad(xs.astype(np.float64), ys.astype(np.float64), gamma)
            ohp = d.hpwl(xs.astype(np.float64), ys.astype(np.float64))
            assert hp == pytest.approx(ohp, rel=2e-6)
            assert wa == pytest.approx(owa, rel=1e-4)
            sc = max(np.abs(ogx).max(), np.abs(ogy).max())
            assert np.abs(gx - ogx).max() <= 1e-4 * sc
            assert np.abs(gy - ogy).max() <= 1e-4 * sc
            # translation invariance: with no fixed pins the per-cell sum vanishes
            if kind != "mesh":
                assert abs(gx.sum()) <= 1e-3 * sc * np.sqrt(d.n_mov)


def test_cell_order_keeps_the_callers_order(oracle):
    """cell_order = 1 stores the movable cells in breadth-first order over the
    netlist.  On a netlist stored in random order, every per-cell array that
    crosses the boundary stays in the caller's order: positions, cell sizes,
    bin ranges and pin cells exactly, the gradient components to FP32
    summation order (a cell's pin partials are added in another order), and
    both match the FP64 oracle built from the caller's netlist."""
    say("case cell_order: shuffled 20k-cell netlist, 128^2 bins")
    nl = A.generate("stdcell", seed=3, n_movable=20_000, shuffle_storage=1)
    out = {}
    for order in (0, 1):
        p = A.placer_params(A.CONFIGS["stdcell_200k"], bins_x=128, bins_y=128, cell_order=order)
        d, _, _ = oracle_design(oracle, nl, p)
        with A.Placer(nl, p) as pl:
            x, y = spread_positions(d, np.random.default_rng(8), 0.4)
            pl.set_positions(x, y)
            xs, ys = pl.get_positions()
            assert np.array_equal(xs, x) and np.array_equal(ys, y)
            bw = (nl.die[2] - nl.die[0]) / 128
            g = pl.gradient(2.0 * bw, 3e-3)
            bins = pl.cell_bins(True)
            w, h, _, _ = pl.cells()
            ns, pc, _, _, perm = pl.nets()
        out[order] = (g, bins, w, h, ns, pc, perm)
        o = d.gradient(x.astype(np.float64), y.astype(np.float64), p.target_density, 2.0 * bw, 3e-3)
        for k in ("gx", "gy", "gwx", "gwy"):
            assert np.abs(g[k] - o[k]).max() <= 2e-4 * np.abs(o[k]).max(), (order, k)
    (g0, b0, w0, h0, ns0, pc0, pm0), (g1, b1, w1, h1, ns1, pc1, pm1) = out[0], out[1]
    assert np.array_equal(b0, b1) and np.array_equal(w0, w1) and np.array_equal(h0, h1)
    for k in ("gx", "gy", "gwx", "gwy", "fx", "fy"):
        err = np.abs(g0[k] - g1[k]).max() / np.abs(g0[k]).max()
        say(f"  {k}: order 0 vs 1 rel {err:.2e}")
        assert err <= 1e-5, k
    # every net holds the same cells, in the caller's numbering
    def nets(ns, pc, pm):
        return {int(pm[k]): sorted(pc[ns[k]:ns[k + 1]].tolist()) for k in range(len(pm))}
    assert nets(ns0, pc0, pm0) == nets(ns1, pc1, pm1)


def test_cell_order_run_converges():
    say("case cell_order: full run, shuffled 20k cells")
    nl = A.generate("stdcell", seed=3, n_movable=20_000, shuffle_storage=1)
    res = {}
    for order in (0, 1):
        p = A.placer_params(A.CONFIGS["stdcell_200k"], bins_x=128, bins_y=128, cell_order=order, use_graph=1)
        with A.Placer(nl, p) as pl:
            res[order] = pl.run()
    say(f"  order 0: it {res[0]['iterations']} hpwl {res[0]['hpwl']:.1f}; order 1: it {res[1]['iterations']} "
        f"hpwl {res[1]['hpwl']:.1f}")
    assert res[1]["overflow"] <= 0.10
    assert res[1]["hpwl"] == pytest.approx(res[0]["hpwl"], rel=0.03)


def test_wa_wide_and_degenerate_nets(oracle):
    """Nets wider than a wavefront (strided path), degree-1 and empty nets."""
    say("case wa wide/degenerate nets")
    rng = np.random.default_rng(5)
    n = 600
    deg = [2, 1, 0, 100, 65, 64, 33, 3, 200, 2]
    ns = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    pc = np.concatenate([rng.choice(n, k, replace=False) for k in deg]).astype(np.int32)
    nl = A.Netlist(w=np.ones(n), h=np.ones(n), x=rng.uniform(0, 60, n), y=rng.uniform(0, 60, n),
                   n_mov=n, net_start=ns, pin_cell=pc, pin_dx=rng.uniform(0, 1, len(pc)),
                   pin_dy=rng.uniform(0, 1, len(pc)), die=(0, 0, 64, 64))
    p = A.placer_params(A.CONFIGS["mesh_c1"], use_fillers=0)
    d, _, _ = oracle_design(oracle, nl, p)
    with A.Placer(nl, p) as pl:
        xs, ys = pl.get_positions()
        gx, gy, wa, hp = pl.computeAllPartials(2.0)
        ogx, ogy, owa = d.wa_grad(xs.astype(np.float64), ys.astype(np.float64), 2.0)
        assert hp == pytest.approx(d.hpwl(xs.astype(np.float64), ys.astype(np.float64)), rel=1e-6)
        sc = np.abs(ogx).max()
        assert np.abs(gx - ogx).max() <= 1e-4 * sc and np.abs(gy - ogy).max() <= 1e-4 * sc


@pytest.mark.parametrize("bins", [32, 128, 512, 1024])
def test_field_fp32_vs_fp64(oracle, bins):
    say(f"case field fp32 bins={bins}")
    n = {32: 5_000, 128: 20_000, 512: 100_000, 1024: 200_000}[bins]
    nl, p = make("stdcell", n, bins)
    d, _, _ = oracle_design(oracle, nl, p)
    with A.Placer(nl, p) as pl:
        x, y = spread_positions(d, np.random.default_rng(2), 0.7)
        pl.set_positions(x, y)
        fmap, _, _, _ = pl.computeOverlaps()
        ex, ey = pl.computeElectricFields(api.EP_PREC_FP32)
        bw = (nl.die[2] - nl.die[0]) / bins
        rho = fmap.astype(np.float64) / (bw * bw)
        _, oex, oey = oracle.field(rho, wy_scale=1.0)
        err = max(np.abs(ex - oex).max() / np.abs(oex).max(), np.abs(ey - oey).max() / np.abs(oey).max())
        say(f"  field rel err (max-norm) bins={bins}: {err:.3e}")
        assert err <= 1e-5


@pytest.mark.parametrize("bins", [32, 128, 512, 1024])
def test_field_bf16x3_vs_fp64(oracle, bins):
    """Split-BF16 transform products: error against the FP64 solve, per grid size."""
    say(f"case field bf16x3 bins={bins}")
    n = {32: 5_000, 128: 20_000, 512: 100_000, 1024: 200_000}[bins]
    nl, p = make("stdcell", n, bins)
    d, _, _ = oracle_design(oracle, nl, p)
    with A.Placer(nl, p) as pl:
        x, y = spread_positions(d, np.random.default_rng(2), 0.7)
        pl.set_positions(x, y)
        fmap, _, _, _ = pl.computeOverlaps()
        ex, ey = pl.computeElectricFields(api.EP_PREC_BF16X3)
        bw = (nl.die[2] - nl.die[0]) / bins
        _, oex, oey = oracle.field(fmap.astype(np.float64) / (bw * bw), wy_scale=1.0)
        err = max(np.abs(ex - oex).max() / np.abs(oex).max(), np.abs(ey - oey).max() / np.abs(oey).max())
        say(f"  field bf16x3 rel err (max-norm) bins={bins}: {err:.3e}")
        assert err <= 2e-4


@pytest.mark.parametrize("kind,n,bins", CASES[:2])
def test_full_gradient(oracle, kind, n, bins):
    say(f"case gradient {kind} n={n} bins={bins}")
    nl, p = make(kind, n, bins)
    d, _, _ = oracle_design(oracle, nl, p)
    with A.Placer(nl, p) as pl:
        x, y = spread_positions(d, np.random.default_rng(8), 0.4)
        pl.set_positions(x, y)
        xs, ys = pl.get_positions()
        bw = (nl.die[2] - nl.die[0]) / bins
        gamma, lam = 2.0 * bw, 3e-3
        g = pl.gradient(gamma, lam)
        o = d.gradient(xs.astype(np.float64), ys.astype(np.float64), p.target_density, gamma, lam)
        for k in ("gwx", "gwy", "fx", "fy", "gx", "gy"):
            sc = np.abs(o[k]).max()
            err = np.abs(g[k] - o[k]).max() / sc
            say(f"  {k}: rel err {err:.2e}")
            assert err <= 2e-4, k
        assert g["hpwl"] == pytest.approx(o["hpwl"], rel=2e-6)
        assert g["tau"] == pytest.approx(o["tau"], rel=1e-4)


@pytest.mark.parametrize("kind,n,bins", [("mesh", 0, 32), ("stdcell", 20_000, 128)])
def test_placement_trajectory_parity(oracle, kind, n, bins):
    """Final HPWL at overflow <= 0.10: HIP FP32 vs the FP64 oracle trajectory."""
    say(f"case placement {kind} n={n} bins={bins}")
    nl, p = make(kind, n, bins)
    d, x0, y0 = oracle_design(oracle, nl, p)
    _, _, oinfo = d.place(x0, y0, oracle_params(p, d.bin_size[0], nl.n_nets))
    with A.Placer(nl, p) as pl:
        res = pl.run()
    say(f"  oracle: it={oinfo['iters']} hpwl={oinfo['hpwl']:.2f} tau={oinfo['overflow']:.4f}; "
        f"gpu: it={res['iterations']} hpwl={res['hpwl']:.2f} tau={res['overflow']:.4f} {res['seconds']:.3f}s")
    assert oinfo["overflow"] <= 0.10 and res["overflow"] <= 0.10
    assert res["hpwl"] == pytest.approx(oinfo["hpwl"], rel=5e-3)


def test_stdcell_200k_fp32_matches_oracle_ensemble(oracle):
    """The 200k-cell workload (512x512 bins) in the plain FP32 form against the
    FP64 oracle's placements of it (tests/golden/oracle_stdcell_200k.npz,
    written by make_oracle_stdcell_200k.py in the build container) from the
    same K starts: member 0 is the workload's own start, the others move half
    of the x coordinates by one FP32 ulp (support.nudged_start).

    Member 0's trajectory tracks the oracle's: the same evaluation count in
    every iteration, HPWL and overflow within 1e-3 for the first 400
    iterations (the oracle and its own one-ulp members separate at that level
    near iteration 440-470).  From there a run is chaotic: the oracle's own
    final HPWL moves by ~1.6% between members, because HPWL grows ~0.2% per
    iteration while the overflow creeps across 0.10.  So the final HPWL is
    compared at matched overflow (hpwl_at_overflow) and over the ensemble: the
    device members' mean within 0.5% of the oracle members' mean; member 0's
    final HPWL and every member's iteration count inside the oracle members'
    range (widened by 0.5% and 10 iterations)."""
    path = os.path.join(os.path.dirname(__file__), "golden", "oracle_stdcell_200k.npz")
    fx = np.load(path, allow_pickle=False)
    k = len(fx["hpwl"])
    say(f"case stdcell_200k fp32 vs oracle ensemble ({k} members)")
    nl = A.generate("stdcell", seed=1, n_movable=200_000)
    p = A.placer_params(A.CONFIGS["stdcell_200k"], use_graph=1)
    _, x0, y0 = oracle_design(oracle, nl, p)
    runs = []
    with A.Placer(nl, p) as pl:
        for m in range(k):
            pl.set_positions(nudged_start(x0, m), y0)
            r = pl.run(log_cap=5000)
            r["at"] = hpwl_at_overflow([l["overflow"] for l in r["log"]], [l["hpwl"] for l in r["log"]])
            runs.append(r)
            say(f"  member {m}: gpu it {r['iterations']} hpwl {r['hpwl']:.6e} at 0.10 {r['at']:.6e} | oracle it "
                f"{fx['iters'][m]} hpwl {fx['hpwl'][m]:.6e} at 0.10 {fx['hpwl_at_010'][m]:.6e}")
    log = runs[0]["log"]
    ev = np.array([l["evals"] for l in log])
    hp = np.array([l["hpwl"] for l in log])
    tau = np.array([l["overflow"] for l in log])
    o = fx["log0"]  # evals, hpwl, overflow, lambda, gamma, alpha
    n = 400
    rel = np.abs(hp[:n] / o[:n, 1] - 1)
    say(f"  member 0, first {n} iterations: max |hpwl/oracle - 1| {rel.max():.2e}, "
        f"max |tau - oracle| {np.abs(tau[:n] - o[:n, 2]).max():.2e}")
    assert np.array_equal(ev[:n], o[:n, 0].astype(int))
    assert rel.max() <= 1e-3
    assert np.abs(tau[:n] - o[:n, 2]).max() <= 1e-3
    gmean = float(np.mean([r["at"] for r in runs]))
    omean = float(fx["hpwl_at_010"].mean())
    say(f"  at overflow 0.10: device mean {gmean:.6e}, oracle mean {omean:.6e} ({100 * (gmean / omean - 1):+.3f}%)")
    assert abs(gmean / omean - 1) <= 5e-3
    assert all(r["overflow"] <= 0.10 for r in runs)
    assert fx["hpwl"].min() * 0.995 <= runs[0]["hpwl"] <= fx["hpwl"].max() * 1.005
    assert all(fx["iters"].min() - 10 <= r["iterations"] <= fx["iters"].max() + 10 for r in runs)


def test_placement_bf16x3_matches_fp32():
    """Mixed-precision transforms reach overflow 0.10 with HPWL within 0.5% of FP32."""
    say("case placement bf16x3 vs fp32")
    nl, p = make("stdcell", 20_000, 128)
    res = {}
    for prec in (api.EP_PREC_FP32, api.EP_PREC_BF16X3):
        p.precision = prec
        with A.Placer(nl, p) as pl:
            res[prec] = pl.run()
    a, b = res[api.EP_PREC_FP32], res[api.EP_PREC_BF16X3]
    say(f"  fp32 it={a['iterations']} hpwl={a['hpwl']:.1f}; bf16x3 it={b['iterations']} hpwl={b['hpwl']:.1f}")
    assert b["overflow"] <= 0.10
    assert b["hpwl"] == pytest.approx(a["hpwl"], rel=5e-3)


def test_graph_replay_matches_eager(oracle):
    say("case graph replay vs eager")
    nl, p = make("stdcell", 20_000, 128)
    recs = {}
    for g in (0, 1):
        p.use_graph = g
        with A.Placer(nl, p) as pl:
            pl.initializePlacement()
            recs[g] = [pl.nudgeAllNodes() for _ in range(30)]
    for a, b in zip(recs[0], recs[1]):
        assert a["hpwl"] == pytest.approx(b["hpwl"], rel=1e-3)
        assert a["overflow"] == pytest.approx(b["overflow"], rel=1e-3, abs=1e-4)


@pytest.mark.parametrize("precision", [api.EP_PREC_FP32, api.EP_PREC_BF16X3])
def test_placement_is_deterministic(precision):
    """The density maps accumulate in 64-bit fixed point and every other
    reduction runs in a fixed order, so two runs of the same placement give
    bit-identical trajectories, positions and HPWL."""
    say(f"case determinism precision={precision}")
    nl, p = make("stdcell", 20_000, 128, use_graph=1, precision=precision)
    out = []
    for _ in range(2):
        with A.Placer(nl, p) as pl:
            r = pl.run(log_cap=2000)
            x, y = pl.get_positions()
        out.append((r, x, y))
    (r1, x1, y1), (r2, x2, y2) = out
    say(f"  it {r1['iterations']} / {r2['iterations']}  hpwl {r1['hpwl']!r} / {r2['hpwl']!r}")
    assert r1["iterations"] == r2["iterations"] and r1["evals"] == r2["evals"]
    assert r1["hpwl"] == r2["hpwl"] and r1["overflow"] == r2["overflow"]
    assert [l["hpwl"] for l in r1["log"]] == [l["hpwl"] for l in r2["log"]]
    assert np.array_equal(x1, x2) and np.array_equal(y1, y2)


@pytest.mark.parametrize("precision,learned_step,fork", [(api.EP_PREC_FP32, 0, 0), (api.EP_PREC_BF16X3, 0, 0),
                                                         (api.EP_PREC_BF16X3, 1, 0), (api.EP_PREC_FP32, 0, 1),
                                                         (api.EP_PREC_BF16X3, 1, 1), (api.EP_PREC_FP16X3, 0, 1)])
def test_device_control_matches_host_control(precision, learned_step, fork):
    """ep_run with graphs decides every trial on the device (the step-control
    kernel replays ep_step's FP64 decisions); without graphs ep_run loops over
    the host's ep_step.  Both take the same decisions, so the runs agree
    exactly: iteration and evaluation counts, every logged iteration, the
    step records and the final positions.  fork = 1 lowers the two-branch
    threshold so the captured evaluation runs the density scatter and the field
    on a second branch beside the WA gradient and the per-cell pin sums (the
    form every run of >= 400k cells takes); the host loop never forks."""
    from aieplace_amd import learned
    say(f"case device vs host step control precision={precision} learned={learned_step} fork={fork}")
    nl, p = make("stdcell", 20_000, 128, precision=precision, learned_step=learned_step,
                 fork_min_cells=1 if fork else 0)
    out = {}
    for g in (0, 1):
        p.use_graph = g
        with A.Placer(nl, p) as pl:
            if learned_step:
                pl.set_step_model(learned.load())
            r = pl.run(log_cap=2000)
            x, y = pl.get_positions()
            f, t = pl.take_step_records()
        out[g] = (r, x, y, f, t)
    (r0, x0, y0, f0, t0), (r1, x1, y1, f1, t1) = out[0], out[1]
    say(f"  host: it {r0['iterations']} ev {r0['evals']} hpwl {r0['hpwl']!r} {r0['seconds']:.3f}s; "
        f"device: it {r1['iterations']} ev {r1['evals']} hpwl {r1['hpwl']!r} {r1['seconds']:.3f}s")
    assert r0["iterations"] == r1["iterations"] and r0["evals"] == r1["evals"]
    assert r0["learned_proposals"] == r1["learned_proposals"]
    assert r0["learned_accepted"] == r1["learned_accepted"]
    assert r0["hpwl"] == r1["hpwl"] and r0["overflow"] == r1["overflow"]
    assert len(r0["log"]) == len(r1["log"]) == r0["iterations"]
    for a, b in zip(r0["log"], r1["log"]):
        assert a == b
    assert np.array_equal(x0, x1) and np.array_equal(y0, y1)
    assert np.array_equal(f0, f1) and np.array_equal(t0, t1)


def test_device_run_hands_state_to_host_steps():
    """A device-controlled run that stops at max_iters leaves the optimiser state
    (step lengths, lambda, gamma, histories, the accepted buffers) with the
    host: continuing with ep_step gives exactly the trajectory of a run that
    used ep_step from the start.  The log holds at most log_cap entries."""
    say("case device run -> host steps hand-off")
    nl, p = make("stdcell", 20_000, 128, use_graph=1, max_iters=40)
    with A.Placer(nl, p) as pl:
        r = pl.run(log_cap=2000)
        assert r["iterations"] == 40 and len(r["log"]) == 40
        tail = [pl.nudgeAllNodes() for _ in range(20)]
        xa, ya = pl.get_positions()
    p.use_graph = 0
    with A.Placer(nl, p) as pl:
        pl.initializePlacement()
        ref = [pl.nudgeAllNodes() for _ in range(60)]
        xb, yb = pl.get_positions()
    keys = ("iteration", "evals", "hpwl", "overflow", "lam", "gamma", "alpha")
    for a, b in zip(r["log"], ref[:40]):
        assert all(a[k] == b[k] for k in keys), (a, b)
    for a, b in zip(tail, ref[40:]):
        assert all(a[k] == b[k] for k in keys), (a, b)
    assert np.array_equal(xa, xb) and np.array_equal(ya, yb)
    p.use_graph = 1
    with A.Placer(nl, p) as pl:
        short = pl.run(log_cap=7)
        assert short["iterations"] == 40 and len(short["log"]) == 7
