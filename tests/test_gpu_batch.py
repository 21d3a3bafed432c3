"""Batched engine (include/eplace_batch.h) against the single-instance path and the oracle.

The batch runs the single-instance algorithm on every instance with
device-side control, so per instance:
  * the initial gamma, lambda_0 and preconditioned gradient equal the
    single-instance values at the same point (FP32 rounding) and the FP64
    oracle's gradient (2e-4 of its scale, the single-instance bar);
  * the trajectory of every instance reproduces its single-instance run
    (same iterations and evaluations, HPWL to FP64 summation order): both
    paths are deterministic and take the same decisions.
"""
import sys

import numpy as np
import pytest

import aieplace_amd as A
from aieplace_amd import api
from aieplace_amd.batch import BatchPlacer
from support import oracle_design

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_batch(counts, bins, **kw):
    nls = [A.generate("stdcell", seed=k + 1, n_movable=c) for k, c in enumerate(counts)]
    p = A.placer_params(A.CONFIGS["stdcell_200k"], bins_x=bins[0], bins_y=bins[0], **kw)
    return nls, p


def single_params(p, b):
    q = api.default_params()
    for name, _ in api.EpParams._fields_:
        setattr(q, name, getattr(p, name))
    q.bins_x = q.bins_y = b
    return q


def test_batch_init_matches_single_and_oracle(oracle):
    counts, bins = [5_000, 9_000, 14_000], [128, 128, 256]
    say(f"case batch init counts={counts} bins={bins}")
    nls, p = small_batch(counts, bins)
    with BatchPlacer(nls, bins, p) as bp:
        bp.init()
        for i, (nl, b) in enumerate(zip(nls, bins)):
            st = bp.state(i)
            gx, gy = bp.gradient(i)
            q = single_params(p, b)
            with A.Placer(nl, q) as pl:
                tau0 = pl.overflow()
                bw = (nl.die[2] - nl.die[0]) / b
                gamma = 4.0 * bw * 10 ** ((tau0 - 0.1) * 20 / 9 - 1)
                # FP32 bin width on the device side: agreement to ~1e-6
                assert abs(st["gamma"] - gamma) <= 1e-5 * gamma, (i, st["gamma"], gamma)
                g0 = pl.gradient(st["gamma"], 0.0)
                sw = np.abs(g0["gwx"].astype(np.float64)).sum() + np.abs(g0["gwy"].astype(np.float64)).sum()
                sf = np.abs(g0["fx"].astype(np.float64)).sum() + np.abs(g0["fy"].astype(np.float64)).sum()
                lam0 = p.lambda_eta * sw / sf
                assert abs(st["lambda"] - lam0) <= 1e-4 * lam0, (i, st["lambda"], lam0)
                g = pl.gradient(st["gamma"], st["lambda"])
                x, y = pl.get_positions()
            for k, v in (("gx", gx), ("gy", gy)):
                err = np.abs(v - g[k]).max() / np.abs(g[k]).max()
                say(f"  inst {i} {k} batch vs single {err:.2e}")
                assert err < 1e-4, (i, k, err)
            d, _, _ = oracle_design(oracle, nl, q)
            o = d.gradient(x.astype(np.float64), y.astype(np.float64), q.target_density,
                           st["gamma"], st["lambda"])
            for k, v in (("gx", gx), ("gy", gy)):
                err = np.abs(v - o[k]).max() / np.abs(o[k]).max()
                say(f"  inst {i} {k} batch vs oracle {err:.2e}")
                assert err < 2e-4, (i, k, err)


@pytest.mark.parametrize("precision", [api.EP_PREC_FP32, api.EP_PREC_BF16X3])
def test_batch_trajectory_matches_single(precision):
    """First 30 iterations: the device-side control takes the same decisions
    (backtracking, lambda/gamma schedules) as ep_step, so HPWL, overflow and
    the evaluation count agree with the single-instance run."""
    counts, bins = [6_000, 11_000, 60_000, 8_000], [128, 128, 256, 128]
    say(f"case batch trajectory precision={precision} counts={counts} bins={bins}")
    nls, p = small_batch(counts, bins, precision=precision, max_iters=30)
    with BatchPlacer(nls, bins, p) as bp:
        info = bp.run()
        res = bp.results()
    assert info["finished"] == len(counts)
    # the autotuned K step of the batched products (the single-instance
    # products use another one: equal trajectories show that it changes nothing)
    steps = {16, 32, 64} if precision == api.EP_PREC_FP32 else {32, 64, 128}
    say(f"  transform K step {info['transform_k_step']}, density tile {info['density_cells_per_thread']}")
    assert info["transform_k_step"] in steps
    for i, (nl, b) in enumerate(zip(nls, bins)):
        with A.Placer(nl, single_params(p, b)) as pl:
            r = pl.run()
        say(f"  inst {i}: batch it={res['iterations'][i]} ev={res['evals'][i]} hpwl={res['hpwl'][i]:.6e} "
            f"ovf={res['overflow'][i]:.5f} | single it={r['iterations']} ev={r['evals']} "
            f"hpwl={r['hpwl']:.6e} ovf={r['overflow']:.5f}")
        assert res["iterations"][i] == r["iterations"] == 30
        assert res["evals"][i] == r["evals"], i
        assert abs(res["hpwl"][i] - r["hpwl"]) <= 1e-9 * r["hpwl"], i
        assert abs(res["overflow"][i] - r["overflow"]) <= 1e-9, i


def test_batch_run_matches_single():
    """Full runs.  Both paths are deterministic (fixed-point density maps,
    fixed-order reductions) and take the same decisions, so every instance's
    run reproduces its single-instance run: same iteration and evaluation
    counts, HPWL and overflow to FP64 summation-order differences (the two
    paths reduce the same partial sums in different groupings)."""
    counts, bins = [6_000, 11_000, 60_000, 8_000], [128, 128, 256, 128]
    say(f"case batch run counts={counts} bins={bins}")
    nls, p = small_batch(counts, bins)
    with BatchPlacer(nls, bins, p) as bp:
        info = bp.run()
        res = bp.results()
        say(f"  batch: {info}")
        assert info["finished"] == len(counts)
        # everything done: a further run launches nothing and moves nothing
        x_before = [bp.positions(i)[0].copy() for i in range(len(counts))]
        again = bp.run()
        assert again["launches"] == 0
        for i in range(len(counts)):
            assert np.array_equal(bp.positions(i)[0], x_before[i])
        xb = [bp.positions(i) for i in range(len(counts))]
    for i, (nl, b) in enumerate(zip(nls, bins)):
        with A.Placer(nl, single_params(p, b)) as pl:
            r = pl.run()
            xs, ys = pl.get_positions()
        dx = float(np.abs(xs - xb[i][0]).max())
        say(f"  inst {i}: batch it={res['iterations'][i]} ev={res['evals'][i]} hpwl={res['hpwl'][i]!r} "
            f"| single it={r['iterations']} ev={r['evals']} hpwl={r['hpwl']!r} | max |dx| {dx:.3g}")
        assert res["overflow"][i] <= p.stop_overflow + 1e-9
        assert res["iterations"][i] == r["iterations"] and res["evals"][i] == r["evals"], i
        assert abs(res["hpwl"][i] - r["hpwl"]) <= 1e-9 * r["hpwl"], i
        assert abs(res["overflow"][i] - r["overflow"]) <= 1e-9, i
        # the decisions never differed, so the final placements are identical
        assert np.array_equal(xs, xb[i][0]) and np.array_equal(ys, xb[i][1]), i


def test_batch_partial_runs_continue_exactly():
    """ep_batch_run with max_launches stops after exactly that many graph
    replays, although the host keeps the next replay queued while it reads a
    replay's state back; a later call continues from there.  A run split into
    pieces therefore ends where one uninterrupted run does, bit for bit."""
    counts, bins = [5_000, 8_000, 11_000], [128, 128, 128]
    say(f"case batch partial runs counts={counts}")
    nls, p = small_batch(counts, bins)
    with BatchPlacer(nls, bins, p) as bp:
        whole = bp.run()
        ref = bp.results()
        ref_pos = [bp.positions(i) for i in range(len(counts))]
    with BatchPlacer(nls, bins, p) as bp:
        done = 0
        for k in (1, 2, 5):
            info = bp.run(max_launches=k)
            say(f"  run(max_launches={k}): {info['launches']} replays, "
                f"{info['instance_iterations']} instance-iterations so far")
            assert info["launches"] == k and info["finished"] == 0
            done += k
        rest = bp.run()
        res = bp.results()
        say(f"  rest: {rest['launches']} replays; whole run: {whole['launches']}")
        assert rest["finished"] == len(counts)
        assert done + rest["launches"] == whole["launches"]
        for i in range(len(counts)):
            assert res["iterations"][i] == ref["iterations"][i] and res["evals"][i] == ref["evals"][i], i
            assert res["hpwl"][i] == ref["hpwl"][i] and res["overflow"][i] == ref["overflow"][i], i
            x, y = bp.positions(i)
            assert np.array_equal(x, ref_pos[i][0]) and np.array_equal(y, ref_pos[i][1]), i


def test_batch_reset_restarts_from_initial_placement():
    counts, bins = [5_000, 7_000], [128, 128]
    say(f"case batch reset counts={counts}")
    nls, p = small_batch(counts, bins)
    with BatchPlacer(nls, bins, p) as bp:
        x0 = [bp.positions(i) for i in range(2)]
        first = bp.run()
        moved = [not np.array_equal(bp.positions(i)[0], x0[i][0]) for i in range(2)]
        assert all(moved)
        bp.reset()
        for i in range(2):
            x, y = bp.positions(i)
            assert np.array_equal(x, x0[i][0]) and np.array_equal(y, x0[i][1])
        second = bp.run()
        assert second["finished"] == 2
        r1 = first["instance_iterations"]
        r2 = second["instance_iterations"]
        assert abs(r1 - r2) <= max(6, 0.05 * r1), (r1, r2)



def test_renumbered_batch_matches_single():
    """Instances stored in breadth-first order (cell_order = 1) from netlists in
    random order: the batch reproduces the single-instance runs, and positions
    come back in the caller's order."""
    counts, bins = [6_000, 9_000], [128, 128]
    say(f"case renumbered batch counts={counts}")
    nls = [A.generate("stdcell", seed=k + 5, n_movable=c, shuffle_storage=1) for k, c in enumerate(counts)]
    p = A.placer_params(A.CONFIGS["stdcell_200k"], bins_x=128, bins_y=128, cell_order=1)
    with BatchPlacer(nls, bins, p) as bp:
        x0 = [bp.positions(i) for i in range(len(counts))]
        info = bp.run()
        res = bp.results()
        xb = [bp.positions(i) for i in range(len(counts))]
    assert info["finished"] == len(counts)
    for i, (nl, b) in enumerate(zip(nls, bins)):
        assert np.array_equal(x0[i][0][:nl.n_mov], nl.x[:nl.n_mov])  # caller's order, before the run
        with A.Placer(nl, single_params(p, b)) as pl:
            r = pl.run()
            xs, ys = pl.get_positions()
        assert res["iterations"][i] == r["iterations"] and res["evals"][i] == r["evals"], i
        assert np.array_equal(xs, xb[i][0]) and np.array_equal(ys, xb[i][1]), i


def test_forked_batch_matches_single():
    """The batch graph with its WA gradient on a second captured branch (the
    form the 64-instance workload takes while >= 400k cells run), forced here
    at a few thousand cells: every instance still reproduces its
    single-instance host-controlled run exactly."""
    counts, bins = [6_000, 11_000, 8_000], [128, 128, 128]
    say(f"case forked batch counts={counts}")
    nls, p = small_batch(counts, bins, fork_min_cells=1)
    with BatchPlacer(nls, bins, p) as bp:
        info = bp.run()
        res = bp.results()
        xb = [bp.positions(i) for i in range(len(counts))]
    assert info["finished"] == len(counts)
    for i, (nl, b) in enumerate(zip(nls, bins)):
        with A.Placer(nl, single_params(p, b)) as pl:
            r = pl.run()
            xs, ys = pl.get_positions()
        say(f"  inst {i}: batch it={res['iterations'][i]} ev={res['evals'][i]} | single it={r['iterations']} "
            f"ev={r['evals']}")
        assert res["iterations"][i] == r["iterations"] and res["evals"][i] == r["evals"], i
        assert np.array_equal(xs, xb[i][0]) and np.array_equal(ys, xb[i][1]), i


@pytest.mark.parametrize("precision", [api.EP_PREC_FP32, api.EP_PREC_BF16X3])
def test_batch_64_workload(precision):
    """The batched workload itself: 64 instances of 5k-150k cells (log-uniform,
    seeds 1-64) on 128/256/512 bins, one graph replay per four trials with
    compaction 64 -> 32 -> ... -> 1.  Every instance reaches overflow <= 0.10,
    and a sample (smallest, largest and two mid-size instances) reproduces its
    single-instance host-controlled run exactly."""
    from aieplace_amd.batch import batch_workload
    nls, bins = batch_workload(64)
    counts = [nl.n_mov for nl in nls]
    say(f"case batch 64 precision={precision}: {sum(counts)} movable cells, bins "
        f"{ {b: bins.count(b) for b in sorted(set(bins))} }")
    p = A.placer_params(A.CONFIGS["stdcell_200k"], precision=precision)
    with BatchPlacer(nls, bins, p) as bp:
        info = bp.run()
        res = bp.results()
        order = np.argsort(counts)
        sample = [int(order[0]), int(order[-1]), int(order[len(order) // 3]), int(order[2 * len(order) // 3])]
        xb = {i: bp.positions(i) for i in sample}
    say(f"  {info['instance_iterations']} instance-iterations, {info['launches']} replays, "
        f"{info['seconds']:.3f} s ({info['instance_iterations'] / info['seconds']:.0f} instance-iterations/s), "
        f"iterations {res['iterations'].min()}..{res['iterations'].max()}")
    assert info["finished"] == 64
    assert (res["overflow"] <= p.stop_overflow + 1e-9).all(), np.where(res["overflow"] > p.stop_overflow)
    assert (res["iterations"] < p.max_iters).all()
    for i in sample:
        with A.Placer(nls[i], single_params(p, bins[i])) as pl:
            r = pl.run()
            xs, ys = pl.get_positions()
        say(f"  inst {i} ({counts[i]} cells, {bins[i]}^2): batch it={res['iterations'][i]} ev={res['evals'][i]} "
            f"| single it={r['iterations']} ev={r['evals']}")
        assert res["iterations"][i] == r["iterations"] and res["evals"][i] == r["evals"], i
        assert np.array_equal(xs, xb[i][0]) and np.array_equal(ys, xb[i][1]), i
