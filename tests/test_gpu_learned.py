"""Learned step length: device MLP kernels against the FP64 oracle, training,
and the verified proposal inside the Nesterov loop (acceptance, fallback,
and what the trained network gains over the plain search)."""
import sys

import numpy as np
import pytest

import aieplace_amd as A
from aieplace_amd import api, learned

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def f32_blob(oracle, seed):
    return oracle.mlp_random_weights(np.random.default_rng(seed)).astype(np.float32)


def test_mlp_forward_and_grad_match_oracle(oracle):
    say("case mlp forward/grad vs oracle")
    rng = np.random.default_rng(3)
    w = f32_blob(oracle, 3)
    x = rng.standard_normal((300, 16)).astype(np.float32)
    y = rng.standard_normal(300).astype(np.float32)
    out = learned.forward(w, x)
    oout = oracle.mlp_forward(w.astype(np.float64), x.astype(np.float64))
    assert np.abs(out - oout).max() <= 1e-4 * max(1.0, np.abs(oout).max())
    g, loss = learned.grad(w, x, y)
    og, oloss = oracle.mlp_grad(w.astype(np.float64), x.astype(np.float64), y.astype(np.float64))
    assert loss == pytest.approx(oloss, rel=1e-4)
    err = np.abs(g - og).max() / np.abs(og).max()
    say(f"  grad rel err (max-norm): {err:.2e}")
    assert err <= 1e-4
    assert np.all(g[:34] == 0)


def test_mlp_training_reduces_loss():
    say("case mlp training")
    rng = np.random.default_rng(4)
    x = rng.standard_normal((4000, 16)).astype(np.float32)
    y = (np.tanh(x[:, 0] - 0.5 * x[:, 3]) + 0.3 * x[:, 7] ** 2).astype(np.float32)
    w, curve = learned.train(x, y, epochs=30, batch=256)
    say(f"  loss: {curve[0]:.4f} -> {curve[-1]:.4f}")
    assert curve[-1] < 0.1 * curve[0]
    pred = learned.forward(w, x)
    assert np.mean((pred - y) ** 2) < 0.1 * np.var(y)


def _instance(seed, n=20_000, bins=128, **kw):
    nl = A.generate("stdcell", seed=seed, n_movable=n)
    p = A.placer_params(A.CONFIGS["stdcell_200k"], bins_x=bins, bins_y=bins, **kw)
    return nl, p


@pytest.fixture(scope="module")
def small_model():
    """Trained on records of 4 plain runs (seeds 11-14) by the device kernels."""
    feats, targs = [], []
    for s in (11, 12, 13, 14):
        nl, p = _instance(s, use_graph=1)
        with A.Placer(nl, p) as pl:
            pl.run()
            f, t = pl.take_step_records()
        feats.append(f)
        targs.append(t)
    X, Y = np.concatenate(feats), np.concatenate(targs)
    w, curve = learned.train(X, Y, epochs=25, batch=256)
    return w, curve, len(Y)


def test_records_are_collected():
    nl, p = _instance(21, use_graph=1)
    with A.Placer(nl, p) as pl:
        res = pl.run()
        f, t = pl.take_step_records()
        assert len(t) == res["iterations"] and f.shape == (len(t), 16)
        assert np.all(np.isfinite(f)) and np.all(np.isfinite(t))
        f2, _ = pl.take_step_records()
        assert len(f2) == 0  # taken records are cleared


def _ablation(w, seed, n, bins, precision=api.EP_PREC_BF16X3):
    """The same form with and without the model.  The final HPWL of a run
    varies with the filler placement (and FP32 summation order), so both forms
    run with filler seeds 7-10; `plain_hpwls` and `lrn_hpwls` give the spreads,
    `plain` and `lrn` are the seed-7 runs."""
    nl, p = _instance(seed, n=n, bins=bins, use_graph=1, precision=precision)
    runs = {0: [], 1: []}
    for learned_step in (0, 1):
        p.learned_step = learned_step
        for fs in (7, 8, 9, 10):
            p.filler_seed = fs
            with A.Placer(nl, p) as pl:
                if learned_step:
                    pl.set_step_model(w)
                r = pl.run()
            runs[learned_step].append(r)
            assert r["overflow"] <= 0.10
            if learned_step:  # every iteration proposes unless the model sits out after a rejection
                assert 0 < r["learned_proposals"] <= r["iterations"]
    plain, lrn = runs[0][0], runs[1][0]
    plain_hpwls = [r["hpwl"] for r in runs[0]]
    lrn_hpwls = [r["hpwl"] for r in runs[1]]
    rate = sum(r["learned_accepted"] for r in runs[1]) / max(1, sum(r["learned_proposals"] for r in runs[1]))
    say(f"  same form without model: {plain['iterations']} it {plain['evals']} evals "
        f"hpwl {plain['hpwl']:.1f} {plain['seconds']:.3f}s; with model: {lrn['iterations']} it "
        f"{lrn['evals']} evals hpwl {lrn['hpwl']:.1f} {lrn['seconds']:.3f}s; acceptance {rate:.3f} "
        f"({lrn['evals'] / lrn['iterations']:.3f} vs {plain['evals'] / plain['iterations']:.3f} evals/it); "
        f"hpwl over filler seeds: plain {min(plain_hpwls):.1f}..{max(plain_hpwls):.1f} "
        f"(mean {np.mean(plain_hpwls):.1f}), learned {min(lrn_hpwls):.1f}..{max(lrn_hpwls):.1f} "
        f"(mean {np.mean(lrn_hpwls):.1f})")
    return plain, lrn, rate, plain_hpwls, lrn_hpwls


def test_learned_step_small_model(small_model):
    """A model trained on 4 runs only: every proposal is checked; rejected ones
    fall back to the plain search, and the run still converges."""
    w, curve, nrec = small_model
    say(f"case learned step (4-run model, {nrec} records, loss {curve[0]:.3f} -> {curve[-1]:.3f})")
    plain, lrn, rate, spread, lspread = _ablation(w, 31, 20_000, 128)
    # the mean over filler seeds within 0.5% of the plain form's range
    assert np.mean(lspread) <= max(spread) * 1.005


def test_learned_step_ablation_committed_model():
    """The committed model (200 seeded runs, 50k-500k cells, trained by the
    device kernels): acceptance rate, evaluations per iteration and HPWL against
    the same form without the model, on held-out instances."""
    import os
    if not os.path.exists(learned.DEFAULT_WEIGHTS):
        pytest.skip("no committed step model")
    w = learned.load()
    for seed, n, bins in ((31, 20_000, 128), (32, 200_000, 512)):
        say(f"case ablation seed={seed} n={n} bins={bins}")
        plain, lrn, rate, spread, lspread = _ablation(w, seed, n, bins)
        assert rate > 0.5
        assert np.mean(lspread) <= max(spread) * 1.005


def test_learned_runs_start_fresh():
    """The model's rejection backoff and the proposal counters belong to one run:
    a second run from the same start (after set_positions) repeats the first one
    exactly.  A model that is always rejected ends its first run inside a
    backoff, so any state carried over would change when the second run's
    proposals happen (and with them the evaluation counts)."""
    say("case learned runs start fresh")
    w = np.zeros(api.MLP_NWEIGHTS, np.float32)
    w[16:32] = 1.0
    w[32], w[33] = 6.0, 1.0  # always proposes 10^6 bin widths: always rejected
    for model in (w, learned.load()):
        nl, p = _instance(52, use_graph=1, learned_step=1, precision=api.EP_PREC_BF16X3)
        with A.Placer(nl, p) as pl:
            pl.set_step_model(model)
            x0, y0 = pl.get_positions()
            r1 = pl.run(log_cap=4000)
            pl.set_positions(x0, y0)
            r2 = pl.run(log_cap=4000)
        say(f"  run 1: it {r1['iterations']} ev {r1['evals']} proposals {r1['learned_proposals']} "
            f"accepted {r1['learned_accepted']}; run 2: it {r2['iterations']} ev {r2['evals']} "
            f"proposals {r2['learned_proposals']} accepted {r2['learned_accepted']}")
        for r in (r1, r2):
            assert 0 < r["learned_proposals"] <= r["iterations"]
        for k in ("iterations", "evals", "learned_proposals", "learned_accepted", "hpwl"):
            assert r1[k] == r2[k], k
        assert r1["log"] == r2["log"]


def test_learned_1m_workload():
    """The learned-step workload: 1M cells on 1024x1024 bins, split-BF16
    transforms, with and without the committed model (device-controlled runs).
    Reports acceptance, gradient evaluations per iteration and HPWL against the
    same form without the model.  What the network gains is wirelength: its
    run stops at the same overflow with at least 1% lower HPWL (it proposes
    slightly shorter steps than the Lipschitz estimate, at the cost of a few
    percent more evaluations).  Runs are deterministic, so the bound is a
    fixed outcome, not a sample."""
    import os
    if not os.path.exists(learned.DEFAULT_WEIGHTS):
        pytest.skip("no committed step model")
    say("case learned_1m: 1M cells, 1024^2 bins, split-BF16")
    nl = A.generate("stdcell", seed=1, n_movable=1_000_000)
    p = A.placer_params(A.CONFIGS["learned_1m"], precision=api.EP_PREC_BF16X3, use_graph=1)
    with A.Placer(nl, p) as pl:
        plain = pl.run()
    p.learned_step = 1
    with A.Placer(nl, p) as pl:
        pl.set_step_model(learned.load())
        lrn = pl.run()
    rate = lrn["learned_accepted"] / max(1, lrn["learned_proposals"])
    say(f"  without model: {plain['iterations']} it, {plain['evals']} evals "
        f"({plain['evals'] / plain['iterations']:.3f}/it), hpwl {plain['hpwl']:.5e}, {plain['seconds']:.3f} s")
    say(f"  with model:    {lrn['iterations']} it, {lrn['evals']} evals "
        f"({lrn['evals'] / lrn['iterations']:.3f}/it), hpwl {lrn['hpwl']:.5e}, {lrn['seconds']:.3f} s; "
        f"{lrn['learned_proposals']} proposals, acceptance {rate:.3f}")
    assert plain["overflow"] <= 0.10 and lrn["overflow"] <= 0.10
    assert 0 < lrn["learned_proposals"] <= lrn["iterations"]
    assert rate > 0.9
    assert lrn["hpwl"] <= plain["hpwl"] * 0.99


def test_corrupted_model_falls_back():
    """A model proposing absurd steps is rejected by the check every time and
    the plain search still converges."""
    say("case corrupted step model")
    w = np.zeros(api.MLP_NWEIGHTS, np.float32)
    w[16:32] = 1.0
    w[32], w[33] = 6.0, 1.0  # always proposes 10^6
    nl, p = _instance(41, use_graph=1, learned_step=1)
    with A.Placer(nl, p) as pl:
        pl.set_step_model(w)
        res = pl.run()
    say(f"  proposals {res['learned_proposals']} accepted {res['learned_accepted']} "
        f"overflow {res['overflow']:.4f}")
    assert res["overflow"] <= 0.10
    assert res["learned_accepted"] <= 0.02 * res["learned_proposals"]
    with pytest.raises(api.EplaceError):
        bad = w.copy()
        bad[100] = np.nan
        with A.Placer(nl, p) as pl:
            pl.set_step_model(bad)
