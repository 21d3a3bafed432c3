"""Pins the oracle's Nesterov placer (eo_place, CPU only) by what the
mathematics fixes, independently of the device.

eo_place has no counterpart in the reference (its loop is fixed-step descent,
AIEplace.cpp:44-76, 993-1046); it follows the ePlace paper (README.md:19).  The
gradient it descends is eo_gradient, which test_oracle.py pins against the
reference's golden vectors, its RePlAce field solver and closed forms.  Here:

* the schedules in its log follow their definitions: lambda_0 = eta |gW|_1/|F|_1
  at the start, lambda_{k+1}/lambda_k = clamp(1.05^(1 - dHPWL/Href), 0.95, 1.05)
  (1.05 when HPWL fell), gamma_k = 4 bw 10^((tau_{k-1} - 0.1) 20/9 - 1), and the
  momentum coefficients (a_k - 1)/a_{k+1} with a_0 = 1, a_{k+1} = (1 + sqrt(4a_k^2 + 1))/2;
* every accepted step passes alpha_hat >= 0.95 alpha, where alpha_hat is
  recomputed from eo_gradient at the accepted points (not read from the log);
* the first six iterates equal hand-computed Nesterov steps (coefficient 0,
  then (a_1 - 1)/a_2 with a_1 = (1 + sqrt 5)/2);
* with lambda = 0 a cell tied to four mirrored pads converges to the point
  that symmetry fixes, and fillers (no nets) never move.
"""
import math

import numpy as np
import pytest

import aieplace_amd as A
from support import oracle_design, oracle_params

ETA = 8e-5


@pytest.fixture(scope="module")
def small(oracle):
    """3000 standard cells on 64x64 bins (about the 200k-cell workload's cells per bin)."""
    nl = A.generate("stdcell", seed=3, n_movable=3000)
    p = A.placer_params(A.CONFIGS["stdcell_200k"], bins_x=64, bins_y=64)
    d, x0, y0 = oracle_design(oracle, nl, p)
    op = oracle_params(p, d.bin_size[0], nl.n_nets)
    x0, y0 = x0.astype(np.float64), y0.astype(np.float64)
    _, _, info = d.place(x0, y0, op)
    return d, x0, y0, op, info


def clip(d, x, y):
    n = d.n_place
    return (np.clip(x, d.die[0], d.die[2] - d.w[:n]), np.clip(y, d.die[1], d.die[3] - d.h[:n]))


def gamma_of(op, bw, tau):
    return op["gamma_bins"] * bw * 10.0 ** ((tau - 0.1) * 20.0 / 9.0 - 1.0)


def start_state(d, x0, y0, op):
    """tau, gamma_0, lambda_0 and the preconditioned gradient at the start."""
    bw = d.bin_size[0]
    mov = d.density(x0, y0, "exact")
    tau0 = d.overflow(mov, d.density_fixed(), op["target_density"])
    g0 = gamma_of(op, bw, tau0)
    o = d.gradient(x0, y0, op["target_density"], g0, 0.0)
    sw = np.abs(o["gwx"]).sum() + np.abs(o["gwy"]).sum()
    sf = np.abs(o["fx"]).sum() + np.abs(o["fy"]).sum()
    lam0 = op["lambda_eta"] * sw / sf
    return tau0, g0, lam0, o["hpwl"]


def test_schedules_follow_their_definitions(small):
    d, x0, y0, op, info = small
    log = info["log"]
    bw = d.bin_size[0]
    assert info["overflow"] <= 0.10 and len(log) == info["iters"] > 50
    assert all(l["overflow"] > 0.10 for l in log[:-1])  # stops at the first tau <= 0.10
    tau0, g0, lam0, hp0 = start_state(d, x0, y0, op)
    assert log[0]["gamma"] == pytest.approx(g0, rel=1e-12)
    assert log[0]["lam"] == pytest.approx(lam0, rel=1e-9)
    a = 1.0
    prev_hpwl = hp0
    for k, l in enumerate(log):
        a1 = 0.5 * (1.0 + math.sqrt(4.0 * a * a + 1.0))
        assert l["coef"] == pytest.approx((a - 1.0) / a1, rel=1e-14, abs=0), k
        a = a1
        assert 1 <= l["evals"] <= op["max_backtrack"]
        if k + 1 < len(log):
            dh = l["hpwl"] - prev_hpwl
            mx, mn = op["mu_max"], op["mu_min"]  # 1.05 and 0.95 as FP32 parameters
            mu = mx if dh < 0 else min(mx, max(mn, mx ** (1.0 - dh / op["hpwl_ref"])))
            assert log[k + 1]["lam"] / l["lam"] == pytest.approx(mu, rel=1e-12), k
            assert log[k + 1]["gamma"] == pytest.approx(gamma_of(op, bw, l["overflow"]), rel=1e-12), k
        prev_hpwl = l["hpwl"]
    assert log[0]["coef"] == 0.0
    assert log[1]["coef"] == pytest.approx(((1 + math.sqrt(5)) / 2 - 1) /
                                           (0.5 * (1 + math.sqrt((1 + math.sqrt(5)) ** 2 + 1))), rel=1e-14)
    # lambda grows by at most mu_max per iteration, so the iteration count is at
    # least log(lambda_final / lambda_0) / log(mu_max)
    growth = log[-1]["lam"] / log[0]["lam"]
    assert len(log) >= math.log(growth) / math.log(op["mu_max"]) - 1


def test_accepted_steps_pass_the_backtracking_test(small, oracle):
    """alpha_hat recomputed from eo_gradient at the accepted points v_{k-1}, v_k
    (each with its iteration's lambda and gamma, the stored-gradient rule):
    it equals the logged value and satisfies alpha_hat >= 0.95 alpha_used.  The
    first six iterates equal the hand-computed Nesterov steps."""
    d, x0, y0, op, info = small
    log = info["log"]
    K = 6
    pts = [(x0, y0)]
    for k in range(1, K + 1):
        q = dict(op, max_iters=k)
        x, y, inf = d.place(x0, y0, q)
        assert inf["iters"] == k
        assert [l["alpha_used"] for l in inf["log"]] == [l["alpha_used"] for l in log[:k]]
        pts.append((x, y))
    _, g0, lam0, _ = start_state(d, x0, y0, op)
    rho = op["target_density"]
    grads = [d.gradient(x0, y0, rho, g0, lam0)]
    for k in range(K):
        x, y = pts[k + 1]
        grads.append(d.gradient(x, y, rho, log[k]["gamma"], log[k]["lam"]))
    for k in range(K):
        (xa, ya), (xb, yb) = pts[k], pts[k + 1]
        ga, gb = grads[k], grads[k + 1]
        dv = math.sqrt(((xb - xa) ** 2).sum() + ((yb - ya) ** 2).sum())
        dg = math.sqrt(((gb["gx"] - ga["gx"]) ** 2).sum() + ((gb["gy"] - ga["gy"]) ** 2).sum())
        ahat = dv / dg
        assert ahat == pytest.approx(log[k]["alpha"], rel=1e-9), k
        if log[k]["evals"] < op["max_backtrack"]:
            assert ahat >= 0.95 * log[k]["alpha_used"] * (1 - 1e-12), k
        if k > 0 and log[k]["evals"] == 1:  # first trial starts from the last alpha_hat
            assert log[k]["alpha_used"] == log[k - 1]["alpha"]
    # hand-computed Nesterov iterates: u_k = clip(v_{k-1} - alpha_k g_{k-1}),
    # v_k = clip(u_k + c_k (u_k - u_{k-1})), u_{-1} = v_{-1} = start
    a1 = (1 + math.sqrt(5)) / 2
    assert log[1]["coef"] == pytest.approx((a1 - 1) / (0.5 * (1 + math.sqrt(4 * a1 * a1 + 1))), rel=1e-14)
    ux, uy = x0, y0
    vx, vy = x0, y0
    for k in range(K):
        al, c = log[k]["alpha_used"], log[k]["coef"]
        nux, nuy = clip(d, vx - al * grads[k]["gx"], vy - al * grads[k]["gy"])
        vx, vy = clip(d, nux + c * (nux - ux), nuy + c * (nuy - uy))
        if k == 1:
            # from here on a momentum taken from v_{k-1} instead of u_{k-1} differs
            assert np.abs(vx - nux).max() > 1e-6
        ux, uy = nux, nuy
        np.testing.assert_allclose(pts[k + 1][0], vx, rtol=0, atol=1e-9, err_msg=f"iterate {k}")
        np.testing.assert_allclose(pts[k + 1][1], vy, rtol=0, atol=1e-9, err_msg=f"iterate {k}")


def _pads_design(oracle, start):
    """One movable 1x1 cell tied by 2-pin nets to four fixed 1x1 pads at the
    mirrored points (c +- 8, c) and (c, c +- 8) of a 32x32 die (pins at the
    cell and pad centres)."""
    c = 16.0
    pads = [(c - 8, c), (c + 8, c), (c, c - 8), (c, c + 8)]
    w = np.ones(5)
    d = oracle.Design(w=w, h=w.copy(), n_mov=1, n_fill=0,
                      fix_x=np.array([p[0] - 0.5 for p in pads]), fix_y=np.array([p[1] - 0.5 for p in pads]),
                      net_start=np.arange(0, 9, 2), pin_cell=np.array([0, 1, 0, 2, 0, 3, 0, 4]),
                      pin_dx=np.full(8, 0.5), pin_dy=np.full(8, 0.5), die=(0, 0, 32, 32), bins=(16, 16))
    return d, np.array([start[0]]), np.array([start[1]])


@pytest.mark.parametrize("start", [(3.0, 25.0), (20.5, 9.25), (15.5, 2.0)])
def test_lambda_zero_converges_to_the_symmetric_minimiser(oracle, start):
    """With lambda = 0 only the WA wirelength acts.  By the mirror symmetry of
    the pads its minimiser puts the cell centre at (16, 16), i.e. the lower-left
    corner at (15.5, 15.5)."""
    d, x, y = _pads_design(oracle, start)
    op = dict(target_density=1.0, gamma_bins=0.25, gamma_schedule=0, lambda_eta=0.0, hpwl_ref=1.0,
              mu_max=1.05, mu_min=0.95, max_backtrack=10, stop_overflow=-1.0, max_iters=400,
              init_disp_bins=0.1)
    xf, yf, info = d.place(x, y, op)
    assert info["iters"] == 400
    assert all(l["lam"] == 0.0 for l in info["log"])
    assert abs(xf[0] - 15.5) < 1e-3 and abs(yf[0] - 15.5) < 1e-3, (xf, yf)


def test_fillers_do_not_move_without_density_force(oracle):
    """Fillers carry no nets: with lambda = 0 their gradient is zero and every
    iterate leaves them where they started, while the movable cells move."""
    nl = A.generate("stdcell", seed=4, n_movable=1500)
    p = A.placer_params(A.CONFIGS["stdcell_200k"], bins_x=32, bins_y=32)
    d, x0, y0 = oracle_design(oracle, nl, p)
    assert d.n_fill > 0
    op = dict(oracle_params(p, d.bin_size[0], nl.n_nets), lambda_eta=0.0, stop_overflow=-1.0, max_iters=15)
    x, y, _ = d.place(x0.astype(np.float64), y0.astype(np.float64), op)
    sl = slice(d.n_mov, d.n_place)
    assert np.array_equal(x[sl], x0[sl].astype(np.float64)) and np.array_equal(y[sl], y0[sl].astype(np.float64))
    assert np.abs(x[:d.n_mov] - x0[:d.n_mov]).max() > 1e-3
