"""Size-independent properties at the full 2M-cell, 1024x1024-bin configuration
(stdcell_2m), where the FP64 oracle is too slow to run:

* HPWL equals the span sum computed from the same FP32 pin positions in numpy
  (per-net spans are exact; only the FP64 summation order differs);
* the WA wirelength lies below the HPWL and tends to it as gamma -> 0;
* the movable map conserves the movable area (every cell inside the die) and
  the overflow equals the overflow formula applied to the returned maps;
* maps and gradients are bit-identical across repeated calls (fixed-point
  density accumulation, fixed-order reductions).
"""
import sys

import numpy as np
import pytest

import aieplace_amd as A

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def span_sum(px, py, net_start):
    """sum over nets of (max - min) x + (max - min) y, per-net spans in FP32."""
    deg = np.diff(net_start)
    starts = net_start[:-1][deg > 0]
    tot = 0.0
    for v in (px, py):
        hi = np.maximum.reduceat(v, starts)
        lo = np.minimum.reduceat(v, starts)
        tot += (hi - lo).astype(np.float64).sum()
    return tot


def test_full_size_invariants():
    cfg = A.CONFIGS["stdcell_2m"]
    say("case full size: stdcell_2m, generating")
    nl = A.generate(seed=1, **cfg["gen"])
    p = A.placer_params(cfg)
    assert nl.n_mov == 2_000_000 and p.bins_x == 1024
    xl, yl, xh, yh = nl.die
    n_mov = nl.n_mov
    w, h = nl.w[:n_mov], nl.h[:n_mov]
    rng = np.random.default_rng(11)
    with A.Placer(nl, p) as pl:
        x, y = pl.get_positions()  # movable cells, then fillers (kept where they are)
        x[:n_mov] = (xl + rng.random(n_mov) * (xh - xl - w)).astype(np.float32)
        y[:n_mov] = (yl + rng.random(n_mov) * (yh - yl - h)).astype(np.float32)
        pl.set_positions(x, y)
        xs, ys = pl.get_positions()

        say("case full size: HPWL and WA")
        X = np.concatenate([xs[:n_mov], nl.x[n_mov:].astype(np.float32)])
        Y = np.concatenate([ys[:n_mov], nl.y[n_mov:].astype(np.float32)])
        px = X[nl.pin_cell] + nl.pin_dx.astype(np.float32)
        py = Y[nl.pin_cell] + nl.pin_dy.astype(np.float32)
        hp_np = span_sum(px, py, nl.net_start)
        bw = (xh - xl) / p.bins_x
        gx0 = None
        for gamma in (0.01 * bw, bw, 10.0 * bw):
            gx, gy, wa, hp = pl.computeAllPartials(gamma)
            assert hp == pytest.approx(hp_np, rel=1e-9)
            assert wa <= hp * (1 + 1e-6)
            if gamma < 0.1 * bw:
                assert wa >= 0.999 * hp
                gx0 = gx
            assert np.isfinite(gx).all() and np.isfinite(gy).all()
        gx_again, _, _, _ = pl.computeAllPartials(0.01 * bw)
        assert np.array_equal(gx0, gx_again)

        say("case full size: density and overflow")
        fmap, mmap, fix, tau = pl.computeOverlaps()
        mov_area = float((w * h).sum())
        assert mmap.sum(dtype=np.float64) == pytest.approx(mov_area, rel=1e-6)
        bh = (yh - yl) / p.bins_y
        rho = p.target_density
        v = mmap.astype(np.float64) + rho * fix.astype(np.float64) - rho * bw * bh
        assert tau == pytest.approx(np.maximum(v, 0.0).sum() / mov_area, rel=1e-4)
        assert (mmap >= 0).all() and (fmap >= 0).all()
        fmap2, mmap2, _, tau2 = pl.computeOverlaps()
        assert np.array_equal(mmap, mmap2) and np.array_equal(fmap, fmap2) and tau == tau2
