"""Pins the FP64 oracle (CPU only): reference golden vectors, the reference's own
RePlAce field solver (oracle/_ref, when built), closed forms and invariants.

Golden fixtures are the reference's vck5000/aie/golden_data files, copied as
data into tests/golden/reference_fixtures.
"""
import os

import numpy as np
import pytest
import scipy.fft as sf

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures")


def rd(*parts):
    with open(os.path.join(GOLD, *parts)) as f:
        return np.array([float(l) for l in f if l.strip()])


# ------------------------------------------------------------------ WA (a1-a3)

@pytest.mark.parametrize("bench", range(4))
def test_wa_golden_fast_exp(oracle, bench):
    """createGoldenPartials.py:28-120 layout: lane-major, max pin first; gamma 4."""
    xin = rd("partials", f"x_in{bench}.dat")
    assert xin[0] == 2 and xin[1] == 8
    xs = xin[4:].reshape(2, 8)
    ap, am = rd("partials", f"a_plus{bench}.dat").reshape(2, 8), rd("partials", f"a_minus{bench}.dat").reshape(2, 8)
    bp, cp = rd("partials", f"b_plus{bench}.dat"), rd("partials", f"c_plus{bench}.dat")
    bm, cm = rd("partials", f"b_minus{bench}.dat"), rd("partials", f"c_minus{bench}.dat")
    part = rd("partials", f"partials{bench}.dat").reshape(2, 8)
    bc_stream = rd("partials", f"bc{bench}.dat")
    assert list(bc_stream[:4]) == [2, 8, 0, 0]
    np.testing.assert_array_equal(bc_stream[4:12], bp)
    for lane in range(8):
        x = xs[:, lane]
        a_p, a_m, bc = oracle.wa_terms(x, 4.0, fast=True)
        np.testing.assert_allclose(a_p, ap[:, lane], rtol=1e-13, atol=0)
        np.testing.assert_allclose(a_m, am[:, lane], rtol=1e-13, atol=0)
        np.testing.assert_allclose(bc, [bp[lane], cp[lane], bm[lane], cm[lane]], rtol=1e-13)
        np.testing.assert_allclose(oracle.wa_partials(x, 4.0, fast=True), part[:, lane], rtol=1e-12, atol=1e-13)


def test_wa_golden_true_exp(oracle):
    """True exp differs from the 16-squaring fast_exp by ~a^2/2^17 relative."""
    for bench in range(4):
        xs = rd("partials", f"x_in{bench}.dat")[4:].reshape(2, 8)
        part = rd("partials", f"partials{bench}.dat").reshape(2, 8)
        for lane in range(8):
            np.testing.assert_allclose(oracle.wa_partials(xs[:, lane], 4.0), part[:, lane], rtol=2e-3, atol=2e-4)


def test_fast_exp_closed_form(oracle):
    for a in (0.0, -1.0, -5.0, -18.0):
        assert oracle.fast_exp(a) == pytest.approx((1 + a / 2 ** 16) ** (2 ** 16), rel=1e-12)
        assert oracle.fast_exp(a) == pytest.approx(np.exp(a), rel=a * a / 2 ** 16 + 1e-12)


def test_wa_two_pin_closed_form(oracle):
    rng = np.random.default_rng(3)
    for _ in range(50):
        x = rng.uniform(-100, 100, 2)
        g = rng.uniform(0.3, 20)
        D = abs(x[1] - x[0])
        assert oracle.wa_value(x, g) == pytest.approx(D * np.tanh(D / (2 * g)), rel=1e-9, abs=1e-12)


def test_wa_invariants_and_fd(oracle):
    rng = np.random.default_rng(4)
    for d in (2, 3, 5, 9, 17, 40, 70):
        x = rng.uniform(0, 50, d)
        g = rng.uniform(0.5, 5)
        p = oracle.wa_partials(x, g)
        assert abs(p.sum()) < 1e-9                            # translation invariance
        assert oracle.wa_value(x, g) <= x.max() - x.min() + 1e-12  # WA <= HPWL
        h = 1e-6
        for i in range(d):
            xp, xm = x.copy(), x.copy()
            xp[i] += h
            xm[i] -= h
            fd = (oracle.wa_value(xp, g) - oracle.wa_value(xm, g)) / (2 * h)
            assert p[i] == pytest.approx(fd, abs=1e-6)
        # gamma -> 0 recovers HPWL
        assert oracle.wa_value(x, 1e-3) == pytest.approx(x.max() - x.min(), rel=1e-6)


# ------------------------------------------------------------------ transforms (a8)

def test_transforms_closed_forms(oracle):
    rng = np.random.default_rng(5)
    for n in (4, 16, 32, 33):
        x = rng.standard_normal(n)
        np.testing.assert_allclose(oracle.transform1d("dct", x), sf.dct(x, type=2) / 2, atol=1e-12)
        np.testing.assert_allclose(oracle.transform1d("idct", x), sf.dct(x, type=3) / 2, atol=1e-12)
        xs = np.append(x[1:], 0.0)
        np.testing.assert_allclose(oracle.transform1d("idst", x), sf.dst(xs, type=3) / 2, atol=1e-12)
        k = np.arange(n)
        np.testing.assert_allclose(oracle.transform1d("idxst", x),
                                   (-1.0) ** k * x[0] / 2 + oracle.transform1d("idst", x), atol=1e-12)
        # IDCT(DCT(x)) = N/2 x
        np.testing.assert_allclose(oracle.transform1d("idct", oracle.transform1d("dct", x)), n / 2 * x, atol=1e-10)


def test_dct_golden_rows(oracle):
    """dct_output.dat rows 0-7 (input.dat holds rows 0-7 only; outputs printed %.2f)."""
    inp = rd("density", "input.dat")
    re = inp[0::2][:128].reshape(8, 16)
    out = rd("density", "dct_output.dat").reshape(16, 16, 2)[:, :, 0]
    for r in range(8):
        np.testing.assert_allclose(oracle.transform1d("dct", re[r]), out[r], atol=0.006)
    tpose = rd("density", "dct_output_tpose.dat").reshape(16, 16, 2)[:, :, 0]
    np.testing.assert_array_equal(tpose, out.T)


def test_idct_idxst_golden(oracle):
    """idct/idxst outputs of the 2-D DCT (inputs rounded to 2 decimals in the file)."""
    d2 = rd("density", "dct_2d_output.dat").reshape(16, 16, 2)[:, :, 0]
    idct = rd("density", "idct_output.dat").reshape(16, 16, 2)[:, :, 0]
    idxst = rd("density", "idxst_output.dat").reshape(16, 16, 2)[:, :, 0]
    idct2 = rd("density", "idct_2d_output.dat").reshape(16, 16, 2)[:, :, 0]
    idxst2 = rd("density", "idxst_2d_output.dat").reshape(16, 16, 2)[:, :, 0]
    for r in range(16):
        np.testing.assert_allclose(oracle.transform1d("idct", d2[r]), idct[r], atol=0.1)
        np.testing.assert_allclose(oracle.transform1d("idxst", d2[r]), idxst[r], atol=0.1)
    np.testing.assert_allclose(oracle.transform2d(d2, "idct", "idct"), idct2, atol=0.2)
    np.testing.assert_allclose(oracle.transform2d(d2, "idxst", "idxst"), idxst2, atol=0.2)
    # round trip: IDCT2(DCT2(rho)) = (N/2)^2 rho; input rows 0-7 are known
    inp = rd("density", "input.dat")[0::2][:128].reshape(8, 16)
    np.testing.assert_allclose(idct2[:8] / 64.0, inp, atol=0.01)


# ------------------------------------------------------------------ field (a9)

@pytest.mark.parametrize("m", [4, 8, 16])
def test_field_matches_direct_sum(oracle, m):
    rho = np.random.default_rng(m).random((m, m))
    a = oracle.field(rho)
    b = oracle.field(rho, direct=True)
    for u, v in zip(a, b):
        np.testing.assert_allclose(u, v, atol=1e-12 * max(1, np.abs(v).max()))


@pytest.mark.parametrize("m", [4, 16, 64, 128])
def test_field_matches_reference_replace_solver(oracle, m):
    """replace::FFT::doFFT (python/AIEmath/fft_replace/fft.cpp:118-212), FP32 Ooura FFT."""
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference tree; `make -C oracle ref`)")
    rho = np.random.default_rng(100 + m).random((m, m)) * 3
    mine = oracle.field(rho)
    ref = oracle.ref_fft_field(rho.astype(np.float32))
    for u, v in zip(mine, ref):
        assert np.abs(u - v).max() <= 2e-5 * np.abs(u).max()


def test_field_single_mode_closed_form(oracle):
    m, p, q = 32, 3, 5
    i = np.arange(m) + 0.5
    cx, cy = np.cos(np.pi * p * i / m), np.cos(np.pi * q * i / m)
    rho = np.outer(cx, cy)
    phi, ex, ey = oracle.field(rho)
    wx, wy = np.pi * p / m, np.pi * q / m
    s = wx * wx + wy * wy
    np.testing.assert_allclose(phi, rho / s, atol=1e-12)
    np.testing.assert_allclose(ex, np.outer(np.sin(np.pi * p * i / m), cy) * wx / s, atol=1e-12)
    np.testing.assert_allclose(ey, np.outer(cx, np.sin(np.pi * q * i / m)) * wy / s, atol=1e-12)
    # a constant map has no field (DC removed)
    phi, ex, ey = oracle.field(np.full((m, m), 7.0))
    assert np.abs(ex).max() < 1e-12 and np.abs(ey).max() < 1e-12


def test_field_wy_scale(oracle):
    """wy = pi v / My * wy_scale: a pure y-mode field scales as 1/wy."""
    m, q = 16, 3
    i = np.arange(m) + 0.5
    rho = np.outer(np.ones(m), np.cos(np.pi * q * i / m))
    _, _, ey1 = oracle.field(rho, wy_scale=1.0)
    _, _, ey2 = oracle.field(rho, wy_scale=2.0)
    np.testing.assert_allclose(ey2, ey1 / 2, atol=1e-12)


# ------------------------------------------------------------------ density / overflow (a6, a7)

def _tiny_design(oracle, xs, ys, ws, hs, die=(0, 0, 8, 8), bins=(16, 16), fixed=0):
    n = len(xs)
    nm = n - fixed
    d = oracle.Design(w=np.array(ws, float), h=np.array(hs, float), n_mov=nm, n_fill=0,
                      fix_x=np.array(xs[nm:], float), fix_y=np.array(ys[nm:], float),
                      net_start=np.array([0], np.int32), pin_cell=np.zeros(0, np.int32),
                      pin_dx=np.zeros(0), pin_dy=np.zeros(0), die=die, bins=bins)
    return d, np.array(xs[:nm], float), np.array(ys[:nm], float)


def test_density_exact_area_and_brute_force(oracle):
    rng = np.random.default_rng(7)
    n = 40
    ws, hs = rng.uniform(0.1, 2.0, n), rng.uniform(0.1, 1.5, n)
    xs, ys = rng.uniform(0, 8 - ws), rng.uniform(0, 8 - hs)
    d, x, y = _tiny_design(oracle, list(xs), list(ys), list(ws), list(hs))
    m = d.density(x, y, "exact")
    assert m.sum() == pytest.approx((ws * hs).sum(), rel=1e-12)
    # brute force per bin: closed-form interval overlap
    bw = 0.5
    for b in rng.integers(0, 256, 20):
        i, j = divmod(int(b), 16)
        ox = np.clip(np.minimum(xs + ws, (i + 1) * bw) - np.maximum(xs, i * bw), 0, None)
        oy = np.clip(np.minimum(ys + hs, (j + 1) * bw) - np.maximum(ys, j * bw), 0, None)
        assert m[i, j] == pytest.approx((ox * oy).sum(), abs=1e-12)


def test_density_stretched_conserves_area(oracle):
    rng = np.random.default_rng(8)
    n = 30
    ws, hs = rng.uniform(0.05, 0.4, n), rng.uniform(0.05, 0.4, n)
    xs, ys = rng.uniform(1, 6, n), rng.uniform(1, 6, n)
    d, x, y = _tiny_design(oracle, list(xs), list(ys), list(ws), list(hs))
    m = d.density(x, y, "stretched")
    assert m.sum() == pytest.approx((ws * hs).sum(), rel=1e-12)


def test_overflow_closed_form(oracle):
    # two unit cells stacked in one 1x1 bin of a 4x4 grid on a 4x4 die
    d, x, y = _tiny_design(oracle, [0, 0], [0, 0], [1, 1], [1, 1], die=(0, 0, 4, 4), bins=(4, 4))
    m = d.density(x, y, "exact")
    fix = np.zeros((4, 4))
    assert d.overflow(m, fix, 1.0) == pytest.approx(0.5)      # (2 - 1) / 2
    assert d.overflow(m, fix, 0.5) == pytest.approx(0.75)     # (2 - 0.5) / 2
    fix[0, 0] = 1.0                                           # macro-covered bin
    assert d.overflow(m, fix, 1.0) == pytest.approx(1.0)


def test_bins_f32_contain_rectangle(oracle):
    rng = np.random.default_rng(9)
    n = 2000
    w = rng.uniform(0.5, 8, n).astype(np.float32)
    h = np.ones(n, np.float32)
    x = rng.uniform(0, 1000 - 8, n).astype(np.float32)
    y = rng.uniform(0, 1000 - 1, n).astype(np.float32)
    bw = np.float32(1000 / 256)
    b = oracle.cell_bins_f32(x, y, w, h, 0, 0, bw, bw, 256, 256, stretch=False)
    xe = x.astype(np.float64)
    assert np.all(b[:, 0] * np.float64(bw) <= xe + 1e-3)
    assert np.all((b[:, 0] + 1) * np.float64(bw) > xe - 1e-3)
    assert np.all(b[:, 1] >= b[:, 0]) and np.all(b[:, 1] <= 255)


def test_hpwl_brute_force(oracle):
    rng = np.random.default_rng(10)
    nc, nn = 50, 30
    deg = rng.integers(1, 6, nn)
    ns = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    pc = rng.integers(0, nc, ns[-1]).astype(np.int32)
    dx, dy = rng.uniform(0, 1, ns[-1]), rng.uniform(0, 1, ns[-1])
    d = oracle.Design(w=np.ones(nc), h=np.ones(nc), n_mov=nc, n_fill=0, fix_x=[], fix_y=[],
                      net_start=ns, pin_cell=pc, pin_dx=dx, pin_dy=dy, die=(0, 0, 100, 100), bins=(16, 16))
    x, y = rng.uniform(0, 99, nc), rng.uniform(0, 99, nc)
    tot = 0.0
    for k in range(nn):
        s = slice(ns[k], ns[k + 1])
        px, py = x[pc[s]] + dx[s], y[pc[s]] + dy[s]
        tot += px.max() - px.min() + py.max() - py.min()
    assert d.hpwl(x, y) == pytest.approx(tot, rel=1e-12)
    # netlist WA gradient: per-cell sum of per-pin partials, fd check
    gx, gy, wa = d.wa_grad(x, y, 2.0)
    h = 1e-6
    for c in rng.integers(0, nc, 5):
        xp, xm = x.copy(), x.copy()
        xp[c] += h
        xm[c] -= h
        fd = (d.wa_grad(xp, y, 2.0)[2] - d.wa_grad(xm, y, 2.0)[2]) / (2 * h)
        assert gx[c] == pytest.approx(fd, abs=1e-5)


# ------------------------------------------------------------------ force / placer (a10, a13)

def test_density_force_closed_forms(oracle):
    """F_i = r_i sum_b ov(i,b) E_b: a uniform field gives F = area * E for cells
    inside the die (smoothing conserves area); a field on one bin only gives
    the overlap with that bin."""
    xs, ys, ws, hs = [2.3, 5.1, 0.05], [3.7, 1.2, 7.0], [2.0, 0.3, 0.2], [1.5, 0.2, 0.4]
    d, x, y = _tiny_design(oracle, xs, ys, ws, hs, die=(0, 0, 8, 8), bins=(16, 16))
    ex, ey = np.full((16, 16), 2.0), np.full((16, 16), -3.0)
    fx, fy = d.density_force(x, y, ex, ey)
    area = np.array(ws) * np.array(hs)
    np.testing.assert_allclose(fx[:2], 2.0 * area[:2], rtol=1e-12)
    np.testing.assert_allclose(fy[:2], -3.0 * area[:2], rtol=1e-12)
    assert fx[2] < 2.0 * area[2]  # smoothed rectangle crosses the die edge: part is lost
    e1 = np.zeros((16, 16))
    e1[5, 8] = 1.0                # bin [2.5, 3] x [4, 4.5], inside cell 0
    e1[4, 7] = 10.0               # bin [2, 2.5] x [3.5, 4]: overlap 0.2 x 0.3
    fx, _ = d.density_force(x, y, e1, e1)
    assert fx[0] == pytest.approx(0.25 + 10.0 * 0.06, rel=1e-12)


def test_placer_mesh_converges(oracle):
    import aieplace_amd as A
    from support import oracle_design, oracle_params
    nl = A.generate("mesh", seed=1)
    p = A.placer_params(A.CONFIGS["mesh_c1"])
    d, x0, y0 = oracle_design(oracle, nl, p)
    x, y, info = d.place(x0, y0, oracle_params(p, d.bin_size[0], nl.n_nets))
    assert info["overflow"] <= 0.10
    # 24x24 lattice with pads: optimum HPWL is 1200 (every net spans 1)
    assert 1150 < info["hpwl"] < 1200 * 1.12
    assert info["iters"] < 300
