"""Reference wire format on the device (AIE packet layout) and the split-FP16
transform form with its FP64 error report."""
import os
import sys

import numpy as np
import pytest

import aieplace_amd as A
from aieplace_amd import api
from support import oracle_design

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures", "partials")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rd(name):
    with open(os.path.join(GOLD, name)) as f:
        return np.array([float(l) for l in f if l.strip()], np.float32)


def packet(inp, gamma):
    import ctypes as C
    out = np.zeros(len(inp) - 4, np.float32)
    inp = np.ascontiguousarray(inp, np.float32)
    rc = api.lib.ep_wa_packet(api.fptr(inp), len(inp), C.c_float(gamma), api.fptr(out), len(out))
    return rc, out


@pytest.mark.parametrize("bench", range(4))
def test_wa_packet_golden_replay(oracle, bench):
    """x_in{b}.dat -> partials{b}.dat (gamma 4) with the reference comparer's
    tolerance (test/compare_output.py:13,56-57: rel 0.01 or abs 1e-2), and
    against the FP64 oracle with true exp."""
    say(f"case wa packet golden {bench}")
    xin = rd(f"x_in{bench}.dat")
    rc, out = packet(xin, 4.0)
    assert rc == api.EP_OK
    gold = rd(f"partials{bench}.dat")
    ok = np.isclose(out, gold, rtol=0.01) | (np.abs(out - gold) < 1e-2)
    assert ok.all()
    xs = xin[4:].astype(np.float64).reshape(2, 8)
    ref = np.stack([oracle.wa_partials(xs[:, l], 4.0) for l in range(8)], axis=1).ravel()
    np.testing.assert_allclose(out, ref, rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("bench", range(4))
def test_wa_packet_fast_exp_golden_bit_exact(bench):
    """EP_WA_EXP_FAST64: the AIE kernel's fast_exp with the partial in FP64, in
    the order of the reference's golden generator.  Every output equals the
    golden partial (an FP64 value) rounded once to FP32, bit for bit; the
    hardware-exp mode of the same entry point equals ep_wa_packet."""
    import ctypes as C
    say(f"case wa packet fast_exp golden {bench}")
    xin = np.ascontiguousarray(rd(f"x_in{bench}.dat"), np.float32)
    gold = rd(f"partials{bench}.dat")
    outs = {}
    for mode in (api.EP_WA_EXP_HW, api.EP_WA_EXP_FAST64):
        out = np.zeros(len(xin) - 4, np.float32)
        rc = api.lib.ep_wa_packet_exp(api.fptr(xin), len(xin), C.c_float(4.0), api.fptr(out), len(out), mode)
        assert rc == api.EP_OK
        outs[mode] = out
    np.testing.assert_array_equal(outs[api.EP_WA_EXP_FAST64], gold)
    np.testing.assert_array_equal(outs[api.EP_WA_EXP_HW], packet(xin, 4.0)[1])
    out = np.zeros(len(xin) - 4, np.float32)
    assert api.lib.ep_wa_packet_exp(api.fptr(xin), len(xin), C.c_float(4.0), api.fptr(out), len(out), 7) \
        == api.EP_EINVAL


def test_wa_packet_fast_exp_wider_nets(oracle):
    """EP_WA_EXP_FAST64 on packets the golden set does not hold: nets of 3-9
    pins with padded lanes, against the FP64 oracle's fast_exp partials (one
    FP32 rounding apart)."""
    import ctypes as C
    rng = np.random.default_rng(5)
    for size, n in ((3, 8), (5, 13), (9, 3)):
        groups = (n + 7) // 8
        vals = np.zeros((groups, size, 8), np.float32)
        for g in range(groups):
            for l in range(8):
                if g * 8 + l < n:
                    x = np.sort(rng.integers(0, 10_000, size)).astype(np.float32)
                    vals[g, :, l] = np.concatenate([[x[-1]], x[:-1]])  # pin 0 = max, then ascending
        inp = np.concatenate([[size, n, 0, 0], vals.ravel()]).astype(np.float32)
        out = np.zeros(vals.size, np.float32)
        rc = api.lib.ep_wa_packet_exp(api.fptr(inp), len(inp), C.c_float(4.0), api.fptr(out), len(out),
                                      api.EP_WA_EXP_FAST64)
        assert rc == api.EP_OK
        out = out.reshape(groups, size, 8)
        for g in range(groups):
            for l in range(8):
                if g * 8 + l < n:
                    ref = oracle.wa_partials(vals[g, :, l].astype(np.float64), 4.0, fast=True)
                    np.testing.assert_allclose(out[g, :, l], ref, rtol=2e-7, atol=1e-12)
                else:
                    assert np.all(out[g, :, l] == 0)


def test_wa_packet_wider_nets_and_padding(oracle):
    rng = np.random.default_rng(2)
    size, n = 7, 13  # two groups, 3 padded lanes
    groups = 2
    vals = rng.uniform(0, 5000, (groups, size, 8)).astype(np.float32)
    vals[1, :, 5:] = 0.0
    inp = np.concatenate([[size, n, 0, 0], vals.ravel()]).astype(np.float32)
    rc, out = packet(inp, 3.0)
    assert rc == api.EP_OK
    out = out.reshape(groups, size, 8)
    for g in range(groups):
        for l in range(8):
            if g * 8 + l < n:
                np.testing.assert_allclose(out[g, :, l], oracle.wa_partials(vals[g, :, l], 3.0),
                                           rtol=1e-4, atol=1e-5)
            else:
                assert np.all(out[g, :, l] == 0)
    bad = inp.copy()
    bad[0] = 2.5
    assert packet(bad, 3.0)[0] == api.EP_EINVAL
    assert packet(inp[:20], 3.0)[0] == api.EP_EINVAL


@pytest.mark.parametrize("bins", [32, 128, 512, 1024])
def test_field_fp16x3_and_error_report(oracle, bins):
    """Split-FP16 products with power-of-two operand scaling; the FP64 device
    reference solve reports the error of each form (checked against the oracle)."""
    say(f"case field fp16x3 bins={bins}")
    n = {32: 5_000, 128: 20_000, 512: 100_000, 1024: 200_000}[bins]
    nl = A.generate("stdcell", seed=1, n_movable=n)
    p = A.placer_params(A.CONFIGS["stdcell_200k"], bins_x=bins, bins_y=bins)
    d, _, _ = oracle_design(oracle, nl, p)
    with A.Placer(nl, p) as pl:
        rng = np.random.default_rng(2)
        xl, yl, xh, yh = nl.die
        x = (xl + rng.random(d.n_place) * (xh - xl) * 0.7).astype(np.float32)
        y = (yl + rng.random(d.n_place) * (yh - yl) * 0.7).astype(np.float32)
        pl.set_positions(x, y)
        fmap, _, _, _ = pl.computeOverlaps()
        bw = (xh - xl) / bins
        _, oex, oey = oracle.field(fmap.astype(np.float64) / (bw * bw), wy_scale=1.0)
        errs = {}
        for name, prec in (("fp32", api.EP_PREC_FP32), ("bf16x3", api.EP_PREC_BF16X3),
                           ("fp16x3", api.EP_PREC_FP16X3)):
            rex, rey = pl.field_error(prec)
            ex, ey = pl.computeElectricFields(prec)
            oerr = max(np.abs(ex - oex).max() / np.abs(oex).max(), np.abs(ey - oey).max() / np.abs(oey).max())
            errs[name] = (max(rex, rey), oerr)
            # the device FP64 reference agrees with the oracle's FP64 solve
            assert max(rex, rey) == pytest.approx(oerr, rel=0.05, abs=2e-7)
        say("  " + "  ".join(f"{k}: device-reported {v[0]:.2e} (oracle {v[1]:.2e})" for k, v in errs.items()))
        assert errs["fp32"][1] <= 1e-5
        assert errs["fp16x3"][1] <= 2e-4 and errs["bf16x3"][1] <= 2e-4


def test_placement_fp16x3():
    nl = A.generate("stdcell", seed=1, n_movable=20_000)
    p = A.placer_params(A.CONFIGS["stdcell_200k"], bins_x=128, bins_y=128)
    res = {}
    for prec in (api.EP_PREC_FP32, api.EP_PREC_FP16X3):
        p.precision = prec
        with A.Placer(nl, p) as pl:
            res[prec] = pl.run()
    a, b = res[api.EP_PREC_FP32], res[api.EP_PREC_FP16X3]
    say(f"case placement fp16x3: fp32 hpwl {a['hpwl']:.1f}, fp16x3 hpwl {b['hpwl']:.1f}")
    assert b["overflow"] <= 0.10
    assert b["hpwl"] == pytest.approx(a["hpwl"], rel=5e-3)


# ------------------------------------------------------------------ transforms (ep_transform)

DENS = os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures", "density")


def rd_density(name):
    with open(os.path.join(DENS, name)) as f:
        return np.array([float(l) for l in f if l.strip()])


def transform(x, row_kind, col_kind, precision=api.EP_PREC_FP32):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    api.check(api.lib.ep_transform(row_kind, col_kind, x.shape[0], x.shape[1], api.fptr(x), api.fptr(out),
                                   precision))
    return out


KIND = {"dct": api.EP_TR_DCT, "idct": api.EP_TR_IDCT, "idxst": api.EP_TR_IDXST}


def test_transform_golden_replay():
    """The AIE density graphs' fixtures (vck5000/aie/golden_data/density, fed
    through DCT_in -> DCT_out, DCTGraph.h:41-42 / IDCTGraph.h:40-41 /
    IDXSTGraph.h:42-43) through the MFMA product path, at the fixtures' own
    tolerance (values printed with two decimals; input.dat holds rows 0-7)."""
    say("case transform golden replay (16 x 16)")
    inp = rd_density("input.dat")[0::2][:128].reshape(8, 16)
    dct = rd_density("dct_output.dat").reshape(16, 16, 2)[:, :, 0]
    d2 = rd_density("dct_2d_output.dat").reshape(16, 16, 2)[:, :, 0]
    idct = rd_density("idct_output.dat").reshape(16, 16, 2)[:, :, 0]
    idxst = rd_density("idxst_output.dat").reshape(16, 16, 2)[:, :, 0]
    idct2 = rd_density("idct_2d_output.dat").reshape(16, 16, 2)[:, :, 0]
    idxst2 = rd_density("idxst_2d_output.dat").reshape(16, 16, 2)[:, :, 0]
    for prec in (api.EP_PREC_FP32, api.EP_PREC_BF16X3, api.EP_PREC_FP16X3):
        x = np.zeros((16, 16), np.float32)
        x[:8] = inp
        got = transform(x, api.EP_TR_DCT, api.EP_TR_NONE, prec)
        err = np.abs(got[:8] - dct[:8]).max()
        say(f"  precision {prec}: dct rows max err {err:.4f}")
        assert err <= 0.01
        for kind, ref in (("idct", idct), ("idxst", idxst)):
            got = transform(d2, KIND[kind], api.EP_TR_NONE, prec)
            err = np.abs(got - ref).max()
            say(f"  precision {prec}: {kind} rows max err {err:.4f}")
            assert err <= 0.1
        for kind, ref in (("idct", idct2), ("idxst", idxst2)):
            got = transform(d2, KIND[kind], KIND[kind], prec)
            err = np.abs(got - ref).max()
            say(f"  precision {prec}: {kind} 2-D max err {err:.4f}")
            assert err <= 0.2
        # IDCT2(DCT2(input)) = (N/2)^2 input on the known rows
        back = transform(transform(x, api.EP_TR_DCT, api.EP_TR_DCT, prec), api.EP_TR_IDCT, api.EP_TR_IDCT, prec)
        np.testing.assert_allclose(back[:8] / 64.0, inp, atol=1e-4)


@pytest.mark.parametrize("shape", [(16, 32), (128, 256), (512, 512), (1024, 96)])
def test_transform_matches_oracle(oracle, shape):
    """Every kind pair, one axis or both, against the FP64 oracle transforms
    (eo_transform2d): FP32 within 1e-5 and the split forms within 5e-5 of the
    result's max-norm."""
    say(f"case transform vs oracle {shape}")
    rng = np.random.default_rng(shape[0] + shape[1])
    x = rng.standard_normal(shape).astype(np.float32)
    names = ("dct", "idct", "idxst")
    for rk in names:
        for ck in (None,) + names:
            if ck is None:
                ref = np.stack([oracle.transform1d(rk, r) for r in x.astype(np.float64)])
            else:
                ref = oracle.transform2d(x.astype(np.float64), ck, rk)
            for prec, tol in ((api.EP_PREC_FP32, 1e-5), (api.EP_PREC_BF16X3, 5e-5), (api.EP_PREC_FP16X3, 5e-5)):
                got = transform(x, KIND[rk], api.EP_TR_NONE if ck is None else KIND[ck], prec)
                err = np.abs(got - ref).max() / np.abs(ref).max()
                assert err <= tol, (rk, ck, prec, err)
    # one axis along the columns only
    ref = np.stack([oracle.transform1d("idxst", c) for c in x.astype(np.float64).T]).T
    got = transform(x, api.EP_TR_NONE, api.EP_TR_IDXST)
    assert np.abs(got - ref).max() / np.abs(ref).max() <= 1e-5


def test_transform_rejects_bad_arguments():
    x = np.zeros((16, 16), np.float32)
    out = np.zeros_like(x)
    for args in ((api.EP_TR_NONE, api.EP_TR_NONE, 16, 16), (5, api.EP_TR_DCT, 16, 16),
                 (api.EP_TR_DCT, api.EP_TR_NONE, 24, 16), (api.EP_TR_DCT, api.EP_TR_NONE, 0, 16)):
        assert api.lib.ep_transform(*args, api.fptr(x), api.fptr(out), api.EP_PREC_FP32) == api.EP_EINVAL
    x[3, 3] = np.nan
    assert api.lib.ep_transform(api.EP_TR_DCT, api.EP_TR_NONE, 16, 16, api.fptr(x), api.fptr(out),
                                api.EP_PREC_FP32) == api.EP_ENAN
