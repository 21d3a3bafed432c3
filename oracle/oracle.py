"""ctypes front-end of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; the product
package (aieplace_amd) never imports it.

The C sources (eplace_oracle.c) cite the reference file:line each function
restates.  `ref_fft_field` calls the reference's own RePlAce field solver
(python/AIEmath/fft_replace/fft.cpp, compiled in place by `make ref`) when
oracle/_ref/libref_fft.so exists.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_fft.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(with_ref: bool = True) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if with_ref and os.path.isdir("/root/reference/python/AIEmath/fft_replace"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


class EoDesign(C.Structure):
    _fields_ = [
        ("n_mov", C.c_int32), ("n_fill", C.c_int32), ("n_fix", C.c_int32),
        ("w", C.c_void_p), ("h", C.c_void_p), ("fix_x", C.c_void_p), ("fix_y", C.c_void_p),
        ("n_nets", C.c_int32), ("net_start", C.c_void_p), ("pin_cell", C.c_void_p),
        ("pin_dx", C.c_void_p), ("pin_dy", C.c_void_p),
        ("xl", C.c_double), ("yl", C.c_double), ("xh", C.c_double), ("yh", C.c_double),
        ("bins_x", C.c_int32), ("bins_y", C.c_int32),
    ]


class EoParams(C.Structure):
    _fields_ = [
        ("target_density", C.c_double), ("gamma_bins", C.c_double),
        ("gamma_schedule", C.c_int32), ("lambda_eta", C.c_double),
        ("hpwl_ref", C.c_double), ("mu_max", C.c_double), ("mu_min", C.c_double),
        ("max_backtrack", C.c_int32), ("stop_overflow", C.c_double),
        ("max_iters", C.c_int32), ("init_disp_bins", C.c_double),
    ]


class EoIterLog(C.Structure):
    _fields_ = [
        ("iter", C.c_int32), ("evals", C.c_int32), ("hpwl", C.c_double),
        ("overflow", C.c_double), ("lambda_", C.c_double), ("gamma", C.c_double),
        ("alpha", C.c_double), ("alpha_used", C.c_double), ("coef", C.c_double),
    ]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(with_ref=False)
        L = C.CDLL(LIB_PATH)
        L.eo_fast_exp.restype = C.c_double
        L.eo_fast_exp.argtypes = [C.c_double, C.c_int]
        L.eo_wa_terms.argtypes = [_dp, C.c_int, C.c_double, C.c_int, _dp, _dp, _dp]
        L.eo_wa_partials.argtypes = [_dp, C.c_int, C.c_double, C.c_int, _dp]
        L.eo_wa_value.restype = C.c_double
        L.eo_wa_value.argtypes = [_dp, C.c_int, C.c_double, C.c_int]
        L.eo_wa_grad.argtypes = [C.c_void_p, _dp, _dp, C.c_double, _dp, _dp, C.c_void_p]
        L.eo_hpwl.restype = C.c_double
        L.eo_hpwl.argtypes = [C.c_void_p, _dp, _dp]
        L.eo_cell_bins_f32.argtypes = [C.c_int, _fp, _fp, _fp, _fp] + [C.c_float] * 6 + [
            C.c_int, C.c_int, C.c_int, _ip]
        L.eo_density_exact.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_int, _dp]
        L.eo_density_stretched.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_int, _dp]
        L.eo_density_fixed.argtypes = [C.c_void_p, _dp]
        L.eo_overflow.restype = C.c_double
        L.eo_overflow.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
        for f in ("eo_dct", "eo_idct", "eo_idxst", "eo_idst"):
            getattr(L, f).argtypes = [_dp, C.c_int, _dp]
        L.eo_transform2d.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.eo_field.argtypes = [_dp, C.c_int, C.c_int, C.c_double, C.c_void_p, _dp, _dp]
        L.eo_field_direct.argtypes = [_dp, C.c_int, C.c_int, C.c_double, C.c_void_p, _dp, _dp]
        L.eo_density_force.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
        L.eo_place.restype = C.c_int
        L.eo_place.argtypes = [C.c_void_p, C.POINTER(EoParams), _dp, _dp, C.c_void_p, C.c_int,
                               C.POINTER(C.c_double), C.POINTER(C.c_double),
                               C.POINTER(C.c_int64)]
        L.eo_gradient.restype = C.c_double
        L.eo_gradient.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                  _dp, _dp, C.POINTER(C.c_double)] + [C.c_void_p] * 6
        L.eo_set_threads.argtypes = [C.c_int]
        L.eo_get_threads.restype = C.c_int
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref_fft_field(rho: np.ndarray, bin_size_x: int = 1, bin_size_y: int = 1):
    """Field from the reference's replace::FFT::doFFT (FP32).  rho is (Mx, My)."""
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_PATH)
        _ref.ref_fft_field.restype = C.c_int
        _ref.ref_fft_field.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _fp, _fp, _fp, _fp]
    rho = np.ascontiguousarray(rho, dtype=np.float32)
    mx, my = rho.shape
    phi = np.zeros_like(rho)
    ex = np.zeros_like(rho)
    ey = np.zeros_like(rho)
    rc = _ref.ref_fft_field(mx, my, bin_size_x, bin_size_y, rho, phi, ex, ey)
    if rc != 0:
        raise ValueError(f"ref_fft_field rc={rc}")
    return phi, ex, ey


def set_threads(n: int) -> None:
    lib().eo_set_threads(int(n))


def threads() -> int:
    return int(lib().eo_get_threads())


# ---------------------------------------------------------------- primitives

def fast_exp(a: float, prec: int = 16) -> float:
    return float(lib().eo_fast_exp(float(a), int(prec)))


def wa_terms(x, gamma: float, fast: bool = False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    ap = np.zeros_like(x)
    am = np.zeros_like(x)
    bc = np.zeros(4)
    lib().eo_wa_terms(x, len(x), float(gamma), int(fast), ap, am, bc)
    return ap, am, bc


def wa_partials(x, gamma: float, fast: bool = False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    lib().eo_wa_partials(x, len(x), float(gamma), int(fast), out)
    return out


def wa_value(x, gamma: float, fast: bool = False) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return float(lib().eo_wa_value(x, len(x), float(gamma), int(fast)))


def transform1d(kind: str, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    getattr(lib(), "eo_" + kind)(x, len(x), y)
    return y


_KIND = {"dct": 0, "idct": 1, "idxst": 2, "idst": 3}


def transform2d(x, col_kind: str, row_kind: str):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    lib().eo_transform2d(x, x.shape[0], x.shape[1], _KIND[col_kind], _KIND[row_kind], y)
    return y


def field(rho, wy_scale: float = 1.0, direct: bool = False):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    mx, my = rho.shape
    phi = np.zeros_like(rho)
    ex = np.zeros_like(rho)
    ey = np.zeros_like(rho)
    f = lib().eo_field_direct if direct else lib().eo_field
    f(rho, mx, my, float(wy_scale), phi.ctypes.data, ex, ey)
    return phi, ex, ey


def rows_naive(x, kind: str):
    """Row transforms with one cos() per term (the reference's DCT_naive /
    IDCT_naive / IDXST_naive); rows of a float32 (rows, n) array, n <= 16384."""
    L = lib()
    L.eo_rows_naive.argtypes = [_fp, C.c_int, C.c_int, C.c_int, _fp]
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros_like(x)
    L.eo_rows_naive(x, x.shape[0], x.shape[1], {"dct": 0, "idct": 1, "idxst": 2}[kind], out)
    return out


def cell_bins_f32(x, y, w, h, xl, yl, bw, bh, mx, my, stretch=True):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    h = np.ascontiguousarray(h, dtype=np.float32)
    out = np.zeros(4 * len(x), dtype=np.int32)
    s2bw = np.float32(np.sqrt(2.0) * np.float64(bw))
    s2bh = np.float32(np.sqrt(2.0) * np.float64(bh))
    lib().eo_cell_bins_f32(len(x), x, y, w, h, np.float32(xl), np.float32(yl), np.float32(bw),
                           np.float32(bh), s2bw, s2bh, int(mx), int(my), int(stretch), out)
    return out.reshape(-1, 4)


# ---------------------------------------------------------------- designs

class Design:
    """FP64 view of a placement instance in the oracle's layout.

    Cells are ordered [movable | filler | fixed]; `fix_x/fix_y` hold the fixed
    cells' lower-left corners.  Arrays are kept alive by this object.
    """

    def __init__(self, *, w, h, n_mov, n_fill, fix_x, fix_y, net_start, pin_cell, pin_dx,
                 pin_dy, die, bins):
        self.w = np.ascontiguousarray(w, dtype=np.float64)
        self.h = np.ascontiguousarray(h, dtype=np.float64)
        self.fix_x = np.ascontiguousarray(fix_x, dtype=np.float64)
        self.fix_y = np.ascontiguousarray(fix_y, dtype=np.float64)
        self.net_start = np.ascontiguousarray(net_start, dtype=np.int32)
        self.pin_cell = np.ascontiguousarray(pin_cell, dtype=np.int32)
        self.pin_dx = np.ascontiguousarray(pin_dx, dtype=np.float64)
        self.pin_dy = np.ascontiguousarray(pin_dy, dtype=np.float64)
        self.n_mov, self.n_fill = int(n_mov), int(n_fill)
        self.n_fix = len(self.fix_x)
        self.die = tuple(float(v) for v in die)
        self.bins = (int(bins[0]), int(bins[1]))
        assert len(self.w) == self.n_mov + self.n_fill + self.n_fix
        self.c = EoDesign(
            self.n_mov, self.n_fill, self.n_fix,
            self.w.ctypes.data, self.h.ctypes.data, self.fix_x.ctypes.data,
            self.fix_y.ctypes.data, len(self.net_start) - 1, self.net_start.ctypes.data,
            self.pin_cell.ctypes.data, self.pin_dx.ctypes.data, self.pin_dy.ctypes.data,
            *self.die, *self.bins)

    @property
    def ptr(self):
        return C.addressof(self.c)

    @property
    def n_place(self) -> int:
        return self.n_mov + self.n_fill

    @property
    def bin_size(self):
        return ((self.die[2] - self.die[0]) / self.bins[0], (self.die[3] - self.die[1]) / self.bins[1])

    def wa_grad(self, x, y, gamma):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        gx = np.zeros(self.n_place)
        gy = np.zeros(self.n_place)
        wa = C.c_double(0)
        lib().eo_wa_grad(self.ptr, x, y, float(gamma), gx, gy, C.addressof(wa))
        return gx, gy, wa.value

    def hpwl(self, x, y) -> float:
        return float(lib().eo_hpwl(self.ptr, np.ascontiguousarray(x, dtype=np.float64),
                                   np.ascontiguousarray(y, dtype=np.float64)))

    def density(self, x, y, kind="exact", c0=0, c1=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        m = np.zeros(self.bins[0] * self.bins[1])
        c1 = self.n_mov if c1 is None else c1
        if kind == "exact":
            lib().eo_density_exact(self.ptr, x, y, c0, c1, m)
        else:
            lib().eo_density_stretched(self.ptr, x, y, c0, c1, m)
        return m.reshape(self.bins)

    def density_fixed(self):
        m = np.zeros(self.bins[0] * self.bins[1])
        lib().eo_density_fixed(self.ptr, m)
        return m.reshape(self.bins)

    def overflow(self, map_mov, map_fix, rho_t) -> float:
        return float(lib().eo_overflow(self.ptr, np.ascontiguousarray(map_mov, dtype=np.float64).ravel(),
                                       np.ascontiguousarray(map_fix, dtype=np.float64).ravel(),
                                       float(rho_t)))

    def density_force(self, x, y, ex, ey):
        fx = np.zeros(self.n_place)
        fy = np.zeros(self.n_place)
        lib().eo_density_force(self.ptr, np.ascontiguousarray(x, dtype=np.float64),
                               np.ascontiguousarray(y, dtype=np.float64),
                               np.ascontiguousarray(ex, dtype=np.float64).ravel(),
                               np.ascontiguousarray(ey, dtype=np.float64).ravel(), fx, fy)
        return fx, fy

    def gradient(self, x, y, rho_t, gamma, lam):
        """Full preconditioned gradient; returns dict of components."""
        n = self.n_place
        nb = self.bins[0] * self.bins[1]
        out = {k: np.zeros(n) for k in ("gx", "gy", "gwx", "gwy", "fx", "fy")}
        out["ex"] = np.zeros(nb)
        out["ey"] = np.zeros(nb)
        hp = C.c_double(0)
        tau = lib().eo_gradient(self.ptr, np.ascontiguousarray(x, dtype=np.float64),
                                np.ascontiguousarray(y, dtype=np.float64), float(rho_t),
                                float(gamma), float(lam), out["gx"], out["gy"], C.byref(hp),
                                out["gwx"].ctypes.data, out["gwy"].ctypes.data,
                                out["fx"].ctypes.data, out["fy"].ctypes.data,
                                out["ex"].ctypes.data, out["ey"].ctypes.data)
        out["tau"] = float(tau)
        out["hpwl"] = hp.value
        return out

    def place(self, x, y, params: dict, log_cap: int = 100000):
        """Runs the FP64 Nesterov placer; returns (x, y, info)."""
        x = np.array(x, dtype=np.float64)
        y = np.array(y, dtype=np.float64)
        p = EoParams(**params)
        logs = (EoIterLog * log_cap)()
        fh = C.c_double(0)
        fo = C.c_double(0)
        ev = C.c_int64(0)
        it = lib().eo_place(self.ptr, C.byref(p), x, y, C.addressof(logs), log_cap,
                            C.byref(fh), C.byref(fo), C.byref(ev))
        log = [dict(iter=l.iter, evals=l.evals, hpwl=l.hpwl, overflow=l.overflow,
                    lam=l.lambda_, gamma=l.gamma, alpha=l.alpha, alpha_used=l.alpha_used,
                    coef=l.coef) for l in logs[:min(it, log_cap)]]
        return x, y, dict(iters=it, hpwl=fh.value, overflow=fo.value, evals=ev.value, log=log)


# ---------------------------------------------------------------- step-length MLP

MLP_IN, MLP_HID = 16, 64


def mlp_nweights() -> int:
    L = lib()
    L.eo_mlp_nweights.restype = C.c_int
    return int(L.eo_mlp_nweights())


def mlp_forward(w, x):
    L = lib()
    L.eo_mlp_forward.argtypes = [_dp, _dp, C.c_int, _dp]
    w = np.ascontiguousarray(w, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, MLP_IN)
    out = np.zeros(len(x))
    L.eo_mlp_forward(w, x, len(x), out)
    return out


def mlp_grad(w, x, y):
    L = lib()
    L.eo_mlp_grad.argtypes = [_dp, _dp, _dp, C.c_int, _dp]
    L.eo_mlp_grad.restype = C.c_double
    w = np.ascontiguousarray(w, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, MLP_IN)
    y = np.ascontiguousarray(y, dtype=np.float64)
    g = np.zeros(len(w))
    loss = L.eo_mlp_grad(w, x, y, len(y), g)
    return g, float(loss)


def mlp_random_weights(rng, scale=0.3):
    """A valid blob: normalisation header + random layer weights (tests)."""
    n = mlp_nweights()
    w = rng.standard_normal(n) * scale
    w[0:16] = rng.standard_normal(16)
    w[16:32] = rng.uniform(0.5, 2.0, 16)
    w[32], w[33] = rng.standard_normal(), rng.uniform(0.5, 2.0)
    return w
