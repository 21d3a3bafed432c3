"""Host-side checks that need no GPU: the C ABI library loads and exports every
declared symbol, generators produce the specified instances, filler planning,
and argument validation (which runs before any device call)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import aieplace_amd as A
from aieplace_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(ep_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(api.LIB_PATH)
    decl = declared_symbols()
    assert len(decl) >= 25
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(api.EXPORTS) == decl


def test_mesh_generator_structure():
    nl = A.generate("mesh", seed=1)
    assert nl.n_mov == 576 and nl.n_fix == 96
    assert nl.n_nets == 1200 and nl.n_pins == 2400
    assert nl.die == (0.0, 0.0, 26.0, 26.0)
    assert np.all(nl.w == 1) and np.all(nl.h == 1)
    assert np.all(nl.pin_dx == 0) and np.all(nl.pin_dy == 0)
    deg = np.diff(nl.net_start)
    assert np.all(deg == 2)
    # pads on the perimeter ring, 24 per side
    px, py = nl.x[576:], nl.y[576:]
    on_ring = (px == 0) | (px == 25) | (py == 0) | (py == 25)
    assert on_ring.all()
    # every edge cell is wired to one pad (corner cells to two)
    pc = nl.pin_cell.reshape(-1, 2)
    pad_nets = pc[(pc >= 576).any(axis=1)]
    assert len(pad_nets) == 96
    # deterministic per seed, different across seeds
    nl2 = A.generate("mesh", seed=1)
    assert np.array_equal(nl.x, nl2.x)
    assert not np.array_equal(nl.x, A.generate("mesh", seed=2).x)


def test_stdcell_generator_statistics():
    nl = A.generate("stdcell", seed=1, n_movable=50_000)
    n = nl.n_mov
    assert n == 50_000
    w = nl.w[:n]
    assert set(np.unique(w)) == set(range(1, 9)) and np.all(nl.h[:n] == 1)
    deg = np.diff(nl.net_start)
    assert abs(nl.n_nets / n - 1.05) < 0.01
    f = lambda m: np.mean(m)
    assert abs(f(deg == 2) - 0.60) < 0.01
    assert abs(f(deg == 3) - 0.18) < 0.01
    assert abs(f((deg >= 4) & (deg <= 10)) - 0.20) < 0.01
    assert abs(f(deg >= 11) - 0.02) < 0.005 and deg.max() <= 64
    # pins of one net are distinct cells, offsets inside the cell
    for k in range(0, nl.n_nets, 997):
        cells = nl.pin_cell[nl.net_start[k]:nl.net_start[k + 1]]
        assert len(set(cells.tolist())) == len(cells)
    assert np.all(nl.pin_dx >= 0) and np.all(nl.pin_dx <= nl.w[nl.pin_cell])
    assert np.all(nl.pin_dy >= 0) and np.all(nl.pin_dy <= nl.h[nl.pin_cell])
    # locality: pins within a few sigma of each other in the 1-D order
    span = [np.ptp(nl.pin_cell[nl.net_start[k]:nl.net_start[k + 1]]) for k in range(0, nl.n_nets, 101)]
    assert np.median(span) < 600
    # die: utilisation 0.7 of the free area, macros 2% of the die, no overlap
    W = nl.die[2]
    macro = nl.w[n:] * nl.h[n:]
    assert abs(macro.sum() / (W * W) - 0.02) < 0.01
    assert abs(w.sum() / (W * W - macro.sum()) - 0.7) < 0.01
    mx, my, ms = nl.x[n:], nl.y[n:], nl.w[n:]
    assert np.all(ms >= 20) and np.all(ms <= 80) and np.all(nl.h[n:] == ms)
    for i in range(len(ms)):
        for j in range(i + 1, len(ms)):
            assert (mx[i] + ms[i] <= mx[j] or mx[j] + ms[j] <= mx[i] or
                    my[i] + ms[i] <= my[j] or my[j] + ms[j] <= my[i])
    # initial centres: die centre + N(0, (0.01 W)^2)
    cx = nl.x[:n] + w / 2
    assert abs(cx.mean() - W / 2) < 0.01 * W * 0.05
    assert abs(cx.std() - 0.01 * W) < 0.001 * W


def test_filler_plan_closed_form():
    nl = A.generate("mesh", seed=1)
    p = A.placer_params(A.CONFIGS["mesh_c1"])
    n_fill, fw, fh, fx, fy = A.netlist.plan_fillers(nl, p)
    # rho_t * (676 - 96) - 576 = 4 unit fillers
    assert (n_fill, fw, fh) == (4, 1.0, 1.0)
    assert np.all((fx >= 0) & (fx <= 25)) and np.all((fy >= 0) & (fy <= 25))
    p.use_fillers = 0
    assert A.netlist.plan_fillers(nl, p)[0] == 0
    nl2 = A.generate("stdcell", seed=3, n_movable=10_000)
    p2 = A.placer_params(A.CONFIGS["stdcell_200k"], bins_x=128, bins_y=128)
    n_fill, fw, fh, _, _ = A.netlist.plan_fillers(nl2, p2)
    W = nl2.die[2]
    fix = (nl2.w[10_000:] * nl2.h[10_000:]).sum()
    mov = (nl2.w[:10_000] * nl2.h[:10_000]).sum()
    expect = np.floor((0.9 * (W * W - fix) - mov) / (fw * fh))
    assert n_fill == expect
    assert fw == pytest.approx(nl2.w[:10_000].mean(), rel=1e-6)


def _create(nl, p):
    ctx = C.c_void_p()
    rc = api.lib.ep_create(C.byref(nl.c_struct()), C.byref(p), C.byref(ctx))
    return rc, ctx


def test_validation_errors_before_device():
    nl = A.generate("mesh", seed=1)
    p = A.placer_params(A.CONFIGS["mesh_c1"], bins_x=30)
    rc, _ = _create(nl, p)
    assert rc == api.EP_EINVAL and b"multiples of 16" in api.lib.ep_last_error()
    p = A.placer_params(A.CONFIGS["mesh_c1"], target_density=1.5)
    assert _create(nl, p)[0] == api.EP_EINVAL
    bad = A.generate("mesh", seed=1)
    bad.pin_cell[5] = 10_000
    assert _create(bad, A.placer_params(A.CONFIGS["mesh_c1"]))[0] == api.EP_EINVAL
    nan = A.generate("mesh", seed=1)
    nan.x[3] = np.nan
    assert _create(nan, A.placer_params(A.CONFIGS["mesh_c1"]))[0] == api.EP_ENAN
    ns = A.generate("mesh", seed=1)
    ns.net_start[10] = ns.net_start[12]
    assert _create(ns, A.placer_params(A.CONFIGS["mesh_c1"]))[0] == api.EP_EINVAL


def test_create_without_device_reports_edevice():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    nl = A.generate("mesh", seed=1)
    rc, _ = _create(nl, A.placer_params(A.CONFIGS["mesh_c1"]))
    assert rc == api.EP_EDEVICE


def test_null_arguments():
    assert api.lib.ep_destroy(None) == api.EP_EINVAL
    assert api.lib.ep_step(None, None) == api.EP_EINVAL
    assert api.lib.ep_create(None, None, None) == api.EP_EINVAL


def _batch_create(nls, bins, p):
    structs = [nl.c_struct() for nl in nls]
    arr = (C.POINTER(api.EpNetlist) * len(nls))(*[C.pointer(s) for s in structs])
    b = np.asarray(bins, np.int32)
    h = C.c_void_p()
    rc = api.lib.ep_batch_create(arr, api.iptr(b), len(nls), C.byref(p), C.byref(h))
    return rc, h


def test_batch_validation_errors_before_device():
    nls = [A.generate("stdcell", seed=s, n_movable=2_000) for s in (1, 2)]
    p = A.placer_params(A.CONFIGS["stdcell_200k"])
    rc, _ = _batch_create(nls, [128, 100], p)
    assert rc == api.EP_EINVAL and b"instance 1" in api.lib.ep_last_error()
    assert api.lib.ep_batch_create(None, None, 0, C.byref(p), C.byref(C.c_void_p())) == api.EP_EINVAL
    q = A.placer_params(A.CONFIGS["stdcell_200k"], learned_step=1)
    assert _batch_create(nls, [128, 128], q)[0] == api.EP_ENOTSUP
    bad = A.generate("stdcell", seed=3, n_movable=2_000)
    bad.pin_cell[7] = 10 ** 7
    rc, _ = _batch_create([nls[0], bad], [128, 128], p)
    assert rc == api.EP_EINVAL and b"instance 1" in api.lib.ep_last_error()
    assert api.lib.ep_batch_destroy(None) == api.EP_EINVAL
    assert api.lib.ep_batch_run(None, 1, None) == api.EP_EINVAL


def test_batch_create_without_device_reports_edevice():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    nls = [A.generate("stdcell", seed=s, n_movable=2_000) for s in (1, 2)]
    rc, _ = _batch_create(nls, [128, 128], A.placer_params(A.CONFIGS["stdcell_200k"]))
    assert rc == api.EP_EDEVICE


def test_batch_workload_sizes_log_uniform():
    from aieplace_amd.batch import batch_cell_counts
    from aieplace_amd.configs import bins_for_cells

    counts = batch_cell_counts(64)
    assert len(counts) == 64 and min(counts) >= 5_000 and max(counts) <= 150_000
    logs = np.log(counts)
    assert logs.max() - logs.min() > 0.8 * (np.log(150_000) - np.log(5_000))
    assert {bins_for_cells(c) for c in counts} == {128, 256, 512}


def test_integration_quotes_the_compiled_cpp_example():
    """INTEGRATION.md shows the C++ Placer example that the build compiles
    (aieplace_amd/csrc/placer_example.cpp, between its markers) verbatim, and
    the compiled example rejects bad arguments before touching a device."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = open(os.path.join(root, "aieplace_amd", "csrc", "placer_example.cpp")).read()
    body = src.split("// ---- BEGIN INTEGRATION EXAMPLE\n")[1].split("    // ---- END INTEGRATION EXAMPLE")[0]
    doc = open(os.path.join(root, "INTEGRATION.md")).read()
    assert body in doc
    exe = os.path.join(root, "aieplace_amd", "placer_example")
    assert subprocess.run([exe, "torus"], capture_output=True).returncode == 2
    assert subprocess.run([exe, "mesh", "--bogus", "1"], capture_output=True).returncode == 2
