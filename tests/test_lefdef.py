"""LEF/DEF reader and DEF writer (include/eplace_io.h): host code, no GPU.

Pins:
  * the Limbo DEF parser test designs the reference ships (tests/golden/lefdef,
    copied by tests/golden/make_lefdef_fixtures.py), read with a LEF of this
    test's own macros: counts, statuses, locations, IO pins and connections as
    the files state them, and the unknown-component error the original file
    provokes (its nets name `u4`, its components `u4[0]`; the reference
    asserts on that, DataBase.cpp:392-394);
  * orientations: pin offsets against rotation/mirror matrices applied to the
    macro's corners (a computation independent of the reader's table);
  * a generated netlist written as LEF/DEF text, read back exactly (to database
    units), then written by ep_lefdef_write_def and read again;
  * block-structure skipping (PROPERTYDEFINITIONS with DESIGN lines, VIAS,
    SPECIALNETS, routed nets, ROWs split over lines, comments, escaped names).
"""
import os

import numpy as np
import pytest

import aieplace_amd as A
from aieplace_amd import api
from aieplace_amd.lefdef import LefDef
from support import netlist_lefdef, round_half_away

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "lefdef")

# The cells simple.def uses, with this test's own sizes and pin shapes (microns).
SIMPLE_LEF = """
VERSION 5.7 ;
BUSBITCHARS "[]" ;
DIVIDERCHAR "/" ;
UNITS
  DATABASE MICRONS 1000 ;
END UNITS
SITE core
  CLASS CORE ;
  SIZE 0.2 BY 2.0 ;
END core
LAYER metal1
  TYPE ROUTING ;
  DIRECTION HORIZONTAL ;
END metal1
MACRO NAND2_X1
  CLASS CORE ;
  ORIGIN 0 0 ;
  SIZE 0.8 BY 2.0 ;
  PIN A1
    DIRECTION INPUT ;
    PORT
      LAYER metal1 ;
        RECT 0.1 0.9 0.2 1.1 ;
    END
  END A1
  PIN A2
    PORT
      LAYER metal1 ;
        RECT 0.3 0.9 0.4 1.1 ;
    END
  END A2
  PIN ZN
    PORT
      LAYER metal1 ;
        POLYGON 0.6 0.2 0.7 0.2 0.7 1.8 0.6 1.8 ;
    END
  END ZN
  OBS
    LAYER metal1 ;
      RECT 0 0 0.8 0.1 ;
  END
END NAND2_X1
MACRO NOR2_X1
  SIZE 0.8 BY 2.0 ;
  PIN A1 PORT LAYER metal1 ; RECT 0.1 0.4 0.3 0.6 ; END END A1
  PIN A2 PORT LAYER metal1 ; RECT 0.3 0.4 0.5 0.6 ; END END A2
  PIN ZN PORT LAYER metal1 ; RECT 0.6 1.0 0.8 1.2 ; END END ZN
END NOR2_X1
MACRO INV_X1
  SIZE 0.4 BY 2.0 ;
  PIN A PORT LAYER metal1 ; RECT 0.0 0.8 0.1 1.0 ; END END A
  PIN ZN PORT LAYER metal1 ; RECT 0.3 0.8 0.4 1.0 ; END END ZN
END INV_X1
MACRO DFF_X1
  SIZE 3.8 BY 2.0 ;
  PIN D PORT LAYER metal1 ; RECT 0.2 0.9 0.4 1.1 ; END END D
  PIN CK PORT LAYER metal1 ; RECT 1.0 0.9 1.2 1.1 ; END END CK
  PIN Q PORT LAYER metal1 ; RECT 3.4 0.9 3.6 1.1 ; END END Q
END DFF_X1
END LIBRARY
"""


def write(path, text):
    with open(path, "w") as f:
        f.write(text)
    return str(path)


def read(tmp_path, lef_text, def_text):
    lef = write(tmp_path / "cells.lef", lef_text)
    df = write(tmp_path / "design.def", def_text)
    return LefDef([lef], df)


def pins_of(nl):
    """[(cell, dx, dy) per pin] per net."""
    return [[(int(nl.pin_cell[p]), float(nl.pin_dx[p]), float(nl.pin_dy[p]))
             for p in range(nl.net_start[n], nl.net_start[n + 1])] for n in range(nl.n_nets)]


def test_limbo_simple_def_names_an_unknown_component(tmp_path):
    lef = write(tmp_path / "cells.lef", SIMPLE_LEF)
    with pytest.raises(api.EplaceError) as e:
        LefDef([lef], os.path.join(GOLD, "simple.def"))
    assert e.value.code == api.EP_EINVAL and "u4" in str(e.value)


@pytest.mark.parametrize("name", ["simple.def", "simple_w.def"])
def test_limbo_simple_def(tmp_path, name):
    text = open(os.path.join(GOLD, name)).read().replace("u4[0]", "u4")
    with read(tmp_path, SIMPLE_LEF, text) as d:
        assert d.name == "simple"
        inf = d.info
        assert inf["units"] == 1000 and inf["n_macros"] == 4
        assert inf["n_components"] == 6 and inf["n_io_pins"] == 4 and inf["n_nets"] == 8
        assert inf["die"] == (0.0, 0.0, 8.0, 8.0)
        # movable: f1 (PLACED), u2 (UNPLACED), u3 (PLACED), u5\(0\) (PLACED), in file order;
        # fixed: u1 (COVER), u4 (FIXED), then the four IO pins
        assert inf["n_mov"] == 4 and inf["n_fix"] == 6
        nl = d.netlist()
    f1, u2, u3, u5, u1, u4 = range(6)
    out, inp1, inp2, clk = range(6, 10)
    np.testing.assert_array_equal(nl.w[:6], np.float32([3.8, 0.4, 0.4, 0.8, 0.8, 0.8]))
    assert (nl.x[f1], nl.y[f1]) == (np.float32(0.4), 0.0)
    assert (nl.x[u3], nl.y[u3]) == (7.0, 4.0)
    assert (nl.x[u1], nl.y[u1]) == (2.0, 6.0) and (nl.x[u4], nl.y[u4]) == (2.0, 2.0)
    if name == "simple.def":  # UNPLACED ( 4000 2000 )
        assert (nl.x[u2], nl.y[u2]) == (4.0, 2.0)
    else:                     # UNPLACED without a location: the die centre
        assert (nl.x[u2] + 0.2, nl.y[u2] + 1.0) == (4.0, 4.0)
    # IO pins: FIXED location + LAYER box ( 0 0 ) ( 100 100 )
    for p, (x, y) in zip((out, inp1, inp2, clk), ((8, 6), (2, 8), (4, 8), (0, 2))):
        assert (nl.x[p], nl.y[p]) == (x, y)
        assert (nl.w[p], nl.h[p]) == (np.float32(0.1), np.float32(0.1))
    c = lambda v: pytest.approx(v, abs=1e-6)
    nets = pins_of(nl)
    assert nets[0] == [(u1, c(0.65), c(1.0)), (u4, c(0.2), c(0.5))]                    # n1
    assert nets[2] == [(f1, c(3.5), c(1.0)), (u2, c(0.05), c(0.9)), (u4, c(0.4), c(0.5))]  # n3
    assert nets[4] == [(u3, c(0.35), c(0.9)), (out, c(0.05), c(0.05))]                 # out
    assert nets[7] == [(clk, c(0.05), c(0.05)), (f1, c(1.1), c(1.0))]                  # iccad_clk


ORIENT_LEF = """
MACRO ASYM
  SIZE 2.0 BY 1.0 ;
  PIN P PORT LAYER m1 ; RECT 0.2 0.7 0.3 0.8 ; END END P
  PIN Q PORT LAYER m1 ; RECT 1.7 0.1 1.9 0.3 ; END END Q
END ASYM
END LIBRARY
"""

# DEF orientation -> (mirror first: about the "x" axis (y -> -y), the "y" axis
# (x -> -x) or none; then a counter-clockwise rotation in degrees), as the
# LEF/DEF orientation table names them (FN = MY, FS = MX, FW = MX R90, FE = MY R90)
ORIENTS = {"N": ("", 0), "W": ("", 90), "S": ("", 180), "E": ("", 270),
           "FN": ("y", 0), "FS": ("x", 0), "FW": ("x", 90), "FE": ("y", 90)}


def transform(orient, pts):
    """Mirror, rotate; the caller shifts the box's lower-left to 0."""
    flip, deg = ORIENTS[orient]
    pts = np.asarray(pts, dtype=np.float64)
    if flip == "x":
        pts = pts * np.array([1.0, -1.0])
    elif flip == "y":
        pts = pts * np.array([-1.0, 1.0])
    t = np.deg2rad(deg)
    R = np.array([[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]])
    pts = np.rint(pts @ R.T * 1e9) / 1e9
    return pts


def test_orientations(tmp_path):
    comps, nets = [], []
    for k, o in enumerate(ORIENTS):
        comps.append(f"- c{k} ASYM + PLACED ( {1000 * (3 * k + 1)} 2000 ) {o} ;")
        nets.append(f"- n{k} ( c{k} P ) ( c{k} Q ) ;")
    text = ("DESIGN orient ;\nUNITS DISTANCE MICRONS 1000 ;\nDIEAREA ( 0 0 ) ( 30000 10000 ) ;\n"
            f"COMPONENTS {len(comps)} ;\n" + "\n".join(comps) + "\nEND COMPONENTS\n"
            f"NETS {len(nets)} ;\n" + "\n".join(nets) + "\nEND NETS\nEND DESIGN\n")
    with read(tmp_path, ORIENT_LEF, text) as d:
        nl = d.netlist()
    corners = [(0, 0), (2, 0), (0, 1), (2, 1)]
    for k, o in enumerate(ORIENTS):
        t = transform(o, corners + [(0.25, 0.75), (1.8, 0.2)])
        lo = t[:4].min(axis=0)
        size = t[:4].max(axis=0) - lo
        assert (nl.w[k], nl.h[k]) == pytest.approx(tuple(size)), o
        want = t[4:] - lo
        got = [(nl.pin_dx[2 * k + j], nl.pin_dy[2 * k + j]) for j in range(2)]
        np.testing.assert_allclose(got, want, atol=1e-6, err_msg=o)


def check_same(nl, back, units=1000):
    dbu = lambda v: (round_half_away(np.asarray(v, np.float64) * units) / units).astype(np.float32)
    assert back.n_mov == nl.n_mov and back.n_fix == nl.n_fix
    np.testing.assert_array_equal(back.w, nl.w)
    np.testing.assert_array_equal(back.h, nl.h)
    np.testing.assert_array_equal(back.x, dbu(nl.x))
    np.testing.assert_array_equal(back.y, dbu(nl.y))
    np.testing.assert_array_equal(back.net_start, nl.net_start)
    np.testing.assert_array_equal(back.pin_cell, nl.pin_cell)
    np.testing.assert_allclose(back.pin_dx, nl.pin_dx, atol=1e-5)
    np.testing.assert_allclose(back.pin_dy, nl.pin_dy, atol=1e-5)
    assert back.die == pytest.approx(nl.die, abs=1.0 / units)


def test_generated_design_round_trip(tmp_path):
    nl = A.generate("stdcell", seed=3, n_movable=400)
    lef, df = netlist_lefdef(nl)
    with read(tmp_path, lef, df) as d:
        back = d.netlist()
        check_same(nl, back)
        # new movable positions written, read again: fixed cells and nets unchanged
        rng = np.random.default_rng(0)
        x = back.x[:nl.n_mov] + rng.uniform(-3, 3, nl.n_mov).astype(np.float32)
        y = back.y[:nl.n_mov] + rng.uniform(-3, 3, nl.n_mov).astype(np.float32)
        x[0], y[0] = 10.0625, 10.1875  # 10062.5 and 10187.5 units: ties round away from zero
        path = d.writeDEF(str(tmp_path / "out"), np.concatenate([x, np.zeros(7, np.float32)]), y)
        assert path.endswith("gen.def")
    with LefDef([str(tmp_path / "cells.lef")], path) as d2:
        again = d2.netlist()
    assert (again.x[0], again.y[0]) == (np.float32(10.063), np.float32(10.188))
    moved = back
    moved.x = np.concatenate([x, back.x[nl.n_mov:]])
    moved.y = np.concatenate([y, back.y[nl.n_mov:]])
    check_same(moved, again)


def test_read_dir(tmp_path):
    nl = A.generate("mesh", seed=1)
    lef, df = netlist_lefdef(nl, name="mesh")
    half = lef.index("MACRO M300")
    write(tmp_path / "a_cells.lef", lef[:half] + "END LIBRARY\n")
    write(tmp_path / "b_cells.lef", lef[half:])
    write(tmp_path / "floorplan.def", df)
    write(tmp_path / "other.def", "DESIGN broken ;\n")
    with LefDef.from_dir(str(tmp_path)) as d:
        assert d.name == "mesh"
        check_same(nl, d.netlist())
    os.remove(tmp_path / "floorplan.def")
    write(tmp_path / "third.def", df)
    with pytest.raises(api.EplaceError) as e:
        LefDef.from_dir(str(tmp_path))
    assert e.value.code == api.EP_EIO


ROBUST_DEF = r"""
# a comment line
VERSION 5.8 ;
DIVIDERCHAR "/" ;
BUSBITCHARS "[]" ;
DESIGN robust ;
PROPERTYDEFINITIONS
    COMPONENTPIN designRuleWidth REAL ;
    DESIGN FE_CORE_BOX_LL_X REAL 20.140 ;
END PROPERTYDEFINITIONS
UNITS DISTANCE MICRONS 2000 ;
DIEAREA ( 0 0 ) ( 20000 0 ) ( 20000 10000 ) ( 0 10000 ) ;
ROW CORE_ROW_0 core 0 0 N DO 50 BY 1 STEP 400 0
 ;
TRACKS X 190 DO 541 STEP 380 LAYER metal3 ;
GCELLGRID Y 0 DO 2 STEP 140 ;
VIAS 1 ;
- Via1Array-0_1
 + VIARULE Via1Array-0
 + CUTSIZE 140 140
 + ROWCOL 3 3
 ;
END VIAS
COMPONENTS 3 ;
- a\[0\] INV_X1 + SOURCE NETLIST + PLACED ( 2000 2000 ) FS + WEIGHT 1 ;
- b INV_X1
  + FIXED ( 6000 4000 ) N ;
- c INV_X1 ;
END COMPONENTS
PINS 1 ;
- io1 + NET n2 + DIRECTION INPUT + USE SIGNAL
  + LAYER metal2 ( -70 0 ) ( 70 140 )
  + PLACED ( 0 5000 ) E ;
END PINS
SPECIALNETS 1 ;
- VDD ( * VDD ) + ROUTED metal1 200 + SHAPE STRIPE ( 0 0 ) ( 20000 * ) ;
END SPECIALNETS
NETS 3 ;
- n1 ( a\[0\] ZN ) ( b A + SYNTHESIZED ) + USE SIGNAL
  + ROUTED metal2 ( 100 200 ) ( * 900 ) via1 ( b A ) ;
- n2 ( PIN io1 ) ( c A ) ;
- n3 ( c ZN ) ;
END NETS
END DESIGN
"""


def test_block_structure_skipping(tmp_path):
    with read(tmp_path, SIMPLE_LEF, ROBUST_DEF) as d:
        assert d.name == "robust"
        inf = d.info
        assert inf["units"] == 2000 and inf["die"] == (0.0, 0.0, 10.0, 5.0)
        assert inf["n_components"] == 3 and inf["n_io_pins"] == 1 and inf["n_nets"] == 3
        assert inf["n_mov"] == 2 and inf["n_fix"] == 2   # a, c | b, io1
        assert inf["view_nets"] == 2                      # n3 has one pin
        nl = d.netlist()
    a, c, b, io1 = range(4)
    assert (nl.x[a], nl.y[a]) == (1.0, 1.0)
    assert (nl.x[b], nl.y[b]) == (3.0, 2.0)
    # FS: pin ZN (0.35, 0.9) of a 0.4 x 2.0 cell -> (0.35, 1.1)
    nets = pins_of(nl)
    assert nets[0][0] == (a, pytest.approx(0.35), pytest.approx(1.1))
    assert nets[0][1][0] == b and len(nets[0]) == 2
    # E-oriented IO pin box ( -70 0 ) ( 70 140 ) -> ( 0 -70 ) .. ( 140 70 ) around ( 0 5000 )
    assert (nl.x[io1], nl.y[io1]) == (0.0, pytest.approx(4930 / 2000))
    assert (nl.w[io1], nl.h[io1]) == (pytest.approx(0.07), pytest.approx(0.07))
    assert nets[1][0][0] == io1 and nets[1][1][0] == c


def test_errors(tmp_path):
    lef = write(tmp_path / "cells.lef", SIMPLE_LEF)
    with pytest.raises(api.EplaceError) as e:
        LefDef([lef], str(tmp_path / "missing.def"))
    assert e.value.code == api.EP_EIO
    write(tmp_path / "bad.def", "DESIGN x ;\nDIEAREA ( 0 0 ) ( 10 10 ) ;\nCOMPONENTS 1 ;\n- u NOPE ;\nEND COMPONENTS\nEND DESIGN\n")
    with pytest.raises(api.EplaceError) as e:
        LefDef([lef], str(tmp_path / "bad.def"))
    assert e.value.code == api.EP_EINVAL and "NOPE" in str(e.value)
    write(tmp_path / "bad.lef", "MACRO X\n SIZE 1 2 ;\nEND X\n")
    with pytest.raises(api.EplaceError) as e:
        LefDef([str(tmp_path / "bad.lef")], str(tmp_path / "bad.def"))
    assert e.value.code == api.EP_EIO
    write(tmp_path / "nodie.def", "DESIGN x ;\nEND DESIGN\n")
    with pytest.raises(api.EplaceError) as e:
        LefDef([lef], str(tmp_path / "nodie.def"))
    assert e.value.code == api.EP_EIO


def test_mutated_inputs_never_crash(tmp_path):
    """Truncated, token-dropped and token-duplicated LEF/DEF text: the reader
    returns OK or an error status, and a design it accepts has a consistent
    netlist view (the native parser must not crash or read out of bounds)."""
    rng = np.random.default_rng(7)
    base_def = open(os.path.join(GOLD, "simple.def")).read().replace("u4[0]", "u4")
    texts = [(SIMPLE_LEF, base_def), (SIMPLE_LEF, ROBUST_DEF)]
    codes = set()
    for trial in range(300):
        lef, df = texts[trial % 2]
        which = rng.integers(0, 2)
        tok = (lef if which == 0 else df).split(" ")
        op = rng.integers(0, 3)
        if op == 0:
            tok = tok[:rng.integers(0, len(tok))]
        elif op == 1:
            for _ in range(rng.integers(1, 6)):
                if tok:
                    del tok[rng.integers(0, len(tok))]
        else:
            for _ in range(rng.integers(1, 6)):
                k = rng.integers(0, len(tok))
                tok.insert(k, tok[rng.integers(0, len(tok))])
        text = " ".join(tok)
        lef_t, def_t = (text, df) if which == 0 else (lef, text)
        try:
            with read(tmp_path, lef_t, def_t) as d:
                nl = d.netlist()
                n_all = nl.n_mov + nl.n_fix
                assert len(nl.w) == n_all and (nl.pin_cell < max(n_all, 1)).all()
                assert nl.net_start[0] == 0 and nl.net_start[-1] == nl.n_pins
            codes.add(0)
        except api.EplaceError as e:
            assert e.code in (api.EP_EIO, api.EP_EINVAL), str(e)
            codes.add(e.code)
    assert 0 in codes and len(codes) >= 2


RUNNER = os.path.join(os.path.dirname(HERE), "aieplace_amd", "eplace_run")


def test_command_line_errors(tmp_path):
    """eplace_run (the AIEplace.exe counterpart): bad arguments exit 2, an
    unreadable design exits 1 with the library's message (no GPU is touched)."""
    import subprocess
    r = subprocess.run([RUNNER], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([RUNNER, str(tmp_path), "--precision", "fp8"], capture_output=True, text=True)
    assert r.returncode == 2
    r = subprocess.run([RUNNER, str(tmp_path / "missing")], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr
    (tmp_path / "cells.lef").write_text(SIMPLE_LEF)
    r = subprocess.run([RUNNER, str(tmp_path)], capture_output=True, text=True)  # no .def
    assert r.returncode == 1 and "floorplan.def" in r.stderr
