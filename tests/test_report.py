"""Run reporting (aieplace_amd/report.py): host code, no GPU.

The PNG encoder is checked by decoding its files here (signature, chunk CRCs,
zlib stream, filter bytes) back to the array it was given; the snapshot
raster by the pixels a known layout must colour (die outline, a cell in the
lower-left corner of the die drawn at the bottom of the image, IO pins green,
the minimum drawn size); statistics.md by its table rows.
"""
import os
import struct
import zlib

import numpy as np

from aieplace_amd import api, report
from aieplace_amd.netlist import Netlist


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    at, idat, ihdr = 8, b"", None
    while at < len(data):
        n, = struct.unpack(">I", data[at:at + 4])
        tag, body = data[at + 4:at + 8], data[at + 8:at + 8 + n]
        crc, = struct.unpack(">I", data[at + 8 + n:at + 12 + n])
        assert crc == zlib.crc32(tag + body) & 0xFFFFFFFF, tag
        if tag == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif tag == b"IDAT":
            idat += body
        at += 12 + n
    w, h, depth, ctype = ihdr[:4]
    assert (depth, ctype) == (8, 2)
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    assert (raw[:, 0] == 0).all()  # filter type 0 on every row
    return raw[:, 1:].reshape(h, w, 3)


def test_png_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (37, 53, 3), dtype=np.uint8)
    p = str(tmp_path / "a.png")
    report.write_png(p, img)
    np.testing.assert_array_equal(read_png(p), img)


def test_snapshot_raster():
    n = 200
    die = (10.0, 20.0, 110.0, 70.0)   # 100 x 50
    # cell 0: lower-left corner of the die, 10 x 5 (10% of each side);
    # cell 1: a tiny cell (drawn at MIN_SIZE) in the upper-right; cell 2: an IO pin
    x = np.array([10.0, 105.0, 60.0]); y = np.array([20.0, 65.0, 69.0])
    w = np.array([10.0, 0.01, 0.5]); h = np.array([5.0, 0.01, 0.5])
    img = report.draw_placement(x, y, w, h, die, n_io=1, pixels=n)
    span = 0.9 * n
    red = np.all(img == (255, 0, 0), axis=2)
    green = np.all(img == (0, 255, 0), axis=2)
    black = np.all(img == 0, axis=2)
    white = np.all(img == 255, axis=2)
    # the die outline: rows/columns 10 and 189 (margin 0.05 of 200)
    assert black[10:190, 10].all() and black[10:190, 189].all() and black[10, 10:190].all() and black[189, 10:190].all()
    # cell 0 covers canvas x in [10, 28), y in [10, 28) (y up): image rows n-28 .. n-11
    assert red[n - 27:n - 12, 12:27].all()
    assert not red[n - 40:n - 30, 40:60].any()
    # cell 1: at least MIN_SIZE * span = 0.72 px -> one pixel near (10 + 0.95 span, 10 + 0.9 span)
    cx, cy = int(10 + 0.95 * span), int(10 + 0.9 * span)
    assert red[n - 1 - cy, cx]
    # the IO pin is green, not red
    px, py = int(10 + 0.5 * span), int(10 + 0.98 * span)
    assert green[n - 1 - py, px] and not red[n - 1 - py, px]
    assert white[n // 2, n // 2 + 30]


class _StubPlacer:
    def __init__(self, nl):
        self.netlist = nl

    def get_positions(self):
        return self.netlist.x[:self.netlist.n_mov] + 1.0, self.netlist.y[:self.netlist.n_mov]


def test_run_report_files(tmp_path):
    nl = Netlist(w=np.ones(4), h=np.ones(4), x=np.array([1.0, 2.0, 3.0, 0.0]), y=np.array([1.0, 2.0, 3.0, 0.0]),
                 n_mov=3, net_start=np.array([0, 2]), pin_cell=np.array([0, 3]), pin_dx=np.zeros(2),
                 pin_dy=np.zeros(2), die=(0.0, 0.0, 8.0, 8.0))
    pl = _StubPlacer(nl)
    rep = report.RunReport("tiny", out_dir=str(tmp_path / "run"), snapshot_every=2, pixels=64, quiet=True)
    for it in range(5):
        rep.iteration(pl, dict(iteration=it, hpwl=10.0 - it, overflow=0.5 - 0.1 * it, gamma=1.5, lam=0.01,
                               alpha=0.2, evals=it + 1))
    params = api.default_params(bins_x=32, bins_y=32)
    stats = rep.final(pl, dict(iterations=5, evals=6, hpwl=6.0, overflow=0.1, seconds=0.25), params)
    names = sorted(os.listdir(tmp_path / "run"))
    assert names == ["placement_0.png", "placement_2.png", "placement_4.png", "placement_5.png", "statistics.md"]
    text = open(stats).read()
    for row in ("| Benchmark name | tiny |", "| Iterations | 5 |", "| Final HPWL | 6.000000e+00 |",
                "| Final Overflow | 0.1000 |", "| bins | 32 x 32 |", "| gamma | 1.5 |"):
        assert row in text, row
    assert read_png(str(tmp_path / "run" / "placement_5.png")).shape == (64, 64, 3)


def test_output_path_layout(tmp_path):
    import time
    t = time.mktime((2026, 3, 4, 13, 7, 0, 0, 0, -1))
    d = report.output_path("bench", root=str(tmp_path), when=t)
    assert os.path.isdir(d)
    assert os.path.relpath(d, tmp_path) == os.path.join("bench", "run_13:7_62_2026")
