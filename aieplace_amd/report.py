"""Run reporting: the per-iteration table, statistics.md and placement snapshots.

Follows the reference's operator output (vck5000/host/src/AIEplace.cpp:1048-1124
printIterationResults / printFinalResults / getOutputPath, Logger.cpp:31-120
export_markdown, Visualizer.cpp:8-86 drawPlacement):
  * one table row per iteration (iteration, HPWL, overflow);
  * every `snapshot_every` iterations and at the end, `placement_<iteration>.png`:
    white canvas, die outline in black, cells in red, IO pins in green, each
    drawn at least MIN_SIZE of the die wide (Visualizer.h:22-24);
  * at the end, `statistics.md` (run statistics and hyperparameters) and, for a
    LEF/DEF design, `<design>.def`;
  * all under results/<benchmark>/run_<hour>:<minute>_<yday>_<year>.
The snapshot's y axis points up (the reference's cairo canvas has it pointing
down, so its images are mirrored).  PNG files are written by a small encoder
here (RGB, no filtering, zlib) so no imaging library is needed.
"""
from __future__ import annotations

import os
import struct
import time
import zlib
from typing import Optional

import numpy as np

CANVAS_PIXELS = 1048
DIE_START = 0.05   # margin around the die, fraction of the canvas
MIN_SIZE = 0.004   # smallest drawn cell side, fraction of the die


def write_png(path: str, rgb: np.ndarray) -> None:
    """rgb: (height, width, 3) uint8, first row at the top."""
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
    h, w, _ = rgb.shape
    raw = np.concatenate([np.zeros((h, 1), np.uint8), rgb.reshape(h, w * 3)], axis=1).tobytes()

    def chunk(tag: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)))
        f.write(chunk(b"IDAT", zlib.compress(raw, 6)))
        f.write(chunk(b"IEND", b""))


def _cover(boxes: np.ndarray, n: int) -> np.ndarray:
    """Pixels covered by any of the integer boxes [x0, x1) x [y0, y1) (rows = y)."""
    d = np.zeros((n + 1, n + 1), np.int32)
    x0, x1, y0, y1 = boxes.T
    np.add.at(d, (y0, x0), 1)
    np.add.at(d, (y0, x1), -1)
    np.add.at(d, (y1, x0), -1)
    np.add.at(d, (y1, x1), 1)
    return np.cumsum(np.cumsum(d, axis=0), axis=1)[:n, :n] > 0


def draw_placement(x, y, w, h, die, n_io: int = 0, pixels: int = CANVAS_PIXELS) -> np.ndarray:
    """Canvas of the cells (x, y lower-left, w, h); the last n_io cells are IO pins."""
    xl, yl, xh, yh = die
    W, H = xh - xl, yh - yl
    span = (1.0 - 2.0 * DIE_START) * pixels
    x = np.asarray(x, np.float64); y = np.asarray(y, np.float64)
    w = np.asarray(w, np.float64); h = np.asarray(h, np.float64)

    def boxes(sel):
        cx0 = DIE_START * pixels + (x[sel] - xl) / W * span
        cy0 = DIE_START * pixels + (y[sel] - yl) / H * span
        cw = np.maximum(MIN_SIZE, w[sel] / W) * span
        ch = np.maximum(MIN_SIZE, h[sel] / H) * span
        b = np.stack([np.floor(cx0), np.ceil(cx0 + cw), np.floor(cy0), np.ceil(cy0 + ch)], axis=1)
        return np.clip(b, 0, pixels).astype(np.int64)

    img = np.full((pixels, pixels, 3), 255, np.uint8)
    n = len(x)
    cells = _cover(boxes(slice(0, n - n_io)), pixels)
    img[cells] = (255, 0, 0)
    if n_io:
        img[_cover(boxes(slice(n - n_io, n)), pixels)] = (0, 255, 0)
    # die outline, 0.004 of the canvas thick
    t = max(1, int(round(0.004 * pixels)))
    a, b = int(round(DIE_START * pixels)), int(round((1.0 - DIE_START) * pixels))
    img[a:b, a:a + t] = 0
    img[a:b, b - t:b] = 0
    img[a:a + t, a:b] = 0
    img[b - t:b, a:b] = 0
    return img[::-1]  # row 0 at the top of the image: y up


def output_path(benchmark: str, root: str = "results", when: Optional[float] = None) -> str:
    """results/<benchmark>/run_<hour>:<minute>_<yday>_<year> (AIEplace.cpp getOutputPath)."""
    t = time.localtime(when)
    d = os.path.join(root, benchmark, f"run_{t.tm_hour}:{t.tm_min}_{t.tm_yday - 1}_{t.tm_year}")
    os.makedirs(d, exist_ok=True)
    return d


class RunReport:
    """Collects a run's operator output (see the module docstring)."""

    def __init__(self, benchmark: str, root: str = "results", snapshot_every: int = 10,
                 pixels: int = CANVAS_PIXELS, n_io: int = 0, quiet: bool = False, out_dir: Optional[str] = None):
        self.benchmark = benchmark
        self.dir = out_dir or output_path(benchmark, root)
        os.makedirs(self.dir, exist_ok=True)
        self.snapshot_every = snapshot_every
        self.pixels = pixels
        self.n_io = n_io
        self.quiet = quiet
        self.rows = []
        self.snapshots = []

    def snapshot(self, placer, iteration: int) -> str:
        nl = placer.netlist
        n_mov = nl.n_mov
        x, y = placer.get_positions()  # movable, then fillers (not drawn)
        X = np.concatenate([x[:n_mov], nl.x[n_mov:]])
        Y = np.concatenate([y[:n_mov], nl.y[n_mov:]])
        path = os.path.join(self.dir, f"placement_{iteration}.png")
        write_png(path, draw_placement(X, Y, nl.w, nl.h, nl.die, self.n_io, self.pixels))
        self.snapshots.append(path)
        return path

    def iteration(self, placer, rec: dict) -> None:
        self.rows.append(rec)
        if not self.quiet:
            print(f"[DATA] Iteration {rec['iteration']:5d} | HPWL {rec['hpwl']:.6e} | Overflow {rec['overflow']:.4f}",
                  flush=True)
        if self.snapshot_every and rec["iteration"] % self.snapshot_every == 0:
            self.snapshot(placer, rec["iteration"])

    def final(self, placer, res: dict, params, design=None) -> str:
        """statistics.md, the final snapshot and (LEF/DEF designs) the DEF; returns statistics.md's path."""
        prec = {0: "FP32", 1: "split BF16", 2: "split FP16"}.get(int(params.precision), str(params.precision))
        last = self.rows[-1] if self.rows else {}
        rows = [("Benchmark name", self.benchmark), ("Iterations", res["iterations"]),
                ("Gradient evaluations", res.get("evals", "")),
                ("GPU runtime (s)", f"{res.get('seconds', 0.0):.4f}"),
                ("Final HPWL", f"{res['hpwl']:.6e}"), ("Final Overflow", f"{res['overflow']:.4f}")]
        hyper = [("gamma", f"{last.get('gamma', 0.0):.6g}"), ("lambda", f"{last.get('lam', 0.0):.6g}"),
                 ("step length", f"{last.get('alpha', 0.0):.6g}"), ("target density", f"{params.target_density:g}"),
                 ("bins", f"{params.bins_x} x {params.bins_y}"), ("transform precision", prec),
                 ("learned step", "on" if params.learned_step else "off")]
        lines = ["# AIEplace Run Statistics", "", "| | |", "|---:|:---|"]
        lines += [f"| {k} | {v} |" for k, v in rows]
        lines += ["", "## Hyperparameters", "", "| | |", "|---:|:---|"]
        lines += [f"| {k} | {v} |" for k, v in hyper]
        path = os.path.join(self.dir, "statistics.md")
        with open(path, "w") as f:
            f.write("\n".join(lines) + "\n")
        self.snapshot(placer, res["iterations"])
        if design is not None:
            x, y = placer.get_positions()
            design.writeDEF(self.dir, x, y)
        if not self.quiet:
            print(f"[DATA] Iterations {res['iterations']} | Final HPWL {res['hpwl']:.6e} | "
                  f"Final Overflow {res['overflow']:.4f} | output {self.dir}", flush=True)
        return path
