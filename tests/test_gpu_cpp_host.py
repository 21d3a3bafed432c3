"""The C++ host Placer (include/eplace_placer.hpp, the reference's Placer
method names over the C ABI) run end to end by aieplace_amd/placer_example,
whose marked region is the example INTEGRATION.md shows."""
import os
import re
import subprocess
import sys

import pytest

import aieplace_amd as A

pytestmark = pytest.mark.gpu
EXAMPLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "aieplace_amd",
                       "placer_example")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_example(*args):
    r = subprocess.run([EXAMPLE, *args], capture_output=True, text=True, timeout=120)
    say(r.stdout[-600:])
    assert r.returncode == 0, r.stderr
    hp = float(re.search(r"Final HPWL\s+(\S+)", r.stdout).group(1))
    ovf = float(re.search(r"Final Overflow\s+(\S+)", r.stdout).group(1))
    it = int(re.search(r"Iterations\s+(\d+)", r.stdout).group(1))
    assert r.stdout.count("Iteration ") >= 4  # three performIteration lines + run's
    return it, hp, ovf


def python_run(kind, n, bins, graph):
    if kind == "mesh":
        nl = A.generate("mesh", seed=1)
        p = A.placer_params(A.CONFIGS["mesh_c1"], use_graph=graph)
    else:
        nl = A.generate("stdcell", seed=1, n_movable=n)
        p = A.placer_params(A.CONFIGS["stdcell_200k"], bins_x=bins, bins_y=bins, use_graph=graph)
    with A.Placer(nl, p) as pl:
        pl.initializePlacement()
        for _ in range(3):
            pl.nudgeAllNodes()
        r = pl.run()
    return r["iterations"], r["hpwl"], r["overflow"]


def test_cpp_placer_example_mesh():
    say("case C++ Placer example, mesh")
    it, hp, ovf = run_example("mesh")
    pit, php, povf = python_run("mesh", 0, 32, 1)
    assert ovf <= 0.10 and it == pit
    assert hp == pytest.approx(php, rel=1e-6)  # printed with 7 significant digits


def test_cpp_placer_example_stdcell_graph_and_host():
    """Host-driven steps then a device-controlled run, and the same flow with
    host control throughout: identical results (the device replays the host's
    decisions), equal to the Python flow's."""
    say("case C++ Placer example, 20k cells, 128^2 bins, graph 1 vs 0")
    a = run_example("stdcell", "20000", "--bins", "128", "--graph", "1")
    b = run_example("stdcell", "20000", "--bins", "128", "--graph", "0")
    assert a == b
    pit, php, _ = python_run("stdcell", 20_000, 128, 1)
    assert a[0] == pit and a[1] == pytest.approx(php, rel=1e-6)
    assert a[2] <= 0.10
