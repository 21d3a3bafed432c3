"""Writes tests/golden/oracle_stdcell_200k.npz: the FP64 oracle's placements of
the 200k-cell workload (generator seed 1, 512x512 bins, rho_t 0.9, FP32
parameters as the device gets them), run in this container.

The final HPWL of one trajectory at overflow 0.10 is chaotic: a one-ulp change
of the start moves it by about 1%.  So the oracle is run from K starts
(support.nudged_start): the unperturbed one (member 0) and K-1 where a seeded
half of the movable and filler x coordinates is moved to the next FP32 value up.  Stored per member:
final HPWL, overflow, iterations, evaluations and the HPWL interpolated at
overflow 0.10 (support.hpwl_at_overflow); for member 0 also its per-iteration
log (evaluations, HPWL, overflow, lambda, gamma, alpha_hat).

Usage: python tests/golden/make_oracle_stdcell_200k.py [K]   (about 2-3 min per
member on 8 cores)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import aieplace_amd as A  # noqa: E402
from oracle import oracle as O  # noqa: E402
from support import hpwl_at_overflow, nudged_start, oracle_design, oracle_params  # noqa: E402

OUT = os.path.join(HERE, "oracle_stdcell_200k.npz")


def main():
    k = int(sys.argv[1]) if len(sys.argv) > 1 else 6
    threads = os.cpu_count() or 1
    O.set_threads(threads)
    nl = A.generate("stdcell", seed=1, n_movable=200_000)
    p = A.placer_params(A.CONFIGS["stdcell_200k"])
    d, x0, y0 = oracle_design(O, nl, p)
    op = oracle_params(p, d.bin_size[0], nl.n_nets)
    res = {key: [] for key in ("hpwl", "overflow", "iters", "evals", "hpwl_at_010", "seconds")}
    log0 = None
    for m in range(k):
        t = time.time()
        _, _, info = d.place(nudged_start(x0, m), y0, op, log_cap=5000)
        log = np.array([[l["evals"], l["hpwl"], l["overflow"], l["lam"], l["gamma"], l["alpha"]]
                        for l in info["log"]])
        if m == 0:
            log0 = log
        res["hpwl"].append(info["hpwl"])
        res["overflow"].append(info["overflow"])
        res["iters"].append(info["iters"])
        res["evals"].append(info["evals"])
        res["hpwl_at_010"].append(hpwl_at_overflow(log[:, 2], log[:, 1]))
        res["seconds"].append(time.time() - t)
        print(f"member {m}: {info['iters']} it, {info['evals']} evals, hpwl {info['hpwl']:.6e}, "
              f"at 0.10 {res['hpwl_at_010'][-1]:.6e}, {res['seconds'][-1]:.1f} s", flush=True)
    np.savez_compressed(OUT, log0=log0, threads=threads, **{key: np.array(v) for key, v in res.items()})
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
