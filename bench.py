"""Benchmark: seconds to overflow <= 0.10 on the 2M-cell synthetic placement.

A step is one complete global placement of the workload (initial placement ->
overflow <= 0.10) with the netlist already resident in HBM.  Each form is timed
separately over --steps runs after --warmup untimed runs:
  plain           FP32 transforms, Lipschitz step search (the baseline form)
  mixed_learned   split-BF16 transforms + MLP-proposed step verified by the
                  backtracking condition (when built)
`value` is the headline form's seconds (mixed_learned if present, else plain).
The batched form (64 instances of 5k-150k cells, 128/256/512 bins, one graph
replay per 4 Nesterov trials of all unfinished instances) is reported beside it
as instance-iterations/s in FP32 and split-BF16 transforms, with the same
instances run one at a time through the single-instance path on a sample.
Prints ONE JSON line.  Runs on one GPU; --gpus > 1 exits with a message.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sec to overflow<=0.10 @2M cells, mixed+learned vs plain FP32; HPWL gap %"
PEAK_HBM_GBS = 8000.0
PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_BF16_MFMA_TFLOPS = 2500.0
FILLER_SEEDS = (7, 8, 9, 10)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_form(A, nl, cfg, form, steps, warmup):
    from aieplace_amd import api
    kw = dict(precision=form.get("precision", api.EP_PREC_FP32), use_graph=form.get("use_graph", 1),
              learned_step=form.get("learned_step", 0))
    p = A.placer_params(cfg, **kw)
    weights = None
    if kw["learned_step"]:
        from aieplace_amd import learned
        if not os.path.exists(learned.DEFAULT_WEIGHTS):
            raise api.EplaceError(api.EP_ENOTSUP, f"no trained step model at {learned.DEFAULT_WEIGHTS}")
        weights = learned.load()
    pl = A.Placer(nl, p)
    if weights is not None:
        pl.set_step_model(weights)
    x0, y0 = pl.get_positions()
    for w in range(warmup):
        pl.set_positions(x0, y0)
        r = pl.run()
        log(f"[{form['name']}] warmup {w}: {r['iterations']} it, {r['evals']} evals, "
            f"{r['seconds']:.3f}s, hpwl {r['hpwl']:.6e}, tau {r['overflow']:.4f}")
    runs = []
    for k in range(steps):
        pl.set_positions(x0, y0)
        t0 = time.perf_counter()
        r = pl.run()
        r["wall"] = time.perf_counter() - t0
        runs.append(r)
        log(f"[{form['name']}] step {k}: {r['iterations']} it, {r['evals']} evals, "
            f"{r['seconds']:.3f}s (wall {r['wall']:.3f}s), hpwl {r['hpwl']:.6e}, tau {r['overflow']:.4f}")
    return pl, runs


def seed_spread(A, nl, cfg, pl, seeds=FILLER_SEEDS):
    """Untimed: the same form from the initial placement with the fillers of
    each seed (the final HPWL of one run moves by ~1% with them); HPWL at
    overflow exactly 0.10 of each run (hpwl_at_overflow)."""
    from aieplace_amd.netlist import plan_fillers
    from aieplace_amd.placer import hpwl_at_overflow
    n_mov = nl.n_mov
    out = []
    for fs in seeds:
        _, _, _, fx, fy = plan_fillers(nl, A.placer_params(cfg, filler_seed=fs))
        pl.set_positions(np.concatenate([nl.x[:n_mov], fx]).astype(np.float32),
                         np.concatenate([nl.y[:n_mov], fy]).astype(np.float32))
        r = pl.run(log_cap=5000)
        at = hpwl_at_overflow([l["overflow"] for l in r["log"]], [l["hpwl"] for l in r["log"]])
        out.append(dict(seed=fs, iterations=r["iterations"], hpwl=r["hpwl"], hpwl_at_010=at))
    return out


def precision_report(A, nl, cfg, every=50):
    """Untimed: the split-BF16 form run step by step; every `every` iterations the
    field of the current map is solved in each precision and compared with an
    FP64 device solve (max-norm relative error)."""
    from aieplace_amd import api
    p = A.placer_params(cfg, precision=api.EP_PREC_BF16X3, use_graph=1)
    rep = {"every": every, "iteration": [], "fp32": [], "bf16x3": [], "fp16x3": []}
    with A.Placer(nl, p) as pl:
        pl.initializePlacement()
        while True:
            r = pl.nudgeAllNodes()
            if r["iteration"] % every == 0 or r["overflow"] <= p.stop_overflow:
                rep["iteration"].append(r["iteration"])
                for name, prec in (("fp32", api.EP_PREC_FP32), ("bf16x3", api.EP_PREC_BF16X3),
                                   ("fp16x3", api.EP_PREC_FP16X3)):
                    rep[name].append(float("%.3g" % max(pl.field_error(prec))))
            if r["overflow"] <= p.stop_overflow or r["iteration"] >= p.max_iters - 1:
                break
    for name in ("fp32", "bf16x3", "fp16x3"):
        rep[name + "_max"] = max(rep[name])
    return rep


def profile(pl, reps=20):
    from aieplace_amd import api
    stats = (api.EpKernelStat * 16)()
    n = C.c_int32()
    api.check(api.lib.ep_profile(pl._ctx, reps, stats, 16, C.byref(n)))
    out = []
    for s in stats[:n.value]:
        out.append(dict(name=s.name.decode(), launches=s.launches, avg_us=s.avg_us,
                        bytes=s.bytes, flops=s.flops))
    return out


def roofline_of(stats, traffic=None, precision_bf16=False, item=None):
    """Dominant item of ep_profile (longest call), or the item named `item`,
    against its roofline.  Transform products count their algorithmic flops:
    the six dense products of a solve, 6 Mx My (Mx + My) (12 M^3 square); the
    split forms are held to the 16-bit peak.  `traffic`: {item: measured HBM
    bytes per call} from the committed rocprofv3 --pmc summary of the same
    workload, or None."""
    if item is not None:
        dom = next(s for s in stats if s["name"] == item)
    else:
        dom = max((s for s in stats if "_cpt" not in s["name"]), key=lambda s: s["avg_us"])
    sec = dom["avg_us"] * 1e-6
    tr = (traffic or {}).get(dom["name"])
    tr = None if tr is None else round(tr)
    if dom["flops"] > 0:
        peak = PEAK_BF16_MFMA_TFLOPS if precision_bf16 else PEAK_FP32_MFMA_TFLOPS
        ach = dom["flops"] / sec / 1e12
        return dom, dict(bound="mfma", achieved=round(ach, 3), peak=peak, unit="TFLOP/s",
                         frac=round(ach / peak, 4), traffic=tr)
    ach = dom["bytes"] / sec / 1e9
    return dom, dict(bound="hbm", achieved=round(ach, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                     frac=round(ach / PEAK_HBM_GBS, 4), traffic=tr)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads() -> int:
    """Host threads for the CPU baseline: the cores this job is given
    (OMP_NUM_THREADS, 16 per GPU on the measurement hosts), else every CPU."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def cpu_baseline(A, nl, cfg, gpu_evals, budget_s):
    """CPU timings of the same workload on this host, each a bounded sample
    projected to the GPU plain form's gradient-evaluation count:
      * the FP64 oracle (oracle/eplace_oracle.c, OpenMP): a few Nesterov
        iterations, and one field solve (separable products with tables);
      * the field by the reference's own RePlAce FFT solver (oracle/_ref, built
        from python/AIEmath/fft_replace when the reference tree was present):
        O(M^2 log M), one thread;
      * the field as the reference's host path evaluates it (DCT_naive /
        IDCT_naive / IDXST_naive, one cos() per term): a sample of rows of one
        pass, scaled to the six passes of M rows of a solve.
    `value` is the fastest CPU evaluation: the oracle's evaluation with the
    faster of its own field solve and the FFT solve."""
    from oracle import oracle as O
    from aieplace_amd.netlist import plan_fillers
    threads = cpu_threads()
    O.set_threads(threads)
    p = A.placer_params(cfg)
    n_fill, fw, fh, fx, fy = plan_fillers(nl, p)
    n_mov = nl.n_mov
    w = np.concatenate([nl.w[:n_mov], np.full(n_fill, fw, np.float32), nl.w[n_mov:]])
    h = np.concatenate([nl.h[:n_mov], np.full(n_fill, fh, np.float32), nl.h[n_mov:]])
    pc = nl.pin_cell.copy()
    pc[pc >= n_mov] += n_fill
    d = O.Design(w=w, h=h, n_mov=n_mov, n_fill=n_fill, fix_x=nl.x[n_mov:], fix_y=nl.y[n_mov:],
                 net_start=nl.net_start, pin_cell=pc, pin_dx=nl.pin_dx, pin_dy=nl.pin_dy,
                 die=nl.die, bins=(p.bins_x, p.bins_y))
    x0 = np.concatenate([nl.x[:n_mov], fx]).astype(np.float64)
    y0 = np.concatenate([nl.y[:n_mov], fy]).astype(np.float64)
    bw = (nl.die[2] - nl.die[0]) / p.bins_x
    op = dict(target_density=p.target_density, gamma_bins=p.gamma_bins,
              gamma_schedule=p.gamma_schedule, lambda_eta=p.lambda_eta,
              hpwl_ref=0.05 * nl.n_nets * np.float32(bw), mu_max=p.mu_max, mu_min=p.mu_min,
              max_backtrack=p.max_backtrack, stop_overflow=p.stop_overflow, max_iters=1,
              init_disp_bins=p.init_disp_bins)
    t0 = time.perf_counter()
    _, _, info = d.place(x0, y0, op)
    t1 = time.perf_counter() - t0
    per_eval = t1 / max(1, info["evals"])
    iters = min(max(1, int(budget_s / per_eval) - 3), 20)
    if t1 < budget_s / 2 and iters > 1:
        op["max_iters"] = iters
        t0 = time.perf_counter()
        _, _, info = d.place(x0, y0, op)
        t1 = time.perf_counter() - t0
        per_eval = t1 / max(1, info["evals"])
    # the field alone, three ways, on the initial map
    M = p.bins_x
    rho = (d.density(x0, y0, "stretched", 0, d.n_place) + p.target_density * d.density_fixed()) / (bw * bw)
    def best_of(fn, reps=3):
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            t.append(time.perf_counter() - t0)
        return min(t)

    t_field = best_of(lambda: O.field(rho))
    t_fft = best_of(lambda: O.ref_fft_field(rho.astype(np.float32))) if O.ref_available() else None
    sample = np.ascontiguousarray(rho[:max(threads, 64)], np.float32)
    rows = sample.shape[0]
    t_naive = best_of(lambda: O.rows_naive(sample, "dct"), 1) * 6.0 * M / rows
    rest = max(per_eval - t_field, 0.0)  # the evaluation without its field solve
    best = rest + (min(t_field, t_fft) if t_fft is not None else t_field)
    proj = best * gpu_evals
    comp = dict(oracle_eval_s=round(per_eval, 4), oracle_field_s=round(t_field, 4),
                fft_field_s=None if t_fft is None else round(t_fft, 4), naive_cos_field_s=round(t_naive, 2),
                oracle_projection_s=round(per_eval * gpu_evals, 1),
                naive_cos_projection_s=round((rest + t_naive) * gpu_evals, 1))
    return dict(value=round(proj, 2), unit="s", cores=threads, kind="port", cpu_model=cpu_model(),
                host_cpus=os.cpu_count(), components=comp,
                sample=(f"same {nl.n_mov}-cell instance on {threads} host threads: FP64 oracle "
                        f"{op['max_iters']} Nesterov iterations = {info['evals']} gradient evaluations in "
                        f"{t1:.1f}s ({per_eval:.3f}s/eval, field {t_field:.3f}s of it); field by the "
                        f"reference's RePlAce FFT solver "
                        + (f"{t_fft:.3f}s (one thread)" if t_fft is not None else "not built")
                        + f"; naive-cos field (reference DCT_naive structure) {t_naive:.1f}s from {rows} rows of "
                        f"one pass.  value = the evaluation with the faster field x the GPU plain form's {gpu_evals} evaluations"))


def batched_report(A, n=64, sample=16):
    """Instance-iterations/s of the batched engine (FP32 and split-BF16 transforms)
    and of the single-instance path on the first `sample` instances."""
    from aieplace_amd import api
    from aieplace_amd.batch import BatchPlacer, batch_workload
    t0 = time.perf_counter()
    nls, bins = batch_workload(n)
    log(f"[batched] generated {n} instances ({sum(nl.n_mov for nl in nls)} movable cells) "
        f"in {time.perf_counter() - t0:.1f}s")
    cfg = A.CONFIGS["stdcell_200k"]
    rep = dict(instances=n, movable_cells=int(sum(nl.n_mov for nl in nls)),
               bins={str(b): bins.count(b) for b in sorted(set(bins))})
    hpwl32 = None
    for tag, prec in (("fp32", api.EP_PREC_FP32), ("bf16x3", api.EP_PREC_BF16X3)):
        p = A.placer_params(cfg, precision=prec)
        with BatchPlacer(nls, bins, p) as bp:
            bp.run()                      # warm-up (graph capture, caches)
            bp.reset()
            r = bp.run()
            res = bp.results()
        rate = r["instance_iterations"] / r["seconds"]
        log(f"[batched] {tag}: {r['instance_iterations']} instance-iterations in {r['seconds']:.3f}s "
            f"= {rate:.0f}/s, {r['launches']} replays, density tile {r['density_cells_per_thread']} cells/thread, "
            f"transform K step {r['transform_k_step']}")
        rep[tag] = dict(instance_iterations_per_s=round(rate, 1), seconds=round(r["seconds"], 4),
                        instance_iterations=int(r["instance_iterations"]), evals=int(r["evals"]),
                        finished=int(r["finished"]), replays=int(r["launches"]),
                        density_cells_per_thread=int(r["density_cells_per_thread"]),
                        transform_k_step=int(r["transform_k_step"]),
                        max_overflow=round(float(res["overflow"].max()), 4))
        if prec == api.EP_PREC_FP32:
            hpwl32 = res["hpwl"].copy()
    tot_it, tot_s, gaps = 0, 0.0, []
    for i in range(min(sample, n)):
        q = A.placer_params(cfg, bins_x=bins[i], bins_y=bins[i], use_graph=1)
        with A.Placer(nls[i], q) as pl:
            r = pl.run()
        tot_it += r["iterations"]
        tot_s += r["seconds"]
        gaps.append(abs(hpwl32[i] - r["hpwl"]) / r["hpwl"])
    rep["single_instance_sample"] = dict(
        instances=min(sample, n), instance_iterations_per_s=round(tot_it / tot_s, 1),
        median_abs_hpwl_gap_vs_batch_pct=round(100 * float(np.median(gaps)), 3))
    rep["speedup_vs_single_fp32"] = round(rep["fp32"]["instance_iterations_per_s"] /
                                          rep["single_instance_sample"]["instance_iterations_per_s"], 2)
    return rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="stdcell_2m")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-precision-report", action="store_true")
    ap.add_argument("--forms", default="plain,mixed,mixed_learned")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-seed-spread", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only for this path)", file=sys.stderr)
        sys.exit(2)

    import aieplace_amd as A
    from aieplace_amd import api
    cfg = A.CONFIGS[args.workload]
    t0 = time.perf_counter()
    nl = A.generate(seed=1, **cfg["gen"])
    log(f"generated {args.workload}: {nl.n_mov} movable, {nl.n_fix} fixed, {nl.n_nets} nets, "
        f"{nl.n_pins} pins in {time.perf_counter() - t0:.1f}s")

    forms_all = {
        "plain": dict(name="plain", precision=api.EP_PREC_FP32, learned_step=0),
        "mixed": dict(name="mixed", precision=api.EP_PREC_BF16X3, learned_step=0),
        "mixed_learned": dict(name="mixed_learned", precision=api.EP_PREC_BF16X3, learned_step=1),
    }
    results, placers = {}, {}
    for name in args.forms.split(","):
        form = forms_all[name]
        try:
            pl, runs = run_form(A, nl, cfg, form, args.steps, args.warmup)
        except api.EplaceError as e:
            if e.code == api.EP_ENOTSUP:
                log(f"[{name}] not available: {e}")
                continue
            raise
        placers[name] = pl
        sec = [r["seconds"] for r in runs]
        results[name] = dict(seconds=round(float(np.mean(sec)), 4), runs=[round(s, 4) for s in sec],
                             iterations=runs[-1]["iterations"], evals=runs[-1]["evals"],
                             hpwl=runs[-1]["hpwl"], overflow=round(runs[-1]["overflow"], 4))
        if form.get("learned_step"):
            r = runs[-1]
            results[name]["learned_proposals"] = r["learned_proposals"]
            results[name]["learned_acceptance"] = round(
                r["learned_accepted"] / max(1, r["learned_proposals"]), 4)

    head = next((f for f in ("mixed_learned", "mixed", "plain") if f in results), "plain")
    plain = results["plain"]
    spread = {}
    if not args.no_seed_spread:
        for name in ("plain", head) if head != "plain" else ("plain",):
            spread[name] = seed_spread(A, nl, cfg, placers[name])
            log(f"[{name}] hpwl at overflow 0.10 over filler seeds: "
                + ", ".join(f"{q['seed']}: {q['hpwl_at_010']:.5e}" for q in spread[name]))
    out = dict(metric=METRIC, value=results[head]["seconds"], unit="s", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=round(results[head]["seconds"] * 1000, 2),
               higher_is_better=False, scaling="strong", vs_baseline=None,
               dtype="fp32" if head == "plain" else "bf16x3+fp32", data="synthetic",
               config=dict(workload=args.workload, movable_cells=nl.n_mov, fixed_cells=nl.n_fix,
                           nets=nl.n_nets, pins=nl.n_pins, bins=cfg["bins"],
                           target_density=cfg["target_density"], headline_form=head,
                           cell_storage=("random order, renumbered breadth-first by ep_create"
                                         if cfg["gen"].get("shuffle_storage") else "generator order")),
               forms=results)
    if head != "plain":
        m = results[head]
        out["speedup_vs_plain"] = round(plain["seconds"] / m["seconds"], 3)
        out["hpwl_gap_pct"] = round(100.0 * (m["hpwl"] - plain["hpwl"]) / plain["hpwl"], 3)
        if spread:
            a = float(np.mean([q["hpwl_at_010"] for q in spread["plain"]]))
            b = float(np.mean([q["hpwl_at_010"] for q in spread[head]]))
            out["hpwl_gap_pct_seed_mean"] = round(100.0 * (b - a) / a, 3)
            out["hpwl_seed_runs"] = {k: [dict(seed=q["seed"], iterations=q["iterations"],
                                              hpwl_at_010=round(q["hpwl_at_010"], 1)) for q in v]
                                     for k, v in spread.items()}

    if not args.no_precision_report:
        out["field_error_vs_fp64"] = precision_report(A, nl, cfg)
    from aieplace_amd.profiling import read_traffic
    kern = {}
    for name, key in ((head, "roofline"), ("plain", "roofline_plain")) if head != "plain" else (("plain", "roofline"),):
        bf16 = forms_all[name]["precision"] != api.EP_PREC_FP32
        stats = profile(placers[name])
        tag = "bf16x3" if bf16 else "fp32"
        traffic = read_traffic(args.workload, forms_all[name]["precision"])
        dom, roof = roofline_of(stats, traffic, bf16)
        out[key] = dict(kernel=dom["name"], form=name, **roof)
        if roof["traffic"] is not None:
            out[key]["traffic_source"] = f"profiles/traffic_{args.workload}_{tag}.md"
        if key == "roofline" and dom["name"] != "field_transforms":
            # the products and the cell gradient take about as long in the
            # headline form; the products' roofline is reported either way
            _, roof_t = roofline_of(stats, traffic, bf16, item="field_transforms")
            out["roofline_transforms"] = dict(kernel="field_transforms", form=name, **roof_t)
            if roof_t["traffic"] is not None:
                out["roofline_transforms"]["traffic_source"] = f"profiles/traffic_{args.workload}_{tag}.md"
        kern[name] = {q["name"]: dict(avg_us=round(q["avg_us"], 2), launches=q["launches"]) for q in stats}
    out["roofline_flops"] = ("transform products count algorithmic flops: the six dense products of a "
                             "solve, 6 Mx My (Mx + My) = 12 M^3; split-BF16 held to the BF16 peak")
    out["kernels"] = kern
    if not args.no_batched:
        out["batched"] = batched_report(A)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(A, nl, cfg, plain["evals"], args.cpu_seconds)
        out["gpu_vs_cpu_plain"] = round(out["cpu_baseline"]["value"] / plain["seconds"], 1)
    for pl in placers.values():
        pl.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
