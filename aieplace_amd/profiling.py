"""rocprofv3 summaries of the placement kernels (written under profiles/).

  python -m aieplace_amd.profiling stats   [--workload stdcell_2m] [--precision 0]
      kernel-trace statistics of one converged placement plus ep_profile
      repetitions -> profiles/kernels_<workload>_<prec>.md
  python -m aieplace_amd.profiling traffic [--workload stdcell_2m] [--precision 0]
      FETCH_SIZE and WRITE_SIZE per dispatch (two --pmc passes) for the
      placement kernels, scaled by the same counters read on a device copy of
      known size -> profiles/traffic_<workload>_<prec>.md
  python -m aieplace_amd.profiling mfma    [--workload stdcell_2m] [--precision 0]
      matrix-core utilisation of the transform products (one --pmc pass of SQ
      counters): SQ_VALU_MFMA_BUSY_CYCLES / (4 SIMDs x SQ_BUSY_CU_CYCLES) per
      kernel, with the products' achieved TFLOP/s -> profiles/mfma_<workload>_<prec>.md
  python -m aieplace_amd.profiling timeline [--workload stdcell_2m] [--precision 0]
      kernel trace of one placement split at each Nesterov trial: trial wall
      time, kernels' busy time (so the idle gaps between them) and each
      kernel's start offset and duration -> profiles/timeline_<workload>_<prec>.md
  python -m aieplace_amd.profiling batch   [--instances 64] [--precision 0]
      kernel-trace statistics of the batched engine on the 64-instance
      workload (two full runs: warm-up and measured) -> profiles/kernels_batch<n>_<prec>.md
  python -m aieplace_amd.profiling drive ...
      the profiled program itself (run under rocprofv3 by the two commands
      above): a full placement, then `reps` ep_profile repetitions of each
      kernel at the final placement; with --copy-mib N it instead copies an
      N MiB device buffer `reps` times (the calibration kernel).

The parent process never touches the GPU; rocprofv3 runs the driver as a child.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROFILES = os.path.join(ROOT, "profiles")
PREC_NAME = {0: "fp32", 1: "bf16x3", 2: "fp16x3"}
# kernel families of one gradient evaluation (name prefix -> ep_profile item)
FAMILIES = [
    ("k_sort_", "cell_sort"), ("k_density", "density_scatter"), ("k_overflow", "overflow"),
    ("k_map_stats", "overflow"), ("k_wa", "wa_grad"), ("k_sym_", "field_transforms"),
    ("k_clear_max", "field_transforms"), ("k_pin_sum", "pin_sum"), ("k_cell_grad", "cell_grad"),
    ("k_trial", "nesterov_trial"),
    ("k_finalize", "finalize"),
]


def _short(name: str) -> str:
    n = name.replace("ep::(anonymous namespace)::", "").replace("void ", "")
    return re.sub(r"\(.*", "", n)


# ---------------------------------------------------------------- driver (GPU)

def drive(args) -> None:
    sys.path.insert(0, ROOT)
    if args.copy_mib:
        # device-to-device hipMemcpy: the runtime's blit kernel moves exactly
        # copy_mib MiB each way
        import ctypes as C
        hip = C.CDLL("libamdhip64.so")
        nbytes = args.copy_mib * (1 << 20)
        src, dst = C.c_void_p(), C.c_void_p()
        for ptr in (src, dst):
            if hip.hipMalloc(C.byref(ptr), C.c_size_t(nbytes)) != 0:
                raise RuntimeError("hipMalloc failed")
        hip.hipMemset(src, 1, C.c_size_t(nbytes))
        for _ in range(args.reps):
            if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        hip.hipDeviceSynchronize()
        hip.hipFree(src)
        hip.hipFree(dst)
        print(json.dumps({"copy_bytes": nbytes, "reps": args.reps}), flush=True)
        return
    import aieplace_amd as A
    cfg = A.CONFIGS[args.workload]
    nl = A.generate(seed=1, **cfg["gen"])
    p = A.placer_params(cfg, precision=args.precision, use_graph=1)
    with A.Placer(nl, p) as pl:
        r = pl.run()
        prof = pl.profile(args.reps) if args.reps > 0 else []
    print(json.dumps({"run": {k: r[k] for k in ("iterations", "evals", "seconds", "hpwl", "overflow")},
                      "profile": prof}), flush=True)


def drive_batch(args) -> None:
    """The batched engine on `instances` generated instances: a warm-up run
    (graph capture, tile autotuning) and a measured run."""
    sys.path.insert(0, ROOT)
    import aieplace_amd as A
    from aieplace_amd.batch import BatchPlacer, batch_workload
    nls, bins = batch_workload(args.instances)
    p = A.placer_params(A.CONFIGS["stdcell_200k"], precision=args.precision)
    with BatchPlacer(nls, bins, p) as bp:
        bp.run()
        bp.reset()
        r = bp.run()
    run = {k: float(r[k]) for k in ("instance_iterations", "seconds", "evals", "launches", "finished")}
    run["movable_cells"] = int(sum(nl.n_mov for nl in nls))
    print(json.dumps({"run": run}), flush=True)


def _drive_cmd(args, extra=()):
    cmd = [sys.executable, "-u", "-m", "aieplace_amd.profiling", "drive", "--workload", args.workload,
           "--precision", str(args.precision), "--reps", str(args.reps)]
    return cmd + list(extra)


def _rocprof(pre, cmd, out_dir, timeout):
    full = ["rocprofv3"] + pre + ["--output-format", "csv", "-d", out_dir, "-o", "run", "--"] + cmd
    env = dict(os.environ, TMPDIR=os.environ.get("TMPDIR", "/tmp"))
    r = subprocess.run(full, cwd=ROOT, capture_output=True, text=True, timeout=timeout, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(full)} failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    res = None
    for line in r.stdout.splitlines():
        if line.startswith('{"run"') or line.startswith('{"copy_bytes"'):
            res = json.loads(line)
    return full, res


def _cmd_text(cmd) -> str:
    """A command as the summaries print it: run from the repository root."""
    return " ".join(cmd).replace(sys.executable, "python3").replace(ROOT + os.sep, "")


def _csv(out_dir, suffix):
    files = glob.glob(os.path.join(out_dir, "**", f"*{suffix}"), recursive=True)
    if not files:
        raise RuntimeError(f"no *{suffix} under {out_dir}")
    with open(files[0]) as f:
        return list(csv.DictReader(f))


def _family(kname: str):
    for pre, fam in FAMILIES:
        if kname.startswith(pre):
            return fam
    return None


# ---------------------------------------------------------------- stats

def stats(args) -> str:
    out_dir = os.path.join(args.scratch, f"stats_{args.workload}_{args.precision}")
    cmd = _drive_cmd(args)
    full, res = _rocprof(["--kernel-trace", "--stats"], cmd, out_dir, args.timeout)
    rows = _csv(out_dir, "kernel_stats.csv")
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    prec = PREC_NAME[args.precision]
    lines = [f"# Kernel statistics: {args.workload}, {prec} transforms", "",
             "Command (from the repository root):", "", "```", _cmd_text(full),
             "```", "",
             f"One full placement (graph replays) followed by {args.reps} ep_profile repetitions of each "
             "kernel at the final placement.  In the placement the WA gradient and the per-cell pin sums "
             "(k_wa, k_pin_sum) run on a second captured branch beside the density scatter and the "
             "transform products, so those kernels' durations include the sharing of the CUs and the "
             "shares add up to more than the wall time; ep_profile (below) times each item alone.", ""]
    if res:
        r = res["run"]
        lines += [f"Run: {r['iterations']} iterations, {r['evals']} gradient evaluations, "
                  f"{r['seconds']:.3f} s, HPWL {r['hpwl']:.6e}, overflow {r['overflow']:.4f}.", ""]
    lines += ["| kernel | calls | avg us | total ms | share |", "|---|---:|---:|---:|---:|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        lines.append(f"| {_short(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                     f"{float(r['TotalDurationNs']) / 1e6:.2f} | {100 * float(r['TotalDurationNs']) / tot:.1f}% |")
    if res:
        lines += ["", "ep_profile (HIP events, same process, per call of each item):", "",
                  "| item | launches per call | avg us | algorithmic bytes | flops |", "|---|---:|---:|---:|---:|"]
        for q in res["profile"]:
            lines.append(f"| {q['name']} | {q['launches']} | {q['avg_us']:.2f} | {q['bytes']:.4g} | {q['flops']:.4g} |")
    path = os.path.join(PROFILES, f"kernels_{args.workload}_{prec}.md")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def timeline(args) -> str:
    """Kernel trace of one placement (no ep_profile repetitions), split at the
    Nesterov trial kernel: per trial the wall time from one k_trial start to
    the next, the union of the kernels' busy intervals in it (both captured
    branches), so the idle time between kernels, and each kernel's mean
    duration and start offset within the trial."""
    out_dir = os.path.join(args.scratch, f"timeline_{args.workload}_{args.precision}")
    cmd = [sys.executable, "-u", "-m", "aieplace_amd.profiling", "drive", "--workload", args.workload,
           "--precision", str(args.precision), "--reps", "0"]
    if args.from_trace:  # summarise a trace this command wrote before (e.g. on another host)
        full = ["rocprofv3", "--kernel-trace", "--output-format", "csv", "-d", out_dir, "-o", "run", "--"] + cmd
    else:
        full, _ = _rocprof(["--kernel-trace"], cmd, out_dir, args.timeout)
    rows = _csv(out_dir, "kernel_trace.csv")
    ev = sorted(((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), _short(r["Kernel_Name"])) for r in rows))
    starts = [k for k, e in enumerate(ev) if e[2].startswith("k_trial")]
    trials = []
    for a, b in zip(starts[:-1], starts[1:]):
        win = ev[a:b]
        t0, t1 = win[0][0], ev[b][0]
        busy, cur_s, cur_e = 0, None, None
        for s0, e0, _ in win:
            s0, e0 = max(s0, t0), min(e0, t1)
            if cur_e is None or s0 > cur_e:
                if cur_e is not None:
                    busy += cur_e - cur_s
                cur_s, cur_e = s0, e0
            else:
                cur_e = max(cur_e, e0)
        if cur_e is not None:
            busy += cur_e - cur_s
        trials.append((t1 - t0, busy, [(n, s0 - t0, e0 - s0) for s0, e0, n in win]))
    # the converged part of the run: the last half of the trials, without the
    # trials of the final replay that return at once after the stop
    med = sorted(t[0] for t in trials)[len(trials) // 2]
    tail = [t for t in trials[len(trials) // 2:] if t[0] > 0.3 * med]
    wall = sum(t[0] for t in tail) / len(tail) / 1e3
    busy = sum(t[1] for t in tail) / len(tail) / 1e3
    # trials that end with the host's read-back of the control state (a runtime
    # copy kernel) against the others: the difference is what the host's
    # round trip between two batches of graph replays costs the GPU
    rb = [t[0] for t in tail if any(n.startswith("__amd_rocclr_copy") for n, _, _ in t[2])]
    other = [t[0] for t in tail if not any(n.startswith("__amd_rocclr_copy") for n, _, _ in t[2])]
    per = {}
    for _, _, ks in tail:
        seen = {}
        for n, off, dur in ks:
            k = n + ("" if n not in seen else f" #{seen[n] + 1}")
            seen[n] = seen.get(n, 0) + 1
            per.setdefault(k, []).append((off, dur))
    prec = PREC_NAME[args.precision]
    lines = [f"# Kernel timeline: {args.workload}, {prec} transforms", "",
             "Command (from the repository root):", "", "```", _cmd_text(full),
             "```", "",
             f"One device-controlled placement ({len(trials)} trials), split at each Nesterov trial kernel; "
             f"the figures are means over the last {len(tail)} trials.", "",
             f"Trial wall time {wall:.1f} us; kernels busy (union of both captured branches) {busy:.1f} us; "
             f"idle between kernels {wall - busy:.1f} us ({100 * (wall - busy) / wall:.0f}%).", ""]
    if rb and other:
        lines += [f"Trials ending with the control-state read-back: {len(rb)}, mean wall "
                  f"{sum(rb) / len(rb) / 1e3:.1f} us; the other {len(other)}: {sum(other) / len(other) / 1e3:.1f} us.", ""]
    lines += ["| kernel | start offset us | duration us |", "|---|---:|---:|"]
    for k, v in sorted(per.items(), key=lambda kv: sum(o for o, _ in kv[1]) / len(kv[1])):
        lines.append(f"| {k} | {sum(o for o, _ in v) / len(v) / 1e3:.1f} | {sum(d for _, d in v) / len(v) / 1e3:.1f} |")
    lines.append(f"\ntimeline wall_us {wall:.2f} busy_us {busy:.2f}")
    path = os.path.join(PROFILES, f"timeline_{args.workload}_{prec}.md")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def batch_stats(args) -> str:
    out_dir = os.path.join(args.scratch, f"stats_batch{args.instances}_{args.precision}")
    cmd = [sys.executable, "-u", "-m", "aieplace_amd.profiling", "drive-batch", "--precision",
           str(args.precision), "--instances", str(args.instances)]
    full, res = _rocprof(["--kernel-trace", "--stats"], cmd, out_dir, args.timeout)
    rows = _csv(out_dir, "kernel_stats.csv")
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    prec = PREC_NAME[args.precision]
    lines = [f"# Kernel statistics: batched engine, {args.instances} instances, {prec} transforms", "",
             "Command (from the repository root):", "", "```", _cmd_text(full),
             "```", "",
             "Two full batched runs (the first captures the graphs and autotunes the tiles; the second is the "
             "measured one), so every count is about twice one run's.  The WA gradient runs on a second captured "
             "branch while the running instances hold >= 400k cells, so the shares add up to more than the "
             "wall time.", ""]
    if res:
        r = res["run"]
        lines += [f"Measured run (under the profiler): {r['movable_cells']} movable cells, "
                  f"{r['instance_iterations']:.0f} instance-iterations, {r['evals']:.0f} evaluations, "
                  f"{r['launches']:.0f} graph replays, {r['finished']:.0f} instances finished, "
                  f"{r['seconds']:.3f} s ({r['instance_iterations'] / r['seconds']:.0f} instance-iterations/s).", ""]
    lines += ["| kernel | calls | avg us | total ms | share |", "|---|---:|---:|---:|---:|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        lines.append(f"| {_short(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                     f"{float(r['TotalDurationNs']) / 1e6:.2f} | {100 * float(r['TotalDurationNs']) / tot:.1f}% |")
    path = os.path.join(PROFILES, f"kernels_batch{args.instances}_{prec}.md")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


# ---------------------------------------------------------------- traffic

# kernel dispatches per call of each ep_profile item (field: 4 products, plus
# the maxima reset of the split-FP16 form)
DISPATCHES = {"cell_sort": 3, "overflow": 2, "field_transforms": 4}


def _dispatches(rows, counter):
    """[(dispatch id, kernel, KiB)] of one counter, in dispatch order."""
    out = [(int(r["Dispatch_Id"]), _short(r["Kernel_Name"]), float(r["Counter_Value"]))
           for r in rows if r["Counter_Name"] == counter]
    return sorted(out)


def _per_kernel(disp, last_n):
    by = {}
    for _, k, v in disp:
        by.setdefault(k, []).append(v)
    return {k: sum(v[-last_n:]) / len(v[-last_n:]) for k, v in by.items()}


def _per_call(disp, fam, n_disp, reps):
    """KiB of one call of item `fam`: the last reps * n_disp dispatches of its
    kernels (the ep_profile repetitions at the end of the driver) / reps."""
    mine = [v for _, k, v in disp if _family(k) == fam]
    take = mine[-reps * n_disp:]
    return sum(take) / reps if take else 0.0


def traffic(args) -> str:
    prec = PREC_NAME[args.precision]
    disp, drv, copy, cmds = {}, None, {}, []
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        od = os.path.join(args.scratch, f"pmc_{counter}_{args.workload}_{args.precision}")
        full, d = _rocprof(["--pmc", counter], _drive_cmd(args), od, args.timeout)
        cmds.append(full)
        drv = drv or d
        disp[counter] = _dispatches(_csv(od, "counter_collection.csv"), counter)
        od = os.path.join(args.scratch, f"pmc_{counter}_copy")
        full, _ = _rocprof(["--pmc", counter], _drive_cmd(args, ["--copy-mib", str(args.calib_mib)]), od,
                           args.timeout)
        cmds.append(full)
        copy[counter] = max(_per_kernel(_dispatches(_csv(od, "counter_collection.csv"), counter),
                                        args.reps).values())
    # cross-check: read requests by size (one pass, four TCC counters)
    req = ["TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum"]
    od = os.path.join(args.scratch, f"pmc_RDREQ_{args.workload}_{args.precision}")
    full, _ = _rocprof(["--pmc"] + req, _drive_cmd(args), od, args.timeout)
    cmds.append(full)
    rq_rows = _csv(od, "counter_collection.csv")
    rq = {c: _per_kernel(_dispatches(rq_rows, c), args.reps) for c in req}
    copy_bytes = args.calib_mib * (1 << 20)
    fetch_copy, write_copy = copy["FETCH_SIZE"] * 1024.0, copy["WRITE_SIZE"] * 1024.0
    f_fetch, f_write = copy_bytes / fetch_copy, copy_bytes / write_copy
    fetch = _per_kernel(disp["FETCH_SIZE"], args.reps)
    write = _per_kernel(disp["WRITE_SIZE"], args.reps)
    prof = {q["name"]: q for q in drv["profile"]} if drv else {}
    lines = [f"# HBM traffic per call: {args.workload}, {prec} transforms", "",
             "Two rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) over the same "
             f"driver; the last {args.reps} calls of each item are the ep_profile repetitions at the final "
             "placement.  Each counter is scaled by the factor that maps its reading on a "
             f"device-to-device hipMemcpy of {args.calib_mib} MiB (the runtime's copy kernel) to the bytes "
             "that copy moves.", "",
             "Commands (from the repository root):", ""]
    lines += ["```"] + [_cmd_text(c) for c in cmds] + ["```", ""]
    lines += [f"Calibration: the copy of {copy_bytes} B reads FETCH_SIZE {fetch_copy:.0f} B and WRITE_SIZE "
              f"{write_copy:.0f} B: factors {f_fetch:.4f} (fetch) and {f_write:.4f} (write).", ""]
    lines += ["| kernel | FETCH KiB/dispatch | WRITE KiB/dispatch | scaled MB/dispatch (fetch+write) |",
              "|---|---:|---:|---:|"]
    for k in sorted(fetch):
        fb = fetch[k] * 1024 * f_fetch
        wb = write.get(k, 0.0) * 1024 * f_write
        lines.append(f"| {k} | {fetch[k]:.1f} | {write.get(k, 0.0):.1f} | {(fb + wb) / 1e6:.3f} |")
    lines += ["", "Read requests by size (third pass; bytes = 32 n32 + 64 n64 + 128 n128), per dispatch:", "",
              "| kernel | requests | 32B | 64B | 128B | read MB by request size | calibrated FETCH MB |",
              "|---|---:|---:|---:|---:|---:|---:|"]
    for k in sorted(rq[req[0]]):
        n, n32, n64, n128 = (rq[c].get(k, 0.0) for c in req)
        rb = 32 * n32 + 64 * n64 + 128 * n128
        lines.append(f"| {k} | {n:.4g} | {n32:.4g} | {n64:.4g} | {n128:.4g} | {rb / 1e6:.3f} | "
                     f"{fetch.get(k, 0.0) * 1024 * f_fetch / 1e6:.3f} |")
    summary = {}
    for fam, q in prof.items():
        n_disp = DISPATCHES.get(fam, 1) + (1 if fam == "field_transforms" and args.precision == 2 else 0)
        fb = _per_call(disp["FETCH_SIZE"], fam, n_disp, args.reps) * 1024 * f_fetch
        wb = _per_call(disp["WRITE_SIZE"], fam, n_disp, args.reps) * 1024 * f_write
        if fb + wb > 0:
            summary[fam] = dict(traffic_bytes=fb + wb, fetch=fb, write=wb, algorithmic_bytes=q["bytes"],
                                dispatches=n_disp)
    lines += ["", "Per ep_profile item (one call = `dispatches` kernel dispatches):", "",
              "| item | dispatches | measured fetch B | measured write B | measured total B | "
              "algorithmic B | measured / algorithmic |",
              "|---|---:|---:|---:|---:|---:|---:|"]
    for fam, s in summary.items():
        ratio = s["traffic_bytes"] / s["algorithmic_bytes"] if s["algorithmic_bytes"] else float("nan")
        lines.append(f"| {fam} | {s['dispatches']} | {s['fetch']:.4g} | {s['write']:.4g} | "
                     f"{s['traffic_bytes']:.4g} | {s['algorithmic_bytes']:.4g} | {ratio:.2f} |")
    lines += ["", "Machine-readable (bench.py reads the `traffic` lines):", ""]
    for fam, s in summary.items():
        lines.append(f"traffic {fam} {s['traffic_bytes']:.6g}")
    path = os.path.join(PROFILES, f"traffic_{args.workload}_{prec}.md")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


# ---------------------------------------------------------------- MFMA utilisation

MFMA_COUNTERS = ["SQ_VALU_MFMA_BUSY_CYCLES", "SQ_BUSY_CU_CYCLES", "SQ_INSTS_MFMA", "SQ_INSTS_VALU",
                 "SQ_WAVES", "SQ_WAVE_CYCLES", "SQ_WAIT_INST_ANY", "SQ_ACTIVE_INST_VALU"]
PEAK_TFLOPS = {0: 157.3, 1: 2500.0, 2: 2500.0}  # dense MFMA: FP32, BF16, FP16


def mfma(args) -> str:
    prec = PREC_NAME[args.precision]
    od = os.path.join(args.scratch, f"pmc_mfma_{args.workload}_{args.precision}")
    full, _ = _rocprof(["--pmc"] + MFMA_COUNTERS, _drive_cmd(args), od, args.timeout)
    rows = _csv(od, "counter_collection.csv")
    per = {c: _per_kernel(_dispatches(rows, c), args.reps) for c in MFMA_COUNTERS}
    # timings from the same driver run without the profiler (counter collection
    # serialises the dispatches)
    plain = subprocess.run(_drive_cmd(args), cwd=ROOT, capture_output=True, text=True, timeout=args.timeout)
    drv = None
    for line in plain.stdout.splitlines():
        if line.startswith('{"run"'):
            drv = json.loads(line)
    if plain.returncode != 0 or drv is None:
        raise RuntimeError(f"driver run failed ({plain.returncode}):\n{plain.stderr[-2000:]}")
    prof = {q["name"]: q for q in drv["profile"]}
    lines = [f"# Matrix-core utilisation: {args.workload}, {prec} transforms", "",
             "One rocprofv3 --pmc pass of SQ counters over the driver (a full placement, then "
             f"{args.reps} ep_profile repetitions of each item; per-dispatch values are the mean of each "
             "kernel's last dispatches).  MFMA busy = SQ_VALU_MFMA_BUSY_CYCLES / (4 SIMDs x "
             "SQ_BUSY_CU_CYCLES): the share of SIMD cycles, while the CUs are busy, in which the matrix core "
             "works.  VALU active and waiting are per wave: SQ_ACTIVE_INST_VALU (quad-cycles, 4 cycles "
             "each) and SQ_WAIT_INST_ANY over SQ_WAVE_CYCLES.", "",
             f"Written by `python -m aieplace_amd.profiling mfma --workload {args.workload} --precision "
             f"{args.precision}`, which runs (from the repository root):", "", "```",
             _cmd_text(full), "```", "",
             "| kernel | waves | MFMA instr / wave | MFMA busy | VALU instr / wave | VALU active | waiting |",
             "|---|---:|---:|---:|---:|---:|---:|"]
    busy_m = busy_c = 0.0
    for k in sorted(per["SQ_WAVES"]):
        w = per["SQ_WAVES"][k]
        if w <= 0:
            continue
        mb, cb = per["SQ_VALU_MFMA_BUSY_CYCLES"].get(k, 0.0), per["SQ_BUSY_CU_CYCLES"].get(k, 0.0)
        wc = per["SQ_WAVE_CYCLES"].get(k, 0.0)
        util = mb / (4.0 * cb) if cb else 0.0
        va = per["SQ_ACTIVE_INST_VALU"].get(k, 0.0) / wc if wc else 0.0
        wt = per["SQ_WAIT_INST_ANY"].get(k, 0.0) / wc if wc else 0.0
        lines.append(f"| {k} | {w:.0f} | {per['SQ_INSTS_MFMA'].get(k, 0.0) / w:.1f} | {100 * util:.1f}% | "
                     f"{per['SQ_INSTS_VALU'].get(k, 0.0) / w:.0f} | {100 * va:.1f}% | {100 * wt:.1f}% |")
        if _family(k) == "field_transforms":
            busy_m += mb
            busy_c += cb
    q = prof.get("field_transforms")
    if q and busy_c:
        tf = q["flops"] / (q["avg_us"] * 1e-6) / 1e12
        lines += ["", f"Transform products together (the four dispatches of one field solve): MFMA busy "
                      f"{100 * busy_m / (4 * busy_c):.1f}% of SIMD cycles; {q['avg_us']:.1f} us per solve "
                      f"(ep_profile, HIP events, the driver run again without the profiler) for "
                      f"{q['flops']:.4g} algorithmic flops (the six dense transform products, "
                      f"6 Mx My (Mx + My)) = {tf:.1f} TFLOP/s, "
                      f"{100 * tf / PEAK_TFLOPS[args.precision]:.1f}% of the dense "
                      f"{'FP32' if args.precision == 0 else '16-bit'} MFMA peak "
                      f"({PEAK_TFLOPS[args.precision]:g} TFLOP/s).  The symmetric half-products execute "
                      "half of these flops in FP32; the split forms execute three products each, 1.5x "
                      "of them at the 16-bit rate.", ""]
        lines.append(f"mfma field_transforms busy {busy_m / (4 * busy_c):.4f} tflops {tf:.2f}")
    path = os.path.join(PROFILES, f"mfma_{args.workload}_{prec}.md")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def read_traffic(workload: str, precision: int):
    """{item: measured HBM bytes per call} from a committed traffic summary, or {}."""
    path = os.path.join(PROFILES, f"traffic_{workload}_{PREC_NAME[precision]}.md")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic":
                out[parts[1]] = float(parts[2])
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m aieplace_amd.profiling")
    ap.add_argument("cmd", choices=["drive", "drive-batch", "stats", "traffic", "mfma", "batch", "timeline"])
    ap.add_argument("--workload", default="stdcell_2m")
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--instances", type=int, default=64, help="batch: instances of the batched workload")
    ap.add_argument("--copy-mib", type=int, default=0, help="drive: copy calibration instead of placement")
    ap.add_argument("--calib-mib", type=int, default=256, help="traffic: size of the calibration copy")
    ap.add_argument("--scratch", default=os.path.join(tempfile.gettempdir(), "eplace_rocprof"),
                    help="where rocprofv3 writes its raw output (the summaries go to profiles/)")
    ap.add_argument("--timeout", type=int, default=300)
    ap.add_argument("--from-trace", action="store_true",
                    help="timeline: summarise the trace already under --scratch instead of tracing")
    args = ap.parse_args(argv)
    if args.cmd == "drive":
        drive(args)
        return 0
    if args.cmd == "drive-batch":
        drive_batch(args)
        return 0
    os.makedirs(PROFILES, exist_ok=True)
    t0 = time.time()
    path = {"stats": stats, "traffic": traffic, "mfma": mfma, "batch": batch_stats,
            "timeline": timeline}[args.cmd](args)
    print(f"wrote {path} in {time.time() - t0:.1f}s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
