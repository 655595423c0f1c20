#!/usr/bin/env python3
"""Throughput of the phase-field time step on one MI355X.

Workload: 256^3, 40 seeded particles (radii U(5,12), tanh width 4), heterogeneous
Ag2Te/PbTe elasticity solved to a relative residual of 1e-8, zero mode on,
dt = 0.01 (the 256^3 heterogeneous workload the metric is quoted on).
A step is one call of the time step on the resident fields (elastic fixed point,
phi update, composition update with the mass-conserving zero mode).

Reported: timesteps/s of the mixed-precision path (FP32 transforms, FP64
state/reductions/zero mode/residual, the learned network at the cold start of
the elastic fixed point) as ``value``; the plain FP64 path on the
same inputs; max |x_B mixed - x_B FP64| after the same steps; elastic updates
per step; the roofline of the dominant kernel (HIP events inside the library);
the step's timed kernels on their own bytes; the cold-start gain of the
learned warm start; and the CPU FP64 oracle timed on one warm 256^3 step as
``cpu_baseline``.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
       python bench.py --accuracy --workload 2|3|5 [--steps K]
The second form is the long-horizon accuracy run: the FP64 path with the
reference iteration and the mixed + accelerated path side by side on the
same inputs for the workload's stated number of steps (500 at 128^3, 2000 at
246^3), reporting max |dx_B|, max |dphi|, the mass drift of both and the
differences of the free-energy and elastic-energy densities every 100 steps.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "timesteps/s at 256^3 mixed+learned vs FP64 plain at max|delta c|<=1e-5; elastic iters"
HBM_PEAK_GBS = 8000.0


# accelerated elastic solver of the fast path: quadratic extrapolation of the
# last three converged displacements as warm start, Anderson depth 3
ACCEL = dict(elastic_warm_start_order=3, elastic_anderson_depth=3)


def _roctx():
    """The ROCTx range API when the profiler's library is present (a range
    around the timed steps lets a kernel trace select their dispatches);
    without a profiler attached the calls do nothing."""
    import ctypes
    try:
        lib = ctypes.CDLL("librocprofiler-sdk-roctx.so.1")
        lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
        return lib
    except (OSError, AttributeError):
        return None


ROCTX = _roctx()

# the accelerated solve's kernel classes the in-step probe can bracket
SOLVE_CLASSES = ("zpass_hij", "ypass_fwd6", "ypass_inv6", "xpass_fwd6", "xpass_inv6", "green", "aa_update")


def probe_table(pf, sim, names, steps):
    """Per-class in-step figures from the launch probe: launches per step,
    mean duration over all launches (the gated ones of iterations after the
    solve stopped included, as a kernel trace averages them), and the
    algorithmic bytes of the working launches over the summed duration."""
    out = {}
    for name in names:
        n, w, us, b = sim.probe_read(pf.KERNELS[name])
        if n == 0:
            continue
        out[name] = {"launches": n, "working_launches": w, "per_step": n / steps, "us_mean": us / n,
                     "us_per_step": us / steps, "bytes_per_working_launch": b / w if w else 0.0,
                     "GBps": b / us * 1e-3 if us > 0 else 0.0,
                     "frac_hbm": (b / us * 1e-3 if us > 0 else 0.0) / HBM_PEAK_GBS}
    return out


def run_path(pf, pd, phi, xB, cap, prec, warmup, steps, net=None, probe=False):
    """Warm-up and timed steps.  With `probe` (the accelerated solver), the
    warm-up steps after the first bracket every solve kernel class with HIP
    events, the class with the most in-step time is picked, and the timed
    steps bracket that class alone; returns its in-step figures too."""
    sim = pf.Simulation(pd, prec)
    sim.set_fields(phi, xB, xB_cap=cap)
    if net is not None:
        sim.attach_net(net)
    sim.mass0 = sim.scalars()["mass_xBtot"]
    probed = None
    if probe and warmup >= 2:
        sim.step(1)
        sim.probe([pf.KERNELS[k] for k in SOLVE_CLASSES])
        sim.step(warmup - 1)
        warm = probe_table(pf, sim, SOLVE_CLASSES, warmup - 1)
        dom = max(warm, key=lambda k: warm[k]["us_per_step"]) if warm else None
        probed = {"warmup_all_classes": warm, "dominant": dom}
        sim.probe([pf.KERNELS[dom]] if dom else [])
    else:
        sim.step(warmup)
    sim.sync()
    it0 = sim.scalars()["elastic_iterations_total"]
    if ROCTX is not None:
        ROCTX.roctxRangePushA(b"timed steps")
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    t1 = time.perf_counter()
    if ROCTX is not None:
        ROCTX.roctxRangePop()
    sc = sim.scalars()
    if probed is not None and probed["dominant"]:
        probed["timed"] = probe_table(pf, sim, [probed["dominant"]], steps)[probed["dominant"]]
        sim.probe([])
    return sim, (t1 - t0), (sc["elastic_iterations_total"] - it0) / steps, sc, probed


def kernel_table(pf, sim, iters_per_step, mixed):
    """Per-kernel mean duration, bytes and share of the step (launch counts
    from the stage sequence with I elastic updates per step)."""
    updates = max(iters_per_step - 1.0, 0.0)   # every update after the first of a solve
    counts = {
        "zpass_hij": iters_per_step,
        "ypass_fwd6": iters_per_step,
        "ypass_inv6": iters_per_step + 1,
        "xpass_fwd6": iters_per_step,
        "xpass_inv6": iters_per_step + 1,
        "green": iters_per_step,
        "zpass_phirhs": 1,
        "zpass_flux": 1,
        "aa_update": updates,
    }
    table = {}
    for name, kid in pf.KERNELS.items():
        us, b = sim.time_kernel(kid, reps=10)
        if b == 0.0:   # not part of this solver
            continue
        table[name] = (us, b)
    for name, (us, b) in list(table.items()):
        table[name] = {"us": us, "bytes": b, "GBps": b / us * 1e-3, "frac_hbm": b / us * 1e-3 / HBM_PEAK_GBS,
                       "per_step": counts[name], "us_per_step": us * counts[name]}
    return table


def learned_cold_start(pf, learned, net, pd, pd_fast, phi, xB):
    """First (cold) step of the workload with and without the learned warm
    start: elastic updates to the 1e-8 residual, step time, and the change in
    x_B (the tolerance-level difference of two converged solves).  Each pair
    runs once untimed first, so the times exclude the one-time loading of the
    kernels' code at their first launch in the process."""
    out = {}
    for label, p, prec in (("mixed_accelerated", pd_fast, pf.PF_MIXED), ("fp64_reference_iteration", pd, pf.PF_FP64)):
        learned.cold_start(p, phi, xB, prec)
        learned.cold_start(p, phi, xB, prec, net)
        i0, t0, x0 = learned.cold_start(p, phi, xB, prec)
        i1, t1, x1 = learned.cold_start(p, phi, xB, prec, net)
        out[label] = {"updates_without": i0, "updates_with": i1, "step_ms_without": 1e3 * t0,
                      "step_ms_with": 1e3 * t1, "max_abs_dxB": float(np.abs(x1 - x0).max())}
        print(f"[bench] learned cold start {label}: {i0} -> {i1} updates, {1e3 * t0:.1f} -> {1e3 * t1:.1f} ms",
              file=sys.stderr, flush=True)
    # the network runs on the grid averaged over pool^3 voxels
    pool = net.pool()
    grid = tuple(n // pool for n in phi.shape)
    us = net.time_inference(grid, reps=5)
    # algorithmic multiply-adds of the 3^3 convolutions, 7 -> 16 -> 16 -> 16 -> 16 -> 3 channels; the
    # MFMA work pads the first layer's input and the last layer's output to 16 channels and the 27
    # taps to 14 pairs of the K = 32 instruction
    flop = 2.0 * 27 * (7 * 16 + 3 * 16 * 16 + 16 * 3) * float(np.prod(grid))
    padded = 2.0 * 14 * 32 * 16 * 5 * float(np.prod(grid))
    out["inference"] = {"grid": list(grid), "pool": pool, "us": us, "flop": flop, "flop_padded": padded,
                        "achieved_tflops": flop / us * 1e-6, "peak_tflops_bf16": 2500.0,
                        "frac": flop / us * 1e-6 / 2500.0}
    return out


def step_kernels(table, ms_per_step):
    """The step's HBM accounting on the kernels' own bytes: each timed kernel's
    algorithmic bytes per launch (every input read once, every output written
    once, only the retained modes the pruned passes move) times its launches
    per step.  `achieved_GBps` is those bytes over the time those launches
    take; `share_of_step` is that time over the measured step time (the rest
    is the first full map of each solve, the one-component phi and Y
    transforms, the finish and zero-mode kernels and the gaps)."""
    b = sum(v["bytes"] * v["per_step"] for v in table.values())
    us = sum(v["us_per_step"] for v in table.values())
    ach = b / us * 1e-3
    return {"bytes_per_step": b, "us_per_step": us, "achieved_GBps": ach, "peak_GBps": HBM_PEAK_GBS,
            "frac": ach / HBM_PEAK_GBS, "share_of_step": us / (ms_per_step * 1e3)}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_256cube_mixed.json")


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE passes of this benchmark (calibrated against a
    streaming kernel of known size), or None when that record is absent."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        rec = json.load(f)
    k = rec.get("kernels", {}).get(kernel)
    return None if k is None else k["traffic_bytes"]


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_step_times(n, workers):
    from cuda_sto_pf_amd.configs import config
    from oracle import step_ref
    step_ref.FFT_WORKERS = workers
    try:
        pd, phi, xB, meta = config(5, n=n)
        o = step_ref.Oracle(step_ref.Params(**pd), phi, xB, xB_cap=meta["xB_max_safe"])
        t0 = time.perf_counter()
        o.step()
        cold = time.perf_counter() - t0
        t0 = time.perf_counter()
        o.step()
        warm = time.perf_counter() - t0
    finally:
        step_ref.FFT_WORKERS = None
    return cold, warm, o.last_elastic_iterations


def gpu_state_snapshot(pf, sim):
    """The FP64 handle's state after its timed steps: fields, the converged
    displacement of the last solve (compact retained spectrum, expanded to
    the oracle's masked half spectrum) and the frozen zero-mode target."""
    from oracle.step_ref import KSpace, Params, dealias_axis_mask
    p = sim.params
    nx, ny, nz = p.Nx, p.Ny, p.Nz
    ix = np.nonzero(dealias_axis_mask(nx, p.dx))[0]
    iy = np.nonzero(dealias_axis_mask(ny, p.dy))[0]
    iz = np.nonzero(dealias_axis_mask(nz, p.dz, wrap=False))[0]
    u = sim.elastic_state()
    azs = u.shape[1] // (ix.size * iy.size)
    u = u.reshape(3, ix.size, iy.size, azs)[..., :iz.size]
    full = []
    for c in range(3):
        f = np.zeros((nx, ny, nz // 2 + 1), dtype=np.complex128)
        f[np.ix_(ix, iy, iz)] = u[c]
        full.append(f)
    sc = sim.scalars()
    return {"phi": sim.field(pf.PF_PHI), "Y": sim.field(pf.PF_Y), "xB": sim.field(pf.PF_XB),
            "dYdt": sim.field(pf.PF_DYDT), "u_k": full, "target_mass": sc["zero_mode_target_mass"],
            "steps_done": sc["steps_done"]}


def cpu_baseline_256(pd, cap, snap, workers):
    """One warm step of the FP64 numpy oracle at the full 256^3 shape, started
    from the state the GPU's FP64 reference iteration reached (the same
    fields and converged displacement the oracle's own previous step would
    hold), with the transforms on `workers` scipy.fft threads and the
    point-wise numpy work on one thread."""
    from oracle import step_ref
    step_ref.FFT_WORKERS = workers
    try:
        o = step_ref.Oracle(step_ref.Params(**pd), snap["phi"], snap["xB"], dYdt0=snap["dYdt"], xB_cap=cap)
        o.Y = snap["Y"].copy()
        o.xB = snap["xB"].copy()
        o.u_k = [c.copy() for c in snap["u_k"]]
        o.warm_valid = True
        o.target_mass = snap["target_mass"]
        t0 = time.perf_counter()
        o.step()
        dt = time.perf_counter() - t0
    finally:
        step_ref.FFT_WORKERS = None
    return dt, o.last_elastic_iterations


def cpu_baseline(pd=None, cap=None, snap=None, workers=None):
    """The FP64 numpy oracle timed on the GPU box's host.  Headline: one warm
    step at 256^3 from the FP64 GPU state (cpu_baseline_256).  Secondary: the
    40-particle geometry generated at 80^3, one cold step (untimed) then one
    timed warm step, single-threaded and with the transforms on `workers`
    scipy.fft threads, extrapolated to 256^3 by R log R."""
    workers = workers or min(16, os.cpu_count() or 1)
    n = 80
    r_s, r_t = n ** 3, 256 ** 3
    scale = (r_t * math.log(r_t)) / (r_s * math.log(r_s))
    c1, w1, it1 = _oracle_step_times(n, None)
    cw, ww, itw = _oracle_step_times(n, workers)
    single = {"value": 1.0 / (w1 * scale), "cores": 1, "warm_step_s": w1, "cold_step_s": c1}
    threaded = {"value": 1.0 / (ww * scale), "cores": workers, "warm_step_s": ww, "cold_step_s": cw}
    best = threaded if threaded["value"] > single["value"] else single
    small = {"value": best["value"], "cores": best["cores"],
             "sample": f"oracle/step_ref.py FP64 numpy, 1 warm step at {n}^3 of the same geometry "
                       f"({best['warm_step_s']:.2f} s, {it1} elastic updates; cold step {best['cold_step_s']:.2f} s "
                       f"untimed), extrapolated to 256^3 by R log R (x{scale:.1f}); transforms on "
                       f"{best['cores']} thread(s)",
             "single_thread": single, "fft_threads": threaded}
    out = {"unit": "timesteps/s", "kind": "port", "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
           "extrapolated_80cube": small}
    if snap is not None:
        t, it = cpu_baseline_256(pd, cap, snap, workers)
        out.update({"value": 1.0 / t, "cores": workers, "warm_step_s": t, "elastic_updates": it,
                    "sample": f"oracle/step_ref.py FP64 numpy, 1 warm step at 256^3 of the workload from the "
                              f"FP64 GPU state after {snap['steps_done']} steps ({t:.1f} s, {it} elastic "
                              f"updates); transforms on {workers} scipy.fft threads, point-wise work on one"})
    else:
        out.update({"value": small["value"], "cores": small["cores"], "sample": small["sample"]})
    return out


def accuracy_run(pf, idx, steps, pair="mixed"):
    """Both paths in lockstep over `steps` steps of workload `idx`: the FP64
    reference iteration against the mixed accelerated path (pair "mixed") or
    against the FP64 accelerated path (pair "fp64-accel", which separates the
    solver tolerance's share of the difference from the precision's)."""
    from cuda_sto_pf_amd.configs import config
    pd, phi, xB, meta = config(idx)
    steps = steps or meta["steps"]
    cap = meta["xB_max_safe"]
    ref = pf.Simulation(pd, pf.PF_FP64)
    fast = pf.Simulation(dict(pd, **ACCEL), pf.PF_MIXED if pair == "mixed" else pf.PF_FP64)
    ref.set_fields(phi, xB, xB_cap=cap)
    fast.set_fields(phi, xB, xB_cap=cap)
    m0 = ref.scalars()["mass_xBtot"]
    rows = []
    it_r = it_f = 0
    t_r = t_f = 0.0
    every = 100 if steps >= 100 else max(1, steps // 5)
    for s in range(1, steps + 1):
        t0 = time.perf_counter()
        ref.step()
        ref.sync()
        t1 = time.perf_counter()
        fast.step()
        fast.sync()
        t_r += t1 - t0
        t_f += time.perf_counter() - t1
        sr, sf = ref.scalars(), fast.scalars()
        it_r += sr["last_elastic_iterations"]
        it_f += sf["last_elastic_iterations"]
        if s % every == 0 or s == steps:
            xr, xf = ref.field(pf.PF_XB), fast.field(pf.PF_XB)
            pr, pm = ref.field(pf.PF_PHI), fast.field(pf.PF_PHI)
            dxa = np.abs(xr - xf)
            at = np.unravel_index(int(np.argmax(dxa)), dxa.shape)
            row = {"step": s, "max_abs_dxB": float(dxa.max()),
                   "p999_abs_dxB": float(np.quantile(dxa, 0.999)),
                   "max_at": [int(v) for v in at], "phi_at_max": float(pr[at]), "xB_at_max": float(xr[at]),
                   "max_abs_dphi": float(np.abs(pr - pm).max()),
                   "mass_drift_rel_fp64": abs(sr["mass_xBtot"] / m0 - 1.0),
                   "mass_drift_rel_mixed": abs(sf["mass_xBtot"] / m0 - 1.0),
                   "f_total_fp64": sr["f_total"], "f_total_rel_diff": abs(sf["f_total"] / sr["f_total"] - 1.0),
                   "f_elastic_fp64": sr["f_elastic"],
                   "f_elastic_rel_diff": abs(sf["f_elastic"] / sr["f_elastic"] - 1.0),
                   "volume_fraction": sr["volume_fraction"],
                   "elastic_updates_fp64": it_r / s, "elastic_updates_mixed": it_f / s}
            rows.append(row)
            print(f"[accuracy] workload {idx} step {s}: max|dxB|={row['max_abs_dxB']:.2e} "
                  f"max|dphi|={row['max_abs_dphi']:.2e} mass drift {row['mass_drift_rel_fp64']:.1e}/"
                  f"{row['mass_drift_rel_mixed']:.1e} dF/F={row['f_total_rel_diff']:.1e} "
                  f"dFel/Fel={row['f_elastic_rel_diff']:.1e}", file=sys.stderr, flush=True)
    return {"workload": meta["workload"], "grid": list(phi.shape), "steps": steps, "pair": pair,
            "fp64_reference_iteration": {"steps_per_s": steps / t_r, "elastic_updates_per_step": it_r / steps},
            "mixed_accelerated": {"steps_per_s": steps / t_f, "elastic_updates_per_step": it_f / steps},
            "gate_max_abs_dxB": 1e-5, "max_abs_dxB": max(r["max_abs_dxB"] for r in rows), "rows": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 20; with --accuracy the workload's stated horizon)")
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-learned-eval", action="store_true",
                    help="skip the cold-start comparison with/without the network (profiling runs)")
    ap.add_argument("--workload", type=int, default=5,
                    help="workload of the run (5: the 256^3 headline; 2 and 3 for --accuracy and profiling)")
    ap.add_argument("--accuracy", action="store_true",
                    help="long-horizon FP64 vs mixed comparison of --workload (no throughput line)")
    ap.add_argument("--pair", choices=("mixed", "fp64-accel"), default="mixed",
                    help="with --accuracy: the path compared with the FP64 reference iteration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (--gpus {args.gpus} requested); "
              "the multi-GPU decomposition is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)

    from cuda_sto_pf_amd import pfcore as pf
    from cuda_sto_pf_amd.configs import config

    if args.accuracy:
        print(json.dumps(accuracy_run(pf, args.workload, args.steps, args.pair)))
        return
    if args.steps is None:
        args.steps = 20
    pd, phi, xB, meta = config(args.workload)
    cap = meta["xB_max_safe"]
    print(f"[bench] workload {meta['workload']} {phi.shape} warmup={args.warmup} steps={args.steps}",
          file=sys.stderr, flush=True)

    pd_fast = dict(pd, **ACCEL)
    from cuda_sto_pf_amd import learned
    net = learned.load() if os.path.exists(learned.DEFAULT_WEIGHTS) else None
    sim_m, t_m, it_m, sc_m, probed = run_path(pf, pd_fast, phi, xB, cap, pf.PF_MIXED, args.warmup, args.steps, net,
                                              probe=True)
    print(f"[bench] mixed: {args.steps / t_m:.2f} steps/s, {it_m:.2f} elastic updates/step",
          file=sys.stderr, flush=True)
    xb_m = sim_m.field(pf.PF_XB)
    result_fp64 = None
    max_dc = None
    if not args.no_fp64:
        sim_d, t_d, it_d, sc_d, _ = run_path(pf, pd, phi, xB, cap, pf.PF_FP64, args.warmup, args.steps)
        xb_d = sim_d.field(pf.PF_XB)
        snap = None if args.no_cpu else gpu_state_snapshot(pf, sim_d)
        max_dc = float(np.abs(xb_m - xb_d).max())
        result_fp64 = {"steps_per_s": args.steps / t_d, "ms_per_step": 1e3 * t_d / args.steps,
                       "elastic_updates_per_step": it_d,
                       "mass_drift_rel": abs(sc_d["mass_xBtot"] / sim_d.mass0 - 1.0)}
        print(f"[bench] fp64: {args.steps / t_d:.2f} steps/s, {it_d:.2f} updates/step, "
              f"max|dc| mixed vs fp64 = {max_dc:.3e}", file=sys.stderr, flush=True)
        table_d = kernel_table(pf, sim_d, it_d, mixed=False)
        sim_d.close()
        # FP64 with the same accelerated solver, to separate the gain of the
        # solver from the gain of precision
        sim_a, t_a, it_a, sc_a, _ = run_path(pf, pd_fast, phi, xB, cap, pf.PF_FP64, args.warmup, args.steps)
        result_fp64["accelerated_solver"] = {"steps_per_s": args.steps / t_a,
                                             "elastic_updates_per_step": it_a,
                                             "max_abs_dc_vs_fp64_plain": float(np.abs(sim_a.field(pf.PF_XB) - xb_d).max())}
        print(f"[bench] fp64 + accelerated solver: {args.steps / t_a:.2f} steps/s, {it_a:.2f} updates/step",
              file=sys.stderr, flush=True)
        sim_a.close()

    learned_stats = (learned_cold_start(pf, learned, net, pd, pd_fast, phi, xB)
                     if net is not None and not args.no_learned_eval else None)

    table = kernel_table(pf, sim_m, it_m, mixed=True)
    # roofline of the kernel with the most in-step time, measured on its
    # launches inside the timed steps (HIP events on the handle's stream)
    if probed and probed.get("timed"):
        dom_name = probed["dominant"]
        tm = probed["timed"]
        roofline = {"kernel": dom_name, "bound": "hbm", "achieved": tm["GBps"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": tm["GBps"] / HBM_PEAK_GBS, "traffic": pmc_traffic(dom_name),
                    "bytes_per_launch": tm["bytes_per_working_launch"], "us_per_launch": tm["us_mean"],
                    "launches": tm["launches"], "working_launches": tm["working_launches"],
                    "measured": "HIP events around each launch of the class in the timed steps; achieved = "
                                "algorithmic bytes of the working launches / summed duration of all launches"}
    else:
        dom_name = max(table, key=lambda k: table[k]["us_per_step"])
        dom = table[dom_name]
        roofline = {"kernel": dom_name, "bound": "hbm", "achieved": dom["GBps"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": dom["GBps"] / HBM_PEAK_GBS, "traffic": pmc_traffic(dom_name),
                    "bytes_per_launch": dom["bytes"], "us_per_launch": dom["us"],
                    "measured": "HIP events around repeated launches of the kernel alone"}
    for k, v in table.items():
        print(f"[bench] kernel {k:14s} {v['us']:8.1f} us  {v['GBps']:7.0f} GB/s  x{v['per_step']:.1f}/step",
              file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pd, cap, None if args.no_fp64 else snap)
        print(f"[bench] cpu baseline {cpu['value']:.4g} steps/s ({cpu['sample']})", file=sys.stderr, flush=True)

    value = args.steps / t_m
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_m / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 transforms / f64 state, reductions, zero mode, residual",
        "data": "synthetic",
        "config": {"workload": ("256^3, 40 particles R~U(5,12), hetero elastic tol 1e-8, zero mode, dt 0.01"
                                if args.workload == 5 else f"{meta['workload']} {phi.shape[0]}^3"),
                   "grid": list(phi.shape), "learned_warm_start": net is not None,
                   "kz_stride": sc_m["kz_stride"],
                   "tile_plan": {"zpass_rows": sc_m["zpass_rows"], "ypass_cols": sc_m["ypass_cols"],
                                 "xpass_cols": sc_m["xpass_cols"]},
                   "path": "mixed precision, learned cold start + extrapolated warm start (order 3) + "
                           "Anderson depth 3"},
        "elastic_updates_per_step": it_m,
        "mass_drift_rel": abs(sc_m["mass_xBtot"] / sim_m.mass0 - 1.0),
        "max_abs_dc_vs_fp64": max_dc,
        "fp64_plain": result_fp64,
        "mixed_over_fp64": (value / result_fp64["steps_per_s"]) if result_fp64 else None,
        "roofline": roofline,
        "step_kernels": {
            "mixed": step_kernels(table, 1e3 * t_m / args.steps),
            "fp64_plain": step_kernels(table_d, result_fp64["ms_per_step"]) if result_fp64 else None},
        "learned_warm_start": learned_stats,
        "kernels_mixed": table,
        "kernels_mixed_in_step": probed["warmup_all_classes"] if probed else None,
        "kernels_fp64": table_d if not args.no_fp64 else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
