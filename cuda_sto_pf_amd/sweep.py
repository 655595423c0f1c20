"""Workload 4: the batched minimize sweep.

48 phi-only minimize instances of uneven size (24 at 64^3, 16 at 96^3, 8 at
128^3), each one sphere or oblate spheroid with a Haar-rotated eigenstrain
(configs.sweep_instances), run with --minimize-dt 0.05, --minimize-max-iter
500, --V0 0.1 (the minimize branch of main_cuda.cu:28341-32600).  They run
once one at a time, once as a batch of streams (pf_batch_minimize: one
stream per instance, host workers overlapping the instances on the GPU) and
once fused (pf_batch_minimize_fused: every stage of an iteration one batched
launch over the instances); the final phi of every instance is compared with
the solo run and the throughput reported.

    python -m cuda_sto_pf_amd.sweep --threads 8 --report sweep.json
"""
import argparse
import json
import sys
import time

import numpy as np

from . import pfcore as pf
from .configs import sweep_instances

ACCEL = dict(elastic_warm_start_order=3, elastic_anderson_depth=3)


def make_sims(instances, prec, accel=True, net=None):
    sims = []
    for inst in instances:
        pd = dict(inst["params"], **ACCEL) if accel else inst["params"]
        s = pf.Simulation(pd, prec)
        s.set_fields(inst["phi"], np.zeros_like(inst["phi"]))
        if net is not None:
            s.attach_net(net)
        sims.append(s)
    return sims


def run(instances, threads, opts, prec, accel=True, net=None, fused=False, launches=None, graphs=True):
    """Wall time, per-instance stats and final phi of one sweep: one instance
    at a time (threads <= 1), on one stream per instance with `threads` host
    workers, or fused (every stage one batched launch over the instances);
    `launches` (a dict) receives the fused run's launch counts."""
    sims = make_sims(instances, prec, accel, net)
    for s in sims:
        s.sync()
    t0 = time.perf_counter()
    if fused:
        b = pf.Batch(sims, 1)
        b.set_graphs(graphs)
        stats = b.minimize_fused(opts)
        if launches is not None:
            launches["batched"], launches["recorded"], launches["graphs"], launches["graph_launches"] = \
                b.fused_launches()
        b.close()
    elif threads <= 1:
        stats = [s.minimize(opts) for s in sims]
    else:
        b = pf.Batch(sims, threads)
        stats = b.minimize(opts)
        b.close()
    for s in sims:
        s.sync()
    wall = time.perf_counter() - t0
    phis = [s.field(pf.PF_PHI) for s in sims]
    for s in sims:
        s.close()
    return wall, stats, phis


def dynamics(n, count, steps, prec, threads=8):
    """Wall time of `steps` time steps of `count` handles of the workload-2
    geometry at n^3 (accelerated solver): one at a time, on one stream per
    handle, and fused (with and without graph replay); the fused fields are
    compared with the solo ones."""
    from .configs import config

    def make():
        out = []
        for _ in range(count):
            pd, phi, xB, meta = config(2, n)
            s = pf.Simulation(dict(pd, **ACCEL), prec)
            s.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
            s.step(1)   # past the cold start
            s.sync()
            out.append(s)
        return out

    res = {}
    sims = make()
    t0 = time.perf_counter()
    for s in sims:
        s.step(steps)
    for s in sims:
        s.sync()
    res["alone_s"] = time.perf_counter() - t0
    ref = [s.field(pf.PF_XB) for s in sims]
    sims = make()
    b = pf.Batch(sims, threads)
    t0 = time.perf_counter()
    b.step(steps)
    res["streams_s"] = time.perf_counter() - t0
    b.close()
    for graphs in (False, True):
        sims = make()
        b = pf.Batch(sims, 1)
        b.set_graphs(graphs)
        t0 = time.perf_counter()
        b.step_fused(steps)
        res["fused_graphs_s" if graphs else "fused_s"] = time.perf_counter() - t0
        lc = b.fused_launches()
        b.close()
        res["max_abs_dxB_fused"] = max(float(np.abs(s.field(pf.PF_XB) - r).max()) for s, r in zip(sims, ref))
    res.update(batched_launches=lc[0], recorded_launches=lc[1], graphs=lc[2], graph_launches=lc[3])
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--max-iter", type=int, default=500)
    ap.add_argument("--dt", type=float, default=0.05)
    ap.add_argument("--V0", type=float, default=0.1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--learned", action="store_true", help="learned cold start of every instance")
    ap.add_argument("--subset", type=int, default=0, help="only the first n instances")
    ap.add_argument("--report", default=None)
    ap.add_argument("--modes", default="alone,streams,fused",
                    help="comma list of alone, streams, fused (fused alone: for profiling)")
    ap.add_argument("--dynamics", default=None, metavar="N,COUNT,STEPS",
                    help="time steps instead: COUNT handles at N^3, STEPS steps each")
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    if a.dynamics:
        n, count, steps = (int(v) for v in a.dynamics.split(","))
        prec = pf.PF_FP64 if a.precision == "fp64" else pf.PF_MIXED
        r = dynamics(n, count, steps, prec, a.threads)
        log(f"[dynamics] {count} x {n}^3, {steps} steps, {a.precision}: " + json.dumps(r))
        if a.report:
            with open(a.report, "w") as f:
                json.dump(dict(n=n, count=count, steps=steps, precision=a.precision, **r), f, indent=1)
        return 0
    inst = sweep_instances()
    if a.subset:
        inst = inst[:a.subset]
    prec = pf.PF_FP64 if a.precision == "fp64" else pf.PF_MIXED
    net = None
    if a.learned:
        from . import learned
        net = learned.load()
    opts = pf.minimize_opts(dt=a.dt, V0=a.V0, max_iter=a.max_iter)
    log(f"[sweep] {len(inst)} instances, {a.precision}, max_iter {a.max_iter}, learned={a.learned}")
    modes = set(a.modes.split(","))
    if modes == {"fused"}:
        lc = {}
        t_fus, st_fus, _ = run(inst, 1, opts, prec, fused=True, launches=lc)
        it_fus = sum(s["iteration"] for s in st_fus)
        log(f"[sweep] fused: {t_fus:.2f} s, {it_fus / t_fus:.0f} iterations/s, {lc['batched']} batched launches "
            f"for {lc['recorded']} recorded, {lc['graph_launches']} graph launches of {lc['graphs']} graphs")
        return 0
    t_seq, st_seq, phi_seq = run(inst, 1, opts, prec, net=net)
    it_seq = sum(s["iteration"] for s in st_seq)
    log(f"[sweep] one at a time: {t_seq:.2f} s, {it_seq} iterations, {it_seq / t_seq:.0f} iterations/s")
    t_bat, st_bat, phi_bat = run(inst, a.threads, opts, prec, net=net)
    it_bat = sum(s["iteration"] for s in st_bat)
    log(f"[sweep] batch ({a.threads} workers): {t_bat:.2f} s, {it_bat} iterations, {it_bat / t_bat:.0f} iterations/s")
    dmax = max(float(np.abs(x - y).max()) for x, y in zip(phi_seq, phi_bat))
    same_iters = all(x["iteration"] == y["iteration"] for x, y in zip(st_seq, st_bat))
    conv = sum(s["converged"] for s in st_bat)
    log(f"[sweep] max|phi batch - phi alone| = {dmax:.1e}, same iteration counts: {same_iters}, converged {conv}")
    fused = {}
    if not a.learned:
        lc, lc0 = {}, {}
        t_fng, st_fng, phi_fng = run(inst, 1, opts, prec, fused=True, launches=lc0, graphs=False)
        log(f"[sweep] fused without graphs: {t_fng:.2f} s")
        t_fus, st_fus, phi_fus = run(inst, 1, opts, prec, fused=True, launches=lc)
        it_fus = sum(s["iteration"] for s in st_fus)
        dmax_f = max(float(np.abs(x - y).max()) for x, y in zip(phi_seq, phi_fus))
        same_f = all(x["iteration"] == y["iteration"] for x, y in zip(st_seq, st_fus))
        same_f = same_f and all(np.array_equal(x, y) for x, y in zip(phi_fus, phi_fng))
        log(f"[sweep] fused: {t_fus:.2f} s, {it_fus / t_fus:.0f} iterations/s, {lc['batched']} batched launches "
            f"for {lc['recorded']} recorded, {lc['graph_launches']} graph launches of {lc['graphs']} graphs, "
            f"max|phi fused - phi alone| = {dmax_f:.1e}, same iteration counts: {same_f}")
        fused = dict(fused_s=t_fus, fused_no_graphs_s=t_fng, graphs=lc["graphs"], graph_launches=lc["graph_launches"],
                     speedup_fused=t_seq / t_fus, iterations_per_s_fused=it_fus / t_fus,
                     fused_launches=lc["batched"], recorded_launches=lc["recorded"], max_abs_dphi_fused=dmax_f,
                     same_iteration_counts_fused=same_f)
    rep = dict(instances=len(inst), precision=a.precision, learned=a.learned, max_iter=a.max_iter, dt=a.dt, V0=a.V0,
               threads=a.threads, one_at_a_time_s=t_seq, batch_s=t_bat, speedup=t_seq / t_bat,
               iterations=it_bat, iterations_per_s_batch=it_bat / t_bat, iterations_per_s_alone=it_seq / t_seq,
               converged=conv, max_abs_dphi=dmax, same_iteration_counts=same_iters, **fused,
               per_instance=[dict(name=i["name"], radius=i["radius"], aspect=i["aspect"], iterations=s["iteration"],
                                  converged=s["converged"], mean_h=s["mean_h"], rms_res=s["rms_res"])
                             for i, s in zip(inst, st_bat)])
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
