"""Workload 4 in reduced form (GPU): a batch of minimize instances of three
sizes with rotated eigenstrains, run concurrently by pf_batch_minimize,
reproduces every instance's solo run exactly (same phi, same iteration
count); the batched dynamics entry does the same for time steps."""
import numpy as np
import pytest

from cuda_sto_pf_amd import sweep
from cuda_sto_pf_amd.configs import config, sweep_instances


@pytest.mark.gpu
def test_batch_minimize_equals_solo_runs(pf):
    inst = sweep_instances()
    pick = [inst[0], inst[8], inst[9], inst[24], inst[25], inst[47]]   # 128^3, 96^3, 64^3
    opts = pf.minimize_opts(dt=0.05, V0=0.1, max_iter=15)
    t1, s1, p1 = sweep.run(pick, 1, opts, pf.PF_MIXED)
    t2, s2, p2 = sweep.run(pick, 4, opts, pf.PF_MIXED)
    print(f"solo {t1:.2f} s, batch {t2:.2f} s", flush=True)
    for a, b, x, y in zip(s1, s2, p1, p2):
        assert a["iteration"] == b["iteration"] == 15
        assert np.array_equal(x, y)
        assert abs(b["mean_h"] - 0.1) < 1e-3


@pytest.mark.gpu
def test_batch_step_equals_solo_steps(pf):
    sims, solo = [], []
    for n in (16, 20, 24):
        pd, phi, xB, meta = config(2, n)
        for lst in (sims, solo):
            s = pf.Simulation(pd, pf.PF_FP64)
            s.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
            lst.append(s)
    b = pf.Batch(sims, threads=3)
    b.step(3)
    for s in solo:
        s.step(3)
    for s, t in zip(sims, solo):
        assert np.array_equal(s.field(pf.PF_XB), t.field(pf.PF_XB))


@pytest.mark.gpu
@pytest.mark.parametrize("prec,iters", [("mixed", 15), ("fp64", 4)])
def test_fused_batch_minimize_equals_solo_runs(pf, prec, iters):
    """The fused batch (each stage one batched launch over the instances of a
    grid size) reproduces every instance's solo run bit for bit, with fewer
    launches than the instances record."""
    inst = sweep_instances()
    pick = [inst[0], inst[8], inst[9], inst[24], inst[25], inst[47]]   # 128^3, 96^3, 64^3
    p = pf.PF_MIXED if prec == "mixed" else pf.PF_FP64
    opts = pf.minimize_opts(dt=0.05, V0=0.1, max_iter=iters)
    t1, s1, p1 = sweep.run(pick, 1, opts, p)
    for graphs in (False, True):
        lc = {}
        t2, s2, p2 = sweep.run(pick, 1, opts, p, fused=True, launches=lc, graphs=graphs)
        print(f"{prec}: solo {t1:.2f} s, fused {t2:.2f} s (graphs {graphs}), {lc['batched']} batched launches for "
              f"{lc['recorded']} recorded, {lc['graph_launches']} graph launches of {lc['graphs']} graphs",
              flush=True)
        for a, b, x, y in zip(s1, s2, p1, p2):
            assert a["iteration"] == b["iteration"] == iters
            assert a["lambda_vol"] == b["lambda_vol"] and a["rms_res"] == b["rms_res"]
            assert np.array_equal(x, y)
        assert lc["batched"] * 2 < lc["recorded"]
        # graphs are replayed (the same plan recurs across iterations)
        assert (lc["graph_launches"] > lc["graphs"]) if graphs else lc["graph_launches"] == 0


@pytest.mark.gpu
def test_fused_batch_refuses_what_it_cannot_record(pf):
    inst = sweep_instances()[47:]
    sims = sweep.make_sims(inst, pf.PF_MIXED, accel=False)
    b = pf.Batch(sims, 1)
    with pytest.raises(pf.PfError, match="accelerated elastic solver"):
        b.minimize_fused(pf.minimize_opts(dt=0.05, V0=0.1, max_iter=2))
    b.close()


def _dyn_sims(pf, prec, sizes=(16, 20, 24)):
    out = []
    for n in sizes:
        pd, phi, xB, meta = config(2, n)
        pd = dict(pd, elastic_warm_start_order=3, elastic_anderson_depth=3)
        s = pf.Simulation(pd, prec)
        s.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
        out.append(s)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fused_batch_step_equals_solo_steps(pf, prec):
    """Time steps as a fused batch (each phase of a step one batched launch
    per kernel over the instances, recurring phases replayed as HIP graphs)
    reproduce every instance's solo steps with the same elastic and
    zero-mode iteration counts: bit for bit in mixed precision, to rounding in
    FP64 (the batched and the single compilation of an FP64 kernel on the
    step's composition path fuse multiply-adds differently: 1e-14 after 5
    steps); one handle alone runs its steps as graph replays."""
    def same(x, y):
        if prec == "mixed":
            return np.array_equal(x, y)
        return float(np.abs(x - y).max()) <= 1e-12 * max(1.0, float(np.abs(x).max()))
    p = pf.PF_FP64 if prec == "fp64" else pf.PF_MIXED
    solo, fused = _dyn_sims(pf, p), _dyn_sims(pf, p)
    for s in solo:
        s.step(5)
    b = pf.Batch(fused, 1)
    b.step_fused(5)
    batched, recorded, graphs, glaunch = b.fused_launches()
    print(f"{prec}: {batched} batched launches for {recorded} recorded, {glaunch} graph launches of {graphs} graphs",
          flush=True)
    b.close()
    for s, t in zip(solo, fused):
        for f in (pf.PF_PHI, pf.PF_Y, pf.PF_XB):
            assert same(s.field(f), t.field(f))
        a, c = s.scalars(), t.scalars()
        assert a["elastic_iterations_total"] == c["elastic_iterations_total"]
        assert a["zero_mode_iterations"] == c["zero_mode_iterations"] and a["steps_done"] == c["steps_done"] == 5
    assert batched < recorded and glaunch > graphs   # three grid sizes: the transform passes stay one launch per size
    # one handle: its steps as graph replays, then ordinary steps continue it
    one, ref = _dyn_sims(pf, p, (24,)), _dyn_sims(pf, p, (24,))
    b1 = pf.Batch(one, 1)
    b1.step_fused(4)
    assert b1.fused_launches()[3] > 0
    b1.close()
    one[0].step(2)
    ref[0].step(6)
    for f in (pf.PF_PHI, pf.PF_Y, pf.PF_XB):
        assert same(one[0].field(f), ref[0].field(f))
