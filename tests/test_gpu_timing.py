"""pf_time_kernel: every kernel id launches and reports its bytes, in the
form the configured solver runs it; the kernels a solver does not run report
0 bytes."""
import numpy as np
import pytest

from cuda_sto_pf_amd.configs import config


def _sim(pf, pd, prec):
    _, phi, xB, meta = config(2, 32)
    s = pf.Simulation(pd, prec)
    s.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
    s.step(2)
    return s


CASES = [("fp64_reference_iteration", "PF_FP64", 0, ("aa_update",)),
         ("mixed_no_history", "PF_MIXED", 0, ("aa_update",)),
         ("mixed_accelerated", "PF_MIXED", 3, ())]


@pytest.mark.gpu
@pytest.mark.parametrize("name_,prec,depth,absent", CASES, ids=[c[0] for c in CASES])
def test_time_kernel_all_ids(pf, name_, prec, depth, absent):
    """Neither the reference iteration nor the mixed solver without a history
    runs the Anderson update."""
    pd = config(2, 32)[0]
    if depth:
        pd = dict(pd, elastic_warm_start_order=3, elastic_anderson_depth=depth)
    s = _sim(pf, pd, getattr(pf, prec))
    for name, kid in pf.KERNELS.items():
        us, b = s.time_kernel(kid, reps=3)
        print(f"{name}: {us:.1f} us, {b / 1e6:.2f} MB", flush=True)
        if name in absent:
            assert b == 0.0
            continue
        assert b > 0.0 and np.isfinite(us) and us > 0.0, name
    with pytest.raises(pf.PfError):
        s.time_kernel(len(pf.KERNELS), reps=1)
    # the state is invalid afterwards; set_fields restores a runnable handle
    _, phi, xB, meta = config(2, 32)
    s.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
    s.step(1)
    assert np.isfinite(s.field(pf.PF_XB)).all()


@pytest.mark.gpu
def test_probe_counts_the_solve_launches(pf):
    """The in-step probe brackets every launch of the accelerated solve's
    kernel classes: as many working launches as elastic updates (the Anderson
    update one fewer per solve, the inverse passes one more), each working
    launch with the bytes pf_time_kernel counts for its form, and a step
    taken with the probe off adds nothing."""
    pd = dict(config(2, 32)[0], elastic_warm_start_order=3, elastic_anderson_depth=3)
    s = _sim(pf, pd, pf.PF_MIXED)
    names = ("zpass_hij", "ypass_fwd6", "ypass_inv6", "xpass_fwd6", "xpass_inv6", "green", "aa_update")
    s.probe([pf.KERNELS[k] for k in names])
    it0 = s.scalars()["elastic_iterations_total"]
    steps = 4
    s.step(steps)
    iters = s.scalars()["elastic_iterations_total"] - it0
    got = {k: s.probe_read(pf.KERNELS[k]) for k in names}
    s.probe([])
    s.step(1)
    for k in names:
        assert s.probe_read(pf.KERNELS[k])[0] == 0, k
    for k, (n, w, us, b) in got.items():
        print(f"{k}: {n} launches, {w} working, {us:.1f} us, {b / 1e6:.2f} MB", flush=True)
        assert n >= w > 0 and us > 0.0, k
    want = {"zpass_hij": iters, "ypass_fwd6": iters, "xpass_fwd6": iters, "green": iters,
            "ypass_inv6": iters + steps, "xpass_inv6": iters + steps, "aa_update": iters - steps}
    for k, w in want.items():
        assert got[k][1] == w, k
    # fixed-size classes: bytes per working launch equal pf_time_kernel's count
    # (the state is invalid after pf_time_kernel, so it runs last)
    for k in ("ypass_fwd6", "ypass_inv6", "xpass_fwd6", "xpass_inv6"):
        _, b1 = s.time_kernel(pf.KERNELS[k], reps=1)
        assert got[k][3] == pytest.approx(got[k][1] * b1, rel=1e-12), k
    # the z pass: the first update of each solve maps phi and x_B, the others
    # the increment form pf_time_kernel runs; the Green update with a full
    # history is pf_time_kernel's form, fewer slots read less
    _, bz = s.time_kernel(pf.KERNELS["zpass_hij"], reps=1)
    assert got["zpass_hij"][3] > got["zpass_hij"][1] * bz
    _, bg = s.time_kernel(pf.KERNELS["green"], reps=1)
    assert got["green"][3] < got["green"][1] * bg
    with pytest.raises(pf.PfError):
        s.probe_read(len(pf.KERNELS))
