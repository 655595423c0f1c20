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
