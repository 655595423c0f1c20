"""What the trained network gains (GPU): with the committed weights, a cold
start of held-out configurations (seeds not used in training) needs fewer
elastic updates to the 1e-8 residual, on the reference iteration (FP64) and on
the mixed accelerated path, and x_B agrees with the start-from-homogeneous
solve to the solver tolerance."""
import numpy as np
import pytest

from cuda_sto_pf_amd import learned
from cuda_sto_pf_amd.configs import base_params, config


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [777, 4242])
def test_learned_cold_start_needs_fewer_updates(pf, seed):
    net = learned.load()
    phi, xB = learned.training_fields(np.random.default_rng(seed), 64)
    pd = base_params(64, hetero=True)
    for label, p, prec, tol in (("fp64", pd, pf.PF_FP64, 1e-9),
                                ("mixed", dict(pd, **learned.ACCEL), pf.PF_MIXED, 1e-6)):
        i0, _, x0 = learned.cold_start(p, phi, xB, prec)
        i1, _, x1 = learned.cold_start(p, phi, xB, prec, net)
        d = np.abs(x1 - x0).max()
        print(f"seed {seed} {label}: cold-start updates {i0} -> {i1}, max|dxB| {d:.1e}", flush=True)
        assert i1 < i0
        assert d < tol


@pytest.mark.gpu
def test_learned_cold_start_workload2(pf):
    net = learned.load()
    pd, phi, xB, _ = config(2)
    i0, _, x0 = learned.cold_start(pd, phi, xB, pf.PF_FP64)
    i1, _, x1 = learned.cold_start(pd, phi, xB, pf.PF_FP64, net)
    print(f"workload 2 (128^3) fp64: {i0} -> {i1}", flush=True)
    assert i1 <= i0 - 3
    assert np.abs(x1 - x0).max() < 1e-9
