"""The benchmark workloads at their full sizes, through size-independent
properties (the CPU oracle does not finish these sizes in seconds):

* x_B of the mixed-precision accelerated path within 1e-5 of the FP64 path
  running the reference iteration, on identical inputs;
* the zero mode holds the mass of both paths at the t = 0 value to rounding;
* every elastic solve reaches the 1e-8 residual (fail-closed otherwise);
* 246^3 exercises the direct 41-point stage of the mixed-radix transform;
* 256^3 (40 particles) is the grid the throughput is quoted on.
"""
import numpy as np
import pytest

from cuda_sto_pf_amd.configs import config

ACCEL = dict(elastic_warm_start_order=3, elastic_anderson_depth=3)


def _pair(pf, idx, steps):
    pd, phi, xB, meta = config(idx)
    cap = meta["xB_max_safe"]
    ref = pf.Simulation(pd, pf.PF_FP64)
    fast = pf.Simulation(dict(pd, **ACCEL), pf.PF_MIXED)
    ref.set_fields(phi, xB, xB_cap=cap)
    fast.set_fields(phi, xB, xB_cap=cap)
    m0 = ref.scalars()["mass_xBtot"]
    it_r = it_f = 0
    for s in range(steps):
        ref.step()
        fast.step()
        sr, sf = ref.scalars(), fast.scalars()
        it_r += sr["last_elastic_iterations"]
        it_f += sf["last_elastic_iterations"]
        assert sr["last_elastic_residual"] <= 1e-8 and sf["last_elastic_residual"] <= 1e-8
    dx = np.abs(ref.field(pf.PF_XB) - fast.field(pf.PF_XB)).max()
    dphi = np.abs(ref.field(pf.PF_PHI) - fast.field(pf.PF_PHI)).max()
    print(f"workload {idx} {phi.shape}: {steps} steps, max|dxB|={dx:.2e} max|dphi|={dphi:.2e}, "
          f"elastic updates fp64 {it_r} / mixed+accel {it_f}", flush=True)
    return dx, m0, ref.scalars()["mass_xBtot"], fast.scalars()["mass_xBtot"], it_r, it_f


@pytest.mark.gpu
@pytest.mark.parametrize("idx,steps", [(2, 20), (3, 6), (5, 10)])
def test_fullsize_mixed_vs_fp64(pf, idx, steps):
    dx, m0, mr, mf, it_r, it_f = _pair(pf, idx, steps)
    assert dx < 1e-5
    assert abs(mr - m0) <= 1e-12 * m0
    assert abs(mf - m0) <= 1e-12 * m0
    assert it_f < it_r


@pytest.mark.gpu
def test_workload2_stated_horizon(pf):
    """Workload 2 over its stated 500 steps at 128^3: the mixed accelerated path
    stays within the 1e-5 gate of the FP64 reference iteration in x_B, both
    hold the mass to rounding, and the free-energy and elastic-energy
    densities agree (the mixed path updates Y by its increment, so FP32
    rounding does not accumulate in Y)."""
    pd, phi, xB, meta = config(2)
    cap = meta["xB_max_safe"]
    ref = pf.Simulation(pd, pf.PF_FP64)
    fast = pf.Simulation(dict(pd, **ACCEL), pf.PF_MIXED)
    ref.set_fields(phi, xB, xB_cap=cap)
    fast.set_fields(phi, xB, xB_cap=cap)
    m0 = ref.scalars()["mass_xBtot"]
    for block in range(5):
        ref.step(100)
        fast.step(100)
        dx = np.abs(ref.field(pf.PF_XB) - fast.field(pf.PF_XB)).max()
        sr, sf = ref.scalars(), fast.scalars()
        print(f"workload 2 step {100 * (block + 1)}: max|dxB|={dx:.2e} "
              f"dF/F={abs(sf['f_total'] / sr['f_total'] - 1):.1e} "
              f"dFel/Fel={abs(sf['f_elastic'] / sr['f_elastic'] - 1):.1e}", flush=True)
        assert dx < 1e-5
    assert abs(sr["mass_xBtot"] - m0) <= 1e-12 * m0 and abs(sf["mass_xBtot"] - m0) <= 1e-12 * m0
    assert abs(sf["f_total"] / sr["f_total"] - 1) < 1e-6
    assert abs(sf["f_elastic"] / sr["f_elastic"] - 1) < 1e-3
