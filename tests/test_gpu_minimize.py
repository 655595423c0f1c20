"""Phi-only minimize mode, HIP (libpfcore) against the FP64 CPU oracle
(oracle/minimize_ref.py) on identical inputs: phi to 1e-10 in FP64 mode after
every iteration, the multiplier, residuals, volume and energies of each
iteration, the elastic iteration counts and the converged iteration."""
import math

import numpy as np
import pytest

from cuda_sto_pf_amd.configs import base_params, spheres_phi
from oracle import thermo as th
from oracle.minimize_ref import MinimizeOptions, MinimizeOracle
from oracle.step_ref import Params

FP64_TOL = 1e-10


def _case(shape, elastic=True, hetero=True, radius=4.0, V0_scale=1.05):
    pd = base_params(shape, hetero=hetero, elastic=elastic, zero_mode=False)
    phi = spheres_phi(shape, [(shape[0] / 2, shape[1] / 2, shape[2] / 2)], [radius], width=1.5)
    V0 = V0_scale * th.h_of_phi(th.clamp01(phi)).mean() if V0_scale > 0 else 0.0
    return pd, phi, V0


CASES = [
    ("sphere16_hetero", ) + _case((16, 16, 16)),
    ("rect_noelastic_lock", ) + _case((16, 12, 8), elastic=False, radius=3.0, V0_scale=0.0),
    ("rect_homog_shrink", ) + _case((12, 16, 8), hetero=False, radius=3.5, V0_scale=0.9),
]


def _pair(pf, pd, phi, V0, prec, **kw):
    opts = dict(dt=0.05, V0=V0, **kw)
    o = MinimizeOracle(Params(**pd), phi, MinimizeOptions(**opts))
    sim = pf.Simulation(pd, prec)
    sim.set_fields(phi, np.zeros_like(phi))
    sim.minimize_begin(pf.minimize_opts(**opts))
    return o, sim


def _close(a, b, rel, abs_):
    return abs(a - b) <= max(rel * max(abs(a), abs(b)), abs_)


@pytest.mark.gpu
@pytest.mark.parametrize("name,pd,phi,V0", CASES, ids=[c[0] for c in CASES])
def test_fp64_minimize_parity(pf, name, pd, phi, V0):
    o, sim = _pair(pf, pd, phi, V0, pf.PF_FP64)
    for it in range(6):
        so = o.iterate()
        sg = sim.minimize_iterate()
        dphi = np.abs(sim.field(pf.PF_PHI) - o.phi).max()
        print(f"{name} iter {it + 1}: dphi={dphi:.2e} lambda gpu={sg['lambda_vol']:.12e} cpu={so['lambda_vol']:.12e} "
              f"rms_res {sg['rms_res']:.6e}/{so['rms_res']:.6e} rms_dphi {sg['rms_dphi']:.6e}/{so['rms_dphi']:.6e} "
              f"el iters {sg['elastic_iterations']}/{so['elastic_iterations']}", flush=True)
        assert dphi < FP64_TOL
        assert sg["iteration"] == it + 1
        assert sg["elastic_iterations"] == so["elastic_iterations"]
        for k, kg, rel, ab in (("lambda_vol", "lambda_vol", 1e-9, 1e-12), ("lambda_raw", "lambda_raw", 1e-9, 1e-12),
                               ("lambda_kkt", "lambda_kkt", 1e-9, 1e-12), ("rms_res", "rms_res", 1e-8, 1e-12),
                               ("rms_dphi", "rms_dphi", 1e-6, 1e-9), ("mean_h", "mean_h", 1e-12, 1e-15),
                               ("V0", "V0", 1e-13, 0.0), ("vol_err_rel", "vol_err_rel", 1e-6, 1e-12),
                               ("F_surf", "F_surf_hat", 1e-10, 1e-15), ("F_el", "F_el_hat", 1e-10, 1e-15),
                               ("el_core_sum", "el_core_sum", 1e-10, 1e-13),
                               ("interface_sum", "interface_sum", 1e-10, 1e-12)):
            assert _close(sg[kg], so[k], rel, ab), (k, sg[kg], so[k])


@pytest.mark.gpu
def test_fp64_minimize_converges_like_oracle(pf):
    """Run to the convergence rule (condA held for 10 iterations) with
    thresholds the 12^3 case reaches quickly: same converged iteration."""
    pd, phi, V0 = _case((12, 12, 12), elastic=False, radius=3.0, V0_scale=1.02)
    kw = dict(rms_dphi_threshold=1e-3, rms_res_threshold=1e-2, max_iter=200)
    o, sim = _pair(pf, pd, phi, V0, pf.PF_FP64, **kw)
    o.run()
    st = sim.minimize(pf.minimize_opts(dt=0.05, V0=V0, **kw))
    print(f"oracle converged={o.converged} at {o.iteration}; gpu converged={st['converged']} at {st['iteration']}",
          flush=True)
    assert o.converged and st["converged"] == 1
    assert st["iteration"] == o.iteration
    assert np.abs(sim.field(pf.PF_PHI) - o.phi).max() < 1e-9


@pytest.mark.gpu
def test_mixed_minimize_accuracy(pf):
    pd, phi, V0 = CASES[0][1:]
    o, sim = _pair(pf, pd, phi, V0, pf.PF_MIXED)
    for _ in range(6):
        o.iterate()
        sg = sim.minimize_iterate()
    dphi = np.abs(sim.field(pf.PF_PHI) - o.phi).max()
    print(f"mixed minimize: dphi={dphi:.2e} vol_err_rel={sg['vol_err_rel']:.2e}", flush=True)
    assert dphi < 1e-5
    assert math.isclose(sg["mean_h"], o.stats["mean_h"], rel_tol=1e-6)


@pytest.mark.gpu
def test_minimize_requires_begin(pf):
    pd, phi, V0 = CASES[1][1:]
    sim = pf.Simulation(pd, pf.PF_FP64)
    sim.set_fields(phi, np.zeros_like(phi))
    with pytest.raises(pf.PfError):
        sim.minimize_iterate()


@pytest.mark.gpu
def test_fp64_minimize_two_post_projections(pf):
    """Two post-projection sub-steps and a different relaxation: the second
    sub-step re-evaluates g_exp and the multiplier at the projected phi."""
    pd, phi, V0 = CASES[0][1:]
    o, sim = _pair(pf, pd, phi, V0, pf.PF_FP64, post_projection_iters=2, eta_lambda_vol=0.5)
    for it in range(3):
        so = o.iterate()
        sg = sim.minimize_iterate()
        dphi = np.abs(sim.field(pf.PF_PHI) - o.phi).max()
        print(f"pp2 iter {it + 1}: dphi={dphi:.2e} lambda {sg['lambda_vol']:.12e}/{so['lambda_vol']:.12e}", flush=True)
        assert dphi < FP64_TOL
        assert _close(sg["lambda_vol"], so["lambda_vol"], 1e-9, 1e-12)


# ---------------------------------------------------------------- full model
def _full_case(n=16, elastic=True, hetero=True):
    from cuda_sto_pf_amd.configs import config
    pd, phi, xB, meta = config(2, n)
    pd = dict(pd, zero_mode=0, elastic_enabled=1 if elastic else 0)
    if not hetero:
        pd["dC"] = np.zeros((6, 6))
    return pd, phi, xB


FULL_CASES = [("cfg2geom16_hetero",) + _full_case(), ("cfg2geom16_noelastic",) + _full_case(elastic=False)]


@pytest.mark.gpu
@pytest.mark.parametrize("name,pd,phi,xB", FULL_CASES, ids=[c[0] for c in FULL_CASES])
def test_fp64_full_model_minimize_parity(pf, name, pd, phi, xB):
    """--minimize-full-model: per iteration phi and Y to 1e-10, the multiplier,
    residuals, rms_dY, the excess energies and the elastic iteration counts."""
    from oracle.minimize_ref import MinimizeFullOracle
    opts = dict(dt=0.05, V0=0.0)
    o = MinimizeFullOracle(Params(**pd), phi, xB, MinimizeOptions(**opts))
    sim = pf.Simulation(dict(pd, dt=0.05), pf.PF_FP64)
    sim.set_fields(phi, xB, xB_cap=0.07)
    sim.minimize_begin(pf.minimize_opts(full_model=1, **opts))
    for it in range(6):
        so = o.iterate()
        sg = sim.minimize_iterate()
        dphi = np.abs(sim.field(pf.PF_PHI) - o.phi).max()
        dY = np.abs(sim.field(pf.PF_Y) - o.o.Y).max()
        print(f"{name} iter {it + 1}: dphi={dphi:.2e} dY={dY:.2e} F_tot {sg['F_total_excess_hat']:.12e}/"
              f"{so['F_total_excess']:.12e} rms_dY {sg['rms_dY']:.6e}/{so['rms_dY']:.6e} "
              f"el iters {sg['elastic_iterations']}/{so['elastic_iterations']}", flush=True)
        assert dphi < FP64_TOL and dY < FP64_TOL
        assert sg["elastic_iterations"] == so["elastic_iterations"]
        for k, kg, rel, ab in (("lambda_vol", "lambda_vol", 1e-9, 1e-12), ("lambda_kkt", "lambda_kkt", 1e-9, 1e-12),
                               ("rms_res", "rms_res", 1e-8, 1e-12), ("rms_dphi", "rms_dphi", 1e-6, 1e-9),
                               ("rms_dY", "rms_dY", 1e-7, 1e-10), ("mean_h", "mean_h", 1e-12, 1e-15),
                               ("F_surf", "F_surf_hat", 1e-10, 1e-15), ("F_el", "F_el_hat", 1e-10, 1e-15),
                               ("F_chem_excess", "F_chem_excess_hat", 1e-9, 1e-13),
                               ("F_total_excess", "F_total_excess_hat", 1e-9, 1e-13),
                               ("el_core_sum", "el_core_sum", 1e-10, 1e-13)):
            assert _close(sg[kg], so[k], rel, ab), (k, sg[kg], so[k])
        if it > 0:
            assert _close(sg["energy_diff_rel"], so["energy_diff_rel"], 1e-5, 1e-13)
        else:
            assert math.isnan(sg["energy_diff_rel"])


@pytest.mark.gpu
def test_full_model_minimize_refuses_the_zero_mode(pf):
    pd, phi, xB = _full_case()
    sim = pf.Simulation(dict(pd, zero_mode=1), pf.PF_FP64)
    sim.set_fields(phi, xB, xB_cap=0.07)
    with pytest.raises(pf.PfError, match="zero mode"):
        sim.minimize_begin(pf.minimize_opts(full_model=1, dt=0.05))
