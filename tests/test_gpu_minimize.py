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


# ------------------------------------------------- constraints of the full model
MASS_KEYS = (("mass_err_rel", "mass_err_rel", 0.0, 1e-12), ("mass_lambda", "mass_lambda", 1e-8, 1e-13),
             ("mass_current", "mass_current", 1e-14, 1e-12), ("mass_target", "mass_target", 1e-15, 1e-12))


@pytest.mark.gpu
@pytest.mark.parametrize("name,pd,phi,xB", FULL_CASES, ids=[c[0] for c in FULL_CASES])
def test_fp64_full_model_mass_constraint_parity(pf, name, pd, phi, xB):
    """--minimize-mass-constraint: after the phi projections the zero-mode
    search shifts the clamped Y onto the initial sum of x_Btot, Y bounded by
    the minimize cap; phi and Y per iteration to 1e-10, the shift, rms_dY with
    the shift included, and the conserved sum."""
    from oracle.minimize_ref import MinimizeFullOracle
    opts = dict(dt=0.05, V0=0.0, mass_constraint=1)
    o = MinimizeFullOracle(Params(**pd), phi, xB, MinimizeOptions(**opts))
    sim = pf.Simulation(dict(pd, dt=0.05), pf.PF_FP64)
    sim.set_fields(phi, xB, xB_cap=0.07)
    sim.minimize_begin(pf.minimize_opts(full_model=1, **opts))
    ycap = math.log(0.07 / 0.93)
    for it in range(5):
        so = o.iterate()
        sg = sim.minimize_iterate()
        Yg = sim.field(pf.PF_Y)
        dphi = np.abs(sim.field(pf.PF_PHI) - o.phi).max()
        dY = np.abs(Yg - o.o.Y).max()
        print(f"{name} mass iter {it + 1}: dphi={dphi:.2e} dY={dY:.2e} lambda {sg['mass_lambda']:.12e}/"
              f"{so['mass_lambda']:.12e} err {sg['mass_err_rel']:.2e} steps {sg['mass_iterations']} "
              f"(newton {sg['mass_newton_steps']} bisection {sg['mass_bisection_steps']})", flush=True)
        assert dphi < FP64_TOL and dY < FP64_TOL
        assert Yg.max() <= ycap + 1e-12
        for k, kg, rel, ab in MASS_KEYS + (("rms_dY", "rms_dY", 1e-7, 1e-10), ("lambda_vol", "lambda_vol", 1e-9, 1e-12),
                                           ("F_total_excess", "F_total_excess_hat", 1e-9, 1e-13)):
            assert _close(sg[kg], so[k], rel, ab), (k, sg[kg], so[k])
        assert sg["mass_err_rel"] < 1e-12
        assert sg["mass_newton_steps"] + sg["mass_bisection_steps"] == sg["mass_iterations"]
        assert sg["mass_lambda_lower"] <= sg["mass_lambda"] <= sg["mass_lambda_upper"]
        assert sg["mass_Ystar_max"] <= ycap + 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name,pd,phi,xB", FULL_CASES, ids=[c[0] for c in FULL_CASES])
def test_fp64_freeze_phi_parity(pf, name, pd, phi, xB):
    """--minimize-freeze-phi (V4): phi stays byte-identical, the composition
    relaxes under the mass constraint; Y per iteration to 1e-10, the fixed-phi
    stopping quantities, NaN phi residuals."""
    from oracle.minimize_ref import MinimizeFullOracle
    opts = dict(dt=0.05, V0=0.0, mass_constraint=1, freeze_phi=1, post_projection_iters=0)
    o = MinimizeFullOracle(Params(**pd), phi, xB, MinimizeOptions(**opts))
    sim = pf.Simulation(dict(pd, dt=0.05), pf.PF_FP64)
    sim.set_fields(phi, xB, xB_cap=0.07)
    phi0 = sim.field(pf.PF_PHI)
    sim.minimize_begin(pf.minimize_opts(full_model=1, **opts))
    for it in range(5):
        so = o.iterate()
        sg = sim.minimize_iterate()
        dY = np.abs(sim.field(pf.PF_Y) - o.o.Y).max()
        print(f"{name} frozen iter {it + 1}: dY={dY:.2e} rms_dY {sg['rms_dY']:.6e}/{so['rms_dY']:.6e} "
              f"lambda {sg['mass_lambda']:.6e}", flush=True)
        assert sim.field(pf.PF_PHI).tobytes() == phi0.tobytes()
        assert dY < FP64_TOL
        for k, kg, rel, ab in MASS_KEYS + (("rms_dY", "rms_dY", 1e-7, 1e-10),
                                           ("F_total_excess", "F_total_excess_hat", 1e-9, 1e-13)):
            assert _close(sg[kg], so[k], rel, ab), (k, sg[kg], so[k])
        assert sg["vol_err_rel"] == 0.0 and sg["V0"] == sg["mean_h"]
        assert math.isnan(sg["rms_res"]) and math.isnan(sg["rms_dphi"]) and math.isnan(sg["lambda_kkt"])
        assert sg["convergence_count"] == 0 or (sg["rms_dY"] < 5 * 5e-5)


@pytest.mark.gpu
def test_full_model_constraint_contracts(pf):
    """main_cuda.cu:24562-24623, 28130-28146: freeze_phi needs the mass
    constraint, no post-projection and V0 <= 0; the constraint needs the full
    model and a positive tolerance; an explicit target must match the fields."""
    pd, phi, xB = _full_case()
    sim = pf.Simulation(dict(pd, dt=0.05), pf.PF_FP64)
    sim.set_fields(phi, xB, xB_cap=0.07)
    bad = [dict(full_model=1, freeze_phi=1, post_projection_iters=0),
           dict(full_model=1, mass_constraint=1, freeze_phi=1, post_projection_iters=1),
           dict(full_model=1, mass_constraint=1, freeze_phi=1, post_projection_iters=0, V0=0.1),
           dict(full_model=0, mass_constraint=1),
           dict(full_model=1, mass_constraint=1, mass_tolerance_relative=0.0),
           dict(full_model=1, mass_constraint=1, mass_max_iterations=0),
           dict(full_model=1, mass_constraint=1, target_mass_code=1.0)]
    for kw in bad:
        with pytest.raises(pf.PfError):
            sim.minimize_begin(pf.minimize_opts(dt=0.05, **kw))
    m = sim.scalars()["mass_xBtot"]
    sim.minimize_begin(pf.minimize_opts(dt=0.05, full_model=1, mass_constraint=1, target_mass_code=m))
    st = sim.minimize_iterate()
    assert st["mass_target"] == m and st["mass_err_rel"] < 1e-12


def _v5_inputs(scale=1.01):
    from cuda_sto_pf_amd.configs import config
    from cuda_sto_pf_amd.raw_init import component_h_sums, ownership_labels
    pd, phi, xB = _full_case()
    meta = config(2, 16)[3]
    lab = ownership_labels(phi.shape, meta["centres"])
    K = len(meta["centres"])
    return pd, phi, xB, lab, K, component_h_sums(phi, lab, K) * scale


@pytest.mark.gpu
def test_fp64_component_constraints_parity(pf):
    """--minimize-component-volume-constraints (V5): one multiplier per
    ownership label, the per-label KKT residual and the per-label projection
    onto targets 1% above the initial h-volumes; phi and Y per iteration to
    1e-10, the residuals, the largest label error and the conserved mass."""
    import ctypes as C
    from oracle.minimize_ref import MinimizeFullOracle
    pd, phi, xB, lab, K, T = _v5_inputs()
    opts = dict(dt=0.05, V0=0.0, mass_constraint=1, post_projection_iters=0)
    o = MinimizeFullOracle(Params(**pd), phi, xB, MinimizeOptions(**opts, component_labels=lab,
                                                                  component_targets=T))
    sim = pf.Simulation(dict(pd, dt=0.05), pf.PF_FP64)
    sim.set_fields(phi, xB, xB_cap=0.07)
    lab_c = np.ascontiguousarray(lab, dtype=np.int32)
    T_c = np.ascontiguousarray(T, dtype=np.float64)
    mo = pf.minimize_opts(full_model=1, component_count=K,
                          component_labels=lab_c.ctypes.data_as(C.POINTER(C.c_int)),
                          component_target_h_sums=T_c.ctypes.data_as(C.POINTER(C.c_double)), **opts)
    sim.minimize_begin(mo)
    for it in range(4):
        so = o.iterate()
        sg = sim.minimize_iterate()
        dphi = np.abs(sim.field(pf.PF_PHI) - o.phi).max()
        dY = np.abs(sim.field(pf.PF_Y) - o.o.Y).max()
        print(f"v5 iter {it + 1}: dphi={dphi:.2e} dY={dY:.2e} vol_err {sg['vol_err_rel']:.2e}/{so['vol_err_rel']:.2e} "
              f"rms_res {sg['rms_res']:.6e}/{so['rms_res']:.6e}", flush=True)
        assert dphi < FP64_TOL and dY < FP64_TOL
        for k, kg, rel, ab in (("rms_res", "rms_res", 1e-8, 1e-12), ("rms_dphi", "rms_dphi", 1e-6, 1e-9),
                               ("rms_dY", "rms_dY", 1e-7, 1e-10), ("mean_h", "mean_h", 1e-12, 1e-15),
                               ("vol_err_rel", "vol_err_rel", 0.0, 1e-12), ("mass_err_rel", "mass_err_rel", 0.0, 1e-12),
                               ("F_total_excess", "F_total_excess_hat", 1e-9, 1e-13)):
            assert _close(sg[kg], so[k], rel, ab), (k, sg[kg], so[k])
        assert math.isnan(sg["lambda_kkt"]) and math.isnan(sg["lambda_raw"])


@pytest.mark.gpu
def test_component_constraint_contracts(pf):
    """Labels out of range, a label owning no cell, and V5 without the mass
    constraint are refused at pf_minimize_begin (main_cuda.cu:2805-2833,
    24638-24670, 27276-27290)."""
    import ctypes as C
    pd, phi, xB, lab, K, T = _v5_inputs(1.0)
    sim = pf.Simulation(dict(pd, dt=0.05), pf.PF_FP64)
    sim.set_fields(phi, xB, xB_cap=0.07)
    T_c = np.ascontiguousarray(T, dtype=np.float64)
    empty = lab.copy()
    empty[lab == K - 1] = 0
    outside = lab.copy()
    outside[0, 0, 0] = K
    for labels, extra in ((empty, {}), (outside, {}), (lab, dict(mass_constraint=0)),
                          (lab, dict(post_projection_iters=1))):
        l_c = np.ascontiguousarray(labels, dtype=np.int32)
        kw = dict(dict(dt=0.05, full_model=1, mass_constraint=1, post_projection_iters=0), **extra)
        with pytest.raises(pf.PfError):
            sim.minimize_begin(pf.minimize_opts(component_count=K,
                                                component_labels=l_c.ctypes.data_as(C.POINTER(C.c_int)),
                                                component_target_h_sums=T_c.ctypes.data_as(C.POINTER(C.c_double)),
                                                **kw))
