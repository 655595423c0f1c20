"""Pins of the full-model minimize oracle (oracle/minimize_ref.py,
MinimizeFullOracle; CPU only).

The iteration composes pieces pinned elsewhere: the phi right-hand side
(tests/test_oracle_variational.py), the composition update
(tests/test_oracle_pins.py), the multiplier step and volume projection
(tests/test_oracle_minimize.py).  These tests pin what the composition adds:
a uniform matrix at the baseline composition is an exact stationary point
with zero chemical excess; the minimize Y cap keeps x_B at or below
minimize_xB_max_safe; and the pseudo-time flow lowers the total excess free
energy F_chem_excess + F_surf + F_el it reports.

The constraints (main_cuda.cu:32297-32366, 32508-32522): the mass constraint
is a uniform shift of the iteration's clamped Y, taken at the final phi, that
puts sum (1-h) x_B + h v_B back on the initial sum and keeps Y under the
minimize cap; with freeze_phi, phi never changes and only Y relaxes."""
import math

import numpy as np

from cuda_sto_pf_amd.configs import base_params, config
from oracle import thermo as th
from oracle.minimize_ref import MinimizeFullOracle, MinimizeOptions
from oracle.step_ref import Params


def test_uniform_matrix_is_stationary_with_zero_excess():
    shape = (8, 8, 8)
    pd = base_params(shape, hetero=True, zero_mode=False, elastic=False)
    x0 = 0.021
    m = MinimizeFullOracle(Params(**pd), np.zeros(shape), np.full(shape, x0),
                           MinimizeOptions(dt=0.05, V0=0.0, xB_ref=0.0))
    assert m.x_ref == x0
    for _ in range(3):
        st = m.iterate()
        assert st["rms_dphi"] == 0.0 and st["rms_dY"] == 0.0
        assert abs(st["F_chem_excess"]) < 1e-15
        assert np.array_equal(m.o.phi, np.zeros(shape))
        assert np.abs(m.o.xB - x0).max() < 1e-15


def test_composition_stays_below_the_minimize_cap():
    pd, phi, xB, meta = config(2, 16)
    pd = dict(pd, zero_mode=0)
    cap = 0.0205   # just above the matrix composition: the flow presses against it
    opts = MinimizeOptions(dt=0.05, V0=0.0, xB_max_safe=cap)
    m = MinimizeFullOracle(Params(**pd), phi, xB, opts)
    assert m.o.xB.max() <= cap
    ycap = math.log(cap / (1.0 - cap))
    assert abs(m.y_cap - ycap) < 1e-15
    touched = 0
    for _ in range(4):
        m.iterate()
        assert m.o.Y.max() <= ycap and m.o.xB.max() <= cap * (1 + 1e-14)
        touched += int(np.count_nonzero(m.o.Y == ycap))
    assert touched > 0   # the cap was active


def test_pseudo_time_flow_lowers_the_total_excess_free_energy():
    pd, phi, xB, meta = config(2, 16)
    pd = dict(pd, zero_mode=0)
    m = MinimizeFullOracle(Params(**pd), phi, xB, MinimizeOptions(dt=0.05, V0=0.0))
    F = []
    for _ in range(8):
        F.append(m.iterate()["F_total_excess"])
    assert all(b < a for a, b in zip(F, F[1:])), F
    # the baseline is the far-field composition when given (ic_23d_xB_out > 0)
    m2 = MinimizeFullOracle(Params(**pd), phi, xB, MinimizeOptions(dt=0.05, xB_ref=0.03))
    g0 = (0.97 * float(th.mu_A_dimless(0.03, m2.o.T, pd["mu_reference_scale"]))
          + 0.03 * float(th.mu_B_dimless(0.03, m2.o.T, pd["mu_reference_scale"])))
    assert m2.x_ref == 0.03 and abs(m2.g_bulk0 - g0) < 1e-15


def _pair_runs(**kw):
    pd, phi, xB, meta = config(2, 16)
    pd = dict(pd, zero_mode=0)
    base = dict(dt=0.05, V0=0.0)
    free = MinimizeFullOracle(Params(**pd), phi, xB, MinimizeOptions(**base))
    held = MinimizeFullOracle(Params(**pd), phi, xB, MinimizeOptions(**base, mass_constraint=1, **kw))
    return free, held


def test_mass_constraint_is_a_uniform_shift_onto_the_initial_sum():
    free, held = _pair_runs()
    m0 = held.mass_target
    assert m0 == held.o.mass()
    free.iterate()
    sh = held.iterate()
    # the constraint runs last: phi is the unconstrained iteration's
    assert np.array_equal(held.o.phi, free.o.phi)
    shift = held.o.Y - free.o.Y
    lam = sh["mass_lambda"]
    assert abs(lam) > 1e-6                      # the free flow does move the sum
    assert np.abs(shift - lam).max() < 1e-13    # one constant shift of Y
    assert abs(free.o.mass() - m0) > 1e3 * 1e-12 * max(abs(m0), 1.0)
    # the sum at the final phi, recomputed independently of the search
    h = th.h_of_phi(held.o.phi)
    direct = math.fsum(((1.0 - h) * th.sigmoid_unclipped(held.o.Y) + h * held.p.v_B).ravel())
    assert abs(direct - m0) <= 1e-12 * max(abs(m0), 1.0)
    assert held.o.Y.max() <= held.y_cap + 1e-15
    # rms_dY counts the shift, the clamp's dY/dt does not
    n = held.o.Y.size
    assert abs(sh["rms_dY"] - math.sqrt(math.fsum(((free.o.dYdt + lam / 0.05) ** 2).ravel()) / n)) < 1e-12
    assert np.array_equal(held.o.dYdt, free.o.dYdt)
    for _ in range(3):
        st = held.iterate()
        assert st["mass_err_rel"] < 1e-12


def test_freeze_phi_relaxes_only_the_composition():
    pd, phi, xB, meta = config(2, 16)
    pd = dict(pd, zero_mode=0)
    m = MinimizeFullOracle(Params(**pd), phi, xB,
                           MinimizeOptions(dt=0.05, V0=0.0, mass_constraint=1, freeze_phi=1,
                                           post_projection_iters=0))
    phi0 = m.o.phi.copy()
    m0 = m.mass_target
    rms = []
    for _ in range(6):
        st = m.iterate()
        assert np.array_equal(m.o.phi, phi0)
        assert st["vol_err_rel"] == 0.0
        assert math.isnan(st["rms_res"]) and math.isnan(st["rms_dphi"])
        assert abs(m.o.mass() - m0) <= 1e-12 * max(abs(m0), 1.0)
        rms.append(st["rms_dY"])
    assert rms[-1] < 0.05 * rms[0]


def test_freeze_phi_contract():
    pd, phi, xB, meta = config(2, 16)
    pd = dict(pd, zero_mode=0)
    for kw in (dict(freeze_phi=1, post_projection_iters=0),
               dict(freeze_phi=1, mass_constraint=1),
               dict(freeze_phi=1, mass_constraint=1, post_projection_iters=0, V0=0.2)):
        try:
            MinimizeFullOracle(Params(**pd), phi, xB, MinimizeOptions(**dict(dict(dt=0.05, V0=0.0), **kw)))
        except ValueError:
            continue
        raise AssertionError(kw)


def _v5_case(scale=1.0):
    from cuda_sto_pf_amd.raw_init import component_h_sums, ownership_labels
    pd, phi, xB, meta = config(2, 16)
    pd = dict(pd, zero_mode=0)
    lab = ownership_labels(phi.shape, meta["centres"])
    K = len(meta["centres"])
    return pd, phi, xB, lab, K, component_h_sums(phi, lab, K) * scale


def test_component_constraints_hold_every_label_sum():
    """V5: each ownership region keeps its own h-volume (here 1% above the
    initial one); the single global constraint keeps only the total and lets
    the regions exchange volume."""
    from cuda_sto_pf_amd.raw_init import component_h_sums
    pd, phi, xB, lab, K, T = _v5_case(1.01)
    v5 = MinimizeFullOracle(Params(**pd), phi, xB,
                            MinimizeOptions(dt=0.05, V0=0.0, mass_constraint=1, post_projection_iters=0,
                                            component_labels=lab, component_targets=T))
    glob = MinimizeFullOracle(Params(**pd), phi, xB,
                              MinimizeOptions(dt=0.05, V0=0.0, mass_constraint=1, post_projection_iters=0))
    for _ in range(4):
        st = v5.iterate()
        glob.iterate()
        sums = component_h_sums(v5.o.phi, lab, K)
        assert np.abs(sums - T).max() <= 1e-12 * T.max()
        assert st["vol_err_rel"] <= 1e-13
        assert math.isnan(st["lambda_kkt"]) and math.isfinite(st["rms_res"])
    drift = component_h_sums(glob.o.phi, lab, K) - component_h_sums(phi, lab, K)
    # the global first-order projection holds the total to its 1e-4 tolerance,
    # while single regions move by far more
    assert abs(drift.sum()) < 1e-4 * T.sum() and np.abs(drift).max() > 20 * abs(drift.sum())


def test_component_multipliers_reduce_to_the_global_one():
    """Two labels over mirror-image halves of a symmetric field: each label's
    multiplier equals the global -sum h'g / sum h'^2 (pinned for the
    phi-only minimize), and the per-label projection matches the global one."""
    pd, phi, xB, lab, K, T = _v5_case()
    half = phi[:8]
    sym = np.concatenate([half, half[::-1]], axis=0)
    xs = np.concatenate([xB[:8], xB[:8][::-1]], axis=0)
    lab2 = np.zeros(sym.shape, dtype=np.int32)
    lab2[8:] = 1
    from cuda_sto_pf_amd.raw_init import component_h_sums
    T2 = component_h_sums(sym, lab2, 2)
    m = MinimizeFullOracle(Params(**pd), sym, xs,
                           MinimizeOptions(dt=0.05, V0=0.0, mass_constraint=1, post_projection_iters=0,
                                           component_labels=lab2, component_targets=T2))
    g = np.random.default_rng(4).normal(size=sym.shape)
    g = g + g[::-1]
    lam = m.comp_multipliers(sym, g)
    glob = m.multiplier(sym, g)
    assert abs(lam[0] - glob) < 1e-12 * abs(glob) and abs(lam[1] - glob) < 1e-12 * abs(glob)


def test_component_contract():
    pd, phi, xB, lab, K, T = _v5_case()
    base = dict(dt=0.05, V0=0.0, mass_constraint=1, post_projection_iters=0, component_labels=lab,
                component_targets=T)
    bad_lab = lab.copy()
    bad_lab[lab == 3] = 2          # label 3 owns nothing
    for kw in (dict(mass_constraint=0), dict(post_projection_iters=1), dict(V0=0.1), dict(component_labels=bad_lab),
               dict(component_targets=np.where(np.arange(K) == 0, -1.0, T)), dict(component_targets=T[:1]),
               dict(freeze_phi=1)):
        try:
            MinimizeFullOracle(Params(**pd), phi, xB, MinimizeOptions(**dict(base, **kw)))
        except ValueError:
            continue
        raise AssertionError(kw)
