"""Pins of the minimize-mode oracle against what the mathematics fixes (CPU
only): the phi RHS is the phi-derivative of the energy density at frozen total
strain; the multiplier makes the constrained direction orthogonal to h'; a
uniform field is stationary; the volume projection is first-order exact; the
pseudo-time flow lowers the free energy."""
import math

import numpy as np

from cuda_sto_pf_amd.configs import base_params, spheres_phi
from oracle import thermo as th
from oracle.minimize_ref import MinimizeOptions, MinimizeOracle
from oracle.step_ref import Params, voigt_apply_field


def _case(n=16, elastic=True, hetero=True, radius=4.0):
    pd = base_params(n, hetero=hetero, elastic=elastic, zero_mode=False)
    p = Params(**pd)
    phi = spheres_phi((n, n, n), [(n / 2, n / 2, n / 2)], [radius], width=1.5)
    return p, phi


def _gel(p, eps, phi):
    """0.5 (eps - h eps00) : (C + h dC) : (eps - h eps00), an independent
    statement of the elastic energy density at fixed total strain."""
    h = th.h_of_phi(phi)
    el = [eps[i] - h * p.eps00[i] for i in range(6)]
    s = voigt_apply_field(p.C, p.dC, h, el)
    return 0.5 * (s[0] * el[0] + s[1] * el[1] + s[2] * el[2] + 2.0 * (s[3] * el[3] + s[4] * el[4] + s[5] * el[5]))


def test_phi_rhs_is_derivative_of_energy_density():
    p, phi = _case()
    m = MinimizeOracle(p, phi, MinimizeOptions())
    m.o.elastic_solve()
    phi0 = m.o.phi
    g = m.g_explicit(phi0) - p.W * th.g_prime_of_phi(phi0)
    d = 1e-6
    fd = (_gel(p, m.o.eps, phi0 + d) - _gel(p, m.o.eps, phi0 - d)) / (2 * d)
    scale = np.abs(fd).max()
    assert scale > 1e-3
    assert np.abs(g - fd).max() <= 1e-7 * scale


def test_multiplier_orthogonality_and_sign():
    p, phi = _case()
    m = MinimizeOracle(p, phi, MinimizeOptions())
    m.o.elastic_solve()
    g = m.g_explicit(m.o.phi) - p.kappa_phi * m.laplacian(np.fft.rfftn(m.o.phi))
    lam = m.multiplier(m.o.phi, g)
    hp = th.h_prime_of_phi(m.o.phi)
    r = hp * (g + lam * hp)
    assert abs(math.fsum(r.ravel())) <= 1e-12 * math.fsum(np.abs(hp * g).ravel())


def test_uniform_field_is_stationary():
    """phi = c: g_full = W g'(c), lambda = -W g'(c)/h'(c); with the multiplier
    unrelaxed (eta = 1) the constrained RHS vanishes and phi stays at c; the
    KKT residual is zero."""
    n = 8
    p = Params(**base_params(n, elastic=False, zero_mode=False))
    c = 0.3
    m = MinimizeOracle(p, np.full((n, n, n), c), MinimizeOptions(post_projection_iters=1, V0=0.0, eta_lambda_vol=1.0))
    st = m.iterate()
    lam = -p.W * th.g_prime_of_phi(c) / th.h_prime_of_phi(c)
    assert math.isclose(st["lambda_raw"], lam, rel_tol=1e-12)
    assert np.abs(m.phi - c).max() <= 1e-13
    assert st["rms_res"] <= 1e-13 and st["rms_dphi"] <= 1e-11
    assert st["vol_err_rel"] <= 1e-13


def test_volume_projection_first_order():
    """After phi <- phi + lc h'(phi), <h> misses V0 only at second order in
    the correction: halving the mismatch quarters the remaining error."""
    p, phi = _case(elastic=False)
    base = th.h_of_phi(th.clamp01(phi)).mean()
    errs = []
    for rel in (2e-2, 1e-2):
        V0 = base * (1 + rel)
        hp = th.h_prime_of_phi(phi)
        lc = (V0 - base) / (hp * hp).mean()
        after = th.h_of_phi(th.clamp01(phi + lc * hp)).mean()
        errs.append(abs(after - V0) / (V0 - base))
    assert errs[0] < 0.05 and errs[1] < 0.05
    assert 1.6 < errs[0] / errs[1] < 2.4


def test_pseudo_time_flow_lowers_free_energy():
    p, phi = _case(elastic=False, radius=5.0)
    m = MinimizeOracle(p, phi, MinimizeOptions(dt=0.02, V0=0.0))
    F = [m.iterate()["F_surf"] for _ in range(12)]
    assert all(F[i + 1] <= F[i] + 1e-12 for i in range(1, len(F) - 1))
    assert F[-1] < F[1]


def test_elastic_minimize_runs_and_holds_volume():
    p, phi = _case(n=16)
    V0 = 1.05 * th.h_of_phi(th.clamp01(phi)).mean()
    m = MinimizeOracle(p, phi, MinimizeOptions(dt=0.05, V0=V0))
    for _ in range(4):
        st = m.iterate()
        assert st["vol_err_rel"] < 1e-3
        assert st["elastic_iterations"] >= 2
