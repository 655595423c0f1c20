"""Pins of the eta-update oracle (CPU only): each term of the eta RHS is the
eta-derivative of the energy density it comes from; a uniform field follows
the explicit k = 0 update; a single Fourier mode decays as
1/(1 + L dt kappa k^2)."""
import numpy as np

from cuda_sto_pf_amd.configs import base_params
from oracle import thermo as th
from oracle.eta_ref import EtaParams, eta_rhs, eta_update
from oracle.step_ref import Params


def _p(n=8, **kw):
    return Params(**base_params(n, elastic=False, zero_mode=False))


def _fields(n=8, seed=3):
    rng = np.random.default_rng(seed)
    eta = rng.uniform(0.05, 0.95, (n, n, n))
    phi = rng.uniform(0.0, 0.6, (n, n, n))
    xa = rng.uniform(0.005, 0.08, (n, n, n))
    trs = rng.normal(0.0, 0.5, (n, n, n))
    return eta, phi, xa, trs


def test_chemical_drive_is_eta_derivative():
    p = _p()
    q = EtaParams(gp_delta_g0=0.01)
    eta, phi, xa, trs = _fields()
    cv = p.thermo_convex_extrapolation_enabled
    c = 1.0 / (p.Vm_alpha_0 + p.dVm_alpha_dxB * xa)
    ga = (1 - xa) * th.mu_A_dimless(xa, p.T, p.mu_reference_scale, cv) + xa * th.mu_B_dimless(
        xa, p.T, p.mu_reference_scale, cv)
    xg = q.gp_xB_fixed
    gg = (1 - xg) * th.mu_A_dimless(xg, p.T, p.mu_reference_scale, cv) + xg * th.mu_B_dimless(
        xg, p.T, p.mu_reference_scale, cv) - q.gp_delta_g0

    def f(e):
        return c * (1 - th.h_of_phi(phi)) * th.h_of_phi(e) * (gg - ga)
    d = 1e-6
    fd = (f(eta + d) - f(eta - d)) / (2 * d)
    got = eta_rhs(p, q, eta, phi, xa, trs)
    assert np.abs(got - fd).max() <= 1e-7 * np.abs(fd).max()


def test_raw_reaction_drive_units():
    """The dimensionless and raw-unit drives differ exactly by the energy scale."""
    p = _p()
    eta, phi, xa, trs = _fields()
    q1 = EtaParams(gp_raw_reaction_drive_only=1, gp_mu_reference_raw=-3.0e4)
    q2 = EtaParams(gp_raw_reaction_drive_only=1, gp_raw_reaction_drive_use_raw_units_debug=1,
                   gp_mu_reference_raw=-3.0e4)
    r1 = eta_rhs(p, q1, eta, phi, xa, trs)
    r2 = eta_rhs(p, q2, eta, phi, xa, trs)
    assert np.allclose(r2, r1 * p.mu_reference_scale, rtol=1e-12, atol=0)


def test_elastic_and_double_well_are_eta_derivatives():
    p = _p()
    q = EtaParams(gp_W_eta=0.7, gp_eps_iso=0.01, gp_elastic=1, gp_elastic_derivative_scale=0.5,
                  gp_raw_reaction_drive_only=1, gp_reaction_nu_A=0.0, gp_reaction_nu_B=0.0)
    eta, phi, xa, trs = _fields()

    def f(e):
        dw = q.gp_W_eta * th.g_of_phi(e)
        el = -trs * (1 - th.h_of_phi(phi)) * th.h_of_phi(e) * (q.gp_eps_iso - xa * p.eps_iso_over_vB)
        return dw + q.gp_elastic_derivative_scale * el
    d = 1e-6
    fd = (f(eta + d) - f(eta - d)) / (2 * d)
    got = eta_rhs(p, q, eta, phi, xa, trs)
    assert np.abs(got - fd).max() <= 1e-7 * np.abs(fd).max()


def test_uniform_field_explicit_update():
    p = _p()
    q = EtaParams(gp_W_eta=1.3, gp_kappa_eta=2.0, gp_L_eta=0.8, gp_delta_g0=0.0)
    n = 8
    c = 0.35
    eta = np.full((n, n, n), c)
    phi = np.full((n, n, n), 0.2)
    xa = np.full((n, n, n), 0.02)
    r = eta_rhs(p, q, eta, phi, xa, np.zeros((n, n, n)))
    out = eta_update(p, q, eta, phi, xa, np.zeros((n, n, n)))
    assert np.abs(out - np.clip(c - q.gp_L_eta * p.dt * r, 0, 1)).max() <= 1e-14


def test_single_mode_decay():
    """phi = 1 removes the chemical and elastic drive ((1-h) = 0); W = 0: the
    mode decays by exactly 1/(1 + L dt kappa k^2)."""
    n = 16
    p = _p(n)
    q = EtaParams(gp_W_eta=0.0, gp_kappa_eta=2.0, gp_L_eta=1.0)
    x = np.arange(n)
    k = 2 * np.pi * 2 / n
    eta = 0.5 + 0.1 * np.cos(k * x)[:, None, None] * np.ones((n, n, n))
    out = eta_update(p, q, eta, np.ones((n, n, n)), np.full((n, n, n), 0.02), np.zeros((n, n, n)))
    expect = 0.5 + 0.1 / (1 + q.gp_L_eta * p.dt * q.gp_kappa_eta * k * k) * np.cos(k * x)[:, None, None]
    assert np.abs(out - expect).max() <= 1e-13
