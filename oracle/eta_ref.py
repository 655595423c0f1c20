"""Oracle (test infrastructure only): CPU FP64 restatement of one semi-implicit
update of the GP-zone order parameter eta, the kernels
``compute_eta_rhs_kernel`` (cuda_kernels.cu:845-946),
``eta_semi_implicit_update_kernel`` (cuda_kernels.cu:1045-1069) and
``eta_normalize_and_clamp_kernel`` (cuda_kernels.cu:1086-1098) in the order
of the step loop (main_cuda.cu:29948-30006).

Only ``tests/`` may import it.  No workload runs the GP-zone model, so this is
kernel-level parity: given eta, phi, x_B(alpha) and tr(sigma), one update.

Pinning (tests/test_oracle_eta.py): the chemical drive against a numerical
eta-derivative of c (1-h(phi)) h(eta) (g_GP - g_alpha); the elastic part
against a numerical eta-derivative of -tr(sigma) (1-h(phi)) h(eta)
(eps_GP - eps_alpha); a uniform field against its explicit k = 0 update;
a single Fourier mode against 1/(1 + L dt kappa k^2) decay.
"""
from dataclasses import dataclass

import numpy as np

from . import thermo as th
from .step_ref import KSpace, Params, fwd, inv


@dataclass
class EtaParams:
    """GP keys of params_default (main_cuda.cu:15793-15828, 16164-16167)."""
    gp_xB_fixed: float = 0.3529411764705882
    gp_delta_g0: float = 0.0
    gp_W_eta: float = 0.0
    gp_kappa_eta: float = 0.0
    gp_L_eta: float = 0.0
    gp_eps_iso: float = 0.0
    gp_elastic_derivative_scale: float = 1.0
    gp_elastic: int = 0                 # gp_elastic_solver_enabled && gp_active_eta_coupling
    gp_raw_reaction_drive_only: int = 0
    gp_raw_reaction_drive_use_raw_units_debug: int = 0
    gp_reaction_nu_A: float = 0.65
    gp_reaction_nu_B: float = 0.35
    gp_mu_reference_raw: float = 0.0


def clamp_eps(x, e):
    return np.clip(x, e, 1.0 - e)


def eta_rhs(p: Params, q: EtaParams, eta, phi, xB_alpha, tr_sigma):
    """compute_eta_rhs_kernel (cuda_kernels.cu:845-946)."""
    eta = th.clamp01(eta)
    phi = th.clamp01(phi)
    xa = clamp_eps(xB_alpha, 1e-12)
    h_phi = th.h_of_phi(phi)
    hp_eta = th.h_prime_of_phi(eta)
    c_ref = 1.0 / (p.Vm_alpha_0 + p.dVm_alpha_dxB * xa)
    cv = p.thermo_convex_extrapolation_enabled
    if q.gp_raw_reaction_drive_only:
        muA = th.mu_PbTe_calphad(p.T, xa)
        muB = th.mu_Ag2Te_calphad(p.T, xa)
        if q.gp_raw_reaction_drive_use_raw_units_debug:
            mdm = q.gp_reaction_nu_A * muA + q.gp_reaction_nu_B * muB - q.gp_mu_reference_raw
        else:
            s = p.mu_reference_scale if abs(p.mu_reference_scale) > 1e-300 else 1.0
            mdm = q.gp_reaction_nu_A * (muA / s) + q.gp_reaction_nu_B * (muB / s) - q.gp_mu_reference_raw / s
        chem = c_ref * (1.0 - h_phi) * hp_eta * (-mdm)
    else:
        muA = th.mu_A_dimless(xa, p.T, p.mu_reference_scale, cv)
        muB = th.mu_B_dimless(xa, p.T, p.mu_reference_scale, cv)
        g_alpha = (1.0 - xa) * muA + xa * muB
        xg = float(clamp_eps(q.gp_xB_fixed, 1e-12))
        muAg = th.mu_A_dimless(xg, p.T, p.mu_reference_scale, cv)
        muBg = th.mu_B_dimless(xg, p.T, p.mu_reference_scale, cv)
        g_gp = (1.0 - xg) * muAg + xg * muBg - q.gp_delta_g0
        chem = c_ref * (1.0 - h_phi) * hp_eta * (g_gp - g_alpha)
    el = 0.0
    if q.gp_elastic:
        # compute_dgel_deta_gp_point (cuda_kernels.cu:293-308)
        eps_alpha = th.clamp01(xa) * p.eps_iso_over_vB
        delta = (1.0 - h_phi) * hp_eta * (q.gp_eps_iso - eps_alpha)
        el = q.gp_elastic_derivative_scale * (-tr_sigma * delta)
    return chem + q.gp_W_eta * th.g_prime_of_phi(eta) + el


def eta_update(p: Params, q: EtaParams, eta, phi, xB_alpha, tr_sigma):
    """main_cuda.cu:29952-30006: dealiased transforms, semi-implicit update
    with L_eta, kappa_eta, then clamp to [0, 1]."""
    ks = KSpace(p)
    shape = (p.Nx, p.Ny, p.Nz)
    ek = fwd(np.asarray(eta, dtype=np.float64), ks)
    rk = fwd(eta_rhs(p, q, eta, phi, xB_alpha, tr_sigma), ks)
    Ldt = q.gp_L_eta * p.dt
    ek = (ek - Ldt * rk) / (1.0 + Ldt * q.gp_kappa_eta * ks.k2) * ks.mask
    return np.clip(inv(ek, shape), 0.0, 1.0)
