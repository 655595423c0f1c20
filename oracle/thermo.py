"""Oracle (test infrastructure only): CPU FP64 restatement of the reference
thermodynamics and interpolation functions.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this package; the product path never does.

Each function follows the named reference source line by line in its
arithmetic, vectorised over numpy arrays (float64).

Pinning (tests/test_oracle_pins.py): the finite-difference Gamma against the
analytic derivative, x_eq against the tangent condition, h and g against
their closed-form values and the logit/sigmoid round trip; the
SGTE/CALPHAD Gibbs energies (which have no known-answer value in the
reference) against an independent restatement of the reference's
physical-unit converter (Unit_Psedobinary.py:441-485), on both sides of every
breakpoint (tests/test_units.py).
"""
import math

import numpy as np

R_GAS = 8.31446261815324          # thermo_utils.h:260
X_LIMIT_CONVEX = 0.09             # thermo_utils.h:296


# ---------------------------------------------------------------- phase_functions.h
def h_of_phi(phi):
    """h = phi^3 (6 phi^2 - 15 phi + 10)   (phase_functions.h:7-11)."""
    p2 = phi * phi
    p3 = p2 * phi
    return p3 * (6.0 * p2 - 15.0 * phi + 10.0)


def h_prime_of_phi(phi):
    """h' = 30 phi^2 (1-phi)^2   (phase_functions.h:14-17)."""
    p2 = phi * phi
    return 30.0 * p2 * (1.0 - phi) * (1.0 - phi)


def g_of_phi(phi):
    """g = phi^2 (1-phi)^2   (phase_functions.h:51-53)."""
    return phi * phi * (1.0 - phi) * (1.0 - phi)


def g_prime_of_phi(phi):
    """g' = 2 phi (1-phi)(1-2 phi)   (phase_functions.h:56-58)."""
    return 2.0 * phi * (1.0 - phi) * (1.0 - 2.0 * phi)


# ---------------------------------------------------------------- thermo_utils.h
def clamp01(x):
    return np.clip(x, 0.0, 1.0)


def clamp_fraction_eps(x):
    """thermo_utils.h:306-310."""
    return np.clip(x, 1e-12, 1.0 - 1e-12)


def GHSER_Pb(T):
    """thermo_utils.h:316-322 (scalar T)."""
    if T < 600.61:
        return (-7650.085 + 101.700244 * T - 24.5242231 * T * math.log(T)
                - 0.00365895 * T * T - 2.4395e-7 * T ** 3)
    return (-10531.095 + 154.243182 * T - 32.4913959 * T * math.log(T)
            + 0.00154613 * T * T + 8.05448e25 * T ** -9)


def GHSER_Ag(T):
    """thermo_utils.h:324-330."""
    if T < 1234.93:
        return (-7209.512 + 118.202013 * T - 23.8463314 * T * math.log(T)
                - 0.001790585 * T * T - 3.98587e-7 * T ** 3 - 12011.0 / T)
    return (-15095.252 + 190.266404 * T - 33.472 * T * math.log(T)
            + 1.411773e29 * T ** -9)


def GHSER_Te(T):
    """thermo_utils.h:332-338."""
    if T < 722.66:
        return (-10544.679 + 183.372894 * T - 35.6687 * T * math.log(T)
                + 0.01583435 * T * T - 5.240417e-6 * T ** 3 + 155015.0 / T)
    return (9160.595 - 129.265373 * T + 13.004 * T * math.log(T)
            - 0.0362361 * T * T + 5.006367e-6 * T ** 3 - 1.28681e30 * T ** -9)


def G_PbTe_Solid(T):
    """thermo_utils.h:341-344."""
    return (-76063.2138 + 9.67716633 * T) + GHSER_Pb(T) + GHSER_Te(T)


def G_Ag2Te_Solid(T):
    """thermo_utils.h:347-351."""
    base_per_atom = -10128.93 - 12.645115 * T
    G_atom = base_per_atom + (2.0 / 3.0) * GHSER_Ag(T) + (1.0 / 3.0) * GHSER_Te(T)
    return 3.0 * G_atom


def get_L_param(T):
    """thermo_utils.h:359-361."""
    return 41212.9 - 18.05 * T


def mu_PbTe_calphad(T, xB):
    """thermo_utils.h:364-369."""
    x = clamp_fraction_eps(xB)
    return G_PbTe_Solid(T) + R_GAS * T * np.log(1.0 - x) + get_L_param(T) * x * x


def mu_Ag2Te_calphad(T, xB):
    """thermo_utils.h:372-377."""
    x = clamp_fraction_eps(xB)
    return G_Ag2Te_Solid(T) + R_GAS * T * np.log(x) + get_L_param(T) * (1.0 - x) * (1.0 - x)


def dmu_PbTe_calphad_dx(T, xB):
    """thermo_utils.h:380-386."""
    x = clamp_fraction_eps(xB)
    return -R_GAS * T / (1.0 - x) + 2.0 * get_L_param(T) * x


def dmu_Ag2Te_calphad_dx(T, xB):
    """thermo_utils.h:389-395."""
    x = clamp_fraction_eps(xB)
    return R_GAS * T / x - 2.0 * get_L_param(T) * (1.0 - x)


def mu_PbTe_raw(T, xB, convex):
    """thermo_utils.h:400-414: CALPHAD, or the quadratic continuation above
    X_LIMIT_CONVEX when the convex extrapolation switch is on."""
    base = mu_PbTe_calphad(T, xB)
    if not convex:
        return base
    xc = X_LIMIT_CONVEX
    dx = xB - xc
    ext = mu_PbTe_calphad(T, xc) + dmu_PbTe_calphad_dx(T, xc) * dx + 5000.0 * dx * dx
    return np.where(xB <= xc, base, ext)


def mu_Ag2Te_raw(T, xB, convex):
    """thermo_utils.h:417-430."""
    base = mu_Ag2Te_calphad(T, xB)
    if not convex:
        return base
    xc = X_LIMIT_CONVEX
    dx = xB - xc
    ext = mu_Ag2Te_calphad(T, xc) + dmu_Ag2Te_calphad_dx(T, xc) * dx + 5000.0 * dx * dx
    return np.where(xB <= xc, base, ext)


def _scale(energy_scale):
    if abs(energy_scale) < 1e-30:
        return 1e-30 if energy_scale >= 0.0 else -1e-30
    return energy_scale


def mu_A_dimless(xB, T, energy_scale, convex=False):
    """thermo_utils.h:437-440."""
    return mu_PbTe_raw(T, xB, convex) / _scale(energy_scale)


def mu_B_dimless(xB, T, energy_scale, convex=False):
    """thermo_utils.h:443-446."""
    return mu_Ag2Te_raw(T, xB, convex) / _scale(energy_scale)


def solve_x_eq(T):
    """thermo_utils.h:453-481 (Newton on RT ln x + L (1-x)^2 = 0)."""
    L = get_L_param(T)
    RT = R_GAS * T
    x = math.exp(-L / RT)
    if x < 1e-9:
        x = 1e-9
    if x > 0.99:
        x = 0.5
    for _ in range(20):
        f = RT * math.log(x) + L * (1.0 - x) * (1.0 - x)
        df = RT / x - 2.0 * L * (1.0 - x)
        delta = f / df
        x -= delta
        if x < 1e-10:
            x = 1e-10
        if x > 0.999:
            x = 0.999
        if abs(delta) < 1e-8:
            break
    return x


def c_xB_phi(xB, Vm_alpha_0, dVm_alpha_dxB, Vm_compound, h):
    """thermo_utils.h:497-504."""
    Vm_alpha = Vm_alpha_0 + dVm_alpha_dxB * xB
    denom = Vm_alpha * (1.0 - h) + Vm_compound * h
    denom = np.where(np.abs(denom) < 1e-12, np.where(denom >= 0.0, 1e-12, -1e-12), denom)
    return 1.0 / denom


def mu_tot_mix(mu_A, mu_B, xB, mu0_compound, h):
    """thermo_utils.h:507-511."""
    mu_mix = mu_A * (1.0 - xB) + mu_B * xB
    return (1.0 - h) * mu_mix + h * mu0_compound


def v_weighted_mu_compound(muA_eq, muB_eq, v_A, v_B):
    """thermo_utils.h:514-519."""
    v_sum = v_A + v_B
    if abs(v_sum) < 1e-30:
        v_sum = 1e-30 if v_sum >= 0.0 else -1e-30
    return (v_A * muA_eq + v_B * muB_eq) / v_sum


def sigmoid_unclipped(Y):
    """Stable logistic used by the Y kernels (cuda_kernels.cu:2201-2208,
    pf_zero_mode_sigmoid cuda_kernels.cu:3391-3399)."""
    Y = np.asarray(Y, dtype=np.float64)
    pos = Y >= 0.0
    e = np.exp(np.where(pos, -Y, Y))
    return np.where(pos, 1.0 / (1.0 + e), e / (1.0 + e))


def sigmoid_from_logit(Y, Y_clip, eps):
    """thermo_utils.h:522-536: clip Y, stable logistic, clip x to [eps,1-eps]."""
    Yc = np.clip(Y, -Y_clip, Y_clip)
    return np.clip(sigmoid_unclipped(Yc), eps, 1.0 - eps)


def logit_from_fraction(xB, eps, Y_clip):
    """thermo_utils.h:538-545."""
    x = np.clip(xB, eps, 1.0 - eps)
    return np.clip(np.log(x / (1.0 - x)), -Y_clip, Y_clip)


def D_mix(h, D_alpha, D_compound):
    """thermo_utils.h:548-550."""
    return (1.0 - h) * D_alpha + h * D_compound


def gamma_thermo_nonlinear(xB, h, Vm_alpha_0, dVm_alpha_dxB, Vm_compound, T,
                           energy_scale, convex=False):
    """thermo_utils.h:553-577: central difference (step 1e-5) of mu_B - mu_A."""
    xB_calc = np.minimum(xB, 0.08) if convex else xB
    c_bulk = c_xB_phi(xB_calc, Vm_alpha_0, dVm_alpha_dxB, Vm_compound, h)
    dx = 1e-5
    xp = clamp_fraction_eps(xB_calc + dx)
    xm = clamp_fraction_eps(xB_calc - dx)
    dmuA = (mu_A_dimless(xp, T, energy_scale, convex)
            - mu_A_dimless(xm, T, energy_scale, convex)) / (xp - xm)
    dmuB = (mu_B_dimless(xp, T, energy_scale, convex)
            - mu_B_dimless(xm, T, energy_scale, convex)) / (xp - xm)
    return c_bulk * (dmuB - dmuA)


def stabilized_meff(Dm, gamma):
    """cuda_kernels.cu:30-38: D/max(Gamma,1e-4), capped at 1000 D."""
    gamma_safe = np.where(np.isfinite(gamma) & (gamma > 1.0e-4), gamma, 1.0e-4)
    D_safe = np.maximum(Dm, 0.0)
    cap = D_safe * 1000.0
    value = D_safe / gamma_safe
    return np.where(np.isfinite(value) & (value < cap), value, cap)
