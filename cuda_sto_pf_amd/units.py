"""Physical inputs -> code-unit ``.params`` (the reference's parameter front end).

Restates ``Unit_Psedobinary.py``: ``PFParamConverter.convert`` (:601-712),
``build_main_cuda_overrides`` (:741-818), ``load_physical_inputs`` (:820-831)
and ``write_main_cuda_override_file`` (:833-843), with the thermodynamic,
kinetic and interface helpers they call (:441-586).  The arithmetic follows
the reference expression by expression (same operations in the same order,
plain Python floats), because the output is compared value for value with
files the reference wrote; the committed qualification parameters
(tests/golden/warm_start_residual.params) are reproduced bit for bit.

Command line::

    python -m cuda_sto_pf_amd.units --input physical.json --out run.params

Keys absent from the JSON take the reference's built-in inputs
(``USER_PHYSICAL_INPUTS``, Unit_Psedobinary.py:380-437); unknown keys are
rejected; keys starting with ``_`` are comments.
"""
import argparse
import json
import math
from dataclasses import asdict, dataclass, field, replace
from typing import List, Optional

R_GAS = 8.31446261815324          # Unit_Psedobinary.py:439
SCHEMA_VERSION = 1                # :26
# numerical elastic-solver controls the converter emits (:28-38)
ELASTIC_SOLVER_DEFAULTS = {
    "elastic_warm_start_enabled": 1,
    "elastic_residual_control_enabled": 1,
    "elastic_iter_min": 2,
    "elastic_iter_max": 32,
    "elastic_residual_tolerance": 1.0e-6,
    "elastic_residual_absolute_floor": 1.0e-30,
    "elastic_fail_on_nonconvergence": 1,
}

_I3 = [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]


@dataclass
class PhysicalInputs:
    """SI inputs (PhysicalInputs, Unit_Psedobinary.py:230-282) with the
    reference's built-in values (USER_PHYSICAL_INPUTS) as defaults."""
    gamma: float = 0.05                 # J/m^2
    lambda_sm: float = 6.0e-10          # m
    dx: float = 1.0e-10                 # m, fallback of pf_dx and phys_dx_ref
    dt: float = 1.0e-2                  # code units
    v_A: float = 0.0
    v_B: float = 1.0
    temperature_C: float = 380.0
    Vm_compound: float = 4.1009e-5      # m^3/mol
    Vm_alpha_0: float = 4.1009e-5
    pf_dx: Optional[float] = 1.0e-10    # m: the PF grid spacing
    phys_dx_ref: Optional[float] = 1.0e-9   # m: the calibration length
    D_ratio: float = 0.01
    gp_gamma_alpha_gp: float = 0.05
    gp_l_eta: float = 1.0e-9
    gp_D_ratio_eta: float = 0.01
    gp_reaction_nu_A: float = 0.65
    gp_reaction_nu_B: float = 0.35
    gp_M_eta_ratio_to_crit: Optional[float] = 1.0e-2
    gp_M_eta_phys: Optional[float] = None
    gp_xB_eq_alpha_for_eta: Optional[float] = 0.03
    vf_init: float = 0.0
    vf_target: float = 0.04
    L_ref_factor: float = 5.0
    gel_shift_Jm3: float = 0.0
    eps_iso: float = 0.00233
    eigenstrain_tensor: Optional[List[List[float]]] = None
    eigenstrain_principal: Optional[List[float]] = field(default_factory=lambda: [0.046, -0.022, -0.017])
    eigenstrain_rotation_matrix: Optional[List[List[float]]] = field(default_factory=lambda: [r[:] for r in _I3])
    C_tensor_PbTe_GPa: Optional[List[List[float]]] = field(default_factory=lambda: [
        [108.0, 6.0, 6.0, 0.0, 0.0, 0.0],
        [6.0, 71.0, 43.0, -0.0, 0.0, 0.0],
        [6.0, 43.0, 71.0, 0.0, 0.0, 0.0],
        [0.0, -0.0, 0.0, 51.0, 0.0, 0.0],
        [0.0, 0.0, 0.0, 0.0, 14.0, 0.0],
        [0.0, 0.0, 0.0, 0.0, 0.0, 14.0]])
    C_tensor_Ag2Te_GPa: Optional[List[List[float]]] = field(default_factory=lambda: [
        [64.53932315, 42.71742108, 40.48003361, 2.17554086, -0.06871424, 1.43079278],
        [42.71742108, 90.00000000, 46.28257892, 2.50650732, 4.13187759, 1.64846021],
        [40.48003361, 46.28257892, 68.50060964, 2.83747377, 4.65968934, 1.86612765],
        [2.17554086, 2.50650732, 2.83747377, 11.50967838, 0.76715418, -2.29548755],
        [-0.06871424, 4.13187759, 4.65968934, 0.76715418, 17.48003361, 0.50453599],
        [1.43079278, 1.64846021, 1.86612765, -2.29548755, 0.50453599, 13.49032162]])
    E0_xx: float = 0.0
    E0_yy: float = 0.0
    E0_zz: float = 0.0
    E0_yz: float = 0.0
    E0_xz: float = 0.0
    E0_xy: float = 0.0

    @property
    def grid_dx(self):        # resolve_pf_dx (:181-184)
        return self.pf_dx if self.pf_dx is not None else self.dx

    @property
    def ref_dx(self):         # resolve_phys_dx_ref (:187-190)
        return self.phys_dx_ref if self.phys_dx_ref is not None else self.dx


def load_physical_inputs(path=None, **overrides):
    """load_physical_inputs (:820-831): defaults updated by a JSON file and
    keyword overrides; unknown keys raise KeyError."""
    d = asdict(PhysicalInputs())
    given = {}
    if path:
        with open(path, encoding="utf-8") as f:
            given = {k: v for k, v in json.load(f).items() if not k.startswith("_")}
    given.update(overrides)
    unknown = sorted(set(given) - set(d))
    if unknown:
        raise KeyError("unknown input field(s): " + ", ".join(unknown))
    d.update(given)
    return PhysicalInputs(**d)


# ---------------------------------------------------------------- thermodynamics
# SGTE unary Gibbs energies and the pseudo-binary model (:441-485)
def _ghser_pb(T):
    if T < 600.61:
        return -7650.085 + 101.700244 * T - 24.5242231 * T * math.log(T) - 0.00365895 * T ** 2 - 2.4395e-7 * T ** 3
    return -10531.095 + 154.243182 * T - 32.4913959 * T * math.log(T) + 0.00154613 * T ** 2 + 8.05448e25 * T ** (-9)


def _ghser_ag(T):
    if T < 1234.93:
        return (-7209.512 + 118.202013 * T - 23.8463314 * T * math.log(T) - 0.001790585 * T ** 2
                - 3.98587e-7 * T ** 3 - 12011 * T ** (-1))
    return -15095.252 + 190.266404 * T - 33.472 * T * math.log(T) + 1.411773e29 * T ** (-9)


def _ghser_te(T):
    if T < 722.66:
        return (-10544.679 + 183.372894 * T - 35.6687 * T * math.log(T) + 0.01583435 * T ** 2
                - 5.240417e-6 * T ** 3 + 155015 * T ** (-1))
    return (9160.595 - 129.265373 * T + 13.004 * T * math.log(T) - 0.0362361 * T ** 2 + 5.006367e-6 * T ** 3
            - 1.28681e30 * T ** (-9))


def _g_pbte(T):
    return -76063.2138 + 9.67716633 * T + _ghser_pb(T) + _ghser_te(T)


def _g_ag2te(T):
    per_atom = -10128.93 - 12.645115 * T
    return 3.0 * (per_atom + (2.0 / 3.0) * _ghser_ag(T) + (1.0 / 3.0) * _ghser_te(T))


def _l0(T):
    return 41212.9 - 18.05 * T


def _check_x(x):
    if not (0.0 < x < 1.0):
        raise ValueError("x (x_Ag2Te) must lie in (0, 1)")


def mu_ag2te(T, x):
    _check_x(x)
    return _g_ag2te(T) + R_GAS * T * math.log(x) + _l0(T) * (1.0 - x) ** 2


def mu_pbte(T, x):
    _check_x(x)
    return _g_pbte(T) + R_GAS * T * math.log(1.0 - x) + _l0(T) * x ** 2


def x_eq(T_K):
    """Matrix solubility: Newton on RT ln x + L (1-x)^2 = 0 from exp(-L/RT),
    clamped to [1e-9, 0.5] at the start and [1e-9, 0.99] after (:487-507)."""
    L = _l0(T_K)
    RT = R_GAS * T_K
    x = max(1e-9, min(0.5, math.exp(-L / RT)))
    for _ in range(50):
        om = 1.0 - x
        f = RT * math.log(x) + L * om * om
        df = RT / x - 2.0 * L * om
        if abs(df) < 1e-30:
            break
        xn = max(1e-9, min(0.99, x - f / df))
        if abs(xn - x) < 1e-14:
            x = xn
            break
        x = xn
    return max(1e-9, min(0.99, x))


def mu_slope(func, T_K, x0, dx=1e-6):
    """Central-difference slope of mu at x0 (linearize_mu, :509-516)."""
    dx = min(dx, 0.25 * min(x0, 1.0 - x0))
    xp = min(1.0 - 1e-9, x0 + dx)
    xm = max(1e-9, x0 - dx)
    return (func(T_K, xp) - func(T_K, xm)) / (xp - xm)


def diffusivity_ag_in_pbte(T_K):
    """Arrhenius D (m^2/s): D0 = 4.251e-11 cm^2/s, Q = 34.03 kJ/mol (:518-523)."""
    return 4.251e-11 * math.exp(-3.403e+04 / (R_GAS * T_K)) * 1.0e-4


# ---------------------------------------------------------------- interface kinetics
def _zeta0(l, w, kappa, Da, Dc, n=2001):
    """Interface-kinetic integral of the eta profile (compute_gp_eta_zeta0, :557-571):
    -2 Da / l * int h'(xi) xi'(x) [int_0^x (1-h)/D_mix] dx over [-l/2, l/2],
    trapezoid rules on n equally spaced points."""
    a, b = -0.5 * l, 0.5 * l
    step = (b - a) / (n - 1)
    xs = [a + i * step for i in range(n)]
    h = xs[1] - xs[0]
    s = math.sqrt(w / (2.0 * kappa))
    xi = [0.5 * (1.0 - math.tanh(s * x)) for x in xs]
    dxi = [-0.5 * s * (1.0 / math.cosh(s * x) ** 2) for x in xs]
    hv = [6.0 * q ** 5 - 15.0 * q ** 4 + 10.0 * q ** 3 for q in xi]
    hp = [30.0 * q ** 4 - 60.0 * q ** 3 + 30.0 * q ** 2 for q in xi]
    inner = [(1.0 - hv[i]) / max((1.0 - hv[i]) * Da + hv[i] * Dc, 1e-32) for i in range(n)]
    cum = [0.0] * n
    for i in range(1, n):
        cum[i] = cum[i - 1] + 0.5 * (inner[i] + inner[i - 1]) * h
    outer = [hp[i] * dxi[i] * cum[i] for i in range(n)]
    tot = 0.0
    for i in range(1, n):
        tot += 0.5 * (outer[i] + outer[i - 1]) * h
    return -2.0 * Da / l * tot


def _eigenstrain(inp):
    """resolve_eigenstrain_tensor (:135-166): the tensor, or R diag(p) R^T."""
    def mm(A, B):
        return [[sum(A[i][k] * B[k][j] for k in range(3)) for j in range(3)] for i in range(3)]

    def sym(A):
        return [[0.5 * (A[i][j] + A[j][i]) for j in range(3)] for i in range(3)]

    if inp.eigenstrain_tensor is not None:
        t = [[float(v) for v in row] for row in inp.eigenstrain_tensor]
        if len(t) != 3 or any(len(r) != 3 for r in t):
            raise ValueError("eigenstrain_tensor must be 3x3")
        if any(abs(t[i][j] - t[j][i]) > 1e-10 for i in range(3) for j in range(3)):
            raise ValueError("eigenstrain_tensor must be symmetric")
        return sym(t)
    if inp.eigenstrain_principal is None or len(inp.eigenstrain_principal) != 3:
        raise ValueError("give eigenstrain_tensor or a 3-vector eigenstrain_principal")
    p = [float(v) for v in inp.eigenstrain_principal]
    R = _I3 if inp.eigenstrain_rotation_matrix is None else [[float(v) for v in r]
                                                              for r in inp.eigenstrain_rotation_matrix]
    RtR = mm([[R[j][i] for j in range(3)] for i in range(3)], R)
    if any(abs(RtR[i][j] - (1.0 if i == j else 0.0)) > 1e-8 for i in range(3) for j in range(3)):
        raise ValueError("eigenstrain_rotation_matrix must be orthogonal")
    D = [[p[0], 0.0, 0.0], [0.0, p[1], 0.0], [0.0, 0.0, p[2]]]
    return sym(mm(mm(R, D), [[R[j][i] for j in range(3)] for i in range(3)]))


VOIGT21 = ["11", "12", "13", "14", "15", "16", "22", "23", "24", "25", "26", "33", "34", "35", "36",
           "44", "45", "46", "55", "56", "66"]


# ---------------------------------------------------------------- converter
def convert(inp):
    """PFParamConverter.convert (:601-712): code-unit parameter set (dict)."""
    T_K = inp.temperature_C + 273.15
    pf_dx, ref_dx = inp.grid_dx, inp.ref_dx
    if inp.lambda_sm <= 0.0:
        raise ValueError("lambda_sm must be positive")
    w = 12.0 * inp.gamma / inp.lambda_sm              # J/m^3, also the elastic scale
    kappa = 1.5 * inp.gamma * inp.lambda_sm
    if not (inp.gp_gamma_alpha_gp > 0.0) or not (inp.gp_l_eta > 0.0):
        raise ValueError("gp_gamma_alpha_gp and gp_l_eta must be positive")
    gp_w = 12.0 * inp.gp_gamma_alpha_gp / inp.gp_l_eta
    gp_kappa = 1.5 * inp.gp_gamma_alpha_gp * inp.gp_l_eta
    gel_shift_hat = inp.gel_shift_Jm3 / w if abs(w) > 1e-30 else 0.0
    xe = x_eq(T_K)
    xe_eta = inp.gp_xB_eq_alpha_for_eta if inp.gp_xB_eq_alpha_for_eta is not None else xe
    muA1 = mu_slope(mu_pbte, T_K, xe)
    muB1 = mu_slope(mu_ag2te, T_K, xe)
    Da = diffusivity_ag_in_pbte(T_K)
    Dc = Da * inp.D_ratio
    # compute_zeta (:537-540)
    zeta = (inp.gp_reaction_nu_A * muA1 + inp.gp_reaction_nu_B * muB1) * (
        inp.gp_reaction_nu_B / (inp.gp_reaction_nu_A + inp.gp_reaction_nu_B) - xe_eta)
    zeta0 = _zeta0(inp.gp_l_eta, gp_w, gp_kappa, Da, Da * inp.gp_D_ratio_eta)
    zprod = zeta * zeta0
    if abs(zprod) < 1e-30:
        raise ValueError("zeta0 * zeta is zero")
    c_tot = 1.0 / inp.Vm_alpha_0
    Mcrit = 2.0 * Da / (abs(zprod) * inp.gp_l_eta)
    M_eta = inp.gp_M_eta_phys
    if M_eta is None:
        M_eta = (inp.gp_M_eta_ratio_to_crit if inp.gp_M_eta_ratio_to_crit is not None else 1.0e-2) * Mcrit
    nu = inp.gp_reaction_nu_A + inp.gp_reaction_nu_B
    L_eta = (2.0 * nu / (3.0 * c_tot * inp.gp_l_eta)) / ((1.0 / M_eta) + (abs(zprod) * inp.gp_l_eta / (2.0 * Da)))
    L_ref = inp.L_ref_factor * inp.lambda_sm
    t0 = L_ref ** 2 / Da
    L_phi_phys = (4.0 * (inp.v_A + inp.v_B) * Da / (3.0 * c_tot * (inp.lambda_sm) ** 2 * abs(zprod)))
    if inp.C_tensor_PbTe_GPa is None or inp.C_tensor_Ag2Te_GPa is None:
        raise ValueError("both 6x6 stiffness tensors are required")

    def hat(C):
        return [[C[i][j] * 1.0e9 / w for j in range(6)] for i in range(6)]

    return dict(
        W=1.0,
        kappa_phi=kappa / (w * ref_dx ** 2),
        L_phi=L_phi_phys * w * t0,
        D_alpha=Da * t0 / (ref_dx ** 2),
        D_compound=Dc * t0 / (ref_dx ** 2),
        v_A=inp.v_A, v_B=inp.v_B,
        Vm_compound=inp.Vm_compound / inp.Vm_alpha_0,
        Vm_alpha_0=1.0,
        dVm_alpha_dxB=((inp.Vm_compound - inp.Vm_alpha_0) / inp.v_B) / inp.Vm_alpha_0,
        ic_xB_eq_matrix=xe,
        ic_vf_init_phi=inp.vf_init, ic_vf_target_phi=inp.vf_target,
        ic_phi_iface_w=(inp.lambda_sm / pf_dx) / 2.0,
        mu_reference=w / c_tot,
        gp_W_eta_code=gp_w / w if abs(w) > 1e-30 else float("nan"),
        gp_kappa_eta_code=gp_kappa / (w * ref_dx ** 2) if abs(w) > 1e-30 else float("nan"),
        gp_L_eta_code=L_eta * w * t0,
        gp_W_eta_phys=gp_w, gp_kappa_eta_phys=gp_kappa, gp_L_eta_phys=L_eta,
        gp_zeta_eta=zeta, gp_zeta0_eta=zeta0, Mcrit_eta_phys=Mcrit,
        C_hat=hat(inp.C_tensor_PbTe_GPa), Cp_hat=hat(inp.C_tensor_Ag2Te_GPa),
        gel_shift_hat=gel_shift_hat, xAg2Te_eq=xe, temperature_C=inp.temperature_C,
        t_real_unit=t0,
    )


def main_cuda_overrides(inp, pfset=None):
    """The key=value set of the .params file (build_main_cuda_overrides, :741-818)."""
    pfset = pfset or convert(inp)
    if abs(inp.v_B) < 1e-30:
        raise ValueError("v_B must be non-zero (eps_iso_over_vB = eps_iso / v_B)")
    e = _eigenstrain(inp)
    gdx = inp.grid_dx / inp.ref_dx
    xe_eta = inp.gp_xB_eq_alpha_for_eta if inp.gp_xB_eq_alpha_for_eta is not None else pfset["ic_xB_eq_matrix"]
    o = {
        "pf_params_schema_version": SCHEMA_VERSION,
        "dx": gdx, "dy": gdx, "dz": gdx,
        "dt": inp.dt,
        "t_real_unit": pfset["t_real_unit"],
        "temperature_C": inp.temperature_C,
        "mu_reference_scale": pfset["mu_reference"],
    }
    for k in ("W", "kappa_phi", "L_phi", "D_alpha", "D_compound", "v_A", "v_B", "Vm_compound", "Vm_alpha_0",
              "dVm_alpha_dxB", "gp_W_eta_code", "gp_kappa_eta_code", "gp_L_eta_code", "gp_W_eta_phys",
              "gp_kappa_eta_phys", "gp_L_eta_phys"):
        o[k] = pfset[k]
    o.update({
        "gp_gamma_alpha_gp": inp.gp_gamma_alpha_gp,
        "gp_l_eta_nm": inp.gp_l_eta * 1e9,
        "gp_D_ratio": inp.gp_D_ratio_eta,
        "gp_reaction_nu_A": inp.gp_reaction_nu_A,
        "gp_reaction_nu_B": inp.gp_reaction_nu_B,
        "gp_xB_eq_alpha_for_eta": xe_eta,
        "ic_vf_init_phi": pfset["ic_vf_init_phi"],
        "ic_vf_target_phi": pfset["ic_vf_target_phi"],
        "ic_phi_iface_w": pfset["ic_phi_iface_w"],
        "gamma_Jm2": inp.gamma,
        "lambda_sm_m": inp.lambda_sm,
        "Vm_alpha_0_phys_m3mol": inp.Vm_alpha_0,
        "elastic_shift_dimless": pfset["gel_shift_hat"],
        "eps_iso_over_vB": inp.eps_iso / inp.v_B,
        "eps_xx00": e[0][0], "eps_yy00": e[1][1], "eps_zz00": e[2][2],
        "eps_yz00": e[1][2], "eps_xz00": e[0][2], "eps_xy00": e[0][1],
        "E0_xx": inp.E0_xx, "E0_yy": inp.E0_yy, "E0_zz": inp.E0_zz,
        "E0_yz": inp.E0_yz, "E0_xz": inp.E0_xz, "E0_xy": inp.E0_xy,
    })
    o.update(ELASTIC_SOLVER_DEFAULTS)
    if inp.gp_M_eta_ratio_to_crit is not None:
        o["gp_M_eta_ratio_to_crit"] = inp.gp_M_eta_ratio_to_crit
    if inp.gp_M_eta_phys is not None:
        o["gp_M_eta_phys"] = inp.gp_M_eta_phys
    C, Cp = pfset["C_hat"], pfset["Cp_hat"]
    for s in VOIGT21:
        i, j = int(s[0]) - 1, int(s[1]) - 1
        o[f"S_{s}"] = C[i][j]
        o[f"S_p_{s}"] = Cp[i][j] - C[i][j]
    return o


def format_params(overrides):
    """write_main_cuda_override_file (:833-843): sorted keys, ints as is,
    floats %.16e."""
    lines = ["# Auto-generated by Unit_Psedobinary.py",
             "# Format: key=value, consumable by ./main_cuda --pf-param-file <file>"]
    for k in sorted(overrides):
        v = overrides[k]
        lines.append(f"{k}={v}" if isinstance(v, int) else f"{k}={v:.16e}")
    return "\n".join(lines) + "\n"


def write_params(path, inp):
    text = format_params(main_cuda_overrides(inp))
    with open(path, "w", encoding="utf-8") as f:
        f.write(text)
    return text


def main(argv=None):
    ap = argparse.ArgumentParser(description="physical inputs (JSON) -> code-unit .params")
    ap.add_argument("--input", default=None, help="physical-inputs JSON (absent keys take the built-in values)")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    write_params(a.out, load_physical_inputs(a.input))


if __name__ == "__main__":
    main()
