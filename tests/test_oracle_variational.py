"""Pins of the oracle's phi right-hand side, the composition potential mu_x
and the energy scalars against what the free energy fixes (CPU only).

The reference's phi drive (compute_phi_rhs_kernel, cuda_kernels.cu:390-552)
is the phi-derivative of the free-energy density

    e(phi, X) = W g(phi) + mu_tot(x, h) / V(x, h) + 1/2 eps_el : (C + h dC) : eps_el

taken at fixed conserved content X = (1 - h) x + h v_B (so x = (X - h v_B)/(1 - h))
and frozen total strain, with eps_el = eps - eps0, eps0 = (1 - h) x e_iso 1 + h eps00
(eigenstrain_phi_xB_point, cuda_kernels.cu:193-233), mu_tot = (1 - h)((1 - x) mu_A +
x mu_B) + h mu0 and V = V_alpha(x)(1 - h) + V_c h (thermo_utils.h:497-511).  That is
where the kernel's d eps0_ii/d phi = h' (eps00_ii - x e_iso + e_iso (x - v_B))
comes from, and it holds when v_A + v_B = 1 (Gibbs-Duhem turns
(x - v_B)(mu_B - mu_A) into the v-weighted potential difference).  The
composition potential (compute_mu_x_kernel, cuda_kernels.cu:1266-1324) is the
X-derivative of the same density at fixed phi.

The densities below are written here, term by term, from those definitions
with full 3x3x3x3 stiffness tensors and tensor (not engineering) shear
strains; the oracle's functions are compared with their numerical
derivatives.  A wrong sign, a dropped c factor, a missing factor 2 on the
shear or 1/2 on Q, or the eigenstrain derivative at fixed x instead of fixed
X fails by percents or more.
"""
import itertools
import math

import numpy as np

from oracle import thermo as th
from oracle.step_ref import Oracle, Params

T_C = 380.0
SCALE = 20668.536
VOIGT_OF = [[0, 5, 4], [5, 1, 3], [4, 3, 2]]   # (i, j) -> xx yy zz yz xz xy


def _stiffness4(Cv):
    C4 = np.empty((3, 3, 3, 3))
    for i, j, k, m in itertools.product(range(3), repeat=4):
        C4[i, j, k, m] = Cv[VOIGT_OF[i][j], VOIGT_OF[k][m]]
    return C4


def _tensor(v):
    """xx yy zz yz xz xy (each an array) -> [3][3][...] symmetric tensor."""
    xx, yy, zz, yz, xz, xy = v
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])


def _voigt(t):
    return [t[0, 0], t[1, 1], t[2, 2], t[1, 2], t[0, 2], t[0, 1]]


def _h(phi):
    return phi ** 3 * (6.0 * phi ** 2 - 15.0 * phi + 10.0)


def _random_symmetric(rng, scale):
    a = rng.standard_normal((6, 6))
    return scale * (a + a.T) / 2.0


def _params(**kw):
    rng = np.random.default_rng(42)
    C = np.zeros((6, 6))
    C[:3, :3] = 15.28
    for i in range(3):
        C[i, i] = 214.29
        C[i + 3, i + 3] = 26.59
    # a positive-definite, fully anisotropic perturbation (every Voigt entry non-zero)
    dC = _random_symmetric(rng, 12.0) + np.diag([-95.0, -90.0, -92.0, 3.0, 4.0, 5.0])
    base = dict(Nx=4, Ny=4, Nz=6, temperature_C=T_C, mu_reference_scale=SCALE,
                v_A=0.25, v_B=0.75, Vm_compound=1.3, Vm_alpha_0=1.0, dVm_alpha_dxB=0.2,
                W=1.7, kappa_phi=2.0, eps_iso_over_vB=0.011, elastic_shift_dimless=0.0,
                C=C, dC=dC, eps00=(0.046, -0.022, -0.017, 0.006, -0.004, 0.009),
                elastic_enabled=1)
    base.update(kw)
    return Params(**base)


def _fields(p, rng):
    shape = (p.Nx, p.Ny, p.Nz)
    phi = rng.uniform(0.08, 0.85, shape)
    x = rng.uniform(0.004, 0.06, shape)
    eps = [rng.uniform(-0.03, 0.03, shape) for _ in range(6)]
    return phi, x, eps


def _eigenstrain_t(p, h, x):
    iso = (1.0 - h) * x * p.eps_iso_over_vB
    e00 = _tensor([np.full_like(h, v) for v in p.eps00])
    return e00 * h + np.eye(3)[:, :, None, None, None] * iso


def _elastic_density(p, phi, x, eps):
    h = _h(phi)
    el = _tensor(eps) - _eigenstrain_t(p, h, x)
    C4 = _stiffness4(p.C)[..., None, None, None] + _stiffness4(p.dC)[..., None, None, None] * h
    return 0.5 * np.einsum("ij...,ijkl...,kl...->...", el, C4, el)


def _stress(p, phi, x, eps):
    h = _h(phi)
    el = _tensor(eps) - _eigenstrain_t(p, h, x)
    C4 = _stiffness4(p.C)[..., None, None, None] + _stiffness4(p.dC)[..., None, None, None] * h
    return _voigt(np.einsum("ijkl...,kl...->ij...", C4, el))


def _chem_density(p, mu0, phi, x):
    h = _h(phi)
    T = p.T
    muA = th.mu_A_dimless(x, T, SCALE)
    muB = th.mu_B_dimless(x, T, SCALE)
    mol = (1.0 - h) * ((1.0 - x) * muA + x * muB) + h * mu0
    vol = (p.Vm_alpha_0 + p.dVm_alpha_dxB * x) * (1.0 - h) + p.Vm_compound * h
    return p.W * phi ** 2 * (1.0 - phi) ** 2 + mol / vol


def _x_of(p, phi, X):
    h = _h(phi)
    return (X - h * p.v_B) / (1.0 - h)


def _richardson(f, a, d):
    """Central difference with one Richardson step (O(d^4))."""
    d1 = (f(a + d) - f(a - d)) / (2.0 * d)
    d2 = (f(a + d / 2) - f(a - d / 2)) / d
    return (4.0 * d2 - d1) / 3.0


def _oracle(p, phi, x, eps):
    o = Oracle(p, phi, x)
    # the oracle clamps x on entry; the pins use values strictly inside
    assert np.array_equal(o.xB, x)
    o.eps = [e.copy() for e in eps]
    o.sigma = _stress(p, phi, x, eps)
    return o


def test_phi_rhs_chemical_part_is_the_energy_derivative_at_fixed_content():
    p = _params(elastic_enabled=0)
    rng = np.random.default_rng(1)
    phi, x, eps = _fields(p, rng)
    o = _oracle(p, phi, x, eps)
    got = o.phi_rhs()
    X = (1.0 - _h(phi)) * x + _h(phi) * p.v_B
    num = _richardson(lambda f: _chem_density(p, o.mu0, f, _x_of(p, f, X)), phi, 2e-5)
    scale = np.abs(num).max()
    assert scale > 0.1
    assert np.abs(got - num).max() < 1e-8 * scale, np.abs(got - num).max()


def test_phi_rhs_chemical_part_vanishes_at_x_eq_and_has_the_potential_slope():
    """With V_c = V_alpha, v_A = 0 the chemical drive is h' (mu0 - mu_B(x)): zero at
    x = x_eq for every phi, and its x-derivative is -h' mu_B'(x) =
    -h' (RT/x - 2L(1-x))/scale (the regular-solution potential of beta)."""
    p = _params(elastic_enabled=0, v_A=0.0, v_B=1.0, Vm_compound=1.0, dVm_alpha_dxB=0.0, W=0.0)
    phi = np.linspace(0.05, 0.95, p.Nx * p.Ny * p.Nz).reshape(p.Nx, p.Ny, p.Nz)
    xeq = th.solve_x_eq(p.T)
    o = _oracle(p, phi, np.full(phi.shape, xeq), [np.zeros(phi.shape)] * 6)
    hp = 30.0 * phi ** 2 * (1.0 - phi) ** 2
    assert np.abs(o.phi_rhs()).max() < 1e-12 * hp.max()
    R, L, T = th.R_GAS, th.get_L_param(p.T), p.T
    for x0 in (0.003, 0.02, 0.05):
        d = 1e-6
        o.xB = np.full(phi.shape, x0 + d)
        fp = o.phi_rhs()
        o.xB = np.full(phi.shape, x0 - d)
        fm = o.phi_rhs()
        slope = (fp - fm) / (2.0 * d)
        ana = -hp * (R * T / x0 - 2.0 * L * (1.0 - x0)) / SCALE
        assert np.abs(slope - ana).max() < 1e-6 * np.abs(ana).max()


def test_phi_rhs_elastic_part_is_the_energy_derivative_at_fixed_content():
    """-sigma : d eps0/d phi + 1/2 h' eps_el : dC : eps_el at frozen total
    strain and fixed X, with every stiffness entry, shear eigenstrain and
    the isotropic composition strain non-zero."""
    p = _params()
    rng = np.random.default_rng(2)
    phi, x, eps = _fields(p, rng)
    o = _oracle(p, phi, x, eps)
    with_el = o.phi_rhs()
    p0 = _params(elastic_enabled=0)
    o0 = _oracle(p0, phi, x, eps)
    got = with_el - o0.phi_rhs()
    X = (1.0 - _h(phi)) * x + _h(phi) * p.v_B
    num = _richardson(lambda f: _elastic_density(p, f, _x_of(p, f, X), eps), phi, 2e-5)
    scale = np.abs(num).max()
    assert scale > 0.05
    assert np.abs(got - num).max() < 1e-8 * scale, np.abs(got - num).max()
    # the pin separates a wrong level: the derivative at fixed x differs by percents
    num_fixed_x = _richardson(lambda f: _elastic_density(p, f, x, eps), phi, 1e-4)
    assert np.abs(num_fixed_x - num).max() > 1e-3 * scale


def test_elastic_shift_enters_the_drive_as_a_potential_offset():
    """elastic_shift_dimless is subtracted from the driving potential:
    d rhs / d shift = -c h' (cuda_kernels.cu:459-461)."""
    rng = np.random.default_rng(3)
    p = _params(elastic_enabled=1)
    phi, x, eps = _fields(p, rng)
    a = _oracle(p, phi, x, eps).phi_rhs()
    b = _oracle(_params(elastic_enabled=1, elastic_shift_dimless=0.37), phi, x, eps).phi_rhs()
    h = _h(phi)
    c = 1.0 / ((p.Vm_alpha_0 + p.dVm_alpha_dxB * x) * (1.0 - h) + p.Vm_compound * h)
    hp = 30.0 * phi ** 2 * (1.0 - phi) ** 2
    assert np.abs((b - a) - (-0.37 * c * hp)).max() < 1e-12


def test_mu_x_is_the_energy_derivative_in_the_conserved_content():
    """compute_mu_x_kernel: c (mu_B - mu_A - c mu_tot dV) - e_iso tr(sigma) is
    d e/d X at fixed phi and frozen strain (chemical and stress terms)."""
    p = _params()
    rng = np.random.default_rng(4)
    phi, x, eps = _fields(p, rng)
    o = _oracle(p, phi, x, eps)
    o.Y = np.log(x / (1.0 - x))
    got = o.mu_x(phi)
    assert np.abs(o.xB - x).max() < 1e-15
    h = _h(phi)
    X = (1.0 - h) * x + h * p.v_B

    def e(XX):
        xx = (XX - h * p.v_B) / (1.0 - h)
        return _chem_density(p, o.mu0, phi, xx) + _elastic_density(p, phi, xx, eps)

    num = _richardson(e, X, 1e-6)
    scale = np.abs(num).max()
    assert np.abs(got - num).max() < 1e-7 * scale, np.abs(got - num).max()
    # the stress term alone (h = 0 pins it against the x-derivative of the elastic density)
    p0 = _params(elastic_enabled=0)
    o0 = _oracle(p0, phi, x, eps)
    o0.Y = o.Y.copy()
    stress_term = got - o0.mu_x(phi)
    num_el = _richardson(lambda XX: _elastic_density(p, phi, (XX - h * p.v_B) / (1.0 - h), eps), X, 1e-6)
    assert np.abs(stress_term - num_el).max() < 1e-7 * np.abs(num_el).max()
    assert np.abs(num_el).max() > 1e-3


def test_energies_gradient_term_single_mode_closed_form():
    """Central-difference gradient energy of phi = 1/2 + a sin(k.r) on an
    anisotropic grid: <1/2 kappa |grad phi|^2> = kappa a^2/4 sum_i sin^2(k_i d_i)/d_i^2."""
    shape = (8, 10, 12)
    d = (0.7, 1.0, 1.3)
    m = (1, 2, 3)
    p = Params(Nx=shape[0], Ny=shape[1], Nz=shape[2], dx=d[0], dy=d[1], dz=d[2], kappa_phi=2.3,
               elastic_enabled=0, W=0.0)
    ii = [np.arange(n) for n in shape]
    th_ = (2 * np.pi * m[0] * ii[0][:, None, None] / shape[0] + 2 * np.pi * m[1] * ii[1][None, :, None] / shape[1]
           + 2 * np.pi * m[2] * ii[2][None, None, :] / shape[2])
    a = 0.2
    o = Oracle(p, 0.5 + a * np.sin(th_), np.full(shape, 0.02))
    ref = p.kappa_phi * a * a / 4.0 * sum(math.sin(2 * math.pi * m[i] / shape[i]) ** 2 / d[i] ** 2
                                           for i in range(3))
    assert abs(o.energies()["grad"] - ref) < 1e-14 * ref * 10


def test_energies_double_well_uniform():
    """W g(1/2) = W/16 at the barrier top; 0 at both wells."""
    p = Params(Nx=4, Ny=4, Nz=4, W=1.7, elastic_enabled=0)
    for val, ref in ((0.5, 1.7 / 16.0), (0.0, 0.0), (1.0, 0.0)):
        o = Oracle(p, np.full((4, 4, 4), val), np.full((4, 4, 4), 0.02))
        assert abs(o.energies()["dw"] - ref) < 1e-15
        assert o.energies()["grad"] == 0.0


def test_energies_elastic_term_is_the_tensor_contraction():
    """<1/2 sigma : eps_el> with tensor shear strains: the full 3x3x3x3
    contraction of a fully anisotropic stiffness, non-uniform fields."""
    p = _params()
    rng = np.random.default_rng(5)
    phi, x, eps = _fields(p, rng)
    o = _oracle(p, phi, x, eps)
    ref = _elastic_density(p, phi, x, eps).mean()
    assert abs(o.energies()["elastic"] - ref) < 1e-13 * abs(ref)


def test_energies_chemical_term_limits():
    """Bulk chemical density: mu0 inside beta (h = 1, also for phi > 1, which is
    clamped); the alpha molar Gibbs energy in the matrix, whose x-derivative
    is mu_B - mu_A (Gibbs-Duhem of the regular-solution potentials)."""
    p = _params(elastic_enabled=0, W=0.0)
    shape = (p.Nx, p.Ny, p.Nz)
    o = Oracle(p, np.ones(shape), np.full(shape, 0.02))
    assert abs(o.energies()["chem"] - o.mu0) < 1e-14 * abs(o.mu0)
    o.phi = np.full(shape, 1.0 + 5e-7)
    assert abs(o.energies()["chem"] - o.mu0) < 1e-14 * abs(o.mu0)
    for x0 in (0.004, 0.02, 0.3):
        dd = 1e-6
        o.phi = np.zeros(shape)
        o.xB = np.full(shape, x0 + dd)
        gp = o.energies()["chem"]
        o.xB = np.full(shape, x0 - dd)
        gm = o.energies()["chem"]
        slope = (gp - gm) / (2 * dd)
        ref = float(th.mu_B_dimless(x0, p.T, SCALE) - th.mu_A_dimless(x0, p.T, SCALE))
        assert abs(slope - ref) < 1e-6 * max(1.0, abs(ref))
