"""Pins of the CPU oracle against what the mathematics fixes (CPU only).

None of these retypes an oracle formula: each compares the oracle with an
independent statement of the same quantity (direct DFT sums, integer
dealias rule, analytic derivatives, Eshelby's interior solution, mechanical
equilibrium, single-mode decay, bisection)."""
import math

import numpy as np
import pytest

from cuda_sto_pf_amd.configs import base_params, composition, config, spheres_phi
from oracle import thermo as th
from oracle.step_ref import (KSpace, Oracle, Params, dealias_axis_mask, fwd, green_apply, hij_field,
                             inv, strain_from_disp, voigt_apply, voigt_apply_field)


# ------------------------------------------------------------------ transforms
def test_rfftn_convention_matches_direct_dft():
    """fwd() is the unnormalised e^{-ikx} sum, inv() includes 1/N."""
    rng = np.random.default_rng(0)
    a = rng.standard_normal((4, 6, 6))
    nx, ny, nz = a.shape
    X = np.zeros((nx, ny, nz // 2 + 1), complex)
    for kx in range(nx):
        for ky in range(ny):
            for kz in range(nz // 2 + 1):
                s = 0.0
                for x in range(nx):
                    for y in range(ny):
                        for z in range(nz):
                            s += a[x, y, z] * np.exp(-2j * np.pi * (kx * x / nx + ky * y / ny + kz * z / nz))
                X[kx, ky, kz] = s
    ks = KSpace(Params(Nx=nx, Ny=ny, Nz=nz))
    assert np.abs(fwd(a, ks) - X * ks.mask).max() < 1e-12
    # inverse of a dealiased spectrum reproduces the band-limited field
    b = inv(X * ks.mask, a.shape)
    assert np.abs(fwd(b, ks) - X * ks.mask).max() < 1e-12


@pytest.mark.parametrize("n", [6, 8, 9, 12, 16, 41, 96, 128, 246, 256])
def test_dealias_mask_integer_rule(n):
    """|k| < (2/3) pi/d  <=>  3|ii| < N, except exactly on 3|ii| = N, where the
    reference's floating-point comparison decides."""
    m = dealias_axis_mask(n, 1.0)
    for i in range(n):
        ii = i if i <= n // 2 else i - n
        if 3 * abs(ii) != n:
            assert m[i] == (3 * abs(ii) < n), (n, i)
    mz = dealias_axis_mask(n, 1.0, wrap=False)
    for k in range(n // 2 + 1):
        if 3 * k != n:
            assert mz[k] == (3 * k < n)
    # symmetric in +/- ii
    for i in range(1, n):
        assert m[i] == m[n - i] or i == n // 2


# ------------------------------------------------------------------ thermo
def test_gamma_fd_matches_analytic_derivative():
    T = 653.15
    scale = 20668.536
    x = np.array([0.003, 0.01, 0.02, 0.05, 0.08])
    h = np.zeros_like(x)
    G = th.gamma_thermo_nonlinear(x, h, 1.0, 0.0, 1.0, T, scale)
    R = th.R_GAS
    L = th.get_L_param(T)
    dA = -R * T / (1 - x) + 2 * L * x
    dB = R * T / x - 2 * L * (1 - x)
    ana = (dB - dA) / scale
    # central-difference truncation: dx^2/6 |f'''|, f''' = 2RT/x^3 + 2RT/(1-x)^3
    bound = (1e-5) ** 2 / 6 * (2 * R * T / x ** 3 + 2 * R * T / (1 - x) ** 3) / scale
    assert np.all(np.abs(G - ana) <= 1.05 * bound + 1e-9 * np.abs(ana))


def test_gamma_closed_form_equals_the_difference_quotient():
    """The device evaluates the reference's difference quotient of mu_B - mu_A
    in closed form, c_bulk (RT log1p(d / (xm (1 - xp))) / d - 2 L) / scale
    with d = xp - xm (G(T) cancels, the logarithm differences combine).  It
    must equal the literal quotient to that form's own rounding noise, over
    the whole clamped range, including the clamped ends."""
    T = 653.15
    scale = 20668.536
    R = th.R_GAS
    L = th.get_L_param(T)
    x = np.concatenate([np.logspace(-13, -1, 60), np.linspace(0.1, 0.9, 41), 1.0 - np.logspace(-1, -13, 60)])
    h = np.linspace(0.0, 1.0, x.size)
    lit = th.gamma_thermo_nonlinear(x, h, 1.0, 0.1, 1.2, T, scale)
    c_bulk = th.c_xB_phi(x, 1.0, 0.1, 1.2, h)
    xp = th.clamp_fraction_eps(x + 1e-5)
    xm = th.clamp_fraction_eps(x - 1e-5)
    d = xp - xm
    closed = c_bulk * ((R * T) * np.log1p(d / (xm * (1.0 - xp))) / d - 2.0 * L) / scale
    # literal form: mu values O(|G|/scale) differenced over d ~ 2e-5
    noise = 2 * np.finfo(float).eps * (abs(th.G_Ag2Te_Solid(T)) + abs(th.G_PbTe_Solid(T)) + 60 * R * T) / scale / d
    assert np.all(np.abs(closed - lit) <= noise * c_bulk + 1e-13 * np.abs(lit))


def test_potential_difference_closed_form():
    """The phi finish evaluates mu_B - mu_A on the CALPHAD branch as
    (G_B - G_A + RT ln(x/(1-x)) + L (1 - 2x)) / scale (one logarithm).  It
    must equal the difference of the oracle's two potentials to their
    rounding (a few ulps of |G|/scale), over the clamped range."""
    T = 653.15
    scale = 20668.536
    x = np.concatenate([np.logspace(-12, -1, 50), np.linspace(0.1, 0.9, 41), 1.0 - np.logspace(-1, -12, 50)])
    lit = th.mu_B_dimless(x, T, scale) - th.mu_A_dimless(x, T, scale)
    xc = th.clamp_fraction_eps(x)
    closed = (th.G_Ag2Te_Solid(T) - th.G_PbTe_Solid(T) + th.R_GAS * T * np.log(xc / (1.0 - xc))
              + th.get_L_param(T) * (1.0 - 2.0 * xc)) / scale
    tol = 8 * np.finfo(float).eps * (abs(th.G_Ag2Te_Solid(T)) + abs(th.G_PbTe_Solid(T))
                                     + th.R_GAS * T * 28.0) / scale
    assert np.all(np.abs(closed - lit) <= tol)


def test_x_eq_solves_the_tangent_condition():
    T = 653.15
    x = th.solve_x_eq(T)
    L = th.get_L_param(T)
    assert abs(th.R_GAS * T * math.log(x) + L * (1 - x) ** 2) < 1e-6 * L


def test_interpolation_functions():
    p = np.linspace(-0.2, 1.2, 57)
    e = 1e-6
    assert np.allclose((th.h_of_phi(p + e) - th.h_of_phi(p - e)) / (2 * e), th.h_prime_of_phi(p), atol=1e-8)
    assert np.allclose((th.g_of_phi(p + e) - th.g_of_phi(p - e)) / (2 * e), th.g_prime_of_phi(p), atol=1e-8)
    assert th.h_of_phi(0.0) == 0.0 and th.h_of_phi(1.0) == 1.0


def test_logit_sigmoid_roundtrip():
    x = np.array([1e-9, 1e-5, 0.02, 0.5, 0.98, 1 - 1e-9])
    y = th.logit_from_fraction(x, 1e-8, 20.0)
    back = th.sigmoid_from_logit(y, 20.0, 1e-8)
    assert np.allclose(back, np.clip(x, 1e-8, 1 - 1e-8), rtol=1e-12)


# ------------------------------------------------------------------ elasticity
def _iso(lam, mu):
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    for i in range(3):
        C[i, i] = lam + 2 * mu
        C[i + 3, i + 3] = mu
    return C


def _div_sigma_k(ks, sig_k):
    """k_j sigma_ij(k) for the three rows."""
    sxx, syy, szz, syz, sxz, sxy = sig_k
    return [ks.kx * sxx + ks.ky * sxy + ks.kz * sxz,
            ks.kx * sxy + ks.ky * syy + ks.kz * syz,
            ks.kx * sxz + ks.ky * syz + ks.kz * szz]


def test_homogeneous_green_gives_equilibrium():
    """Single-mode and random eigenstrain: C:(eps(u) - eps0) is divergence-free."""
    n = 16
    p = Params(Nx=n, Ny=n, Nz=n, C=_iso(100.0, 30.0))
    ks = KSpace(p)
    rng = np.random.default_rng(3)
    e0 = [rng.standard_normal((n, n, n)) for _ in range(6)]
    e0k = [fwd(c, ks) for c in e0]
    u = green_apply(p.C, ks, voigt_apply(p.C, e0k))
    ek = strain_from_disp(ks, u)
    sig = voigt_apply(p.C, [ek[i] - e0k[i] for i in range(6)])
    div = _div_sigma_k(ks, sig)
    scale = max(np.abs(s).max() for s in sig)
    for d in div:
        d = d * ks.mask
        d[0, 0, 0] = 0
        assert np.abs(d).max() < 1e-11 * scale


def test_eshelby_dilatational_sphere_interior_stress():
    """Isotropic matrix, sphere with eigenstrain e*I: the interior hydrostatic
    stress is uniform, p = -4 mu K e / (K + 4 mu/3) (Eshelby).  With a
    periodic 96^3 box (volume fraction 0.5 %) and a spectrally truncated
    interface the interior mean (r < R/2) is within 2.5 %; shear vanishes."""
    n, R, e = 96, 10.0, 0.01
    lam, mu = 60.0, 30.0
    K = lam + 2 * mu / 3
    p = Params(Nx=n, Ny=n, Nz=n, C=_iso(lam, mu), eps00=(e, e, e, 0, 0, 0), eps_iso_over_vB=0.0,
               elastic_residual_tolerance=1e-12)
    phi = spheres_phi((n, n, n), [(n / 2, n / 2, n / 2)], [R], width=1.0)
    o = Oracle(p, phi, np.full(phi.shape, 0.02))
    o.elastic_solve()
    g = np.arange(n) - n / 2
    r = np.sqrt(g[:, None, None] ** 2 + g[None, :, None] ** 2 + g[None, None, :] ** 2)
    inside = r < R / 2
    p_num = ((o.sigma[0] + o.sigma[1] + o.sigma[2]) / 3)[inside].mean()
    p_ana = -4 * mu * K * e / (K + 4 * mu / 3)
    assert abs(p_num / p_ana - 1) < 0.025, (p_num, p_ana)
    assert np.abs(o.sigma[3][inside]).mean() < 1e-2 * abs(p_ana)


def test_homogeneous_fixed_point_converges_in_one_update():
    pd, phi, xB, meta = config(1, 16)
    o = Oracle(Params(**pd), phi, xB, xB_cap=meta["xB_max_safe"])
    o.elastic_solve()
    assert o.last_elastic_iterations == 2
    assert o.last_elastic_residual == 0.0


def test_heterogeneous_fixed_point_is_equilibrium_and_affine():
    pd, phi, xB, meta = config(2, 16)
    p = Params(**pd)
    o = Oracle(p, phi, xB, xB_cap=meta["xB_max_safe"])
    o.elastic_solve()
    assert o.last_elastic_residual <= p.elastic_residual_tolerance
    ks = o.ks
    # converged: div((C + h dC):(eps - eps0)) = 0 on the retained modes
    sig_k = [fwd(s, ks) for s in o.sigma]
    div = _div_sigma_k(ks, sig_k)
    scale = max(np.abs(s).max() for s in sig_k)
    for d in div:
        d = d * ks.mask
        d[0, 0, 0] = 0
        assert np.abs(d).max() < 1e-6 * scale
    # the map is affine: F(u + d) - F(u) = A d with A from the linear part only
    rng = np.random.default_rng(9)
    u = o.u_k
    d = [(rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)) * ks.mask for c in u]

    def F(v, increment=False):
        eps = [inv(c, o.shape) for c in strain_from_disp(ks, v)]
        hh = hij_field(p, eps, o.phi, o.xB, increment=increment)
        return green_apply(p.C, ks, [fwd(c, ks) for c in hh])

    Fu = F(u)
    Fud = F([u[i] + d[i] for i in range(3)])
    Ad = F(d, increment=True)
    for i in range(3):
        assert np.abs((Fud[i] - Fu[i]) - Ad[i]).max() < 1e-10 * np.abs(Ad[i]).max()


# ------------------------------------------------------------------ updates
def test_phi_update_single_mode_decay():
    """With a zero right-hand side phi^ decays by 1/(1 + L dt kappa k^2)."""
    n = 16
    p = Params(Nx=n, Ny=n, Nz=n, elastic_enabled=0, L_phi=1.3, kappa_phi=2.0, dt=0.05)
    x = np.arange(n)
    phi = 0.5 + 0.1 * np.cos(2 * np.pi * 2 * x / n)[:, None, None] * np.ones((n, n, n))
    o = Oracle(p, phi, np.full((n, n, n), 0.02))
    o.phi_rhs = lambda: np.zeros(o.shape)
    o.phi_update()
    k = 2 * np.pi * 2 / n
    amp = 0.1 / (1 + p.L_phi * p.dt * p.kappa_phi * k * k)
    ref = 0.5 + amp * np.cos(2 * np.pi * 2 * x / n)[:, None, None]
    assert np.abs(o.phi - ref).max() < 1e-14


def test_zero_mode_matches_bisection_and_conserves():
    pd, phi, xB, meta = config(2, 16)
    o = Oracle(Params(**pd), phi, xB, xB_cap=meta["xB_max_safe"])
    m0 = o.target_mass
    for _ in range(3):
        o.step()
        assert abs(o.mass() - m0) <= 1e-12 * max(abs(m0), 1.0)
    # the shift solves the monotone mass equation: bisection gives the same root
    Ystar = o.Y - o.last_lambda
    h = th.h_of_phi(o.phi)

    def f(lam):
        return math.fsum(((1 - h) * th.sigmoid_unclipped(Ystar + lam) + h * o.p.v_B).ravel()) - m0

    lo, hi = -1.0, 1.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) < 0:
            lo = mid
        else:
            hi = mid
    assert abs(0.5 * (lo + hi) - o.last_lambda) < 1e-9


def test_mass_drifts_without_zero_mode():
    """Without the zero mode the lagged-RHS composition update is not
    conservative (the reference's reason for the zero mode)."""
    shape = (16, 12, 8)
    pd = base_params(shape, hetero=True, zero_mode=False)
    rng = np.random.default_rng(11)
    phi = spheres_phi(shape, [(8, 6, 4)], [3.0], width=1.5)
    o = Oracle(Params(**pd), phi, composition(phi, rng), xB_cap=0.5)
    m0 = o.mass()
    o.step(3)
    assert abs(o.mass() - m0) > 1e-12 * m0


def test_composed_y_update_single_mode_recurrence():
    """The whole lagged-RHS composition update (Oracle.Y_update, main_cuda.cu:
    30283-30806: mu_x, the flux M_eff grad mu with the FD Gamma, its
    divergence, the D-bar Laplacian stabiliser, the lagged gamma_local dY/dt
    term, the semi-implicit solve and the dY/dt history) reduces, for a
    small single Fourier mode of Y on a uniform matrix (phi = 0, no
    elasticity, no zero mode), to a closed linear recurrence: M_eff Gamma = D
    makes div J = D q lap Y with q = x(1-x), which the stabiliser cancels, so
        a_1 (1 + dt Db k^2) = a_0,
        a_{n+1} (1 + dt Db k^2) = a_n + (1 - q)(a_n - a_{n-1}),   Db = D q.
    dt = 2 makes dt Db k^2 = 0.14, so a dropped stabiliser, a wrong sign or
    level of the lagged term, or a wrong time step fails by percents."""
    from cuda_sto_pf_amd.configs import base_params
    shape = (16, 4, 4)
    pd = base_params(shape, dt=2.0, hetero=False, zero_mode=False, elastic=False)
    p = Params(**pd)
    x0, eps, m = 0.03, 1e-7, 4
    Y0 = math.log(x0 / (1.0 - x0))
    cosx = np.cos(2.0 * np.pi * m * np.arange(shape[0]) / shape[0])
    Y = Y0 + eps * np.broadcast_to(cosx[:, None, None], shape)
    o = Oracle(p, np.zeros(shape), th.sigmoid_unclipped(Y))

    def amp(f):
        return 2.0 * float(np.mean(f * cosx[:, None, None]))

    a = [amp(o.Y)]
    for _ in range(6):
        o.step()
        a.append(amp(o.Y))
        assert np.abs(o.phi).max() == 0.0
    q = x0 * (1.0 - x0)
    k2 = (2.0 * np.pi * m / shape[0]) ** 2
    den = 1.0 + p.dt * p.D_alpha * q * k2
    assert 0.1 < p.dt * p.D_alpha * q * k2 < 0.2
    b = [a[0], a[0] / den]
    for n in range(1, 6):
        b.append((b[n] + (1.0 - q) * (b[n] - b[n - 1])) / den)
    for n in range(7):
        assert abs(a[n] - b[n]) <= 1e-6 * abs(a[0]), (n, a[n], b[n])
    # the lagged term matters here: without it the third amplitude differs by ~1e-2
    assert abs(b[2] - b[1] / den) > 1e-3 * abs(a[0])
