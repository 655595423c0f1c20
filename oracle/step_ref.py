"""Oracle (test infrastructure only): CPU FP64 restatement of the reference
time step of ``main_cuda`` in two-phase dynamics mode.

It is the checker for the HIP path, never the thing measured or shipped: only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may
import it.

Stage order, buffers and time levels follow ``main_cuda.cu:28427-31623``:

1. elastic solve (``main_cuda.cu:28760-29166``): eigenstrain -> homogeneous
   displacement (cold start) or previous-step u(k) (warm start); Green-tensor
   fixed point on h_ij; residual; final strain (+E0) and stress;
2. phi update (``main_cuda.cu:29752-29825``);
3. Y (logit composition) update (``main_cuda.cu:30283-30806``) with the
   optional mass-conserving zero mode (``main_cuda.cu:4269-4405``);
4. dY/dt history (``main_cuda.cu:31618-31623``).

Deliberate deviation (recorded): the elastic solve runs in FP64 here, where the
reference runs it in FP32 (``main_cuda.cu:28761-28789``); FP64 is the precision
the GPU FP64 mode is checked against.

Every forward transform is followed by the 2/3-rule dealias and every inverse
is applied to dealiased data (``cuda_kernels.cu:103-161``), exactly as in the
reference.  numpy's ``rfftn`` is the unnormalised forward transform and
``irfftn`` includes the 1/N the reference applies after each inverse.

Pinning: the FFT is checked against direct DFT sums; dealias against brute
force; the semi-implicit updates and the elastic Green operator against closed
forms (single Fourier mode, affine fixed point, homogeneous one-step
convergence, div(sigma)=0); the zero mode against bisection; the thermo
against analytic derivatives.  The reference's own tests pin none of the
step arithmetic, so the composed lagged-RHS composition update is pinned by
the closed linear recurrence it reduces to for a small single Fourier mode on
a uniform matrix (tests/test_oracle_pins.py), on top of the component checks.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import thermo as th

VOIGT = ("xx", "yy", "zz", "yz", "xz", "xy")


def voigt_matrix(cvals):
    """21 upper-triangle entries {'11':..,'12':..,...} -> symmetric 6x6."""
    m = np.zeros((6, 6))
    for i in range(6):
        for j in range(i, 6):
            v = float(cvals.get(f"{i + 1}{j + 1}", 0.0))
            m[i, j] = v
            m[j, i] = v
    return m


@dataclass
class Params:
    """Code-unit parameters the hot path reads (``pf_params.h``).  Defaults
    follow ``params_default`` (``main_cuda.cu:15781-16433``)."""
    Nx: int = 32
    Ny: int = 32
    Nz: int = 32
    dx: float = 1.0
    dy: float = 1.0
    dz: float = 1.0
    dt: float = 0.01
    temperature_C: float = 380.0
    mu_reference_scale: float = 20668.536
    v_A: float = 0.0
    v_B: float = 1.0
    Vm_compound: float = 1.0
    Vm_alpha_0: float = 1.0
    dVm_alpha_dxB: float = 0.0
    W: float = 1.0
    kappa_phi: float = 2.0
    L_phi: float = 1.0
    D_alpha: float = 1.0
    D_compound: float = 0.01
    eps_iso_over_vB: float = 0.00233
    elastic_shift_dimless: float = 0.0
    Y_clip: float = 20.0
    xB_eps: float = 1e-8
    thermo_convex_extrapolation_enabled: int = 0
    elastic_enabled: int = 1
    C: np.ndarray = field(default_factory=lambda: np.zeros((6, 6)))
    dC: np.ndarray = field(default_factory=lambda: np.zeros((6, 6)))
    eps00: tuple = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0)   # xx yy zz yz xz xy
    E0: tuple = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0)      # xx yy zz yz xz xy
    elastic_warm_start_enabled: int = 1
    elastic_residual_control_enabled: int = 1
    elastic_iter_min: int = 2
    elastic_iter_max: int = 32
    elastic_residual_tolerance: float = 1e-6
    elastic_residual_absolute_floor: float = 1e-30
    elastic_fail_on_nonconvergence: int = 1
    zero_mode: int = 0
    zero_mode_tolerance_relative: float = 1e-12
    zero_mode_max_iterations: int = 24
    enable_Y_rhs_previous_time_level: int = 0
    disable_Y_rhs_gamma_term: int = 0
    Y_rhs_term_h_scale: float = 1.0
    # build extensions of the accelerated elastic solver (0/1 reproduce the reference):
    # warm start from the last 1, 2 or 3 converged steps (constant, linear or
    # quadratic extrapolation in time) and Anderson mixing depth.
    elastic_warm_start_order: int = 1
    elastic_anderson_depth: int = 0

    @property
    def T(self):
        t = self.temperature_C + 273.15
        return 1.0 if t < 1.0 else t

    @property
    def mu0_compound(self):
        """``pfparams_refresh_thermo`` (main_cuda.cu:15604-15614)."""
        T = self.T
        xeq = th.solve_x_eq(T)
        muA = float(th.mu_A_dimless(xeq, T, self.mu_reference_scale,
                                    self.thermo_convex_extrapolation_enabled))
        muB = float(th.mu_B_dimless(xeq, T, self.mu_reference_scale,
                                    self.thermo_convex_extrapolation_enabled))
        return th.v_weighted_mu_compound(muA, muB, self.v_A, self.v_B)


class ElasticNonConvergence(RuntimeError):
    pass


class ZeroModeRejected(RuntimeError):
    pass


# --------------------------------------------------------------------- k-space
def k_axis(n, d, wrap=True):
    """kx_wrap / ky_wrap / kz_wrap (cuda_common.h:69-83) in the same float order."""
    out = np.empty(n if wrap else n // 2 + 1)
    for i in range(out.size):
        ii = i if (not wrap or i <= n // 2) else i - n
        out[i] = 2.0 * math.pi * ii / (d * n)
    return out


def dealias_axis_mask(n, d, wrap=True):
    """Keep-mask of one axis for ``dealias_23_kernel`` (cuda_kernels.cu:103-132):
    a mode is zeroed when |k| >= (2/3) pi / d on any axis."""
    kmax = math.pi * (2.0 / 3.0) / d
    return np.abs(k_axis(n, d, wrap)) < kmax


class KSpace:
    def __init__(self, p: Params):
        self.kx = k_axis(p.Nx, p.dx)[:, None, None]
        self.ky = k_axis(p.Ny, p.dy)[None, :, None]
        self.kz = k_axis(p.Nz, p.dz, wrap=False)[None, None, :]
        # kspace_build_kernel (cuda_common.cu:55-72)
        self.k2 = self.kx * self.kx + self.ky * self.ky + self.kz * self.kz
        mx = dealias_axis_mask(p.Nx, p.dx)[:, None, None]
        my = dealias_axis_mask(p.Ny, p.dy)[None, :, None]
        mz = dealias_axis_mask(p.Nz, p.dz, wrap=False)[None, None, :]
        self.mask = mx & my & mz
        nzc = p.Nz // 2 + 1
        kidx = np.arange(nzc)
        w = np.where((kidx == 0) | ((p.Nz % 2 == 0) & (kidx == p.Nz // 2)), 1.0, 2.0)
        self.herm_w = np.broadcast_to(w[None, None, :], (p.Nx, p.Ny, nzc))


# Threads for the transforms (None: numpy's single-threaded pocketfft, the
# default everywhere the oracle checks results; an integer runs scipy.fft with
# that many workers, used only to time a multi-threaded CPU baseline).
FFT_WORKERS = None


def fwd(a, ks):
    """Unnormalised R2C (cuFFT D2Z/R2C) followed by dealias."""
    if FFT_WORKERS:
        import scipy.fft
        return scipy.fft.rfftn(a, workers=FFT_WORKERS) * ks.mask
    return np.fft.rfftn(a) * ks.mask


def inv(ak, shape):
    """Z2D/C2R followed by the explicit x(1/N)."""
    if FFT_WORKERS:
        import scipy.fft
        return scipy.fft.irfftn(ak, s=shape, axes=(0, 1, 2), workers=FFT_WORKERS)
    return np.fft.irfftn(ak, s=shape, axes=(0, 1, 2))


# ------------------------------------------------------------------- elasticity
def eigenstrain(p: Params, phi, xB):
    """``eigenstrain_phi_xB_point`` (cuda_kernels.cu:193-233), FP64 stores:
    eps0_ii = (1-h) x eps_iso/v_B + h eps00_ii, eps0_ij = h eps00_ij."""
    h = th.h_of_phi(phi)
    eps_c = th.clamp01(xB) * p.eps_iso_over_vB
    omh = 1.0 - h
    e = []
    for c in range(6):
        if c < 3:
            e.append(omh * eps_c + h * p.eps00[c])
        else:
            e.append(h * p.eps00[c])
    return e


def voigt_apply(Cm, e):
    """sigma_I = sum_J C_IJ e~_J with engineering shear e~ = (xx,yy,zz,2yz,2xz,2xy)."""
    ee = [e[0], e[1], e[2], 2.0 * e[3], 2.0 * e[4], 2.0 * e[5]]
    out = []
    for i in range(6):
        s = 0.0
        for j in range(6):
            if Cm[i, j] != 0.0:
                s = s + Cm[i, j] * ee[j]
        out.append(s if not isinstance(s, float) else np.zeros_like(ee[0]) + s)
    return out


def voigt_apply_field(Cm, dCm, h, e):
    """(C + h dC) : e for a field h."""
    a = voigt_apply(Cm, e)
    b = voigt_apply(dCm, e)
    return [a[i] + h * b[i] for i in range(6)]


def acoustic_tensor(Cm, kx, ky, kz):
    """K_ij = C_{i p j q} k_p k_q (the K11..K33 of cuda_kernels.cu:6818-6870)."""
    # Voigt map of (i,p) pairs
    vi = [[0, 5, 4], [5, 1, 3], [4, 3, 2]]
    k = (kx, ky, kz)
    K = [[0.0] * 3 for _ in range(3)]
    for i in range(3):
        for j in range(3):
            s = 0.0
            for pp in range(3):
                for q in range(3):
                    c = Cm[vi[i][pp], vi[j][q]]
                    if c != 0.0:
                        s = s + c * k[pp] * k[q]
            K[i][j] = s
    return K


def green_apply(Cm, ks, hk):
    """u(k) = -i Omega(k) (k . h(k)), Omega = K^-1 via cofactors
    (cuda_kernels.cu:6484-6633); k=0 and |det|<1e-30 give 0."""
    kx, ky, kz = ks.kx, ks.ky, ks.kz
    K = acoustic_tensor(Cm, kx, ky, kz)
    K11, K12, K13 = K[0]
    K22, K23, K33 = K[1][1], K[1][2], K[2][2]
    det = (K13 * (-K13 * K22 + K23 * K12)
           - K23 * (K23 * K11 - K13 * K12)
           + K33 * (K22 * K11 - K12 * K12))
    bad = np.abs(det) < 1e-30
    det_s = np.where(bad, 1.0, det)
    G11 = (K22 * K33 - K23 * K23) / det_s
    G12 = (K13 * K23 - K33 * K12) / det_s
    G13 = (K12 * K23 - K13 * K22) / det_s
    G22 = (K11 * K33 - K13 * K13) / det_s
    G23 = (K13 * K12 - K11 * K23) / det_s
    G33 = (K11 * K22 - K12 * K12) / det_s
    hxx, hyy, hzz, hyz, hxz, hxy = hk
    t1 = kx * hxx + ky * hxy + kz * hxz
    t2 = kx * hxy + ky * hyy + kz * hyz
    t3 = kx * hxz + ky * hyz + kz * hzz
    ux = -1j * (G11 * t1 + G12 * t2 + G13 * t3)
    uy = -1j * (G12 * t1 + G22 * t2 + G23 * t3)
    uz = -1j * (G13 * t1 + G23 * t2 + G33 * t3)
    out = []
    for u in (ux, uy, uz):
        u = np.where(bad, 0.0, u)
        u[0, 0, 0] = 0.0
        out.append(u * ks.mask)
    return out


def strain_from_disp(ks, u):
    """eps_ij(k) = i (k_i u_j + k_j u_i)/2 (cuda_kernels.cu:6673-6726), dealiased."""
    ux, uy, uz = u
    kx, ky, kz = ks.kx, ks.ky, ks.kz
    e = [1j * kx * ux, 1j * ky * uy, 1j * kz * uz,
         0.5j * (ky * uz + kz * uy), 0.5j * (kz * ux + kx * uz), 0.5j * (ky * ux + kx * uy)]
    return [c * ks.mask for c in e]


def hij_field(p: Params, eps, phi, xB, eps0=None, increment=False):
    """``compute_strain_with_perturbation_kernel`` (cuda_kernels.cu:7002-7153):
    h = -(C + h dC):(eps + E0) + C:eps + (C + h dC):eps0."""
    h = th.h_of_phi(phi)
    if increment:
        d = voigt_apply(p.dC, eps)
        return [-(h * d[i]) for i in range(6)]
    if eps0 is None:
        eps0 = eigenstrain(p, phi, xB)
    tot = [eps[i] + p.E0[i] for i in range(6)]
    a = voigt_apply_field(p.C, p.dC, h, tot)
    b = voigt_apply(p.C, eps)
    c = voigt_apply_field(p.C, p.dC, h, eps0)
    return [-a[i] + b[i] + c[i] for i in range(6)]


def stress_field(p: Params, eps_total, phi, xB):
    """``compute_stress_from_strain_with_effective_stiffness_kernel``
    (cuda_kernels.cu:7381-7480): sigma = (C + h dC):(eps - eps0)."""
    h = th.h_of_phi(phi)
    e0 = eigenstrain(p, phi, xB)
    el = [eps_total[i] - e0[i] for i in range(6)]
    return voigt_apply_field(p.C, p.dC, h, el)


# ----------------------------------------------------------------- the stepper
class Oracle:
    """FP64 CPU time stepper; state arrays are C-order [Nx][Ny][Nz]."""

    def __init__(self, p: Params, phi0, xB0, dYdt0=None, xB_cap=None):
        self.p = p
        self.shape = (p.Nx, p.Ny, p.Nz)
        self.ks = KSpace(p)
        self.T = p.T
        self.mu0 = p.mu0_compound
        # raw_fields initialisation (main_cuda.cu:2937-2965)
        phi = th.clamp01(np.asarray(phi0, dtype=np.float64).reshape(self.shape))
        xb = np.clip(np.asarray(xB0, dtype=np.float64).reshape(self.shape), p.xB_eps, 1.0 - p.xB_eps)
        if xB_cap is not None and np.isfinite(xB_cap):
            xb = np.minimum(xb, xB_cap)
        self.phi = phi
        self.xB = xb
        self.Y = th.logit_from_fraction(xb, p.xB_eps, p.Y_clip)
        self.dYdt = (np.zeros(self.shape) if dYdt0 is None
                     else np.asarray(dYdt0, dtype=np.float64).reshape(self.shape).copy())
        self.u_k = None
        self.warm_valid = False
        self.step_index = 0
        self.eps = [np.zeros(self.shape) for _ in range(6)]
        self.sigma = [np.zeros(self.shape) for _ in range(6)]
        self.last_elastic_iterations = 0
        self.last_elastic_residual = float("nan")
        self.last_lambda = 0.0
        self.target_mass = self.mass() if p.zero_mode else None

    # ------------------------------------------------------------- scalars
    def mass(self, phi=None, xB=None):
        """``reduce_sum_xBtot_kernel`` (cuda_kernels.cu:3792-3828)."""
        phi = self.phi if phi is None else phi
        xB = self.xB if xB is None else xB
        h = th.h_of_phi(phi)
        return math.fsum(((1.0 - h) * xB + self.p.v_B * h).ravel())

    def volume_fraction(self):
        """``reduce_sum_h_kernel`` (cuda_kernels.cu:5580-5607): <h(clamp01 phi)>."""
        return math.fsum(th.h_of_phi(th.clamp01(self.phi)).ravel()) / self.phi.size

    def energies(self):
        """Mean energy densities: gradient (cuda_kernels.cu:773-809), double well
        (811-815), elastic 0.5 sigma:eps_el (1789-1850), bulk chemical
        g = (1-h) mu_mix + h mu0 (730-759, without the baseline shift)."""
        p = self.p
        phi = self.phi
        n = phi.size
        gx = (np.roll(phi, -1, 0) - np.roll(phi, 1, 0)) / (2.0 * p.dx)
        gy = (np.roll(phi, -1, 1) - np.roll(phi, 1, 1)) / (2.0 * p.dy)
        gz = (np.roll(phi, -1, 2) - np.roll(phi, 1, 2)) / (2.0 * p.dz)
        e_grad = 0.5 * p.kappa_phi * (gx * gx + gy * gy + gz * gz)
        e_dw = p.W * th.g_of_phi(phi)
        out = {"grad": math.fsum(e_grad.ravel()) / n, "dw": math.fsum(e_dw.ravel()) / n}
        if p.elastic_enabled:
            e0 = eigenstrain(p, phi, self.xB)
            el = [self.eps[i] - e0[i] for i in range(6)]
            s = self.sigma
            sde = (s[0] * el[0] + s[1] * el[1] + s[2] * el[2]
                   + 2.0 * (s[5] * el[5] + s[4] * el[4] + s[3] * el[3]))
            out["elastic"] = math.fsum((0.5 * sde).ravel()) / n
        else:
            out["elastic"] = 0.0
        phc = th.clamp01(phi)
        h = th.h_of_phi(phc)
        x = np.clip(self.xB, 1e-12, 1.0 - 1e-12)
        cv = p.thermo_convex_extrapolation_enabled
        muA = th.mu_A_dimless(x, self.T, p.mu_reference_scale, cv)
        muB = th.mu_B_dimless(x, self.T, p.mu_reference_scale, cv)
        g = (1.0 - h) * ((1.0 - x) * muA + x * muB) + h * self.mu0
        out["chem"] = math.fsum(g.ravel()) / n
        out["total"] = out["grad"] + out["dw"] + out["elastic"] + out["chem"]
        return out

    # ------------------------------------------------------------- elastic
    def _wdot(self, a, b):
        """Hermitian-weighted inner product of 3-component half spectra (the
        norm of compute_elastic_displacement_residual_kernel, 6436-6467)."""
        w = self.ks.herm_w
        return math.fsum(sum(w * (a[i] * np.conj(b[i])).real for i in range(3)).ravel())

    def elastic_solve_extended(self):
        """Accelerated fixed point with the build's extensions: the same map
        F and the same residual |F(u)-u|/|F(u)| as main_cuda.cu:28887-29046,
        started from a polynomial extrapolation of the last converged steps
        and updated by Anderson mixing of depth m (m = 0: plain iteration).
        The result on convergence is F(u_k), as in the reference."""
        p, ks, shape = self.p, self.ks, self.shape
        m = int(p.elastic_anderson_depth)
        hist = getattr(self, "u_hist", [])
        if not self.warm_valid or not hist:
            e0k = [fwd(c, ks) for c in eigenstrain(p, self.phi, self.xB)]
            U = green_apply(p.C, ks, voigt_apply(p.C, e0k))
        else:
            order = min(int(p.elastic_warm_start_order), len(hist))
            coef = {1: (1.0,), 2: (2.0, -1.0), 3: (3.0, -3.0, 1.0)}[order]
            U = [sum(coef[j] * hist[-1 - j][i] for j in range(order)) for i in range(3)]

        def apply_map(v, increment):
            eps = [inv(c, shape) for c in strain_from_disp(ks, v)]
            hh = hij_field(p, eps, self.phi, self.xB, increment=increment)
            return green_apply(p.C, ks, [fwd(c, ks) for c in hh])

        F = apply_map(U, False)
        g = [F[i] - U[i] for i in range(3)]
        du, dg = [], []
        it = 1
        converged = False
        while True:
            Fu = [U[i] + g[i] for i in range(3)]
            res = math.sqrt(self._wdot(g, g) / max(self._wdot(Fu, Fu), p.elastic_residual_absolute_floor))
            if not math.isfinite(res):
                raise ElasticNonConvergence("elastic residual is non-finite")
            if it >= p.elastic_iter_min and res <= p.elastic_residual_tolerance:
                converged = True
                break
            if it >= p.elastic_iter_max:
                break
            if dg:
                H = np.array([[self._wdot(a, b) for b in dg] for a in dg])
                bv = np.array([self._wdot(g, a) for a in dg])
                H = H + (1e-12 * np.trace(H) / len(dg)) * np.eye(len(dg))
                gam = np.linalg.solve(H, bv)
            else:
                gam = np.zeros(0)
            d = [g[i] - sum(gam[j] * (du[j][i] + dg[j][i]) for j in range(len(gam))) for i in range(3)]
            U = [U[i] + d[i] for i in range(3)]
            Ad = apply_map(d, True)
            gn = [g[i] + Ad[i] - d[i] for i in range(3)]
            if m > 0:
                du.append(d)
                dg.append([gn[i] - g[i] for i in range(3)])
                if len(du) > m:
                    du.pop(0)
                    dg.pop(0)
            g = gn
            it += 1
        U = [U[i] + g[i] for i in range(3)]
        self.last_elastic_iterations = it
        self.last_elastic_residual = res
        self.u_k = U
        self.warm_valid = True
        self.u_hist = (hist + [[c.copy() for c in U]])[-3:]
        if not converged and p.elastic_fail_on_nonconvergence:
            raise ElasticNonConvergence(
                f"elastic solver did not converge: iterations={it} residual={res:.17e}")
        ek = strain_from_disp(ks, U)
        eps = [inv(c, shape) + p.E0[i] for i, c in enumerate(ek)]
        self.eps = eps
        self.sigma = stress_field(p, eps, self.phi, self.xB)

    def elastic_solve(self, increment_mode=False):
        """main_cuda.cu:28760-29166."""
        p, ks, shape = self.p, self.ks, self.shape
        accel = bool(p.elastic_warm_start_enabled or p.elastic_residual_control_enabled)
        if accel and (p.elastic_anderson_depth > 0 or p.elastic_warm_start_order > 1):
            return self.elastic_solve_extended()
        warm = accel and self.warm_valid
        if not warm:
            e0 = eigenstrain(p, self.phi, self.xB)
            e0k = [fwd(c, ks) for c in e0]
            # homogeneous solution: -i K^-1 (C:eps0 . k)  (cuda_kernels.cu:6734-6991)
            u = green_apply(p.C, ks, voigt_apply(p.C, e0k))
        else:
            u = [c.copy() for c in self.u_k]
        it = 0 if accel else 1
        converged = not accel
        res = float("nan")
        e0 = eigenstrain(p, self.phi, self.xB)
        while True:
            ek = strain_from_disp(ks, u)
            eps = [inv(c, shape) for c in ek]
            hh = hij_field(p, eps, self.phi, self.xB, eps0=e0)
            hk = [fwd(c, ks) for c in hh]
            un = green_apply(p.C, ks, hk)
            if accel:
                w = ks.herm_w
                diff = 0.0
                ref = 0.0
                for c in range(3):
                    d = un[c] - u[c]
                    diff = diff + w * (d.real * d.real + d.imag * d.imag)
                    ref = ref + w * (un[c].real ** 2 + un[c].imag ** 2)
                ds = math.fsum(diff.ravel())
                rs = math.fsum(ref.ravel())
                res = math.sqrt(ds / max(rs, p.elastic_residual_absolute_floor))
                u = un
                it += 1
                if it >= p.elastic_iter_min and res <= p.elastic_residual_tolerance:
                    converged = True
                    break
                if it >= p.elastic_iter_max:
                    break
            else:
                u = un
                it += 1
                if it >= p.elastic_iter_max:
                    break
        self.last_elastic_iterations = it
        self.last_elastic_residual = res
        if accel:
            self.u_k = u
            self.warm_valid = True
            if not converged and p.elastic_fail_on_nonconvergence:
                raise ElasticNonConvergence(
                    f"elastic solver did not converge: iterations={it} residual={res:.17e}")
        ek = strain_from_disp(ks, u)
        eps = [inv(c, shape) + p.E0[i] for i, c in enumerate(ek)]
        self.eps = eps
        self.sigma = stress_field(p, eps, self.phi, self.xB)

    # ------------------------------------------------------------- phi
    def phi_rhs(self):
        """``compute_phi_rhs_kernel`` (cuda_kernels.cu:390-552), two-phase branch."""
        p = self.p
        phi = self.phi
        hp = th.h_prime_of_phi(phi)
        gp = th.g_prime_of_phi(phi)
        h = th.h_of_phi(phi)
        xB = th.clamp01(self.xB)
        cv = p.thermo_convex_extrapolation_enabled
        f = p.W * gp
        muA = th.mu_A_dimless(xB, self.T, p.mu_reference_scale, cv)
        muB = th.mu_B_dimless(xB, self.T, p.mu_reference_scale, cv)
        shift = p.elastic_shift_dimless if p.elastic_enabled else 0.0
        delta_mu = self.mu0 - p.v_A * muA - p.v_B * muB - shift
        c_bulk = th.c_xB_phi(xB, p.Vm_alpha_0, p.dVm_alpha_dxB, p.Vm_compound, h)
        mu_total = th.mu_tot_mix(muA, muB, xB, self.mu0, h)
        Vm_alpha = p.Vm_alpha_0 + p.dVm_alpha_dxB * xB
        volume_term = p.Vm_compound - Vm_alpha + p.dVm_alpha_dxB * (xB - p.v_B)
        f = f + c_bulk * hp * (delta_mu - c_bulk * mu_total * volume_term)
        if p.elastic_enabled:
            s = self.sigma
            eps_c = xB * p.eps_iso_over_vB
            diag_iso = -eps_c + p.eps_iso_over_vB * (xB - p.v_B)
            e00 = p.eps00
            d = [hp * (e00[0] + diag_iso), hp * (e00[1] + diag_iso), hp * (e00[2] + diag_iso),
                 hp * e00[3], hp * e00[4], hp * e00[5]]
            dgel_eigen = -(s[0] * d[0] + s[1] * d[1] + s[2] * d[2]
                           + 2.0 * s[5] * d[5] + 2.0 * s[4] * d[4] + 2.0 * s[3] * d[3])
            e0 = eigenstrain(p, phi, xB)
            el = [self.eps[i] - e0[i] for i in range(6)]
            # Q = e^T dC e with engineering shear (cuda_kernels.cu:167-185)
            q = voigt_apply(p.dC, el)
            ee = [el[0], el[1], el[2], 2.0 * el[3], 2.0 * el[4], 2.0 * el[5]]
            Q = sum(ee[i] * q[i] for i in range(6))
            f = f + dgel_eigen + 0.5 * hp * Q
        return f

    def phi_update(self):
        """main_cuda.cu:29752-29825."""
        p, ks = self.p, self.ks
        phik = fwd(self.phi, ks)
        rk = fwd(self.phi_rhs(), ks)
        Ldt = p.L_phi * p.dt
        phik = (phik - Ldt * rk) / (1.0 + Ldt * p.kappa_phi * ks.k2) * ks.mask
        self.phi = np.clip(inv(phik, self.shape), -1e-6, 1.0 + 1e-6)

    # ------------------------------------------------------------- Y
    def mu_x(self, phi_for_Y):
        """``compute_mu_x_kernel`` (cuda_kernels.cu:1266-1324); also refreshes xB."""
        p = self.p
        xB = th.sigmoid_from_logit(self.Y, p.Y_clip, p.xB_eps)
        self.xB = xB
        h = th.h_of_phi(phi_for_Y)
        cv = p.thermo_convex_extrapolation_enabled
        muA = th.mu_A_dimless(xB, self.T, p.mu_reference_scale, cv)
        muB = th.mu_B_dimless(xB, self.T, p.mu_reference_scale, cv)
        c_bulk = th.c_xB_phi(xB, p.Vm_alpha_0, p.dVm_alpha_dxB, p.Vm_compound, h)
        mu_total = th.mu_tot_mix(muA, muB, xB, self.mu0, h)
        mu = c_bulk * (muB - muA - c_bulk * mu_total * p.dVm_alpha_dxB)
        if p.elastic_enabled:
            s = self.sigma
            mu = mu + (-p.eps_iso_over_vB * (s[0] + s[1] + s[2]))
        return mu

    def Y_update(self, phi_n, Y_n, y_cap=None):
        """main_cuda.cu:30283-30806 (lagged RHS) and the zero mode.  y_cap is the
        upper Y bound (Y_clip in dynamics; min(Y_clip, logit(minimize_xB_max_safe))
        in full-model minimize, main_cuda.cu:30639-30646)."""
        p, ks, shape = self.p, self.ks, self.shape
        zm = bool(p.zero_mode)
        phi_np1 = self.phi
        if zm:
            phi_for_Y = phi_n
        else:
            phi_for_Y = phi_n if p.enable_Y_rhs_previous_time_level else phi_np1
        mu = self.mu_x(phi_for_Y)
        muk = fwd(mu, ks)
        xB_prev = self.xB
        h = th.h_of_phi(phi_for_Y)
        Dm = th.D_mix(h, p.D_alpha, p.D_compound)
        G = th.gamma_thermo_nonlinear(xB_prev, h, p.Vm_alpha_0, p.dVm_alpha_dxB, p.Vm_compound,
                                      self.T, p.mu_reference_scale,
                                      p.thermo_convex_extrapolation_enabled)
        Meff = th.stabilized_meff(Dm, G)
        divk = np.zeros_like(muk)
        for ka in (ks.kx, ks.ky, ks.kz):
            g = inv(1j * ka * muk * ks.mask, shape)
            Jk = fwd(Meff * g, ks)
            divk = divk + 1j * ka * Jk
        divJ = inv(divk * ks.mask, shape)
        Yk = fwd(self.Y, ks)
        lapY = inv((-ks.k2 * Yk) * ks.mask, shape)
        # gpu_reduce_sum_DY (cuda_kernels.cu:3870-3921)
        x = np.clip(th.sigmoid_unclipped(self.Y), 1e-12, 1.0 - 1e-12)
        mean_DY = math.fsum((Dm * (x * (1.0 - x))).ravel()) / self.Y.size
        Ynk = fwd(Y_n, ks)
        lg = x * (1.0 - x)
        if zm:
            # compute_Y_rhs_sm_tangent_n_kernel (cuda_kernels.cu:2268-2359)
            h_n = th.h_of_phi(phi_n)
            term_h = th.h_prime_of_phi(phi_n) * ((phi_np1 - phi_n) / p.dt) * (p.v_B - x)
            gamma_local = (1.0 - h_n) * lg - 1.0
            f = divJ - term_h - mean_DY * lapY - gamma_local * self.dYdt
        else:
            # compute_Y_rhs_kernel (cuda_kernels.cu:2185-2266)
            h_new = th.h_of_phi(phi_np1)
            h_old = th.h_of_phi(phi_n)
            term_h = p.Y_rhs_term_h_scale * ((h_new - h_old) / p.dt) * (p.v_B - x)
            hx = h_old if p.enable_Y_rhs_previous_time_level else h_new
            gamma_local = (1.0 - hx) * lg - 1.0
            term_gamma = 0.0 if p.disable_Y_rhs_gamma_term else gamma_local * self.dYdt
            f = divJ + (-term_h - mean_DY * lapY - term_gamma)
        fk = fwd(f, ks)
        Yk = (Ynk + p.dt * fk) / (1.0 + p.dt * mean_DY * ks.k2) * ks.mask
        Ynew = inv(Yk, shape)
        cap = p.Y_clip if y_cap is None else y_cap
        if zm:
            self.zero_mode(Ynew, cap)
        else:
            Ynew = np.clip(np.minimum(Ynew, cap), -p.Y_clip, None)
            self.Y = Ynew
            self.xB = th.sigmoid_from_logit(Ynew, p.Y_clip, p.xB_eps)
        self.mean_DY = mean_DY

    def zero_mode(self, Ystar, Y_cap=None, target=None, tol_rel=None, max_it=None):
        """``run_pf_conserved_y_zero_mode_host`` (main_cuda.cu:4269-4405).  The
        minimize mass constraint calls it with its own target, tolerance and
        iteration cap (main_cuda.cu:32297-32366)."""
        p = self.p
        h = th.h_of_phi(self.phi)
        alpha = 1.0 - h
        Y_cap = p.Y_clip if Y_cap is None else Y_cap
        ymin, ymax = float(Ystar.min()), float(Ystar.max())
        lo = -p.Y_clip - ymin
        hi = Y_cap - ymax
        roundoff = 128.0 * np.finfo(float).eps * max(abs(lo), abs(hi), 1.0)
        if lo > hi + roundoff:
            raise ZeroModeRejected("empty zero-mode bracket")
        target = self.target_mass if target is None else target
        tol_rel = p.zero_mode_tolerance_relative if tol_rel is None else tol_rel
        max_it = p.zero_mode_max_iterations if max_it is None else max_it
        tol = tol_rel * max(abs(target), 1.0)
        hv = h * p.v_B

        def evaluate(lam):
            x = th.sigmoid_unclipped(Ystar + lam)
            return (math.fsum((alpha * x + hv).ravel()),
                    math.fsum((alpha * x * (1.0 - x)).ravel()))

        m_lo, _ = evaluate(lo)
        m_hi, _ = evaluate(hi)
        if m_lo - target > tol or m_hi - target < -tol:
            raise ZeroModeRejected("zero-mode target outside bracket")
        lam = min(max(0.0, lo), hi)
        m, dm = evaluate(lam)
        r = m - target
        it = 0
        while it < max_it and abs(r) > tol:
            if r < 0.0:
                lo = lam
            else:
                hi = lam
            cand = float("nan")
            if dm > 0.0 and math.isfinite(dm):
                cand = lam - r / dm
            if not (math.isfinite(cand) and lo < cand < hi):
                cand = 0.5 * (lo + hi)
            lam = cand
            m, dm = evaluate(lam)
            r = m - target
            it += 1
        if abs(r) > tol:
            raise ZeroModeRejected(f"zero mode did not converge: residual={r:.3e}")
        y = Ystar + lam
        if not np.all(np.isfinite(y)) or y.min() < -p.Y_clip or y.max() > Y_cap:
            raise ZeroModeRejected("zero-mode shift leaves the Y bounds")
        self.Y = y
        self.xB = th.sigmoid_unclipped(y)
        self.last_lambda = lam
        self.zero_mode_iterations = it

    # ------------------------------------------------------------- step
    def step(self, n=1):
        for _ in range(n):
            self.step_index += 1
            phi_n = self.phi.copy()
            Y_n = self.Y.copy()
            if self.p.elastic_enabled:
                self.elastic_solve()
            # energy densities of the start-of-step state with this step's
            # elastic fields (what the minimize diagnostics report,
            # main_cuda.cu:29262-29346)
            self.last_energies = self.energies()
            self.phi_update()
            self.Y_update(phi_n, Y_n)
            self.dYdt = (self.Y - Y_n) / self.p.dt
        return self
