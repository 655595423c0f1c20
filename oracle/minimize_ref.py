"""Oracle (test infrastructure only): CPU FP64 restatement of ``main_cuda``'s
phi-only minimize mode (``--mode=minimize`` without ``--minimize-full-model``).

It is the checker for the HIP minimize path, never the thing measured or
shipped: only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU
leg may import it.

One iteration follows ``main_cuda.cu:28427-32600`` with ``P.mode == 1``:

1. elastic fixed point with composition absent (``xB_r == NULL``), i.e. the
   eigenstrain h(phi) eps00 (``eigenstrain_phi_only_point``,
   ``cuda_kernels.cu:363-383``) -- the dynamics solve with x_B = 0;
2. energies of phi^n (``main_cuda.cu:29264-29346``): F_surf = <grad + dw>,
   F_el = <0.5 sigma:eps_el>, sum(gel * h(clamp01 phi)); V0 is locked to
   <h> at iteration 1 when it is not positive;
3. phi with an under-relaxed volume multiplier (``main_cuda.cu:29357-29750``):
   g_exp = W g'(phi) + dgel/dphi (``compute_phi_rhs_minimize_kernel``,
   ``cuda_kernels.cu:6298-6368``), g_full = g_exp - kappa Lap phi,
   lambda_raw = -sum(h' g_full) / sum(h'^2), lambda = (1-eta) lambda + eta
   lambda_raw, phi^ <- (phi^ - L dt (g_exp + lambda h')^) / (1 + L dt kappa k^2),
   clamp to [-1e-6, 1+1e-6];
4. KKT residual at phi^{n+1} with the stress frozen and the exact multiplier,
   projected at the bounds (``cuda_kernels.cu:684-706``); rms_dphi;
5. post-projection sub-steps with eta = 1 (``main_cuda.cu:32104-32197``);
6. first-order volume projection phi <- clamp01(phi + lc h'(phi)),
   lc = (V0 - <h>) / <h'^2> (``main_cuda.cu:32274-32290``,
   ``cuda_kernels.cu:639-647``);
7. convergence: condA (rms_dphi, rms_res, vol_err_rel below their thresholds)
   held for ``minimize_convergence_steps`` iterations; condB needs the excess
   free energy, which phi-only minimize leaves NaN, so it never fires
   (``main_cuda.cu:32400-32560``).

Pinning (tests/test_oracle_minimize.py): the phi RHS against a numerical
derivative of the elastic energy density at frozen total strain; the
multiplier against the orthogonality sum(h' (g_full + lambda_raw h')) = 0;
the volume projection against its first-order closed form; a uniform field
as an exact stationary point; monotone decrease of the free energy for a
small pseudo-time step without elasticity.
"""
import math
from dataclasses import dataclass, replace

import numpy as np

from . import thermo as th
from .step_ref import Oracle, Params, eigenstrain, fwd, inv, voigt_apply


@dataclass
class MinimizeOptions:
    """``params_default`` minimize entries (``main_cuda.cu:16319-16347``)."""
    dt: float = 0.05
    max_iter: int = 2000
    V0: float = 0.0
    eta_lambda_vol: float = 0.2
    post_projection_iters: int = 1
    rms_dphi_threshold: float = 1.0e-6
    rms_res_threshold: float = 1.0e-4
    vol_err_rel_threshold: float = 1.0e-4
    convergence_steps: int = 10
    min_pseudo_time: float = 0.0
    # full-model minimize (main_cuda.cu:16336-16345)
    rms_dY_threshold: float = 5.0e-5
    energy_diff_rel_threshold: float = 1.0e-9
    xB_max_safe: float = 0.07
    xB_ref: float = 0.03         # ic_23d_xB_out (default 0.03, main_cuda.cu:16215); <= 0: <x_B> of the initial fields
    # constraints of the full model (main_cuda.cu:16321-16328)
    mass_constraint: int = 0     # --minimize-mass-constraint
    freeze_phi: int = 0          # --minimize-freeze-phi (V4, needs the mass constraint)
    mass_tolerance_relative: float = 1.0e-12
    mass_max_iterations: int = 64
    target_mass: float = 0.0     # --minimize-target-mass-code; <= 0: the initial mass
    # V5 (--minimize-component-volume-constraints): one h-volume constraint per
    # ownership label (int array of the grid's shape, values 0..K-1) with the
    # target sums of h(clamp01 phi) per label (main_cuda.cu:24638-24670)
    component_labels: object = None
    component_targets: object = None


class MinimizeOracle:
    def __init__(self, p: Params, phi0, opts: MinimizeOptions):
        self.p = p
        self.opts = opts
        shape = (p.Nx, p.Ny, p.Nz)
        self.o = Oracle(p, phi0, np.zeros(shape))
        self.o.xB = np.zeros(shape)   # composition absent: eps0 = h eps00
        self.shape = shape
        self.lambda_vol = 0.0
        self.V0_target = opts.V0
        self.iteration = 0
        self.conv_count = 0
        self.converged = False
        self.stats = {}

    @property
    def phi(self):
        return self.o.phi

    # ``compute_phi_rhs_minimize_kernel`` (cuda_kernels.cu:6298-6368) with the
    # stress and total strain of the last elastic solve held fixed
    def g_explicit(self, phi):
        p, o = self.p, self.o
        f = p.W * th.g_prime_of_phi(phi)
        if p.elastic_enabled:
            hp = th.h_prime_of_phi(phi)
            s = o.sigma
            e00 = p.eps00
            dgel_eigen = -(s[0] * hp * e00[0] + s[1] * hp * e00[1] + s[2] * hp * e00[2]
                           + 2.0 * (s[5] * hp * e00[5] + s[4] * hp * e00[4] + s[3] * hp * e00[3]))
            h = th.h_of_phi(phi)
            el = [o.eps[i] - h * e00[i] for i in range(6)]
            q = voigt_apply(p.dC, el)
            ee = [el[0], el[1], el[2], 2.0 * el[3], 2.0 * el[4], 2.0 * el[5]]
            Q = sum(ee[i] * q[i] for i in range(6))
            f = f + dgel_eigen + 0.5 * hp * Q
        return f

    def laplacian(self, phik):
        """``compute_laplacian_k_kernel`` (cuda_kernels.cu:3528-3538), dealiased."""
        return inv(-self.o.ks.k2 * phik * self.o.ks.mask, self.shape)

    @staticmethod
    def multiplier(phi, g):
        hp = th.h_prime_of_phi(phi)
        s_hg = math.fsum((hp * g).ravel())
        s_h2 = math.fsum((hp * hp).ravel())
        return (-s_hg / s_h2) if abs(s_h2) > 1e-30 else 0.0

    def semi_implicit(self, phik, rhs):
        p, ks = self.p, self.o.ks
        Ldt = p.L_phi * self.opts.dt
        out = (phik - Ldt * fwd(rhs, ks)) / (1.0 + Ldt * p.kappa_phi * ks.k2) * ks.mask
        return np.clip(inv(out, self.shape), -1e-6, 1.0 + 1e-6)

    def mean_h(self, phi):
        return math.fsum(th.h_of_phi(th.clamp01(phi)).ravel()) / phi.size

    def energies(self):
        p, o = self.p, self.o
        e = o.energies()
        out = {"F_surf": e["grad"] + e["dw"], "F_el": e["elastic"],
               "interface_sum": (e["grad"] + e["dw"]) * o.phi.size, "el_core_sum": 0.0}
        if p.elastic_enabled:
            h = th.h_of_phi(o.phi)
            el = [o.eps[i] - h * p.eps00[i] for i in range(6)]
            s = o.sigma
            gel = 0.5 * (s[0] * el[0] + s[1] * el[1] + s[2] * el[2]
                         + 2.0 * (s[5] * el[5] + s[4] * el[4] + s[3] * el[3]))
            out["el_core_sum"] = math.fsum((gel * th.h_of_phi(th.clamp01(o.phi))).ravel())
        return out

    def iterate(self):
        p, opts, o = self.p, self.opts, self.o
        self.iteration += 1
        n = o.phi.size
        phi_n = o.phi.copy()
        if p.elastic_enabled:
            o.elastic_solve()
        mean_h0 = self.mean_h(phi_n)
        if self.iteration == 1 and self.V0_target <= 0.0:
            self.V0_target = mean_h0
        en = self.energies()
        # phi with the under-relaxed multiplier
        gexp = self.g_explicit(phi_n)
        phik = fwd(phi_n, o.ks)
        gfull = gexp - p.kappa_phi * self.laplacian(phik)
        lam_raw = self.multiplier(phi_n, gfull)
        eta = min(max(opts.eta_lambda_vol, 0.0), 1.0)
        self.lambda_vol = (1.0 - eta) * self.lambda_vol + eta * lam_raw
        phi = self.semi_implicit(phik, gexp + self.lambda_vol * th.h_prime_of_phi(phi_n))
        # KKT residual at phi^{n+1}, exact multiplier, projected at the bounds
        gexp1 = self.g_explicit(phi)
        phik1 = fwd(phi, o.ks)
        gfull1 = gexp1 - p.kappa_phi * self.laplacian(phik1)
        lam_kkt = self.multiplier(phi, gfull1)
        res = gfull1 + lam_kkt * th.h_prime_of_phi(phi)
        res = np.where(((phi <= 1e-12) & (res > 0.0)) | ((phi >= 1.0 - 1e-12) & (res < 0.0)), 0.0, res)
        rms_res = math.sqrt(math.fsum((res * res).ravel()) / n)
        rms_dphi = math.sqrt(math.fsum(((phi - phi_n) ** 2).ravel()) / n) / opts.dt
        # post-projection sub-steps (eta = 1), stress frozen
        for k in range(max(opts.post_projection_iters, 0)):
            if k == 0:
                g_k, phik_k, gf_k = gexp1, phik1, gfull1
            else:
                g_k = self.g_explicit(phi)
                phik_k = fwd(phi, o.ks)
                gf_k = g_k - p.kappa_phi * self.laplacian(phik_k)
            self.lambda_vol = self.multiplier(phi, gf_k)
            phi = self.semi_implicit(phik_k, g_k + self.lambda_vol * th.h_prime_of_phi(phi))
        mean_h = self.mean_h(phi)
        # first-order volume projection
        if self.V0_target > 1e-30 and math.isfinite(mean_h):
            vol_err = self.V0_target - mean_h
            if abs(vol_err) > 1e-14:
                hp = th.h_prime_of_phi(phi)
                mean_hp2 = math.fsum((hp * hp).ravel()) / n
                lc = vol_err / mean_hp2 if mean_hp2 > 1e-30 else 0.0
                phi = th.clamp01(phi + lc * hp)
                mean_h = self.mean_h(phi)
        o.phi = phi
        if self.V0_target > 1e-30:
            vol_err_rel = abs(mean_h - self.V0_target) / (self.V0_target + 1e-30)
        else:
            vol_err_rel = 0.0
        condA = (math.isfinite(rms_dphi) and rms_dphi < opts.rms_dphi_threshold
                 and math.isfinite(rms_res) and rms_res < opts.rms_res_threshold
                 and math.isfinite(vol_err_rel) and vol_err_rel < opts.vol_err_rel_threshold)
        t_ok = self.iteration * opts.dt + 1e-14 >= opts.min_pseudo_time
        self.conv_count = self.conv_count + 1 if (condA and t_ok) else 0
        if self.conv_count >= opts.convergence_steps:
            self.converged = True
        self.stats = dict(mean_h=mean_h, V0=self.V0_target, lambda_vol=self.lambda_vol,
                          F_surf=en["F_surf"], F_el=en["F_el"], interface_sum=en["interface_sum"],
                          el_core_sum=en["el_core_sum"], rms_res=rms_res, rms_dphi=rms_dphi,
                          vol_err_rel=vol_err_rel, lambda_raw=lam_raw, lambda_kkt=lam_kkt,
                          elastic_iterations=o.last_elastic_iterations)
        return self.stats

    def run(self, max_iter=None):
        m = self.opts.max_iter if max_iter is None else max_iter
        while self.iteration < m and not self.converged:
            self.iterate()
        return self


class MinimizeFullOracle(MinimizeOracle):
    """``--mode=minimize --minimize-full-model`` (chemical drive, composition and
    the volume constraint), one iteration after ``main_cuda.cu:28427-32600``
    with ``P.mode == 1`` and ``P.minimize_full_model == 1``:

    1. elastic fixed point with the composition eigenstrain of the current x_B;
    2. energies of phi^n: F_surf, F_el (with the composition eigenstrain), the
       bulk chemical excess <g_bulk> - g_bulk0 with g_bulk0 = mu_mix(x_ref)
       (``compute_gbulk_excess_hat_kernel``, ``cuda_kernels.cu:725-759``;
       baseline ``main_cuda.cu:27400-27432``) and
       F_total_excess = F_chem_excess + F_surf + F_el (``main_cuda.cu:29334-29346``);
    3. x_B refreshed from Y (``compute_mu_x_kernel``), then the dynamics phi
       right-hand side (``compute_phi_rhs_kernel``, chemical drive included)
       as g_exp, and the multiplier update and semi-implicit step of the
       phi-only mode (``main_cuda.cu:29357-29560``);
    4. the KKT residual at phi^{n+1} with the same right-hand side, the stress
       and total strain frozen (``main_cuda.cu:29577-29728``);
    5. the dynamics composition update at the pseudo-time step (``P.dt`` is
       the minimize dt), Y capped at min(Y_clip, logit(minimize_xB_max_safe))
       (``main_cuda.cu:30283-30911``, cap ``30639-30646``);
    6. post-projection sub-steps with the x_B of the updated Y
       (``main_cuda.cu:32104-32190``), the first-order volume projection;
    7. rms_dY = rms(Y^{n+1} - Y^n) / dt, the relative change of F_total_excess
       against the previous iteration, and the full-model condA / condB
       (``main_cuda.cu:32400-32575``).
    """

    def __init__(self, p: Params, phi0, xB0, opts: MinimizeOptions):
        p = replace(p, dt=opts.dt)
        self.p = p
        self.opts = opts
        shape = (p.Nx, p.Ny, p.Nz)
        self.shape = shape
        # raw_fields load with the minimize cap (main_cuda.cu:2940-2965)
        self.o = Oracle(p, phi0, xB0, xB_cap=opts.xB_max_safe)
        o = self.o
        self.x_ref = opts.xB_ref if opts.xB_ref > 0.0 else math.fsum(o.xB.ravel()) / o.xB.size
        muA = float(th.mu_A_dimless(self.x_ref, o.T, p.mu_reference_scale))
        muB = float(th.mu_B_dimless(self.x_ref, o.T, p.mu_reference_scale))
        self.g_bulk0 = (1.0 - self.x_ref) * muA + self.x_ref * muB
        c = min(max(opts.xB_max_safe, p.xB_eps), 1.0 - p.xB_eps)
        self.y_cap = min(p.Y_clip, math.log(c / (1.0 - c)))
        self.lambda_vol = 0.0
        self.V0_target = opts.V0
        self.iteration = 0
        self.conv_count = 0
        self.converged = False
        self.F_prev = float("nan")
        self.stats = {}
        if opts.freeze_phi and (not opts.mass_constraint or opts.post_projection_iters != 0 or opts.V0 > 0.0):
            raise ValueError("--minimize-freeze-phi needs the mass constraint, no post-projection and V0 <= 0 "
                             "(main_cuda.cu:24597-24617)")
        self.labels = None
        if opts.component_labels is not None:
            lab = np.asarray(opts.component_labels)
            tg = np.asarray(opts.component_targets, dtype=np.float64)
            K = tg.size
            if (not opts.mass_constraint or opts.freeze_phi or opts.post_projection_iters != 0 or opts.V0 > 0.0
                    or K < 2 or K > 64 or lab.shape != self.shape or lab.min() < 0 or lab.max() >= K
                    or np.bincount(lab.ravel(), minlength=K).min() == 0 or not np.all(tg > 0.0)
                    or not np.all(np.isfinite(tg))):
                raise ValueError("V5 component constraints: bad labels, targets or options "
                                 "(main_cuda.cu:24638-24670, 27262-27290)")
            self.labels = lab.astype(np.int64)
            self.targets = tg
            self.comp_lambda = np.zeros(K)
            self.comp_err = float("nan")
        self.mass_target = None
        if opts.mass_constraint:
            # main_cuda.cu:28117-28146: the initial ledger, or a matching explicit target
            m0 = o.mass()
            self.mass_target = opts.target_mass if opts.target_mass > 0.0 else m0
            if abs(m0 - self.mass_target) > opts.mass_tolerance_relative * max(abs(self.mass_target), 1.0):
                raise ValueError("explicit minimize target mass does not match the initial ledger")

    def g_explicit(self, phi):
        """compute_phi_rhs_kernel at phi with the current x_B and the stress and
        total strain of the last solve held fixed."""
        o = self.o
        keep = o.phi
        o.phi = phi
        try:
            return o.phi_rhs()
        finally:
            o.phi = keep

    def refresh_xB(self):
        """x_B = sigma_clip(Y) (compute_mu_x_kernel, cuda_kernels.cu:1266-1324)."""
        self.o.xB = th.sigmoid_from_logit(self.o.Y, self.p.Y_clip, self.p.xB_eps)

    def energies(self):
        p, o = self.p, self.o
        e = o.energies()
        out = {"F_surf": e["grad"] + e["dw"], "F_el": e["elastic"],
               "interface_sum": (e["grad"] + e["dw"]) * o.phi.size, "el_core_sum": 0.0,
               "F_chem_excess": e["chem"] - self.g_bulk0}
        if p.elastic_enabled:
            e0 = eigenstrain(p, o.phi, o.xB)
            el = [o.eps[i] - e0[i] for i in range(6)]
            s = o.sigma
            gel = 0.5 * (s[0] * el[0] + s[1] * el[1] + s[2] * el[2]
                         + 2.0 * (s[5] * el[5] + s[4] * el[4] + s[3] * el[3]))
            out["el_core_sum"] = math.fsum((gel * th.h_of_phi(th.clamp01(o.phi))).ravel())
        out["F_total_excess"] = out["F_chem_excess"] + out["F_surf"] + out["F_el"]
        return out

    def iterate(self):
        p, opts, o = self.p, self.opts, self.o
        self.iteration += 1
        n = o.phi.size
        phi_n = o.phi.copy()
        Y_n = o.Y.copy()
        if p.elastic_enabled:
            o.elastic_solve()
        mean_h0 = self.mean_h(phi_n)
        if self.iteration == 1 and self.V0_target <= 0.0:
            self.V0_target = mean_h0
        en = self.energies()
        F_prev = self.F_prev
        self.F_prev = en["F_total_excess"]
        if opts.freeze_phi:
            return self._iterate_fixed_phi(phi_n, Y_n, en, F_prev)
        # phi with the under-relaxed multiplier (one per label under V5,
        # main_cuda.cu:29446-29510)
        self.refresh_xB()
        gexp = self.g_explicit(phi_n)
        phik = fwd(phi_n, o.ks)
        gfull = gexp - p.kappa_phi * self.laplacian(phik)
        eta = min(max(opts.eta_lambda_vol, 0.0), 1.0)
        if self.labels is not None:
            lam_raw = float("nan")
            self.comp_lambda = (1.0 - eta) * self.comp_lambda + eta * self.comp_multipliers(phi_n, gfull)
            phi = self.semi_implicit(phik, gexp + self.comp_lambda[self.labels] * th.h_prime_of_phi(phi_n))
        else:
            lam_raw = self.multiplier(phi_n, gfull)
            self.lambda_vol = (1.0 - eta) * self.lambda_vol + eta * lam_raw
            phi = self.semi_implicit(phik, gexp + self.lambda_vol * th.h_prime_of_phi(phi_n))
        # KKT residual at phi^{n+1} (the exact multipliers there, main_cuda.cu:29660-29728)
        self.refresh_xB()
        gexp1 = self.g_explicit(phi)
        gfull1 = gexp1 - p.kappa_phi * self.laplacian(fwd(phi, o.ks))
        if self.labels is not None:
            lam_kkt = float("nan")
            res = gfull1 + self.comp_multipliers(phi, gfull1)[self.labels] * th.h_prime_of_phi(phi)
        else:
            lam_kkt = self.multiplier(phi, gfull1)
            res = gfull1 + lam_kkt * th.h_prime_of_phi(phi)
        res = np.where(((phi <= 1e-12) & (res > 0.0)) | ((phi >= 1.0 - 1e-12) & (res < 0.0)), 0.0, res)
        rms_res = math.sqrt(math.fsum((res * res).ravel()) / n)
        rms_dphi = math.sqrt(math.fsum(((phi - phi_n) ** 2).ravel()) / n) / opts.dt
        # composition at the pseudo-time step
        o.phi = phi
        o.Y_update(phi_n, Y_n, y_cap=self.y_cap)
        o.dYdt = (o.Y - Y_n) / p.dt
        # post-projection with the updated composition, stress frozen
        for _ in range(max(opts.post_projection_iters, 0)):
            g_k = self.g_explicit(phi)
            phik_k = fwd(phi, o.ks)
            gf_k = g_k - p.kappa_phi * self.laplacian(phik_k)
            self.lambda_vol = self.multiplier(phi, gf_k)
            phi = self.semi_implicit(phik_k, g_k + self.lambda_vol * th.h_prime_of_phi(phi))
        mean_h = self.mean_h(phi)
        if self.labels is not None:
            phi = self.comp_project(phi)
            mean_h = self.mean_h(phi)
        elif self.V0_target > 1e-30 and math.isfinite(mean_h):
            vol_err = self.V0_target - mean_h
            if abs(vol_err) > 1e-14:
                hp = th.h_prime_of_phi(phi)
                mean_hp2 = math.fsum((hp * hp).ravel()) / n
                lc = vol_err / mean_hp2 if mean_hp2 > 1e-30 else 0.0
                phi = th.clamp01(phi + lc * hp)
                mean_h = self.mean_h(phi)
        o.phi = phi
        mass_err_rel = self._mass_constraint()
        rms_dY = math.sqrt(math.fsum(((o.Y - Y_n) ** 2).ravel()) / n) / p.dt
        energy_diff_rel = (abs(en["F_total_excess"] - F_prev) / max(abs(F_prev), 1.0e-20)
                           if math.isfinite(F_prev) else float("nan"))
        if self.labels is not None:
            vol_err_rel = self.comp_err
        else:
            vol_err_rel = (abs(mean_h - self.V0_target) / (self.V0_target + 1e-30)
                           if self.V0_target > 1e-30 else 0.0)
        fin = math.isfinite
        vol_ok = fin(vol_err_rel) and vol_err_rel < opts.vol_err_rel_threshold
        res_ok = fin(rms_res) and rms_res < opts.rms_res_threshold
        mass_ok = mass_err_rel is None or (fin(mass_err_rel) and mass_err_rel < opts.mass_tolerance_relative)
        condA = (fin(rms_dphi) and fin(rms_dY) and rms_dphi < opts.rms_dphi_threshold
                 and rms_dY < opts.rms_dY_threshold and res_ok and vol_ok and mass_ok)
        condB = (fin(energy_diff_rel) and energy_diff_rel < opts.energy_diff_rel_threshold
                 and fin(rms_dphi) and fin(rms_dY) and rms_dphi < 5.0 * opts.rms_dphi_threshold
                 and rms_dY < 5.0 * opts.rms_dY_threshold and res_ok and vol_ok and mass_ok)
        t_ok = self.iteration * opts.dt + 1e-14 >= opts.min_pseudo_time
        self.conv_count = self.conv_count + 1 if ((condA or condB) and t_ok) else 0
        if self.conv_count >= opts.convergence_steps:
            self.converged = True
        self.stats = dict(mean_h=mean_h, V0=self.V0_target, lambda_vol=self.lambda_vol,
                          F_surf=en["F_surf"], F_el=en["F_el"], interface_sum=en["interface_sum"],
                          el_core_sum=en["el_core_sum"], F_chem_excess=en["F_chem_excess"],
                          F_total_excess=en["F_total_excess"], rms_res=rms_res, rms_dphi=rms_dphi,
                          rms_dY=rms_dY, energy_diff_rel=energy_diff_rel, vol_err_rel=vol_err_rel,
                          lambda_raw=lam_raw, lambda_kkt=lam_kkt, elastic_iterations=o.last_elastic_iterations,
                          **self._mass_stats(mass_err_rel))
        return self.stats

    def comp_sums(self, phi, w):
        """Per-label sums of w (compute_component_constraint_moments_kernel,
        cuda_kernels.cu:590-604)."""
        flat = self.labels.ravel()
        w = w.ravel()
        return np.array([math.fsum(w[flat == c]) for c in range(self.targets.size)])

    def comp_multipliers(self, phi, g):
        """-sum h' g / sum h'^2 over each label (main_cuda.cu:29468-29478)."""
        hp = th.h_prime_of_phi(phi)
        s_hg = self.comp_sums(phi, hp * g)
        s_h2 = self.comp_sums(phi, hp * hp)
        return np.where(np.abs(s_h2) > 1e-30, -s_hg / np.where(s_h2 == 0.0, 1.0, s_h2), 0.0)

    def comp_project(self, phi):
        """Up to three first-order corrections phi <- clamp01(phi + c_k h'(phi))
        with c_k = (T_k - sum_k h) / sum_k h'^2 per label, skipped once every
        |T_k - sum_k h| <= 1e-14; then the largest relative error
        (main_cuda.cu:32207-32275, apply_component_volume_projection_kernel
        cuda_kernels.cu:617-627)."""
        for _ in range(3):
            hp = th.h_prime_of_phi(phi)
            s_h2 = self.comp_sums(phi, hp * hp)
            s_h = self.comp_sums(phi, th.h_of_phi(th.clamp01(phi)))
            delta = self.targets - s_h
            corr = np.where(s_h2 > 1e-30, delta / np.where(s_h2 == 0.0, 1.0, s_h2), 0.0)
            if not np.any(np.abs(delta) > 1e-14):
                break
            phi = th.clamp01(phi + corr[self.labels] * hp)
        s_h = self.comp_sums(phi, th.h_of_phi(th.clamp01(phi)))
        self.comp_err = float(np.max(np.abs(s_h - self.targets) / np.maximum(self.targets, 1e-30)))
        return phi

    def _mass_constraint(self):
        """main_cuda.cu:32297-32366: the zero-mode root find on the clamped Y of
        this iteration, at the current phi, with the minimize Y cap, tolerance
        and iteration cap; the relative mass error after it (32443-32452)."""
        if self.mass_target is None:
            return None
        o, opts = self.o, self.opts
        o.zero_mode(o.Y.copy(), self.y_cap, target=self.mass_target, tol_rel=opts.mass_tolerance_relative,
                    max_it=opts.mass_max_iterations)
        self.mass_lambda = o.last_lambda
        return abs(o.mass() - self.mass_target) / max(abs(self.mass_target), 1.0)

    def _mass_stats(self, mass_err_rel):
        if mass_err_rel is None:
            return {}
        return dict(mass_target=self.mass_target, mass_current=self.o.mass(), mass_err_rel=mass_err_rel,
                    mass_lambda=self.mass_lambda)

    def _iterate_fixed_phi(self, phi_n, Y_n, en, F_prev):
        """V4 (--minimize-freeze-phi): phi is held fixed; only the composition
        update and the mass constraint run, and the stopping rule looks at Y,
        the energy, the volume and the mass (main_cuda.cu:32108, 32192-32195,
        32276, 32514-32526)."""
        p, opts, o = self.p, self.opts, self.o
        n = o.phi.size
        o.Y_update(phi_n, Y_n, y_cap=self.y_cap)
        o.dYdt = (o.Y - Y_n) / p.dt
        mean_h = self.mean_h(o.phi)
        mass_err_rel = self._mass_constraint()
        rms_dY = math.sqrt(math.fsum(((o.Y - Y_n) ** 2).ravel()) / n) / p.dt
        energy_diff_rel = (abs(en["F_total_excess"] - F_prev) / max(abs(F_prev), 1.0e-20)
                           if math.isfinite(F_prev) else float("nan"))
        vol_err_rel = (abs(mean_h - self.V0_target) / (self.V0_target + 1e-30)
                       if self.V0_target > 1e-30 else 0.0)
        fin = math.isfinite
        vol_ok = fin(vol_err_rel) and vol_err_rel < opts.vol_err_rel_threshold
        mass_ok = fin(mass_err_rel) and mass_err_rel < opts.mass_tolerance_relative
        condA = fin(rms_dY) and rms_dY < opts.rms_dY_threshold and vol_ok and mass_ok
        condB = (fin(energy_diff_rel) and energy_diff_rel < opts.energy_diff_rel_threshold
                 and fin(rms_dY) and rms_dY < 5.0 * opts.rms_dY_threshold and vol_ok and mass_ok)
        t_ok = self.iteration * opts.dt + 1e-14 >= opts.min_pseudo_time
        self.conv_count = self.conv_count + 1 if ((condA or condB) and t_ok) else 0
        if self.conv_count >= opts.convergence_steps:
            self.converged = True
        nan = float("nan")
        self.stats = dict(mean_h=mean_h, V0=self.V0_target, lambda_vol=self.lambda_vol,
                          F_surf=en["F_surf"], F_el=en["F_el"], interface_sum=en["interface_sum"],
                          el_core_sum=en["el_core_sum"], F_chem_excess=en["F_chem_excess"],
                          F_total_excess=en["F_total_excess"], rms_res=nan, rms_dphi=nan,
                          rms_dY=rms_dY, energy_diff_rel=energy_diff_rel, vol_err_rel=vol_err_rel,
                          lambda_raw=nan, lambda_kkt=nan, elastic_iterations=o.last_elastic_iterations,
                          **self._mass_stats(mass_err_rel))
        return self.stats
