// Point-wise stage operators fused into the z pass, and k-space operators on
// the compact spectrum.
//
// z-pass operators (real space, one voxel per call):
//   OpElasticCold  C:eps0(phi,xB)                       -> 6 spectra
//   OpHij          eps = inv(6)/N -> h_ij               -> 6 spectra
//   OpPhiRhs       eps(+E0) -> sigma, dF/dphi, tr sigma -> g = phi - L dt rhs
//   OpPhiFinish    phi^{n+1} = clamp(inv/N); x_B, mu_x  -> mu spectrum, sum D x(1-x)
//   OpFlux         grad mu -> J = M_eff grad mu; Y^n; s -> 5 spectra
//   OpYStar        Y* = inv/N (zero mode) with min/max
//   OpYFinal       Y = clamp(inv/N), x_B, dY/dt (no zero mode)
// k-space operators (compact spectrum, one retained mode per thread):
//   Green update with residual, strain from displacement, phi and Y
//   semi-implicit updates, spectral gradient.
#pragma once
#include "pf_passes.hpp"
#include "pf_physics.hpp"

namespace pf {

// ----------------------------------------------------------------- elastic
// C : eps0 (the homogeneous-solution source of cuda_kernels.cu:6734-6991).
template <typename T>
struct OpElasticCold {
  struct Args {
    PhysParams P;
    const double* phi;
    const double* xB;
  };
  static constexpr int NRED = 0;
  static constexpr int kinds[1] = {0};
  static constexpr int NPRE = 2;
  __device__ static void load(const Args& a, size_t g, double* pre) {
    pre[0] = a.phi[g];
    pre[1] = a.xB[g];
  }
  __device__ static void apply(const Args& a, size_t g, const double* pre, const T*, T* out, double*) {
    double h = h_of_phi(pre[0]);
    double e0[6], s[6];
    eigenstrain(a.P, h, clamp01(pre[1]), e0);
    voigt_apply<double>(a.P.C, e0, s, a.P.ortho);
#pragma unroll
    for (int c = 0; c < 6; ++c) out[c] = T(s[c]);
  }
};

// h_ij (cuda_kernels.cu:7002-7153).  FP64: all double.  FP32: eps0 and h in
// double rounded to float, the 21-coefficient products in float, as the
// reference does.  `increment` drops the affine part (increments of u).
template <typename T>
struct OpHij {
  struct Args {
    PhysParams P;
    const double* phi;
    const double* xB;
    T invN;
    int increment;
    // optional h(phi) per voxel in T, written by the full map for the
    // increments of the same solve (OpHijInc; T(h) is all they use)
    T* hc;
  };
  static constexpr int NRED = 0;
  static constexpr int kinds[1] = {0};
  static constexpr int NPRE = 2;
  __device__ static void load(const Args& a, size_t g, double* pre) {
    pre[0] = a.phi[g];
    pre[1] = a.increment ? 0.0 : a.xB[g];
  }
  __device__ static void apply(const Args& a, size_t g, const double* pre, const T* vin, T* out, double*) {
    T eps[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) eps[c] = vin[c] * a.invN;
    const double h = h_of_phi(pre[0]);
    if (!a.increment && a.hc) a.hc[g] = T(h);
    if (a.increment) {
      hij_increment_point<T>(a.P, T(h), eps, out);
    } else {
      double e0[6];
      eigenstrain(a.P, h, clamp01(pre[1]), e0);
      T e0t[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) e0t[c] = T(e0[c]);
      hij_point<T>(a.P, T(h), e0t, eps, out);
    }
  }
};

// The linear part of the map on an increment, -h dC : eps(d), with T(h) read
// from the per-voxel cache the full map of the same solve wrote (4 bytes per
// voxel in mixed precision instead of phi's 8).
template <typename T>
struct OpHijInc {
  struct Args {
    PhysParams P;
    const T* hc;
    T invN;
  };
  static constexpr int NRED = 0;
  static constexpr int kinds[1] = {0};
  static constexpr int NPRE = 1;
  __device__ static void load(const Args& a, size_t g, double* pre) { pre[0] = double(a.hc[g]); }
  __device__ static void apply(const Args& a, size_t, const double* pre, const T* vin, T* out, double*) {
    T eps[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) eps[c] = vin[c] * a.invN;
    hij_increment_point<T>(a.P, T(pre[0]), eps, out);
  }
};

// ----------------------------------------------------------------- phi
// dF/dphi explicit part (compute_phi_rhs_kernel, cuda_kernels.cu:390-552,
// two-phase branch) and the combined semi-implicit numerator
// g = phi - L dt rhs (phi^ and rhs^ enter the update linearly,
// phi_semi_implicit_update_kernel cuda_kernels.cu:821-843).
__device__ __forceinline__ double phi_rhs_chem(const PhysParams& P, double phi, double xB,
                                               double& hp_out, double& h_out, double& muA_out,
                                               double& muB_out) {
  const double hp = h_prime_of_phi(phi);
  const double gp = g_prime_of_phi(phi);
  const double h = h_of_phi(phi);
  double f = P.W * gp;
  const double muA = mu_A(P, xB);
  const double muB = mu_B(P, xB);
  muA_out = muA;
  muB_out = muB;
  const double delta_mu = P.mu0 - P.v_A * muA - P.v_B * muB - (P.elastic ? P.shift : 0.0);
  const double c_bulk = c_xB_phi(P, xB, h);
  const double mu_total = mu_tot_mix(muA, muB, xB, P.mu0, h);
  const double Vm_alpha = P.Vm_a0 + P.dVm * xB;
  const double volume_term = P.Vm_c - Vm_alpha + P.dVm * (xB - P.v_B);
  f += c_bulk * hp * (delta_mu - c_bulk * mu_total * volume_term);
  hp_out = hp;
  h_out = h;
  return f;
}

// Energy densities accumulated on the way (mean values after finalisation):
//   red[0] 0.5 sigma:eps_el   (compute_gel_density_kernel, cuda_kernels.cu:1789-1850)
//   red[1] W g(phi)           (compute_dw_energy_density_kernel, 811-815)
//   red[2] (1-h) mu_mix + h mu0 with clamp01(phi), x clamped to [1e-12, 1-1e-12]
//          (compute_gbulk_excess_hat_kernel 730-759, without the baseline shift)
template <typename T, bool ELASTIC>
struct OpPhiRhs {
  struct Args {
    PhysParams P;
    const double* phi;
    const double* xB;
    T* trs;        // out: tr(sigma) per voxel (ELASTIC)
    T invN;
    T* sig6 = nullptr;   // optional out: the 6 stress components (xx,yy,zz,yz,xz,xy), stride n
    size_t n = 0;
    // mixed precision, increment form: the unclamped band-limited phi* of the
    // last update; the transform then carries the clamp correction phi - phi*
    // minus L dt (dF/dphi) (null: phi - L dt (dF/dphi))
    const double* phistar = nullptr;
  };
  static constexpr int NRED = 3;
  static constexpr int kinds[3] = {RED_SUM, RED_SUM, RED_SUM};
  static constexpr int NPRE = sizeof(T) == 4 ? 3 : 2;
  __device__ static void load(const Args& a, size_t g, double* pre) {
    pre[0] = a.phi[g];
    pre[1] = a.xB[g];
    if constexpr (NPRE > 2) pre[2] = a.phistar ? a.phistar[g] : 0.0;
  }
  __device__ static void apply(const Args& a, size_t g, const double* pre, const T* vin, T* out, double* red) {
    const PhysParams& P = a.P;
    const double phi = pre[0];
    const double xraw = pre[1];
    const double xB = clamp01(xraw);
    double hp, h, mA, mB;
    double f = phi_rhs_chem(P, phi, xB, hp, h, mA, mB);
    red[1] += P.W * g_of_phi(phi);
    {
      const double hc = h_of_phi(clamp01(phi));
      const double xc = clamp_fraction_eps(xraw);
      // mu(clamp01(x)) equals mu(clamp_eps(x)) (the calphad branch clamps
      // to [1e-12, 1 - 1e-12] itself) except on the convex continuation
      // for x above 1 - 1e-12
      if (P.convex && xraw > 1.0 - 1e-12) {
        mA = mu_A(P, xc);
        mB = mu_B(P, xc);
      }
      red[2] += (1.0 - hc) * ((1.0 - xc) * mA + xc * mB) + hc * P.mu0;
    }
    if constexpr (ELASTIC) {
      // Strain, stress and Q in the transform precision T (FP32 in mixed
      // mode, as the reference's float stress and strain arrays), the
      // derivative terms and energies in FP64.
      double e0[6];
      T el[6], sig[6];
      eigenstrain(P, h, xB, e0);
#pragma unroll
      for (int c = 0; c < 6; ++c) el[c] = T(double(vin[c] * a.invN) + P.E0[c] - e0[c]);
      // sigma = (C + h dC):(eps - eps0)  (cuda_kernels.cu:7381-7480)
      voigt_apply_h<T>(P, T(h), el, sig);
      const double eps_c = xB * P.eps_iso;
      const double diag_iso = -eps_c + P.eps_iso * (xB - P.v_B);
      const double d0 = hp * (P.eps00[0] + diag_iso);
      const double d1 = hp * (P.eps00[1] + diag_iso);
      const double d2 = hp * (P.eps00[2] + diag_iso);
      const double dgel_eigen = -(sig[0] * d0 + sig[1] * d1 + sig[2] * d2 +
                                  2.0 * sig[5] * (hp * P.eps00[5]) +
                                  2.0 * sig[4] * (hp * P.eps00[4]) +
                                  2.0 * sig[3] * (hp * P.eps00[3]));
      // Q = e~^T dC e~ (cuda_kernels.cu:167-185)
      T q[6];
      voigt_apply<T>(P.dC, el, q, P.ortho);
      const double Q = double(el[0] * q[0] + el[1] * q[1] + el[2] * q[2] + T(2) * el[3] * q[3] +
                              T(2) * el[4] * q[4] + T(2) * el[5] * q[5]);
      f += dgel_eigen + 0.5 * hp * Q;
      a.trs[g] = T(sig[0] + sig[1] + sig[2]);
      if (a.sig6) {
#pragma unroll
        for (int c = 0; c < 6; ++c) a.sig6[c * a.n + g] = sig[c];
      }
      red[0] += 0.5 * double(sig[0] * el[0] + sig[1] * el[1] + sig[2] * el[2] +
                             T(2) * (sig[5] * el[5] + sig[4] * el[4] + sig[3] * el[3]));
    }
    if constexpr (NPRE > 2) {
      if (a.phistar) {
        out[0] = T((phi - pre[2]) - (P.L_phi * P.dt) * f);
        return;
      }
    }
    out[0] = T(phi - (P.L_phi * P.dt) * f);
  }
};

// compute_mu_x_kernel (cuda_kernels.cu:1266-1324) at one voxel: x_B =
// sigma_clip(Y) (into xB_out), the composition potential
// mu = c (mu_B - mu_A - c mu_tot dV) - (eps_iso / v_B) tr(sigma) at phiY, and the
// D_mix(h) x (1 - x) term of reduce_sum_DY_kernel (cuda_kernels.cu:3870-3921,
// x the logistic of Y clamped to [1e-12, 1 - 1e-12]) into dsum.
__device__ __forceinline__ double mu_x_point(const PhysParams& P, double Y, double phiY, double trs,
                                             double& xB_out, double& dsum) {
  // sigmoid_from_logit(Y) evaluates the logistic at the clipped Y; the
  // D_mix sum below needs it at Y itself: one exp unless Y was clipped
  const double Yc = Y > P.Y_clip ? P.Y_clip : (Y < -P.Y_clip ? -P.Y_clip : Y);
  const double sc = sigmoid_unclipped(Yc);
  const double xB = sc < P.xB_eps ? P.xB_eps : (sc > 1.0 - P.xB_eps ? 1.0 - P.xB_eps : sc);
  xB_out = xB;
  const double h = h_of_phi(phiY);
  const double c_bulk = c_xB_phi(P, xB, h);
  double mu;
  if (P.dVm == 0.0 && (!P.convex || xB <= 0.09)) {
    // the volume term vanishes and both potentials are on the CALPHAD
    // branch: mu_B - mu_A with G(T) differenced once and the logarithms
    // combined, (G_B0 - G_A0 + RT ln(x/(1-x)) + L (1-2x)) / scale
    const double x = clamp_fraction_eps(xB);
    mu = c_bulk * ((P.G_B0 - P.G_A0 + P.RT * pf_log(x / (1.0 - x)) + P.Lparam * (1.0 - 2.0 * x)) / P.mu_scale);
  } else {
    const double muA = mu_A(P, xB);
    const double muB = mu_B(P, xB);
    const double mu_total = mu_tot_mix(muA, muB, xB, P.mu0, h);
    mu = c_bulk * (muB - muA - c_bulk * mu_total * P.dVm);
  }
  if (P.elastic) mu += -P.eps_iso * trs;
  double x = (Yc == Y) ? sc : sigmoid_unclipped(Y);
  if (x < 1e-12) x = 1e-12;
  if (x > 1.0 - 1e-12) x = 1.0 - 1e-12;
  dsum = D_mix(P, h) * (x * (1.0 - x));
  return mu;
}

// phi^{n+1} = clamp(inv/N, -1e-6, 1+1e-6) (cuda_kernels.cu:1071-1084), then the
// mu_x evaluation (cuda_kernels.cu:1266-1324) and the D_mix x(1-x) sum of
// reduce_sum_DY_kernel (cuda_kernels.cu:3870-3921).
template <typename T>
struct OpPhiFinish {
  struct Args {
    PhysParams P;
    double* phi_new;
    const double* phi_old;
    const double* Y;
    double* xB;
    const T* trs;
    T invN;
    int phiY_is_old;   // zero mode / previous-time-level: mu uses phi^n
    // mixed precision: the unclamped phi* (1: set it to this update's value;
    // 2: the transform holds the increment, phi* += inv/N)
    double* phistar = nullptr;
    int phi_state = 0;
  };
  static constexpr int NRED = 1;
  static constexpr int kinds[1] = {RED_SUM};
  static constexpr int NPRE = sizeof(T) == 4 ? 4 : 3;
  __device__ static void load(const Args& a, size_t g, double* pre) {
    pre[0] = a.phiY_is_old ? a.phi_old[g] : 0.0;
    pre[1] = a.Y[g];
    pre[2] = a.P.elastic ? double(a.trs[g]) : 0.0;
    if constexpr (NPRE > 3) pre[3] = a.phi_state == 2 ? a.phistar[g] : 0.0;
  }
  __device__ static void apply(const Args& a, size_t g, const double* pre, const T* vin, T* out, double* red) {
    const PhysParams& P = a.P;
    double v = double(vin[0]) * double(a.invN);
    if constexpr (NPRE > 3) {
      if (a.phi_state == 2) v += pre[3];
      if (a.phi_state) a.phistar[g] = v;
    }
    if (v < -1e-6) v = -1e-6;
    if (v > 1.0 + 1e-6) v = 1.0 + 1e-6;
    a.phi_new[g] = v;
    const double phiY = a.phiY_is_old ? pre[0] : v;
    double xB, dsum;
    out[0] = T(mu_x_point(P, pre[1], phiY, pre[2], xB, dsum));
    a.xB[g] = xB;
    red[0] += dsum;
  }
};

// The composition potential alone (compute_mu_x_kernel, cuda_kernels.cu:1266-1324)
// at phiY, x_B = sigma_clip(Y) written, and the D_mix x(1 - x) sum of
// reduce_sum_DY_kernel; the mu field goes to the forward transform.  Used
// where the composition update does not follow a phi update in the same pass
// (full-model minimize).
template <typename T>
struct OpMuX {
  struct Args {
    PhysParams P;
    const double* Y;
    const double* phiY;
    double* xB;
    const T* trs;
  };
  static constexpr int NRED = 1;
  static constexpr int kinds[1] = {RED_SUM};
  static constexpr int NPRE = 3;
  __device__ static void load(const Args& a, size_t g, double* pre) {
    pre[0] = a.Y[g];
    pre[1] = a.phiY[g];
    pre[2] = a.P.elastic ? double(a.trs[g]) : 0.0;
  }
  __device__ static void apply(const Args& a, size_t g, const double* pre, const T*, T* out, double* red) {
    double xB, dsum;
    out[0] = T(mu_x_point(a.P, pre[0], pre[1], pre[2], xB, dsum));
    a.xB[g] = xB;
    red[0] += dsum;
  }
};

// J_a = M_eff grad_a mu (compute_flux_single_component_kernel,
// cuda_kernels.cu:1985-2012) and the explicit source of the Y right-hand side
// s = term_h + term_gamma (compute_Y_rhs_sm_tangent_n_kernel 2268-2359 with the
// zero mode, compute_Y_rhs_kernel 2185-2266 without), plus Y^n itself.
template <typename T>
struct OpFlux {
  struct Args {
    PhysParams P;
    const double* phiY;
    const double* xBp;
    const double* phi_n;
    const double* phi_np1;
    const double* Y;
    const double* dYdt;
    T invN;
    int zero_mode;
    int prev_level;      // enable_Y_rhs_previous_time_level
    int no_gamma;        // disable_Y_rhs_gamma_term
    double term_h_scale;
    int increment;       // emit (J, s) only: the update forms the increment of Y (NCO = 4)
  };
  static constexpr int NRED = 0;
  static constexpr int kinds[1] = {0};
  static constexpr int NPRE = 6;
  static constexpr int LDS_BUDGET = 24 * 1024;   // smaller tiles than the default: measured faster
  __device__ static void load(const Args& a, size_t g, double* pre) {
    pre[0] = a.phiY[g];
    pre[1] = a.xBp[g];
    pre[2] = a.phi_n[g];
    pre[3] = a.phi_np1[g];
    pre[4] = a.Y[g];
    pre[5] = a.dYdt[g];
  }
  __device__ static void apply(const Args& a, size_t g, const double* pre, const T* vin, T* out, double*) {
    const PhysParams& P = a.P;
    const double hY = h_of_phi(pre[0]);
    const double Dm = D_mix(P, hY);
    const double G = gamma_thermo(P, pre[1], hY);
    const double Meff = stabilized_meff(Dm, G);
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = T(Meff * (double(vin[c]) * double(a.invN)));
    const double Y = pre[4];
    double x = sigmoid_unclipped(Y);
    if (x < 1e-12) x = 1e-12;
    if (x > 1.0 - 1e-12) x = 1.0 - 1e-12;
    const double lg = x * (1.0 - x);
    const double pn = pre[2];
    const double pn1 = pre[3];
    double s;
    if (a.zero_mode) {
      const double h_n = h_of_phi(pn);
      const double term_h = h_prime_of_phi(pn) * ((pn1 - pn) / P.dt) * (P.v_B - x);
      const double gamma_local = (1.0 - h_n) * lg - 1.0;
      s = term_h + gamma_local * pre[5];
    } else {
      const double h_new = h_of_phi(pn1);
      const double h_old = h_of_phi(pn);
      const double term_h = a.term_h_scale * ((h_new - h_old) / P.dt) * (P.v_B - x);
      const double hx = a.prev_level ? h_old : h_new;
      const double gamma_local = (1.0 - hx) * lg - 1.0;
      const double term_gamma = a.no_gamma ? 0.0 : gamma_local * pre[5];
      s = term_h + term_gamma;
    }
    if (a.increment) {
      out[3] = T(s);
    } else {
      out[3] = T(Y);
      out[4] = T(s);
    }
  }
};

// Y* = inv/N before the zero-mode shift (copy_scaled_pf_Y_kernel
// cuda_kernels.cu:3401-3407), with its min/max for the bracket.  With Ybase
// the transform holds the increment of the step and Y* = Y^n + inv/N.
template <typename T>
struct OpYStar {
  struct Args {
    double* Ystar;
    T invN;
    const double* Ybase;   // Y^n (increment form) or null
  };
  static constexpr int NRED = 2;
  static constexpr int kinds[2] = {RED_MIN, RED_MAX};
  static constexpr int NPRE = 1;
  __device__ static void load(const Args& a, size_t g, double* pre) { pre[0] = a.Ybase ? a.Ybase[g] : 0.0; }
  __device__ static void apply(const Args& a, size_t g, const double* pre, const T* vin, T*, double* red) {
    const double y = pre[0] + double(vin[0]) * double(a.invN);
    a.Ystar[g] = y;
    red[0] = fmin(red[0], y);
    red[1] = fmax(red[1], y);
  }
};

// Without the zero mode: Y = clamp(inv/N), x_B = sigma_clip(Y)
// (Y_normalize_and_clamp_kernel cuda_kernels.cu:2554-2573) and dY/dt
// (update_dY_dt_prev_kernel cuda_kernels.cu:3545-3556).
template <typename T>
struct OpYFinal {
  struct Args {
    PhysParams P;
    double* Y;
    double* xB;
    double* dYdt;
    T invN;
    double Y_upper_cap;
  };
  static constexpr int NRED = 0;
  static constexpr int kinds[1] = {0};
  static constexpr int NPRE = 1;
  __device__ static void load(const Args& a, size_t g, double* pre) { pre[0] = a.Y[g]; }
  __device__ static void apply(const Args& a, size_t g, const double* pre, const T* vin, T*, double*) {
    double y = double(vin[0]) * double(a.invN);
    if (y > a.Y_upper_cap) y = a.Y_upper_cap;
    if (y < -a.P.Y_clip) y = -a.P.Y_clip;
    const double yn = pre[0];
    a.Y[g] = y;
    a.xB[g] = sigmoid_from_logit(y, a.P.Y_clip, a.P.xB_eps);
    a.dYdt[g] = (y - yn) / a.P.dt;
  }
};

// ----------------------------------------------------------------- k space
struct KDims {
  int Ax, Ay, Az;
  const double* kx;   // [Ax] wrapped k of the compact x index
  const double* ky;   // [Ay]
  const double* kz;   // [Az]
  int Nz_even;
  int Azr;            // retained kz count (Az is the row stride; iz >= Azr is padding)
};

// Compact index of the j-th retained mode (rows of Azr retained kz laid out
// with stride Az).  Arrays the x pass reads or writes live at these slots;
// solver state that only k-space kernels touch (the FP64 displacement and
// residual, the Anderson increments and history, the FP64 phi spectrum) is
// stored by j itself, with no kz padding to read around.
__device__ __forceinline__ long kslot(const KDims& d, long j) {
  const unsigned ju = (unsigned)j;
  const unsigned row = ju / (unsigned)d.Azr;
  return (long)row * d.Az + (long)(ju - row * (unsigned)d.Azr);
}

__device__ __forceinline__ void kvec(const KDims& d, long c, double& kx, double& ky, double& kz,
                                     int& iz) {
  // 32-bit index arithmetic: compact spectra stay below 2^32 entries
  const unsigned cu = (unsigned)c;
  const unsigned r = cu / (unsigned)d.Az;
  iz = (int)(cu - r * (unsigned)d.Az);
  const unsigned ix = r / (unsigned)d.Ay;
  const int iy = (int)(r - ix * (unsigned)d.Ay);
  kx = d.kx[ix];
  ky = d.ky[iy];
  kz = d.kz[iz];
}

// Acoustic tensor K_ij = C_ipjq k_p k_q (cuda_kernels.cu:6818-6870), as
// K = {K11, K12, K13, K22, K23, K33}.
__device__ __forceinline__ void acoustic_tensor(const double* C, double kx, double ky, double kz, double* K) {
  const double kx2 = kx * kx, ky2 = ky * ky, kz2 = kz * kz;
#define CC(i, j) C[vidx(i - 1, j - 1)]
  K[0] = kx2 * CC(1, 1) + 2.0 * kx * (kz * CC(1, 5) + ky * CC(1, 6)) + kz2 * CC(5, 5) +
         2.0 * ky * kz * CC(5, 6) + ky2 * CC(6, 6);
  K[3] = ky2 * CC(2, 2) + 2.0 * ky * (kz * CC(2, 4) + kx * CC(2, 6)) + kz2 * CC(4, 4) +
         2.0 * kx * kz * CC(4, 6) + kx2 * CC(6, 6);
  K[5] = kz2 * CC(3, 3) + 2.0 * kz * (ky * CC(3, 4) + kx * CC(3, 5)) + ky2 * CC(4, 4) +
         2.0 * kx * ky * CC(4, 5) + kx2 * CC(5, 5);
  K[2] = kx2 * CC(1, 5) + kz2 * CC(3, 5) + ky * kz * (CC(3, 6) + CC(4, 5)) + ky2 * CC(4, 6) +
         kx * kz * (CC(1, 3) + CC(5, 5)) + kx * ky * (CC(1, 4) + CC(5, 6));
  K[4] = ky2 * CC(2, 4) + kz2 * CC(3, 4) + ky * kz * (CC(2, 3) + CC(4, 4)) + kx * kz * (CC(3, 6) + CC(4, 5)) +
         kx * ky * (CC(2, 5) + CC(4, 6)) + kx2 * CC(5, 6);
  K[1] = kx2 * CC(1, 6) + ky2 * CC(2, 6) + kz2 * CC(4, 5) + ky * kz * (CC(2, 5) + CC(4, 6)) +
         kx * kz * (CC(1, 4) + CC(5, 6)) + kx * ky * (CC(1, 2) + CC(6, 6));
#undef CC
}

// The acoustic tensor and its inverse by cofactors (cuda_kernels.cu:6484-6633,
// 6818-6878).  Returns false when |det| < 1e-30.
__device__ __forceinline__ bool green_tensor(const double* C, double kx, double ky, double kz,
                                             double* G) {
  double K[6];
  acoustic_tensor(C, kx, ky, kz, K);
  const double K11 = K[0], K12 = K[1], K13 = K[2], K22 = K[3], K23 = K[4], K33 = K[5];
  const double det = K13 * (-K13 * K22 + K23 * K12) - K23 * (K23 * K11 - K13 * K12) +
                     K33 * (K22 * K11 - K12 * K12);
  if (fabs(det) < 1e-30) return false;
  G[0] = (-K23 * K23 + K33 * K22) / det;   // 11
  G[1] = (K23 * K13 - K33 * K12) / det;    // 12
  G[2] = (-K13 * K22 + K23 * K12) / det;   // 13
  G[3] = (-K13 * K13 + K33 * K11) / det;   // 22
  G[4] = (-K23 * K11 + K13 * K12) / det;   // 23
  G[5] = (K22 * K11 - K12 * K12) / det;    // 33
  return true;
}

// eps_ij(k) = i (k_i u_j + k_j u_i)/2 (cuda_kernels.cu:6673-6726)
template <typename T>
__device__ __forceinline__ void strain_of(double kx, double ky, double kz, const cx<double>* u,
                                          cx<T>* const* e, long c) {
  auto ik = [](double k, cx<double> v) { return cx<double>{-k * v.im, k * v.re}; };
  cx<double> exx = ik(kx, u[0]);
  cx<double> eyy = ik(ky, u[1]);
  cx<double> ezz = ik(kz, u[2]);
  cx<double> a = ik(ky, u[2]), b = ik(kz, u[1]);
  cx<double> eyz = {0.5 * (a.re + b.re), 0.5 * (a.im + b.im)};
  a = ik(kz, u[0]);
  b = ik(kx, u[2]);
  cx<double> exz = {0.5 * (a.re + b.re), 0.5 * (a.im + b.im)};
  a = ik(ky, u[0]);
  b = ik(kx, u[1]);
  cx<double> exy = {0.5 * (a.re + b.re), 0.5 * (a.im + b.im)};
  e[0][c] = cx<T>{T(exx.re), T(exx.im)};
  e[1][c] = cx<T>{T(eyy.re), T(eyy.im)};
  e[2][c] = cx<T>{T(ezz.re), T(ezz.im)};
  e[3][c] = cx<T>{T(eyz.re), T(eyz.im)};
  e[4][c] = cx<T>{T(exz.re), T(exz.im)};
  e[5][c] = cx<T>{T(exy.re), T(exy.im)};
}

template <typename T>
struct GreenArgs {
  KDims d;
  long n;
  const cx<T>* h[6];     // h_ij spectra
  const cx<T>* u_old[3]; // previous iterate (residual)
  cx<T>* u_new[3];       // out
  cx<T>* eps[6];         // out: strain of u_new
  double* partials;      // [grid][2] (|du|^2, |u|^2) hermitian-weighted, or null
  double C[21];
  int have_old;
};

// u(k) = -i Omega(k) (k . h(k)); k = 0 and singular det give 0.
template <typename T>
__global__ void __launch_bounds__(NT) kop_green(GreenArgs<T> ga) {
  double acc[2] = {0.0, 0.0};
  const long nact_ = (ga.n / ga.d.Az) * ga.d.Azr;   // retained modes only (no kz padding)
  for (long j_ = blockIdx.x * (long)NT + threadIdx.x; j_ < nact_; j_ += (long)gridDim.x * NT) {
    const long c = kslot(ga.d, j_);
    double kx, ky, kz;
    int iz;
    kvec(ga.d, c, kx, ky, kz, iz);
    cx<double> hh[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) hh[i] = cx<double>{double(ga.h[i][c].re), double(ga.h[i][c].im)};
    double G[6];
    cx<double> u[3] = {{0, 0}, {0, 0}, {0, 0}};
    if (c != 0 && green_tensor(ga.C, kx, ky, kz, G)) {
      // t = k . h   (rows of the symmetric h)
      cx<double> t1 = {kx * hh[0].re + ky * hh[5].re + kz * hh[4].re, kx * hh[0].im + ky * hh[5].im + kz * hh[4].im};
      cx<double> t2 = {kx * hh[5].re + ky * hh[1].re + kz * hh[3].re, kx * hh[5].im + ky * hh[1].im + kz * hh[3].im};
      cx<double> t3 = {kx * hh[4].re + ky * hh[3].re + kz * hh[2].re, kx * hh[4].im + ky * hh[3].im + kz * hh[2].im};
      const double g[3][3] = {{G[0], G[1], G[2]}, {G[1], G[3], G[4]}, {G[2], G[4], G[5]}};
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        cx<double> s = {g[i][0] * t1.re + g[i][1] * t2.re + g[i][2] * t3.re,
                        g[i][0] * t1.im + g[i][1] * t2.im + g[i][2] * t3.im};
        u[i] = cx<double>{s.im, -s.re};   // -i * s
      }
    }
    // the previous iterate is read before u_new is stored (the pointers may
    // alias as far as the compiler knows; a load after a store waits for it)
    cx<T> uo[3];
    if (ga.partials && ga.have_old) {
#pragma unroll
      for (int i = 0; i < 3; ++i) uo[i] = ga.u_old[i][c];
    }
    cx<double> uq[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      cx<T> q = {T(u[i].re), T(u[i].im)};
      ga.u_new[i][c] = q;
      uq[i] = cx<double>{double(q.re), double(q.im)};
    }
    if (ga.partials) {
      // Hermitian weight of compute_elastic_displacement_residual_kernel
      // (cuda_kernels.cu:6436-6467): 1 on the kz = 0 plane, 2 elsewhere (the
      // Nyquist plane is never retained).
      const double w = (iz >= ga.d.Azr) ? 0.0 : ((iz == 0) ? 1.0 : 2.0);
      double dsum = 0.0, rsum = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double dr = uq[i].re, di = uq[i].im;
        if (ga.have_old) {
          dr -= double(uo[i].re);
          di -= double(uo[i].im);
        }
        dsum += dr * dr + di * di;
        rsum += uq[i].re * uq[i].re + uq[i].im * uq[i].im;
      }
      acc[0] += w * dsum;
      acc[1] += w * rsum;
    }
    strain_of<T>(kx, ky, kz, uq, ga.eps, c);
  }
  if (ga.partials) {
    int kinds[2] = {RED_SUM, RED_SUM};
    block_reduce_store<2>(acc, kinds, ga.partials + (size_t)blockIdx.x * 2);
  }
}

// Mixed-precision Green update with an FP64 accumulated displacement U:
//   mode 0 (cold):      U = G(h),             eps = strain(U)
//   mode 1 (full map):  d = G(h) - U, U += d, eps = strain(d)
//   mode 2 (increment): d = G(h),     U += d, eps = strain(d)   (h from the linear part)
//   mode 3 (legacy):    d = G(h) - U, U += d, eps = strain(U)
// The residual sums w|d|^2 and w|U|^2 are FP64, so |d|/|U| has no FP32 floor.
template <typename T>
struct GreenAccArgs {
  KDims d;
  long n;
  const cx<T>* h[6];
  cx<double>* U[3];
  cx<T>* eps[6];
  double* partials;
  double C[21];
  int mode;
};

template <typename T>
__device__ __forceinline__ void kop_green_acc_body(const GreenAccArgs<T>& ga, const B3 blockIdx, const B3 gridDim) {
  double acc[2] = {0.0, 0.0};
  const long nact_ = (ga.n / ga.d.Az) * ga.d.Azr;   // retained modes only (no kz padding)
  for (long j_ = blockIdx.x * (long)NT + threadIdx.x; j_ < nact_; j_ += (long)gridDim.x * NT) {
    const long c = kslot(ga.d, j_);
    double kx, ky, kz;
    int iz;
    kvec(ga.d, c, kx, ky, kz, iz);
    cx<double> hh[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) hh[i] = cx<double>{double(ga.h[i][c].re), double(ga.h[i][c].im)};
    // U is read before any store of the iteration (aliasing as above)
    cx<double> uo3[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) uo3[i] = ga.mode == 0 ? cx<double>{0.0, 0.0} : ga.U[i][j_];
    double G[6];
    cx<double> g[3] = {{0, 0}, {0, 0}, {0, 0}};
    if (c != 0 && green_tensor(ga.C, kx, ky, kz, G)) {
      cx<double> t1 = {kx * hh[0].re + ky * hh[5].re + kz * hh[4].re, kx * hh[0].im + ky * hh[5].im + kz * hh[4].im};
      cx<double> t2 = {kx * hh[5].re + ky * hh[1].re + kz * hh[3].re, kx * hh[5].im + ky * hh[1].im + kz * hh[3].im};
      cx<double> t3 = {kx * hh[4].re + ky * hh[3].re + kz * hh[2].re, kx * hh[4].im + ky * hh[3].im + kz * hh[2].im};
      const double gm[3][3] = {{G[0], G[1], G[2]}, {G[1], G[3], G[4]}, {G[2], G[4], G[5]}};
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        cx<double> s = {gm[i][0] * t1.re + gm[i][1] * t2.re + gm[i][2] * t3.re,
                        gm[i][0] * t1.im + gm[i][1] * t2.im + gm[i][2] * t3.im};
        g[i] = cx<double>{s.im, -s.re};
      }
    }
    cx<double> dlt[3], un[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const cx<double> uo = uo3[i];
      if (ga.mode == 0) {
        un[i] = g[i];
        dlt[i] = g[i];
      } else if (ga.mode == 2) {
        dlt[i] = g[i];
        un[i] = cx<double>{uo.re + g[i].re, uo.im + g[i].im};
      } else {
        dlt[i] = cx<double>{g[i].re - uo.re, g[i].im - uo.im};
        un[i] = g[i];
      }
      ga.U[i][j_] = un[i];
    }
    if (ga.partials) {
      const double w = (iz >= ga.d.Azr) ? 0.0 : ((iz == 0) ? 1.0 : 2.0);
      double dsum = 0.0, rsum = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        dsum += dlt[i].re * dlt[i].re + dlt[i].im * dlt[i].im;
        rsum += un[i].re * un[i].re + un[i].im * un[i].im;
      }
      acc[0] += w * dsum;
      acc[1] += w * rsum;
    }
    strain_of<T>(kx, ky, kz, (ga.mode == 0 || ga.mode == 3) ? un : dlt, ga.eps, c);
  }
  if (ga.partials) {
    int kinds[2] = {RED_SUM, RED_SUM};
    block_reduce_store<2>(acc, kinds, ga.partials + (size_t)blockIdx.x * 2);
  }
}
template <typename T>
__global__ void __launch_bounds__(NT) kop_green_acc(GreenAccArgs<T> ga) {
  kop_green_acc_body<T>(ga, PF_BIDX, PF_GDIM);
}
template <typename T>
struct KGreenAcc {
  struct Args {
    GreenAccArgs<T> ga;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { kop_green_acc_body<T>(a.ga, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL((kop_green_acc<T>), dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.ga);
  }
};


// ---------------------------------------------------------------- accelerated solver
// Anderson-mixed fixed point on the FP64 displacement U with residual g = F(U) - U.
// kop_green_aa evaluates one Green application:
//   mode 0 (full map at U):          g  = G(h) - U
//   mode 1 (linear map on increment d): g' = g + G(h) - d, dg = g' - g  (stored in slot `slot`)
// and reduces (Hermitian-weighted, FP64):
//   [0] |g'|^2, [1] |U + g'|^2, [2+i] <dg, v_i>, [2+AA_MAX+i] <g', v_i>
// with v_i the dg history (slot `slot` replaced by the new dg).
constexpr int AA_MAX = 4;
constexpr int AA_NRED = 2 + 2 * AA_MAX;

// Device-resident control block of the accelerated solver (doubles), written
// by the one-thread kernels that take the loop's decisions; the iteration's
// kernels read their history bookkeeping from it and return at once when
// AC_ACTIVE is 0 (iterations enqueued past convergence).
enum AaCtl {
  AC_ACTIVE = 0, AC_IT, AC_NH, AC_NEXT, AC_RES, AC_CONV, AC_STATUS,
  AC_GAM = 8,                          // AA_MAX mixing coefficients
  AC_H = AC_GAM + AA_MAX,              // AA_MAX x AA_MAX Gram matrix of the dg history
  AC_B = AC_H + AA_MAX * AA_MAX,       // AA_MAX right-hand side <g, dg_j>
  AC_SUMS = AC_B + AA_MAX,             // AA_NRED sums of the last Green application
  AC_COUNT = AC_SUMS + AA_NRED
};

template <typename T>
struct GreenAAArgs {
  KDims d;
  long n;
  const cx<T>* h[6];
  cx<double>* U[3];   // with ctl, mode 1: U += d happens here (kop_aa_update leaves U alone)
  cx<double>* g[3];
  const cx<T>* dinc[3];
  cx<T>* dg[AA_MAX][3];
  int nhist;        // valid dg slots after this update (including `slot`)
  int slot;         // slot receiving the new dg (mode 1)
  int mode;
  double* partials;
  double C[21];
  const double* ctl;   // control block: nhist/slot from it (null: the fields above)
  int m;               // history depth (with ctl)
  // with ctl and m > 0: the increment is read from du[slot], which then
  // receives du + dg (the only combination the Anderson step uses)
  cx<T>* du[AA_MAX][3];
};

__device__ __forceinline__ double wdot(const cx<double>* a, const cx<double>* b) {
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) s += a[i].re * b[i].re + a[i].im * b[i].im;
  return s;
}

template <typename T>
__device__ __forceinline__ void green_apply_point(const double* C, double kx, double ky, double kz,
                                                  const cx<double>* hh, cx<double>* u) {
  double G[6];
  u[0] = u[1] = u[2] = cx<double>{0.0, 0.0};
  if (!green_tensor(C, kx, ky, kz, G)) return;
  cx<double> t1 = {kx * hh[0].re + ky * hh[5].re + kz * hh[4].re, kx * hh[0].im + ky * hh[5].im + kz * hh[4].im};
  cx<double> t2 = {kx * hh[5].re + ky * hh[1].re + kz * hh[3].re, kx * hh[5].im + ky * hh[1].im + kz * hh[3].im};
  cx<double> t3 = {kx * hh[4].re + ky * hh[3].re + kz * hh[2].re, kx * hh[4].im + ky * hh[3].im + kz * hh[2].im};
  const double gm[3][3] = {{G[0], G[1], G[2]}, {G[1], G[3], G[4]}, {G[2], G[4], G[5]}};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    cx<double> s = {gm[i][0] * t1.re + gm[i][1] * t2.re + gm[i][2] * t3.re,
                    gm[i][0] * t1.im + gm[i][1] * t2.im + gm[i][2] * t3.im};
    u[i] = cx<double>{s.im, -s.re};
  }
}

template <typename T>
__device__ __forceinline__ void kop_green_aa_body(const GreenAAArgs<T>& a, const B3 blockIdx, const B3 gridDim) {
  int nhist = a.nhist, slot = a.slot;
  if (a.ctl) {
    if (a.ctl[AC_ACTIVE] == 0.0) return;
    const int nh = (int)a.ctl[AC_NH];
    const bool grow = a.mode == 1 && a.m > 0;
    nhist = grow ? (nh + 1 < a.m ? nh + 1 : a.m) : nh;
    slot = grow ? (int)a.ctl[AC_NEXT] : -1;
  }
  // the increment just applied: the kop_aa_update copy in du[slot] when a
  // history is kept under the control block, dinc otherwise
  const bool hist = a.ctl && slot >= 0;
  const cx<T>* const* dsrc = a.dinc;
  if (hist) {
#pragma unroll
    for (int j = 0; j < AA_MAX; ++j)
      if (j == slot) dsrc = a.du[j];
  }
  double acc[AA_NRED];
#pragma unroll
  for (int r = 0; r < AA_NRED; ++r) acc[r] = 0.0;
  const long nact_ = (a.n / a.d.Az) * a.d.Azr;   // retained modes only (no kz padding)
  for (long j_ = blockIdx.x * (long)NT + threadIdx.x; j_ < nact_; j_ += (long)gridDim.x * NT) {
    const long c = kslot(a.d, j_);
    double kx, ky, kz;
    int iz;
    kvec(a.d, c, kx, ky, kz, iz);
    // Every load of the iteration is issued before its first store: the
    // argument pointers may alias as far as the compiler knows, so a load
    // written after a store would wait for it, one memory round trip each.
    cx<T> hT[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) hT[i] = a.h[i][c];
    cx<double> U[3], go[3];
    cx<T> dT[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) U[i] = a.U[i][j_];
    if (a.mode != 0) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        go[i] = a.g[i][j_];
        dT[i] = dsrc[i][j_];
      }
    } else {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        go[i] = cx<double>{0.0, 0.0};
        dT[i] = cx<T>{T(0), T(0)};
      }
    }
    // the kept dg history (slot `slot` is replaced by this update's dg)
    cx<T> hv[AA_MAX][3];
    if (a.mode == 1) {
#pragma unroll
      for (int j = 0; j < AA_MAX; ++j)
        if (j < nhist && j != slot) {
#pragma unroll
          for (int i = 0; i < 3; ++i) hv[j][i] = a.dg[j][i][j_];
        }
    }
    cx<double> hh[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) hh[i] = cx<double>{double(hT[i].re), double(hT[i].im)};
    cx<double> Gh[3];
    if (c != 0) green_apply_point<T>(a.C, kx, ky, kz, hh, Gh);
    else Gh[0] = Gh[1] = Gh[2] = cx<double>{0.0, 0.0};
    cx<double> gn[3], dgn[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (a.mode == 0) {
        gn[i] = cx<double>{Gh[i].re - U[i].re, Gh[i].im - U[i].im};
        dgn[i] = cx<double>{0.0, 0.0};
      } else {
        const cx<double> di = {double(dT[i].re), double(dT[i].im)};
        if (a.ctl) {
          // the increment applied by this iteration's kop_aa_update
          U[i] = cx<double>{U[i].re + di.re, U[i].im + di.im};
          a.U[i][j_] = U[i];
        }
        gn[i] = cx<double>{go[i].re + Gh[i].re - di.re, go[i].im + Gh[i].im - di.im};
        dgn[i] = cx<double>{gn[i].re - go[i].re, gn[i].im - go[i].im};
      }
      a.g[i][j_] = gn[i];
    }
    const double w = (iz >= a.d.Azr) ? 0.0 : ((iz == 0) ? 1.0 : 2.0);
    cx<double> Fu[3] = {{U[0].re + gn[0].re, U[0].im + gn[0].im},
                        {U[1].re + gn[1].re, U[1].im + gn[1].im},
                        {U[2].re + gn[2].re, U[2].im + gn[2].im}};
    acc[0] += w * wdot(gn, gn);
    acc[1] += w * wdot(Fu, Fu);
    if (a.mode == 1) {
      // the T-rounded residual difference (kept in registers: q0..q2)
      const cx<T> q0 = {T(dgn[0].re), T(dgn[0].im)};
      const cx<T> q1 = {T(dgn[1].re), T(dgn[1].im)};
      const cx<T> q2 = {T(dgn[2].re), T(dgn[2].im)};
      const cx<double> dq[3] = {{double(q0.re), double(q0.im)}, {double(q1.re), double(q1.im)},
                                {double(q2.re), double(q2.im)}};
#pragma unroll
      for (int j = 0; j < AA_MAX; ++j) {
        if (j == slot) {
          a.dg[j][0][j_] = q0;
          a.dg[j][1][j_] = q1;
          a.dg[j][2][j_] = q2;
          if (hist) {
            a.du[j][0][j_] = cx<T>{T(double(dT[0].re) + double(q0.re)), T(double(dT[0].im) + double(q0.im))};
            a.du[j][1][j_] = cx<T>{T(double(dT[1].re) + double(q1.re)), T(double(dT[1].im) + double(q1.im))};
            a.du[j][2][j_] = cx<T>{T(double(dT[2].re) + double(q2.re)), T(double(dT[2].im) + double(q2.im))};
          }
        }
      }
#pragma unroll
      for (int j = 0; j < AA_MAX; ++j) {
        if (j < nhist) {
          cx<double> v[3];
          if (j == slot) {
            v[0] = dq[0];
            v[1] = dq[1];
            v[2] = dq[2];
          } else {
#pragma unroll
            for (int i = 0; i < 3; ++i) v[i] = cx<double>{double(hv[j][i].re), double(hv[j][i].im)};
          }
          acc[2 + j] += w * wdot(dq, v);
          acc[2 + AA_MAX + j] += w * wdot(gn, v);
        }
      }
    }
  }
  int kinds[AA_NRED];
#pragma unroll
  for (int r = 0; r < AA_NRED; ++r) kinds[r] = RED_SUM;
  block_reduce_store<AA_NRED>(acc, kinds, a.partials + (size_t)blockIdx.x * AA_NRED);
}
template <typename T>
__global__ void __launch_bounds__(NT) kop_green_aa(GreenAAArgs<T> a) {
  kop_green_aa_body<T>(a, PF_BIDX, PF_GDIM);
}
template <typename T>
struct KGreenAA {
  struct Args {
    GreenAAArgs<T> a;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { kop_green_aa_body<T>(a.a, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL((kop_green_aa<T>), dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.a);
  }
};


// d = g - sum_j gamma_j (du_j + dg_j);  U += d (the T-rounded d);  du[slot] = d;
// eps = strain(d) for the next linear-map application.
// Under the control block the history slots hold du_j + dg_j (summed by
// kop_green_aa once dg_j is known), so one array is read per history entry,
// and U += d is left to the Green update, which reads U and d anyway.
template <typename T>
struct AAUpdateArgs {
  KDims d;
  long n;
  const cx<double>* g[3];
  cx<double>* U[3];
  cx<T>* du[AA_MAX][3];
  const cx<T>* dg[AA_MAX][3];
  cx<T>* dinc[3];
  cx<T>* eps[6];
  double gamma[AA_MAX];
  int nhist;
  int slot;   // du slot receiving d (or -1 when no history is kept)
  const double* ctl;   // control block: gamma/nhist/slot from it (null: the fields above)
  int m;               // history depth (with ctl)
};

template <typename T>
__device__ __forceinline__ void kop_aa_update_body(const AAUpdateArgs<T>& a, const B3 blockIdx, const B3 gridDim) {
  int nhist = a.nhist, slot = a.slot;
  double gamma[AA_MAX];
#pragma unroll
  for (int j = 0; j < AA_MAX; ++j) gamma[j] = a.gamma[j];
  if (a.ctl) {
    if (a.ctl[AC_ACTIVE] == 0.0) return;
    nhist = (int)a.ctl[AC_NH];
    slot = a.m > 0 ? (int)a.ctl[AC_NEXT] : -1;
#pragma unroll
    for (int j = 0; j < AA_MAX; ++j) gamma[j] = a.ctl[AC_GAM + j];
  }
  const long nact_ = (a.n / a.d.Az) * a.d.Azr;   // retained modes only (no kz padding)
  for (long j_ = blockIdx.x * (long)NT + threadIdx.x; j_ < nact_; j_ += (long)gridDim.x * NT) {
    const long c = kslot(a.d, j_);
    double kx, ky, kz;
    int iz;
    kvec(a.d, c, kx, ky, kz, iz);
    // all loads first (the pointers may alias as far as the compiler knows)
    cx<double> gv[3], uo[3];
    cx<T> hu[AA_MAX][3], hg[AA_MAX][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) gv[i] = a.g[i][j_];
#pragma unroll
    for (int j = 0; j < AA_MAX; ++j)
      if (j < nhist) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          hu[j][i] = a.du[j][i][j_];
          if (!a.ctl) hg[j][i] = a.dg[j][i][j_];
        }
      }
    if (!a.ctl) {
#pragma unroll
      for (int i = 0; i < 3; ++i) uo[i] = a.U[i][j_];
    }
    cx<double> dd[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      cx<double> v = gv[i];
#pragma unroll
      for (int j = 0; j < AA_MAX; ++j) {
        if (j < nhist) {
          if (a.ctl) {
            // the history slot holds du + dg (kop_green_aa)
            v.re -= gamma[j] * double(hu[j][i].re);
            v.im -= gamma[j] * double(hu[j][i].im);
          } else {
            v.re -= gamma[j] * (double(hu[j][i].re) + double(hg[j][i].re));
            v.im -= gamma[j] * (double(hu[j][i].im) + double(hg[j][i].im));
          }
        }
      }
      const cx<T> q = {T(v.re), T(v.im)};
      dd[i] = cx<double>{double(q.re), double(q.im)};
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const cx<T> q = {T(dd[i].re), T(dd[i].im)};
      if (!(a.ctl && slot >= 0)) a.dinc[i][j_] = q;   // with a history, du[slot] holds it
#pragma unroll
      for (int j = 0; j < AA_MAX; ++j)
        if (j == slot) a.du[j][i][j_] = q;
      if (!a.ctl)   // under the control block kop_green_aa adds d to U
        a.U[i][j_] = cx<double>{uo[i].re + dd[i].re, uo[i].im + dd[i].im};
    }
    strain_of<T>(kx, ky, kz, dd, a.eps, c);
  }
}
template <typename T>
__global__ void __launch_bounds__(NT) kop_aa_update(AAUpdateArgs<T> a) {
  kop_aa_update_body<T>(a, PF_BIDX, PF_GDIM);
}
template <typename T>
struct KAaUpdate {
  struct Args {
    AAUpdateArgs<T> a;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { kop_aa_update_body<T>(a.a, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL((kop_aa_update<T>), dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.a);
  }
};


// U <- U + g (the accepted result F(U)), eps = strain(U)
template <typename T>
__device__ __forceinline__ void kop_aa_final_body(KDims d, long n, cx<double>* U0, cx<double>* U1, cx<double>* U2, const cx<double>* g0, const cx<double>* g1, const cx<double>* g2, cx<T>* e0, cx<T>* e1, cx<T>* e2, cx<T>* e3, cx<T>* e4, cx<T>* e5, const B3 blockIdx, const B3 gridDim) {
  cx<double>* U[3] = {U0, U1, U2};
  const cx<double>* g[3] = {g0, g1, g2};
  cx<T>* eps[6] = {e0, e1, e2, e3, e4, e5};
  const long nact_ = (n / d.Az) * d.Azr;   // retained modes only (no kz padding)
  for (long j_ = blockIdx.x * (long)NT + threadIdx.x; j_ < nact_; j_ += (long)gridDim.x * NT) {
    const long c = kslot(d, j_);
    double kx, ky, kz;
    int iz;
    kvec(d, c, kx, ky, kz, iz);
    // all loads before the first store (the pointers may alias as far as the
    // compiler knows, and a load after a store would wait for it)
    cx<double> ua[3], gb[3], u[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      ua[i] = U[i][j_];
      gb[i] = g[i][j_];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      u[i] = cx<double>{ua[i].re + gb[i].re, ua[i].im + gb[i].im};
      U[i][j_] = u[i];
    }
    strain_of<T>(kx, ky, kz, u, eps, c);
  }
}
template <typename T>
__global__ void __launch_bounds__(NT) kop_aa_final(KDims d, long n, cx<double>* U0, cx<double>* U1, cx<double>* U2, const cx<double>* g0, const cx<double>* g1, const cx<double>* g2, cx<T>* e0, cx<T>* e1, cx<T>* e2, cx<T>* e3, cx<T>* e4, cx<T>* e5) {
  kop_aa_final_body<T>(d, n, U0, U1, U2, g0, g1, g2, e0, e1, e2, e3, e4, e5, PF_BIDX, PF_GDIM);
}
template <typename T>
struct KAaFinal {
  struct Args {
    KDims d;
    long n;
    cx<double>* U0;
    cx<double>* U1;
    cx<double>* U2;
    const cx<double>* g0;
    const cx<double>* g1;
    const cx<double>* g2;
    cx<T>* e0;
    cx<T>* e1;
    cx<T>* e2;
    cx<T>* e3;
    cx<T>* e4;
    cx<T>* e5;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { kop_aa_final_body<T>(a.d, a.n, a.U0, a.U1, a.U2, a.g0, a.g1, a.g2, a.e0, a.e1, a.e2, a.e3, a.e4, a.e5, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL((kop_aa_final<T>), dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.d, a.n, a.U0, a.U1, a.U2, a.g0, a.g1, a.g2, a.e0, a.e1, a.e2, a.e3, a.e4, a.e5);
  }
};


template <typename T, typename UT = T>
struct StrainArgs {
  KDims d;
  long n;
  const cx<UT>* u[3];
  cx<T>* eps[6];
  // optional warm start in the same pass (FP64 state): U = sum_k cf[k] H[k]
  // over the nh latest converged states, written to Uout and differentiated
  cx<double>* Uout[3];
  const cx<double>* H[3][3];   // [k][component]
  double cf[3];
  int nh;
  int ustate;   // u is solver state (retained-mode index), not a k-space array
};
template <typename T, typename UT = T>
__device__ __forceinline__ void kop_strain_body(const StrainArgs<T, UT>& sa, const B3 blockIdx, const B3 gridDim) {
  const long nact_ = (sa.n / sa.d.Az) * sa.d.Azr;   // retained modes only (no kz padding)
  for (long j_ = blockIdx.x * (long)NT + threadIdx.x; j_ < nact_; j_ += (long)gridDim.x * NT) {
    const long c = kslot(sa.d, j_);
    double kx, ky, kz;
    int iz;
    kvec(sa.d, c, kx, ky, kz, iz);
    cx<double> u[3];
    if (sa.nh > 0) {
      // every history load before the first store of Uout (aliasing as above)
      cx<double> hk[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        hk[0][i] = sa.H[0][i][j_];
        if (sa.nh > 1) hk[1][i] = sa.H[1][i][j_];
        if (sa.nh > 2) hk[2][i] = sa.H[2][i][j_];
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        cx<double> v = {sa.cf[0] * hk[0][i].re, sa.cf[0] * hk[0][i].im};
        if (sa.nh > 1) {
          v.re += sa.cf[1] * hk[1][i].re;
          v.im += sa.cf[1] * hk[1][i].im;
        }
        if (sa.nh > 2) {
          v.re += sa.cf[2] * hk[2][i].re;
          v.im += sa.cf[2] * hk[2][i].im;
        }
        sa.Uout[i][j_] = v;
        u[i] = v;
      }
    } else {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const long q = sa.ustate ? j_ : c;
        u[i] = cx<double>{double(sa.u[i][q].re), double(sa.u[i][q].im)};
      }
    }
    strain_of<T>(kx, ky, kz, u, sa.eps, c);
  }
}
template <typename T, typename UT = T>
__global__ void __launch_bounds__(NT) kop_strain(StrainArgs<T, UT> sa) {
  kop_strain_body<T, UT>(sa, PF_BIDX, PF_GDIM);
}
template <typename T, typename UT = T>
struct KStrain {
  struct Args {
    StrainArgs<T, UT> sa;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { kop_strain_body<T, UT>(a.sa, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL((kop_strain<T, UT>), dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.sa);
  }
};


// Mixed-precision phi update with the FP64 spectrum Phi of the unclamped,
// band-limited phi*: mode 1 (first update after new fields) phi^ = g^ / den
// and Phi = phi^; mode 2 g^ holds the transform of the clamp correction
// minus L dt (dF/dphi), and the increment
//   d^ = (g^ - L dt kappa k^2 Phi) / den,   den = 1 + L dt kappa k^2
// is written (Phi += d^).  With phi* = inverse(Phi) that is the reference
// update (phi_semi_implicit_update_kernel, cuda_kernels.cu:821-843), and the
// transforms carry only the increment.
template <typename T>
__device__ __forceinline__ void kop_phi_state_body(KDims d, long n, cx<T>* g, double Ldt_kappa, cx<double>* Phi, int mode, const B3 blockIdx, const B3 gridDim) {
  const long nact_ = (n / d.Az) * d.Azr;   // retained modes only (no kz padding)
  for (long j_ = blockIdx.x * (long)NT + threadIdx.x; j_ < nact_; j_ += (long)gridDim.x * NT) {
    const long c = kslot(d, j_);
    double kx, ky, kz;
    int iz;
    kvec(d, c, kx, ky, kz, iz);
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double den = 1.0 + Ldt_kappa * k2;
    const cx<T> v = g[c];
    if (mode == 2) {
      const cx<double> F = Phi[j_];
      const cx<double> dd = {(double(v.re) - Ldt_kappa * k2 * F.re) / den, (double(v.im) - Ldt_kappa * k2 * F.im) / den};
      Phi[j_] = cx<double>{F.re + dd.re, F.im + dd.im};
      g[c] = cx<T>{T(dd.re), T(dd.im)};
    } else {
      const cx<double> w = {double(v.re) / den, double(v.im) / den};
      Phi[j_] = w;
      g[c] = cx<T>{T(w.re), T(w.im)};
    }
  }
}
template <typename T>
__global__ void __launch_bounds__(NT) kop_phi_state(KDims d, long n, cx<T>* g, double Ldt_kappa, cx<double>* Phi, int mode) {
  kop_phi_state_body<T>(d, n, g, Ldt_kappa, Phi, mode, PF_BIDX, PF_GDIM);
}
template <typename T>
struct KPhiState {
  struct Args {
    KDims d;
    long n;
    cx<T>* g;
    double Ldt_kappa;
    cx<double>* Phi;
    int mode;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { kop_phi_state_body<T>(a.d, a.n, a.g, a.Ldt_kappa, a.Phi, a.mode, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL((kop_phi_state<T>), dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.d, a.n, a.g, a.Ldt_kappa, a.Phi, a.mode);
  }
};


// phi^ <- g^ / (1 + L dt kappa k^2)
template <typename T>
__device__ __forceinline__ void kop_phi_body(KDims d, long n, cx<T>* g, double Ldt_kappa, const B3 blockIdx, const B3 gridDim) {
  const long nact_ = (n / d.Az) * d.Azr;   // retained modes only (no kz padding)
  for (long j_ = blockIdx.x * (long)NT + threadIdx.x; j_ < nact_; j_ += (long)gridDim.x * NT) {
    const long c = kslot(d, j_);
    double kx, ky, kz;
    int iz;
    kvec(d, c, kx, ky, kz, iz);
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double den = 1.0 + Ldt_kappa * k2;
    cx<T> v = g[c];
    g[c] = cx<T>{T(double(v.re) / den), T(double(v.im) / den)};
  }
}
template <typename T>
__global__ void __launch_bounds__(NT) kop_phi(KDims d, long n, cx<T>* g, double Ldt_kappa) {
  kop_phi_body<T>(d, n, g, Ldt_kappa, PF_BIDX, PF_GDIM);
}
template <typename T>
struct KPhi {
  struct Args {
    KDims d;
    long n;
    cx<T>* g;
    double Ldt_kappa;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { kop_phi_body<T>(a.d, a.n, a.g, a.Ldt_kappa, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL((kop_phi<T>), dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.d, a.n, a.g, a.Ldt_kappa);
  }
};


// grad_a mu = i k_a mu^ (compute_gradient_single_component_k_kernel 1887-1911)
template <typename T>
__device__ __forceinline__ void kop_grad_body(KDims d, long n, const cx<T>* mu, cx<T>* gx, cx<T>* gy, cx<T>* gz, const B3 blockIdx, const B3 gridDim) {
  const long nact_ = (n / d.Az) * d.Azr;   // retained modes only (no kz padding)
  for (long j_ = blockIdx.x * (long)NT + threadIdx.x; j_ < nact_; j_ += (long)gridDim.x * NT) {
    const long c = kslot(d, j_);
    double kx, ky, kz;
    int iz;
    kvec(d, c, kx, ky, kz, iz);
    const double mr = double(mu[c].re), mi = double(mu[c].im);
    gx[c] = cx<T>{T(-kx * mi), T(kx * mr)};
    gy[c] = cx<T>{T(-ky * mi), T(ky * mr)};
    gz[c] = cx<T>{T(-kz * mi), T(kz * mr)};
  }
}
template <typename T>
__global__ void __launch_bounds__(NT) kop_grad(KDims d, long n, const cx<T>* mu, cx<T>* gx, cx<T>* gy, cx<T>* gz) {
  kop_grad_body<T>(d, n, mu, gx, gy, gz, PF_BIDX, PF_GDIM);
}
template <typename T>
struct KGrad {
  struct Args {
    KDims d;
    long n;
    const cx<T>* mu;
    cx<T>* gx;
    cx<T>* gy;
    cx<T>* gz;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { kop_grad_body<T>(a.d, a.n, a.mu, a.gx, a.gy, a.gz, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL((kop_grad<T>), dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.d, a.n, a.mu, a.gx, a.gy, a.gz);
  }
};


// Y^ <- (Y^n + dt f^)/(1 + dt Dbar k^2),  f^ = i k.J^ + Dbar k^2 Y^n - s^
// (Y_semi_implicit_update_kernel 2528-2550 with divJ_accumulate 2112-2137
// and the spectral Laplacian 3528-3538 folded in)
// Y_semi_implicit_update_kernel (cuda_kernels.cu:2528-2550) with the
// right-hand side f = div J + D-bar k^2 Y^n - s assembled in k space.  With
// yn == null it writes the increment Y^(n+1) - Y^n instead, in which the
// stabiliser cancels: dt (div J - s) / (1 + dt D-bar k^2).  For a band-limited
// Y^n (every state after the first step; the zero-mode shift is the k = 0
// mode) that is the same update, and its FP32 rounding scales with the
// increment instead of with Y.
template <typename T>
__device__ __forceinline__ void kop_yupdate_body(KDims d, long n, const cx<T>* jx, const cx<T>* jy, const cx<T>* jz, const cx<T>* yn, const cx<T>* s, cx<T>* out, const double* dbar, double dt, const B3 blockIdx, const B3 gridDim) {
  const double D = *dbar;
  const long nact_ = (n / d.Az) * d.Azr;   // retained modes only (no kz padding)
  for (long j_ = blockIdx.x * (long)NT + threadIdx.x; j_ < nact_; j_ += (long)gridDim.x * NT) {
    const long c = kslot(d, j_);
    double kx, ky, kz;
    int iz;
    kvec(d, c, kx, ky, kz, iz);
    const double k2 = kx * kx + ky * ky + kz * kz;
    // div J^ = i (kx Jx + ky Jy + kz Jz), accumulated x, y, z in order
    double dr = 0.0, di = 0.0;
    dr += -kx * double(jx[c].im);
    di += kx * double(jx[c].re);
    dr += -ky * double(jy[c].im);
    di += ky * double(jy[c].re);
    dr += -kz * double(jz[c].im);
    di += kz * double(jz[c].re);
    const double den = 1.0 + dt * D * k2;
    if (!yn) {
      out[c] = cx<T>{T(dt * (dr - double(s[c].re)) / den), T(dt * (di - double(s[c].im)) / den)};
      continue;
    }
    const double ynr = double(yn[c].re), yni = double(yn[c].im);
    const double fr = dr + D * k2 * ynr - double(s[c].re);
    const double fi = di + D * k2 * yni - double(s[c].im);
    out[c] = cx<T>{T((ynr + dt * fr) / den), T((yni + dt * fi) / den)};
  }
}
template <typename T>
__global__ void __launch_bounds__(NT) kop_yupdate(KDims d, long n, const cx<T>* jx, const cx<T>* jy, const cx<T>* jz, const cx<T>* yn, const cx<T>* s, cx<T>* out, const double* dbar, double dt) {
  kop_yupdate_body<T>(d, n, jx, jy, jz, yn, s, out, dbar, dt, PF_BIDX, PF_GDIM);
}
template <typename T>
struct KYUpdate {
  struct Args {
    KDims d;
    long n;
    const cx<T>* jx;
    const cx<T>* jy;
    const cx<T>* jz;
    const cx<T>* yn;
    const cx<T>* s;
    cx<T>* out;
    const double* dbar;
    double dt;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { kop_yupdate_body<T>(a.d, a.n, a.jx, a.jy, a.jz, a.yn, a.s, a.out, a.dbar, a.dt, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL((kop_yupdate<T>), dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.d, a.n, a.jx, a.jy, a.jz, a.yn, a.s, a.out, a.dbar, a.dt);
  }
};


}  // namespace pf
