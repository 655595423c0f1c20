// Point-wise physics of the phase-field step (host + device).
//
// Restates phase_functions.h:6-59 and thermo_utils.h:296-577 with the
// temperature-only quantities (G_PbTe, G_Ag2Te, RT, L(T)) hoisted into
// PhysParams on the host, so a voxel evaluates only the composition-dependent
// terms.  Evaluation order of every expression follows the reference so that
// FP64 results agree to rounding.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include "pf_log.hpp"

namespace pf {

#define PF_HD __host__ __device__ __forceinline__

// Voigt order used everywhere: 0=xx 1=yy 2=zz 3=yz 4=xz 5=xy.
// Upper-triangle packing of a symmetric 6x6 matrix: idx(i,j), i<=j.
PF_HD int vidx(int i, int j) {
  if (i > j) {
    int t = i;
    i = j;
    j = t;
  }
  return i * 6 - (i * (i - 1)) / 2 + (j - i);
}

struct PhysParams {
  // thermodynamics (thermo_utils.h), temperature-only terms hoisted
  double T, RT, Lparam, G_A0, G_B0, mu_scale;
  int convex;
  double mu0, v_A, v_B, Vm_c, Vm_a0, dVm;
  // phase field
  double W, kappa, L_phi, D_a, D_c, eps_iso, shift, Y_clip, xB_eps, dt;
  int elastic;
  // elasticity: C and dC as 21 packed upper-triangle Voigt entries
  double C[21], dC[21], eps00[6], E0[6];
  // convex-extrapolation constants at X_LIMIT_CONVEX
  double muA_c, dmuA_c, muB_c, dmuB_c;
  // 1 when C and dC couple no normal to shear component and their shear
  // blocks are diagonal (orthotropic in the grid axes; cubic and isotropic
  // tensors included): the Voigt products then skip the entries that are zero
  int ortho;
};

PF_HD double h_of_phi(double phi) {
  double p2 = phi * phi;
  double p3 = p2 * phi;
  return p3 * (6.0 * p2 - 15.0 * phi + 10.0);
}
PF_HD double h_prime_of_phi(double phi) {
  double p2 = phi * phi;
  return 30.0 * p2 * (1.0 - phi) * (1.0 - phi);
}
PF_HD double g_of_phi(double phi) { return phi * phi * (1.0 - phi) * (1.0 - phi); }
PF_HD double g_prime_of_phi(double phi) { return 2.0 * phi * (1.0 - phi) * (1.0 - 2.0 * phi); }
PF_HD double clamp01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }
PF_HD double clamp_eps(double v, double eps) { return v < eps ? eps : (v > 1.0 - eps ? 1.0 - eps : v); }
PF_HD double clamp_fraction_eps(double x) { return clamp_eps(x, 1e-12); }

// thermo_utils.h:364-377 (CALPHAD regular solution, clamped x; the
// logarithm is pf_log, within one ulp of the correctly rounded value)
PF_HD double mu_PbTe_calphad(const PhysParams& P, double xB) {
  double x = clamp_fraction_eps(xB);
  return P.G_A0 + P.RT * pf_log(1.0 - x) + P.Lparam * x * x;
}
PF_HD double mu_Ag2Te_calphad(const PhysParams& P, double xB) {
  double x = clamp_fraction_eps(xB);
  return P.G_B0 + P.RT * pf_log(x) + P.Lparam * (1.0 - x) * (1.0 - x);
}
// thermo_utils.h:400-430 (optional convex continuation above x = 0.09)
PF_HD double mu_PbTe_raw(const PhysParams& P, double xB) {
  if (!P.convex || xB <= 0.09) return mu_PbTe_calphad(P, xB);
  double d = xB - 0.09;
  return P.muA_c + P.dmuA_c * d + 5000.0 * d * d;
}
PF_HD double mu_Ag2Te_raw(const PhysParams& P, double xB) {
  if (!P.convex || xB <= 0.09) return mu_Ag2Te_calphad(P, xB);
  double d = xB - 0.09;
  return P.muB_c + P.dmuB_c * d + 5000.0 * d * d;
}
PF_HD double mu_A(const PhysParams& P, double xB) { return mu_PbTe_raw(P, xB) / P.mu_scale; }
PF_HD double mu_B(const PhysParams& P, double xB) { return mu_Ag2Te_raw(P, xB) / P.mu_scale; }

// thermo_utils.h:497-511
PF_HD double c_xB_phi(const PhysParams& P, double xB, double h) {
  double Vm_alpha = P.Vm_a0 + P.dVm * xB;
  double denom = Vm_alpha * (1.0 - h) + P.Vm_c * h;
  if (fabs(denom) < 1e-12) denom = (denom >= 0.0) ? 1e-12 : -1e-12;
  return 1.0 / denom;
}
PF_HD double mu_tot_mix(double muA, double muB, double xB, double mu0, double h) {
  double mu_mix = muA * (1.0 - xB) + muB * xB;
  return (1.0 - h) * mu_mix + h * mu0;
}

// stable logistic (cuda_kernels.cu:2201-2208, 3391-3399): 1/(1+e^-Y) for
// Y >= 0, e^Y/(1+e^Y) below, written without a branch so a wave with both
// signs evaluates one exp
PF_HD double sigmoid_unclipped(double Y) {
  const bool pos = Y >= 0.0;
  const double e = exp(pos ? -Y : Y);
  return (pos ? 1.0 : e) / (1.0 + e);
}
// thermo_utils.h:522-536
PF_HD double sigmoid_from_logit(double Y, double Y_clip, double eps) {
  if (Y > Y_clip) Y = Y_clip;
  if (Y < -Y_clip) Y = -Y_clip;
  double x = sigmoid_unclipped(Y);
  if (x < eps) x = eps;
  if (x > 1.0 - eps) x = 1.0 - eps;
  return x;
}
// thermo_utils.h:538-545
PF_HD double logit_from_fraction(double xB, double eps, double Y_clip) {
  if (xB < eps) xB = eps;
  if (xB > 1.0 - eps) xB = 1.0 - eps;
  double Y = log(xB / (1.0 - xB));
  if (Y > Y_clip) return Y_clip;
  if (Y < -Y_clip) return -Y_clip;
  return Y;
}
PF_HD double D_mix(const PhysParams& P, double h) { return (1.0 - h) * P.D_a + h * P.D_c; }

// thermo_utils.h:553-577 (central difference, step 1e-5; kept in FP64)
PF_HD double gamma_thermo(const PhysParams& P, double xB, double h) {
  double xc = xB;
  if (P.convex && xc > 0.08) xc = 0.08;
  double c_bulk = c_xB_phi(P, xc, h);
  const double dx = 1e-5;
  double xp = clamp_fraction_eps(xc + dx);
  double xm = clamp_fraction_eps(xc - dx);
  // The reference's difference quotient of mu_B - mu_A, evaluated in closed
  // form: both points lie on the CALPHAD branch (below the convex
  // continuation at 0.09), G(T) cancels and the logarithm differences combine,
  //   [RT (ln(xp/xm) - ln((1-xp)/(1-xm))) + L ((1-xp)^2 - (1-xm)^2 - xp^2 + xm^2)] / d
  //     = RT log1p(d / (xm (1 - xp))) / d - 2 L,   d = xp - xm.
  // One logarithm instead of four, without the cancellation of the literal
  // differences (whose rounding noise is ~1e-12 of Gamma).
  const double d = xp - xm;
  const double g = P.RT * log1p(d / (xm * (1.0 - xp))) / d - 2.0 * P.Lparam;
  return c_bulk * (g / P.mu_scale);
}
// cuda_kernels.cu:30-38
PF_HD double stabilized_meff(double Dm, double gamma) {
  const double gamma_floor = 1.0e-4;
  const double D_safe = fmax(Dm, 0.0);
  const double gamma_safe = (isfinite(gamma) && gamma > gamma_floor) ? gamma : gamma_floor;
  const double cap = D_safe * 1000.0;
  const double value = D_safe / gamma_safe;
  return (isfinite(value) && value < cap) ? value : cap;
}

// ------------------------------------------------------------------ elasticity
// Eigenstrain (cuda_kernels.cu:193-233): eps0_ii = (1-h) x eps_iso + h eps00_ii,
// eps0_ij = h eps00_ij; x = clamp01(xB).
PF_HD void eigenstrain(const PhysParams& P, double h, double xBc, double* e0) {
  double eps_c = xBc * P.eps_iso;
  double omh = 1.0 - h;
  e0[0] = omh * eps_c + h * P.eps00[0];
  e0[1] = omh * eps_c + h * P.eps00[1];
  e0[2] = omh * eps_c + h * P.eps00[2];
  e0[3] = h * P.eps00[3];
  e0[4] = h * P.eps00[4];
  e0[5] = h * P.eps00[5];
}

// s_I = sum_J M_IJ e~_J with engineering shear e~ = (xx,yy,zz,2yz,2xz,2xy).
// With `ortho` (PhysParams::ortho) the terms of the zero entries are left
// out; each kept term is accumulated in the same order, and a skipped term
// (0 * e) adds an exact zero, so results agree with the full product up to
// the sign of a zero.
template <typename R>
PF_HD void voigt_apply(const double* M, const R* e, R* s, int ortho = 0) {
  R ee[6] = {e[0], e[1], e[2], R(2) * e[3], R(2) * e[4], R(2) * e[5]};
  if (ortho) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      R acc = R(0);
#pragma unroll
      for (int j = 0; j < 3; ++j) acc += R(M[vidx(i, j)]) * ee[j];
      s[i] = acc;
    }
#pragma unroll
    for (int i = 3; i < 6; ++i) s[i] = R(0) + R(M[vidx(i, i)]) * ee[i];
    return;
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    R acc = R(0);
#pragma unroll
    for (int j = 0; j < 6; ++j) acc += R(M[vidx(i, j)]) * ee[j];
    s[i] = acc;
  }
}
// (C + h dC) : e
template <typename R>
PF_HD void voigt_apply_h(const PhysParams& P, R h, const R* e, R* s) {
  R ee[6] = {e[0], e[1], e[2], R(2) * e[3], R(2) * e[4], R(2) * e[5]};
  if (P.ortho) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      R a = R(0), b = R(0);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        a += R(P.C[vidx(i, j)]) * ee[j];
        b += R(P.dC[vidx(i, j)]) * ee[j];
      }
      s[i] = a + h * b;
    }
#pragma unroll
    for (int i = 3; i < 6; ++i) {
      const R a = R(0) + R(P.C[vidx(i, i)]) * ee[i];
      const R b = R(0) + R(P.dC[vidx(i, i)]) * ee[i];
      s[i] = a + h * b;
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    R a = R(0), b = R(0);
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      a += R(P.C[vidx(i, j)]) * ee[j];
      b += R(P.dC[vidx(i, j)]) * ee[j];
    }
    s[i] = a + h * b;
  }
}

// h_ij of the heterogeneous fixed point (cuda_kernels.cu:7002-7153):
//   h = -(C + h dC):(eps + E0) + C:eps + (C + h dC):eps0
//     = -h dC:eps - (C + h dC):E0 + (C + h dC):eps0
template <typename R>
PF_HD void hij_point(const PhysParams& P, R h, const R* e0, const R* eps, R* out) {
  R tot[6], a[6], b[6], c[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) tot[i] = eps[i] + R(P.E0[i]);
  voigt_apply_h<R>(P, h, tot, a);
  voigt_apply<R>(P.C, eps, b, P.ortho);
  voigt_apply_h<R>(P, h, e0, c);
#pragma unroll
  for (int i = 0; i < 6; ++i) out[i] = -a[i] + b[i] + c[i];
}
// linear part only: A(eps) = -h dC : eps (used for increments)
template <typename R>
PF_HD void hij_increment_point(const PhysParams& P, R h, const R* eps, R* out) {
  R d[6];
  voigt_apply<R>(P.dC, eps, d, P.ortho);
#pragma unroll
  for (int i = 0; i < 6; ++i) out[i] = -(h * d[i]);
}

}  // namespace pf
