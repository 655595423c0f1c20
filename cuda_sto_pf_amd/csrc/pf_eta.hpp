// GP-zone order parameter eta: one semi-implicit update as in the step loop
// (main_cuda.cu:29948-30006), fused into two z passes around the pruned
// y/x transforms:
//   z pass 1: rhs = compute_eta_rhs_kernel (cuda_kernels.cu:845-946) at each
//             voxel, output eta - L dt rhs (the update is linear), forward z;
//   k space : divide by 1 + L dt kappa k^2 (eta_semi_implicit_update_kernel,
//             cuda_kernels.cu:1045-1069);
//   z pass 2: inverse z, eta = clamp(v/N, 0, 1) (eta_normalize_and_clamp_kernel,
//             cuda_kernels.cu:1086-1098).
#pragma once
#include "pf_ops.hpp"

namespace pf {

struct EtaConst {
  double xB_gp, delta_g0, W_eta, eps_gp, el_scale, nu_A, nu_B, mu_ref_raw, Ldt;
  int raw_drive, raw_units, elastic;
};

PF_HD double eta_rhs_point(const PhysParams& P, const EtaConst& q, double eta_raw, double phi_raw, double xa_raw,
                           double trs) {
  const double eta = clamp01(eta_raw);
  const double phi = clamp01(phi_raw);
  const double xa = clamp_eps(xa_raw, 1.0e-12);
  const double h_phi = h_of_phi(phi);
  const double hp_eta = h_prime_of_phi(eta);
  const double c_ref = 1.0 / (P.Vm_a0 + P.dVm * xa);
  double chem;
  if (q.raw_drive) {
    const double muA = mu_PbTe_calphad(P, xa);
    const double muB = mu_Ag2Te_calphad(P, xa);
    double mdm;
    if (q.raw_units) {
      mdm = q.nu_A * muA + q.nu_B * muB - q.mu_ref_raw;
    } else {
      const double s = fabs(P.mu_scale) > 1.0e-300 ? P.mu_scale : 1.0;
      mdm = q.nu_A * (muA / s) + q.nu_B * (muB / s) - q.mu_ref_raw / s;
    }
    chem = c_ref * (1.0 - h_phi) * hp_eta * (-mdm);
  } else {
    const double g_alpha = (1.0 - xa) * mu_A(P, xa) + xa * mu_B(P, xa);
    const double xg = q.xB_gp;
    const double g_gp = (1.0 - xg) * mu_A(P, xg) + xg * mu_B(P, xg) - q.delta_g0;
    chem = c_ref * (1.0 - h_phi) * hp_eta * (g_gp - g_alpha);
  }
  double el = 0.0;
  if (q.elastic) {
    // compute_dgel_deta_gp_point (cuda_kernels.cu:293-308)
    const double eps_alpha = clamp01(xa) * P.eps_iso;
    const double delta = (1.0 - h_phi) * hp_eta * (q.eps_gp - eps_alpha);
    el = q.el_scale * (-trs * delta);
  }
  return chem + q.W_eta * g_prime_of_phi(eta) + el;
}

template <typename T>
struct OpEtaRhs {
  struct Args {
    PhysParams P;
    EtaConst q;
    const double* eta;
    const double* phi;
    const double* xa;
    const double* trs;   // may be null
  };
  static constexpr int NRED = 0;
  static constexpr int kinds[1] = {0};
  static constexpr int NPRE = 4;
  __device__ static void load(const Args& a, size_t g, double* pre) {
    pre[0] = a.eta[g];
    pre[1] = a.phi[g];
    pre[2] = a.xa[g];
    pre[3] = a.trs ? a.trs[g] : 0.0;
  }
  __device__ static void apply(const Args& a, size_t, const double* pre, const T*, T* out, double*) {
    const double r = eta_rhs_point(a.P, a.q, pre[0], pre[1], pre[2], pre[3]);
    out[0] = T(pre[0] - a.q.Ldt * r);
  }
};

template <typename T>
struct OpEtaFinish {
  struct Args {
    double* eta_out;
    T invN;
  };
  static constexpr int NRED = 0;
  static constexpr int kinds[1] = {0};
  static constexpr int NPRE = 0;
  __device__ static void load(const Args&, size_t, double*) {}
  __device__ static void apply(const Args& a, size_t g, const double*, const T* vin, T*, double*) {
    double v = double(vin[0]) * double(a.invN);
    if (v < 0.0) v = 0.0;
    if (v > 1.0) v = 1.0;
    a.eta_out[g] = v;
  }
};

}  // namespace pf
