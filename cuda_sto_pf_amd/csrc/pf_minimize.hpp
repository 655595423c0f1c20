// Minimize mode (main_cuda --mode=minimize, phi-only and
// --minimize-full-model): point-wise operators fused into the z passes and
// the element-wise reductions of the volume constraint.
//
// The iteration (main_cuda.cu:29264-29750, 32104-32290, 32400-32560):
//   g_exp  = W g'(phi) + dgel/dphi with stress and total strain held at the
//            values of the last elastic solve (compute_phi_rhs_minimize_kernel,
//            cuda_kernels.cu:6298-6368; composition absent, eps0 = h eps00),
//   g_full = g_exp - kappa Lap phi,
//   lambda = (1 - eta) lambda + eta (-sum h' g_full / sum h'^2),
//   phi^  <- (phi^ - L dt (g_exp + lambda h')^) / (1 + L dt kappa k^2),
// then the KKT residual at the new phi, the post-projection sub-steps with
// eta = 1, and the first-order volume projection phi += lc h'(phi).
//
// The elastic part of g_exp depends on phi only through h and h', so after
// the solve the stress is kept as one scalar per voxel, s00 = sigma : eps00
// (engineering shear), together with the total strain (6 components):
//   dgel/dphi = -h' s00 + 0.5 h' (eps - h eps00) : dC : (eps - h eps00).
#pragma once
#include "pf_ops.hpp"

namespace pf {

// elastic part of the phi-derivative of the energy density at frozen strain
PF_HD double min_dgel(const PhysParams& P, double phi, const double* eps, double s00) {
  const double hp = h_prime_of_phi(phi);
  const double h = h_of_phi(phi);
  double el[6], q[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) el[c] = eps[c] - h * P.eps00[c];
  voigt_apply<double>(P.dC, el, q, P.ortho);
  const double Q = el[0] * q[0] + el[1] * q[1] + el[2] * q[2] + 2.0 * el[3] * q[3] + 2.0 * el[4] * q[4] +
                   2.0 * el[5] * q[5];
  return -hp * s00 + 0.5 * hp * Q;
}

// After the elastic solve: inverse z of the 6 strains -> stress at phi^n,
// keep (eps, s00), g_exp(phi^n), energies; forward z of phi^n.
//   red[0] sum 0.5 sigma:eps_el        (compute_gel_density_kernel, cuda_kernels.cu:1789-1850)
//   red[1] sum W g(phi)                (compute_dw_energy_density_kernel, 811-815)
//   red[2] sum 0.5 sigma:eps_el h(clamp01 phi)  (scale_field_by_h_kernel, 761-771)
template <typename T, bool ELASTIC>
struct OpMinRhs {
  struct Args {
    PhysParams P;
    const double* phi;
    double* gexp;
    T* eps;            // 6 components, stride R
    double* s00;
    size_t R;
    T invN;
  };
  static constexpr int NRED = 3;
  static constexpr int kinds[3] = {RED_SUM, RED_SUM, RED_SUM};
  static constexpr int NPRE = 1;
  __device__ static void load(const Args& a, size_t g, double* pre) { pre[0] = a.phi[g]; }
  __device__ static void apply(const Args& a, size_t g, const double* pre, const T* vin, T* out, double* red) {
    const PhysParams& P = a.P;
    const double phi = pre[0];
    double f = P.W * g_prime_of_phi(phi);
    red[1] += P.W * g_of_phi(phi);
    if constexpr (ELASTIC) {
      double eps[6], el[6], sig[6];
      const double h = h_of_phi(phi);
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        eps[c] = double(vin[c] * a.invN) + P.E0[c];
        el[c] = eps[c] - h * P.eps00[c];
      }
      voigt_apply_h<double>(P, h, el, sig);
      const double s00 = sig[0] * P.eps00[0] + sig[1] * P.eps00[1] + sig[2] * P.eps00[2] +
                         2.0 * (sig[3] * P.eps00[3] + sig[4] * P.eps00[4] + sig[5] * P.eps00[5]);
      const double gel = 0.5 * (sig[0] * el[0] + sig[1] * el[1] + sig[2] * el[2] +
                                2.0 * (sig[5] * el[5] + sig[4] * el[4] + sig[3] * el[3]));
      red[0] += gel;
      red[2] += gel * h_of_phi(clamp01(phi));
#pragma unroll
      for (int c = 0; c < 6; ++c) a.eps[c * a.R + g] = T(eps[c]);
      a.s00[g] = s00;
      f += min_dgel(P, phi, eps, s00);
    }
    a.gexp[g] = f;
    out[0] = T(phi);
  }
};

// inverse z of kappa-free Lap phi -> g_full = g_exp - kappa Lap phi;
//   red[0] sum h'(phi) g_full, red[1] sum h'(phi)^2
// (compute_hprime_times_rhs_kernel / compute_hprime_sq_values_kernel,
//  cuda_kernels.cu:661-682)
template <typename T>
struct OpMinGfull {
  struct Args {
    const double* phi;
    const double* gexp;
    double* gfull;
    double kappa;
    T invN;
  };
  static constexpr int NRED = 2;
  static constexpr int kinds[2] = {RED_SUM, RED_SUM};
  static constexpr int NPRE = 2;
  __device__ static void load(const Args& a, size_t g, double* pre) {
    pre[0] = a.phi[g];
    pre[1] = a.gexp[g];
  }
  __device__ static void apply(const Args& a, size_t g, const double* pre, const T* vin, T*, double* red) {
    const double lap = double(vin[0]) * double(a.invN);
    const double gf = pre[1] - a.kappa * lap;
    const double hp = h_prime_of_phi(pre[0]);
    a.gfull[g] = gf;
    red[0] += hp * gf;
    red[1] += hp * hp;
  }
};

// phi - L dt (g_exp + lambda h'(phi)) -> forward z (the update is linear, so
// the transform of this field is phi^ - L dt rhs^)
template <typename T>
struct OpMinUpd {
  struct Args {
    const double* phi;
    const double* gexp;
    double lambda;
    double Ldt;
    const int* labels;       // per-label multipliers lambdas[labels[i]] when set
    const double* lambdas;   // (add_component_volume_constraint_kernel, cuda_kernels.cu:606-615)
    const double* lambda_p;  // when set, the multiplier is read here (decided on the device)
  };
  static constexpr int NRED = 0;
  static constexpr int kinds[1] = {0};
  static constexpr int NPRE = 3;
  __device__ static void load(const Args& a, size_t g, double* pre) {
    pre[0] = a.phi[g];
    pre[1] = a.gexp[g];
    pre[2] = a.labels ? a.lambdas[a.labels[g]] : (a.lambda_p ? *a.lambda_p : a.lambda);
  }
  __device__ static void apply(const Args& a, size_t, const double* pre, const T*, T* out, double*) {
    const double phi = pre[0];
    out[0] = T(phi - a.Ldt * (pre[1] + pre[2] * h_prime_of_phi(phi)));
  }
};

// inverse z of the updated phi^: phi = clamp(v/N, -1e-6, 1+1e-6)
// (phi_normalize_and_clamp_kernel, cuda_kernels.cu:1071-1084);
//   red[0] sum (phi - phi_ref)^2  (diff_sq_kernel, main_cuda.cu:2236-2240)
// With NEXT: g_exp at the new phi (frozen strain) and forward z of phi.
template <typename T, bool ELASTIC, bool NEXT>
struct OpMinFinish {
  struct Args {
    PhysParams P;
    const double* phi_ref;
    double* phi_out;
    double* gexp;
    const T* eps;
    const double* s00;
    size_t R;
    T invN;
  };
  static constexpr int NRED = 1;
  static constexpr int kinds[1] = {RED_SUM};
  static constexpr int NPRE = 1;
  __device__ static void load(const Args& a, size_t g, double* pre) { pre[0] = a.phi_ref[g]; }
  __device__ static void apply(const Args& a, size_t g, const double* pre, const T* vin, T* out, double* red) {
    double v = double(vin[0]) * double(a.invN);
    if (v < -1e-6) v = -1e-6;
    if (v > 1.0 + 1e-6) v = 1.0 + 1e-6;
    a.phi_out[g] = v;
    const double d = v - pre[0];
    red[0] += d * d;
    if constexpr (NEXT) {
      double f = a.P.W * g_prime_of_phi(v);
      if constexpr (ELASTIC) {
        double eps[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) eps[c] = double(a.eps[c * a.R + g]);
        f += min_dgel(a.P, v, eps, a.s00[g]);
      }
      a.gexp[g] = f;
      out[0] = T(v);
    }
  }
};

// g_exp at the current phi (frozen strain) and forward z of phi: the start
// of a further post-projection sub-step
template <typename T, bool ELASTIC>
struct OpMinEval {
  struct Args {
    PhysParams P;
    const double* phi;
    double* gexp;
    const T* eps;
    const double* s00;
    size_t R;
  };
  static constexpr int NRED = 0;
  static constexpr int kinds[1] = {0};
  static constexpr int NPRE = 1;
  __device__ static void load(const Args& a, size_t g, double* pre) { pre[0] = a.phi[g]; }
  __device__ static void apply(const Args& a, size_t g, const double* pre, const T*, T* out, double*) {
    const double v = pre[0];
    double f = a.P.W * g_prime_of_phi(v);
    if constexpr (ELASTIC) {
      double eps[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) eps[c] = double(a.eps[c * a.R + g]);
      f += min_dgel(a.P, v, eps, a.s00[g]);
    }
    a.gexp[g] = f;
    out[0] = T(v);
  }
};

// ---------------------------------------------------------------- full model
// --minimize-full-model evaluates the dynamics phi right-hand side
// (compute_phi_rhs_kernel, cuda_kernels.cu:390-552: chemical drive, W g' and
// the elastic terms with the composition eigenstrain) in every g_exp, with
// the stress and total strain of the iteration's solve held fixed.  Those are
// kept as the 6 total strains, s00 = sigma : eps00 (engineering shear) and
// tr(sigma), which is all that term needs at a new phi and x_B: with
// d eps0_ii/d phi = h' (eps00_ii - x e_iso + e_iso (x - v_B)) and shears h' eps00_ij,
//   -sigma : d eps0/d phi = -h' (s00 + (-x e_iso + e_iso (x - v_B)) tr sigma),
// plus 1/2 h' Q(eps - eps0(phi, x)).
PF_HD double min_dgel_full(const PhysParams& P, double phi, double xB, const double* eps, double s00,
                           double trs) {
  const double hp = h_prime_of_phi(phi);
  const double h = h_of_phi(phi);
  const double x = clamp01(xB);
  const double diag_iso = -(x * P.eps_iso) + P.eps_iso * (x - P.v_B);
  const double dgel_eigen = -hp * (s00 + diag_iso * trs);
  double e0[6], el[6], q[6];
  eigenstrain(P, h, x, e0);
#pragma unroll
  for (int c = 0; c < 6; ++c) el[c] = eps[c] - e0[c];
  voigt_apply<double>(P.dC, el, q, P.ortho);
  const double Q = el[0] * q[0] + el[1] * q[1] + el[2] * q[2] + 2.0 * el[3] * q[3] + 2.0 * el[4] * q[4] +
                   2.0 * el[5] * q[5];
  return dgel_eigen + 0.5 * hp * Q;
}

// g_exp of the full model at (phi, x_B) with the frozen elastic state
template <bool ELASTIC, typename T>
__device__ __forceinline__ double min_gexp_full(const PhysParams& P, double phi, double xB, const T* eps_s,
                                                size_t R, size_t g, const double* s00, const T* trs) {
  double hp, h, mA, mB;
  double f = phi_rhs_chem(P, phi, clamp01(xB), hp, h, mA, mB);
  if constexpr (ELASTIC) {
    double eps[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) eps[c] = double(eps_s[c * R + g]);
    f += min_dgel_full(P, phi, xB, eps, s00[g], double(trs[g]));
  }
  return f;
}

// After the solve: inverse z of the 6 strains -> stress at phi^n with the
// solve's x_B, kept as (eps, s00, tr sigma); x_B refreshed from Y
// (compute_mu_x_kernel); g_exp(phi^n) at the refreshed x_B; energies with the
// solve's x_B; forward z of phi^n.
//   red[0] sum 0.5 sigma:eps_el, red[1] sum W g(phi), red[2] sum 0.5 sigma:eps_el h(clamp01 phi),
//   red[3] sum (1 - h) mu_mix + h mu0 (compute_gbulk_excess_hat_kernel, cuda_kernels.cu:725-759,
//          before its baseline shift; clamp01 phi, x clamped to [1e-12, 1 - 1e-12])
template <typename T, bool ELASTIC>
struct OpMinRhsFull {
  struct Args {
    PhysParams P;
    const double* phi;
    const double* Y;
    double* xB;
    double* gexp;
    T* eps;            // 6 components, stride R
    double* s00;
    T* trs;
    size_t R;
    T invN;
  };
  static constexpr int NRED = 4;
  static constexpr int kinds[4] = {RED_SUM, RED_SUM, RED_SUM, RED_SUM};
  static constexpr int NPRE = 3;
  __device__ static void load(const Args& a, size_t g, double* pre) {
    pre[0] = a.phi[g];
    pre[1] = a.xB[g];
    pre[2] = a.Y[g];
  }
  __device__ static void apply(const Args& a, size_t g, const double* pre, const T* vin, T* out, double* red) {
    const PhysParams& P = a.P;
    const double phi = pre[0];
    const double xs = pre[1];
    const double hc = h_of_phi(clamp01(phi));
    {
      const double xc = clamp_fraction_eps(xs);
      red[3] += (1.0 - hc) * ((1.0 - xc) * mu_A(P, xc) + xc * mu_B(P, xc)) + hc * P.mu0;
    }
    red[1] += P.W * g_of_phi(phi);
    const double xr = sigmoid_from_logit(pre[2], P.Y_clip, P.xB_eps);
    a.xB[g] = xr;
    if constexpr (ELASTIC) {
      double eps[6], e0[6], el[6], sig[6];
      const double h = h_of_phi(phi);
      eigenstrain(P, h, clamp01(xs), e0);
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        eps[c] = double(vin[c] * a.invN) + P.E0[c];
        el[c] = eps[c] - e0[c];
      }
      voigt_apply_h<double>(P, h, el, sig);
      const double s00 = sig[0] * P.eps00[0] + sig[1] * P.eps00[1] + sig[2] * P.eps00[2] +
                         2.0 * (sig[3] * P.eps00[3] + sig[4] * P.eps00[4] + sig[5] * P.eps00[5]);
      const double gel = 0.5 * (sig[0] * el[0] + sig[1] * el[1] + sig[2] * el[2] +
                                2.0 * (sig[5] * el[5] + sig[4] * el[4] + sig[3] * el[3]));
      red[0] += gel;
      red[2] += gel * hc;
#pragma unroll
      for (int c = 0; c < 6; ++c) a.eps[c * a.R + g] = T(eps[c]);
      a.s00[g] = s00;
      a.trs[g] = T(sig[0] + sig[1] + sig[2]);
    }
    a.gexp[g] = min_gexp_full<ELASTIC, T>(P, phi, xr, a.eps, a.R, g, a.s00, a.trs);
    out[0] = T(phi);
  }
};

// inverse z of the updated phi^: phi = clamp(v/N, -1e-6, 1+1e-6); red[0] sum
// (phi - phi_ref)^2; g_exp at the new phi and the current x_B; forward z of phi
template <typename T, bool ELASTIC>
struct OpMinFinishFull {
  struct Args {
    PhysParams P;
    const double* phi_ref;
    double* phi_out;
    double* gexp;
    const T* eps;
    const double* s00;
    const T* trs;
    const double* xB;
    size_t R;
    T invN;
  };
  static constexpr int NRED = 1;
  static constexpr int kinds[1] = {RED_SUM};
  static constexpr int NPRE = 2;
  __device__ static void load(const Args& a, size_t g, double* pre) {
    pre[0] = a.phi_ref[g];
    pre[1] = a.xB[g];
  }
  __device__ static void apply(const Args& a, size_t g, const double* pre, const T* vin, T* out, double* red) {
    double v = double(vin[0]) * double(a.invN);
    if (v < -1e-6) v = -1e-6;
    if (v > 1.0 + 1e-6) v = 1.0 + 1e-6;
    a.phi_out[g] = v;
    const double d = v - pre[0];
    red[0] += d * d;
    a.gexp[g] = min_gexp_full<ELASTIC, T>(a.P, v, pre[1], a.eps, a.R, g, a.s00, a.trs);
    out[0] = T(v);
  }
};

// g_exp at the current phi and x_B (frozen elastic state) and forward z of phi
template <typename T, bool ELASTIC>
struct OpMinEvalFull {
  struct Args {
    PhysParams P;
    const double* phi;
    double* gexp;
    const T* eps;
    const double* s00;
    const T* trs;
    const double* xB;
    size_t R;
  };
  static constexpr int NRED = 0;
  static constexpr int kinds[1] = {0};
  static constexpr int NPRE = 2;
  __device__ static void load(const Args& a, size_t g, double* pre) {
    pre[0] = a.phi[g];
    pre[1] = a.xB[g];
  }
  __device__ static void apply(const Args& a, size_t g, const double* pre, const T*, T* out, double*) {
    a.gexp[g] = min_gexp_full<ELASTIC, T>(a.P, pre[0], pre[1], a.eps, a.R, g, a.s00, a.trs);
    out[0] = T(pre[0]);
  }
};

// ---------------------------------------------------------------- element-wise
namespace {
// res = g_full + lambda h'(phi), zeroed where it only points out of [0,1]
// (compute_projected_phi_kkt_residual_sq_kernel, cuda_kernels.cu:684-706);
//   red[0] sum res^2
__device__ __forceinline__ void min_kkt_kernel_body(const double* phi, const double* gfull, double lambda, const int* labels, const double* lambdas, long n, double* partials, const double* lambda_p, const B3 blockIdx, const B3 gridDim) {
  if (lambda_p) lambda = *lambda_p;
  double s = 0.0;
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    const double p = phi[i];
    double r = gfull[i] + (labels ? lambdas[labels[i]] : lambda) * h_prime_of_phi(p);
    if ((p <= 1e-12 && r > 0.0) || (p >= 1.0 - 1e-12 && r < 0.0)) r = 0.0;
    s += r * r;
  }
  const int kinds[1] = {RED_SUM};
  block_reduce_store<1>(&s, kinds, partials + blockIdx.x);
}
__global__ void __launch_bounds__(NT) min_kkt_kernel(const double* phi, const double* gfull, double lambda, const int* labels, const double* lambdas, long n, double* partials, const double* lambda_p) {
  min_kkt_kernel_body(phi, gfull, lambda, labels, lambdas, n, partials, lambda_p, PF_BIDX, PF_GDIM);
}
struct KMinKkt {
  struct Args {
    const double* phi;
    const double* gfull;
    double lambda;
    const int* labels;
    const double* lambdas;
    long n;
    double* partials;
    const double* lambda_p;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { min_kkt_kernel_body(a.phi, a.gfull, a.lambda, a.labels, a.lambdas, a.n, a.partials, a.lambda_p, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL(min_kkt_kernel, dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.phi, a.gfull, a.lambda, a.labels, a.lambdas, a.n, a.partials, a.lambda_p);
  }
};


// red[0] sum h(clamp01 phi) (reduce_sum_h_kernel, cuda_kernels.cu:5580-5607),
// red[1] sum h'(phi)^2
__device__ __forceinline__ void min_volume_kernel_body(const double* phi, long n, double* partials, const B3 blockIdx, const B3 gridDim) {
  double v[2] = {0.0, 0.0};
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    const double p = phi[i];
    const double hp = h_prime_of_phi(p);
    v[0] += h_of_phi(clamp01(p));
    v[1] += hp * hp;
  }
  const int kinds[2] = {RED_SUM, RED_SUM};
  block_reduce_store<2>(v, kinds, partials + 2 * (size_t)blockIdx.x);
}
__global__ void __launch_bounds__(NT) min_volume_kernel(const double* phi, long n, double* partials) {
  min_volume_kernel_body(phi, n, partials, PF_BIDX, PF_GDIM);
}
struct KMinVolume {
  struct Args {
    const double* phi;
    long n;
    double* partials;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { min_volume_kernel_body(a.phi, a.n, a.partials, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL(min_volume_kernel, dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.phi, a.n, a.partials);
  }
};


// red[0] sum (a + shift), red[1] sum (a + shift)^2
__global__ void __launch_bounds__(NT) sum_sq_kernel(const double* a, double shift, long n, double* partials) {
  double v[2] = {0.0, 0.0};
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    const double x = a[i] + shift;
    v[0] += x;
    v[1] += x * x;
  }
  const int kinds[2] = {RED_SUM, RED_SUM};
  block_reduce_store<2>(v, kinds, partials + 2 * (size_t)blockIdx.x);
}

// ---------------------------------------------------------------- V5 labels
// Per-label sums over the ownership labels 0..ncomp-1 (at most COMP_MAX):
// sum h'(phi) g, sum h'(phi)^2, sum h(clamp01 phi)
// (compute_component_constraint_moments_kernel, cuda_kernels.cu:590-604).
// A block owns COMP_VPT * NT consecutive cells, so it meets the few labels
// of a compact region; it reduces the labels in its [min, max] range and
// writes zeros for the rest.  partials: 3 ncomp per block.
constexpr int COMP_MAX = 64;
constexpr int COMP_VPT = 8;
__global__ void __launch_bounds__(NT) comp_moments_kernel(const double* phi, const double* g, const int* labels,
                                                          int ncomp, long n, double* partials) {
  __shared__ int range[2];
  const long base = (long)blockIdx.x * NT * COMP_VPT;
  double v[COMP_VPT][3];
  int lab[COMP_VPT];
  int lo = ncomp, hi = -1;
#pragma unroll
  for (int u = 0; u < COMP_VPT; ++u) {
    const long i = base + (long)u * NT + threadIdx.x;
    lab[u] = -1;
    v[u][0] = v[u][1] = v[u][2] = 0.0;
    if (i < n) {
      const double p = phi[i];
      const double hp = h_prime_of_phi(p);
      lab[u] = labels[i];
      v[u][0] = g ? hp * g[i] : 0.0;
      v[u][1] = hp * hp;
      v[u][2] = h_of_phi(clamp01(p));
      lo = min(lo, lab[u]);
      hi = max(hi, lab[u]);
    }
  }
  if (threadIdx.x == 0) {
    range[0] = ncomp;
    range[1] = -1;
  }
  __syncthreads();
  atomicMin(&range[0], lo);
  atomicMax(&range[1], hi);
  __syncthreads();
  const int c0 = range[0], c1 = range[1];
  double* out = partials + (size_t)blockIdx.x * 3 * ncomp;
  for (int c = threadIdx.x; c < ncomp; c += NT)
    if (c < c0 || c > c1) out[3 * c] = out[3 * c + 1] = out[3 * c + 2] = 0.0;
  const int kinds[3] = {RED_SUM, RED_SUM, RED_SUM};
  for (int c = c0; c <= c1; ++c) {
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int u = 0; u < COMP_VPT; ++u)
      if (lab[u] == c) {
        acc[0] += v[u][0];
        acc[1] += v[u][1];
        acc[2] += v[u][2];
      }
    block_reduce_store<3>(acc, kinds, out + 3 * c);
    __syncthreads();   // the reduction's LDS is used again by the next label
  }
}

// out[j] = sum over blocks of partials[b * nval + j], in block order
__global__ void __launch_bounds__(NT) comp_sum_kernel(const double* partials, int nblocks, int nval,
                                                      double* out) {
  for (int j = threadIdx.x; j < nval; j += NT) {
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += partials[(size_t)b * nval + j];
    out[j] = s;
  }
}

// phi <- clamp01(phi + corr[label] h'(phi)) (apply_component_volume_projection_kernel,
// cuda_kernels.cu:617-627)
__global__ void __launch_bounds__(NT) comp_project_kernel(double* phi, const int* labels, const double* corr,
                                                          long n) {
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    const double p = phi[i];
    phi[i] = clamp01(p + corr[labels[i]] * h_prime_of_phi(p));
  }
}

// b = a, red[0] min a, red[1] max a
__global__ void __launch_bounds__(NT) copy_minmax_kernel(const double* a, double* b, long n, double* partials) {
  double v[2] = {1.0e300, -1.0e300};
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    const double x = a[i];
    b[i] = x;
    v[0] = fmin(v[0], x);
    v[1] = fmax(v[1], x);
  }
  const int kinds[2] = {RED_MIN, RED_MAX};
  block_reduce_store<2>(v, kinds, partials + 2 * (size_t)blockIdx.x);
}

// phi <- clamp01(phi + lc h'(phi)) (apply_volume_projection_kernel,
// cuda_kernels.cu:639-647), then red[0] sum h(clamp01 phi).  With mc set,
// lc and whether to project at all come from the minimize control block
// (MC_LC, MC_PROJ); without a projection phi is left alone and the sum is 0.
__global__ void __launch_bounds__(NT) min_project_kernel(double* phi, double lc, long n, double* partials,
                                                         const double* mc);

// ---------------------------------------------------------------- control
// The phi-only iteration's scalar decisions, taken by one thread on the
// device between the passes (the host reads the block once per iteration):
// the under-relaxed and exact multipliers (main_cuda.cu:29468-29500,
// 29596-29610, 32120-32140) and the first-order volume projection
// (32274-32290).  Sums come from the scalar slots the reductions write.
enum MinCtlSlot {
  MC_LAMBDA = 0,   // lambda_vol, carried across iterations
  MC_V0,           // target <h>, set from <h(phi^0)> at the first iteration when not given
  MC_LAM_RAW,      // -sum h' g_full / sum h'^2 at phi^n
  MC_LAM_KKT,      // the same at phi^{n+1}
  MC_LAM_PP,       // the multiplier of the next post-projection sub-step
  MC_LC,           // volume projection coefficient
  MC_PROJ,         // 1 when the projection runs
  MC_MEAN_H,       // <h> after the iteration
  MC_COUNT
};
// scalar slots of the sums: 48/49 sum h, sum h'^2; 50/51, 53/54, 56/57
// sum h' g_full, sum h'^2 at phi^n, phi^{n+1} and the post-projection phi;
// 59 sum h after the projection
enum MinSumSlot { MS_VOL = 48, MS_RAW = 50, MS_KKT = 53, MS_PP = 56, MS_PROJ = 59 };
enum MinCtlStage { MCS_RAW = 0, MCS_KKT, MCS_PP, MCS_PP_COMMIT, MCS_VOLUME, MCS_FINAL, MCS_INIT };
struct MinCtlConst {
  double R;        // voxel count
  double eta;      // under-relaxation of lambda, clamped to [0, 1]
  double V0;       // MCS_INIT: the given target (<= 0: from the first iteration)
  int first;       // MCS_RAW of the first iteration
};

PF_HD inline double min_multiplier_of(double s_hg, double s_h2) {
  return fabs(s_h2) > 1e-30 ? -s_hg / s_h2 : 0.0;
}

__device__ __forceinline__ void min_ctl_kernel_body(const double* s, double* mc, int stage, MinCtlConst c, const B3 blockIdx, const B3 gridDim) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  switch (stage) {
    case MCS_INIT:
      for (int i = 0; i < MC_COUNT; ++i) mc[i] = 0.0;
      mc[MC_V0] = c.V0;
      break;
    case MCS_RAW: {
      const double mean_h0 = s[MS_VOL] / c.R;
      if (c.first && mc[MC_V0] <= 0.0) mc[MC_V0] = mean_h0;
      const double lam_raw = min_multiplier_of(s[MS_RAW], s[MS_RAW + 1]);
      mc[MC_LAM_RAW] = lam_raw;
      mc[MC_LAMBDA] = (1.0 - c.eta) * mc[MC_LAMBDA] + c.eta * lam_raw;
      break;
    }
    case MCS_KKT:
      mc[MC_LAM_KKT] = min_multiplier_of(s[MS_KKT], s[MS_KKT + 1]);
      mc[MC_LAM_PP] = mc[MC_LAM_KKT];
      break;
    case MCS_PP:
      mc[MC_LAM_PP] = min_multiplier_of(s[MS_PP], s[MS_PP + 1]);
      break;
    case MCS_PP_COMMIT:
      mc[MC_LAMBDA] = mc[MC_LAM_PP];
      break;
    case MCS_VOLUME: {
      const double mean_h = s[MS_VOL] / c.R;
      const double V0 = mc[MC_V0];
      double proj = 0.0, lc = 0.0;
      if (V0 > 1e-30 && isfinite(mean_h)) {
        const double vol_err = V0 - mean_h;
        if (fabs(vol_err) > 1e-14) {
          const double mean_hp2 = s[MS_VOL + 1] / c.R;
          lc = mean_hp2 > 1e-30 ? vol_err / mean_hp2 : 0.0;
          proj = 1.0;
        }
      }
      mc[MC_LC] = lc;
      mc[MC_PROJ] = proj;
      mc[MC_MEAN_H] = mean_h;
      break;
    }
    case MCS_FINAL:
      if (mc[MC_PROJ] != 0.0) mc[MC_MEAN_H] = s[MS_PROJ] / c.R;
      break;
  }
}
__global__ void __launch_bounds__(64) min_ctl_kernel(const double* s, double* mc, int stage, MinCtlConst c) {
  min_ctl_kernel_body(s, mc, stage, c, PF_BIDX, PF_GDIM);
}
struct KMinCtl {
  struct Args {
    const double* s;
    double* mc;
    int stage;
    MinCtlConst c;
  };
  static constexpr int BLOCK = 64;
  __device__ static void body(const Args& a, B3 b, B3 g) { min_ctl_kernel_body(a.s, a.mc, a.stage, a.c, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL(min_ctl_kernel, dim3(g.x, g.y, g.z), dim3(64), 0, s, a.s, a.mc, a.stage, a.c);
  }
};


__device__ __forceinline__ void min_project_kernel_body(double* phi, double lc, long n, double* partials, const double* mc, const B3 blockIdx, const B3 gridDim) {
  bool doit = true;
  if (mc) {
    doit = mc[MC_PROJ] != 0.0;
    lc = mc[MC_LC];
  }
  double s = 0.0;
  for (long i = (long)blockIdx.x * NT + threadIdx.x; doit && i < n; i += (long)gridDim.x * NT) {
    const double p = phi[i];
    const double q = clamp01(p + lc * h_prime_of_phi(p));
    phi[i] = q;
    s += h_of_phi(q);
  }
  const int kinds[1] = {RED_SUM};
  block_reduce_store<1>(&s, kinds, partials + blockIdx.x);
}
__global__ void __launch_bounds__(NT) min_project_kernel(double* phi, double lc, long n, double* partials, const double* mc) {
  min_project_kernel_body(phi, lc, n, partials, mc, PF_BIDX, PF_GDIM);
}
struct KMinProject {
  struct Args {
    double* phi;
    double lc;
    long n;
    double* partials;
    const double* mc;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { min_project_kernel_body(a.phi, a.lc, a.n, a.partials, a.mc, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL(min_project_kernel, dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.phi, a.lc, a.n, a.partials, a.mc);
  }
};


}  // namespace

// Lap^ = -k^2 phi^ on the compact spectrum (compute_laplacian_k_kernel,
// cuda_kernels.cu:3528-3538)
template <typename T>
__device__ __forceinline__ void kop_lap_body(KDims d, long n, const cx<T>* in, cx<T>* out, const B3 blockIdx, const B3 gridDim) {
  const long nact_ = (n / d.Az) * d.Azr;   // retained modes only (no kz padding)
  for (long j_ = (long)blockIdx.x * NT + threadIdx.x; j_ < nact_; j_ += (long)gridDim.x * NT) {
    const long i = kslot(d, j_);
    double kx, ky, kz;
    int iz;
    kvec(d, i, kx, ky, kz, iz);
    const double k2 = kx * kx + ky * ky + kz * kz;
    out[i] = cx<T>{T(-k2 * double(in[i].re)), T(-k2 * double(in[i].im))};
  }
}
template <typename T>
__global__ void __launch_bounds__(NT) kop_lap(KDims d, long n, const cx<T>* in, cx<T>* out) {
  kop_lap_body<T>(d, n, in, out, PF_BIDX, PF_GDIM);
}
template <typename T>
struct KLap {
  struct Args {
    KDims d;
    long n;
    const cx<T>* in;
    cx<T>* out;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { kop_lap_body<T>(a.d, a.n, a.in, a.out, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL((kop_lap<T>), dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.d, a.n, a.in, a.out);
  }
};


}  // namespace pf
