// Time-step engine: device buffers, plans and the stage sequence of one step.
//
// The stage order and time levels follow the reference step loop
// (main_cuda.cu:28427-31623):
//   elastic solve (28760-29166) -> phi update (29752-29825) ->
//   Y update (30283-30806) -> zero mode (30833-30911, 4269-4405) -> dY/dt (31618-31623)
// but every transform is a pruned pass sequence and every point-wise stage is
// fused into a z pass (see pf_ops.hpp), so no real-space intermediate other
// than tr(sigma) is written.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <array>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pfcore.h"
#include "pf_engine_api.hpp"
#include "pf_error.hpp"
#include "pf_ops.hpp"
#include "pf_minimize.hpp"
#include "pf_eta.hpp"
#include "pf_regfft.hpp"

#define PF_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return pf::fail(PF_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace pf {

// Supported transform lengths.  Each must factor into 4/2/3/5 and at most
// further primes handled by the direct prime stage.
#define PF_SIZES(X) \
  X(4) X(6) X(8) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(64) X(96) X(128) X(192) X(246) X(256) X(384) X(512)

inline bool size_supported(int n) {
  switch (n) {
#define PF_CASE_OK(v) case v:
    PF_SIZES(PF_CASE_OK)
    return true;
#undef PF_CASE_OK
    default:
      return false;
  }
}

struct AxisPlan {
  int N = 0, A = 0, a = 0;
  bool wrap = true;
  std::vector<double> k;   // k of each compact index, reference expression order
};

// Retained modes of one axis: |k| < (2/3) pi / d with k = 2 pi ii / (d N)
// evaluated exactly as dealias_23_kernel does (cuda_kernels.cu:118-131,
// cuda_common.h:69-83).
inline AxisPlan make_axis(int n, double d, bool wrap) {
  AxisPlan ap;
  ap.N = n;
  ap.wrap = wrap;
  const double kmax = M_PI * (2.0 / 3.0) / d;
  if (wrap) {
    int a = 0;
    while (a <= n / 2) {
      const double k = 2.0 * M_PI * a / (d * n);
      if (!(std::fabs(k) < kmax)) break;
      ++a;
    }
    ap.a = a;
    ap.A = 2 * a - 1;
    for (int c = 0; c < ap.A; ++c) {
      int i = c < a ? c : c + n - ap.A;
      int ii = (i <= n / 2) ? i : i - n;
      ap.k.push_back(2.0 * M_PI * ii / (d * n));
    }
  } else {
    int a = 0;
    while (a <= n / 2) {
      const double k = 2.0 * M_PI * a / (d * n);
      if (!(std::fabs(k) < kmax)) break;
      ++a;
    }
    ap.a = a;
    ap.A = a;
    for (int c = 0; c < a; ++c) ap.k.push_back(2.0 * M_PI * c / (d * n));
  }
  return ap;
}

template <typename T>
std::vector<cx<T>> twiddles(int n) {
  std::vector<cx<T>> w(n);
  for (int t = 0; t < n; ++t) {
    long double ang = 2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)n;
    w[t] = cx<T>{(T)cosl(ang), (T)(-sinl(ang))};
  }
  return w;
}

// ------------------------------------------------------------ misc kernels
struct RedKinds {
  int k[4];
};

namespace {

// Sums (or min/max) of block partials in a fixed order; one workgroup.
// Each thread combines blocks b = tid, tid + NT, ... in that order; the loads
// of FIN_U consecutive terms are issued together (one workgroup reads up to
// ~10^5 partials, so load latency, not bandwidth, sets its duration).
constexpr int FIN_U = 8;

__device__ __forceinline__ void finalize_kernel_body(const double* partials, int nblocks, int nred, RedKinds kinds, double* out, double scale, const B3 blockIdx, const B3 gridDim) {
  double v[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) v[r] = red_identity(kinds.k[r]);
  int b = threadIdx.x;
  for (; b + (FIN_U - 1) * NT < nblocks; b += FIN_U * NT) {
    double t[FIN_U][4];
#pragma unroll
    for (int u = 0; u < FIN_U; ++u)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (r < nred) t[u][r] = partials[(size_t)(b + u * NT) * nred + r];
#pragma unroll
    for (int u = 0; u < FIN_U; ++u)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (r < nred) v[r] = red_combine(kinds.k[r], v[r], t[u][r]);
  }
  for (; b < nblocks; b += NT)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (r < nred) v[r] = red_combine(kinds.k[r], v[r], partials[(size_t)b * nred + r]);
  __shared__ double outv[4];
  block_reduce_store<4>(v, kinds.k, outv);
  __syncthreads();
  if (threadIdx.x == 0)
    for (int r = 0; r < nred; ++r) out[r] = kinds.k[r] == RED_SUM ? outv[r] * scale : outv[r];
}
__global__ void __launch_bounds__(NT) finalize_kernel(const double* partials, int nblocks, int nred, RedKinds kinds, double* out, double scale) {
  finalize_kernel_body(partials, nblocks, nred, kinds, out, scale, PF_BIDX, PF_GDIM);
}
struct KFinalize {
  struct Args {
    const double* partials;
    int nblocks;
    int nred;
    RedKinds kinds;
    double* out;
    double scale;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { finalize_kernel_body(a.partials, a.nblocks, a.nred, a.kinds, a.out, a.scale, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL(finalize_kernel, dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.partials, a.nblocks, a.nred, a.kinds, a.out, a.scale);
  }
};


// dst[0..n) = src[0..n): the device scalars of a handle into a batch
// runner's gather buffer
__global__ void __launch_bounds__(NT) gather_kernel(const double* src, double* dst, int n) {
  for (int i = threadIdx.x; i < n; i += NT) dst[i] = src[i];
}
struct KGather {
  struct Args {
    const double* src;
    double* dst;
    int n;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3, B3) {
    for (int i = threadIdx.x; i < a.n; i += NT) a.dst[i] = a.src[i];
  }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL(gather_kernel, dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.src, a.dst, a.n);
  }
};

// Up to 16 sums of block partials in a fixed order, into the workgroup's
// shared outv[16] (valid after the call for every thread).
__device__ __forceinline__ void sum_partials16(const double* partials, int nblocks, int nred, double* outv) {
  double v[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = 0.0;
  int b = threadIdx.x;
  for (; b + (FIN_U - 1) * NT < nblocks; b += FIN_U * NT) {
    double t[FIN_U][16];
#pragma unroll
    for (int u = 0; u < FIN_U; ++u)
#pragma unroll
      for (int r = 0; r < 16; ++r)
        if (r < nred) t[u][r] = partials[(size_t)(b + u * NT) * nred + r];
#pragma unroll
    for (int u = 0; u < FIN_U; ++u)
#pragma unroll
      for (int r = 0; r < 16; ++r)
        if (r < nred) v[r] += t[u][r];
  }
  for (; b < nblocks; b += NT)
#pragma unroll
    for (int r = 0; r < 16; ++r)
      if (r < nred) v[r] += partials[(size_t)b * nred + r];
  int kinds[16];
  for (int r = 0; r < 16; ++r) kinds[r] = RED_SUM;
  block_reduce_store<16>(v, kinds, outv);
  __syncthreads();
}

// sum (1-h) x + v_B h, sum h(clamp01 phi), count(phi > 0.5), count(0.12 < phi < 0.88)
// (cuda_kernels.cu:3792-3828, 5580-5607, 3746-3790)
__global__ void __launch_bounds__(NT) mass_vf_kernel(const double* phi, const double* xB, double vB,
                                                     long n, double* partials) {
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (long g = blockIdx.x * (long)NT + threadIdx.x; g < n; g += (long)gridDim.x * NT) {
    const double p = phi[g];
    const double h = h_of_phi(p);
    acc[0] += (1.0 - h) * xB[g] + vB * h;
    acc[1] += h_of_phi(clamp01(p));
    acc[2] += (p > 0.5) ? 1.0 : 0.0;
    acc[3] += (p > 0.12 && p < 0.88) ? 1.0 : 0.0;
  }
  int kinds[4] = {RED_SUM, RED_SUM, RED_SUM, RED_SUM};
  block_reduce_store<4>(acc, kinds, partials + (size_t)blockIdx.x * 4);
}

// Logical block index with consecutive indices on one XCD: the 8 XCDs take
// workgroups round-robin, so block b runs on XCD b % 8.  Giving XCD x the
// contiguous run [x nb/8, (x+1) nb/8) keeps neighbouring rows (stencil
// halos) in one L2.  Identity when the grid is not a multiple of 8.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
}

// 0.5 kappa |grad phi|^2 by periodic central differences
// (compute_grad_energy_density_kernel, cuda_kernels.cu:773-809).
// Each block takes a contiguous run of z rows, GE_U rows at a time with all
// their loads issued together; row indices are wave-uniform (scalar unit) and
// the central differences multiply by 1/(2 d), which is exact for the unit
// spacing of the workloads and within one rounding of the reference's
// division otherwise.
constexpr int GE_U = 4;
__device__ __forceinline__ void grad_energy_kernel_body(const double* phi, int Nx, int Ny, int Nz, double dx, double dy, double dz, double kappa, double* partials, const B3 blockIdx, const B3 gridDim) {
  double acc[1] = {0.0};
  const int rows = Nx * Ny;
  const int L = xcd_block(blockIdx.x, gridDim.x);
  const int rb = (rows + gridDim.x - 1) / gridDim.x;
  const int r0 = L * rb;
  const int r1 = r0 + rb < rows ? r0 + rb : rows;
  const double hx = 1.0 / (2.0 * dx), hy = 1.0 / (2.0 * dy), hz = 1.0 / (2.0 * dz);
  for (int r = r0; r < r1; r += GE_U) {
    for (int k = threadIdx.x; k < Nz; k += NT) {
      const int kp = k + 1 == Nz ? 0 : k + 1, km = k == 0 ? Nz - 1 : k - 1;
      double v[GE_U][6];
#pragma unroll
      for (int q = 0; q < GE_U; ++q) {
        const int rr = r + q;
        if (rr < r1) {
          const int i = rr / Ny, j = rr - i * Ny;
          const int ip = i + 1 == Nx ? 0 : i + 1, im = i == 0 ? Nx - 1 : i - 1;
          const int jp = j + 1 == Ny ? 0 : j + 1, jm = j == 0 ? Ny - 1 : j - 1;
          const double* row = phi + (long)rr * Nz;
          v[q][0] = phi[((long)ip * Ny + j) * Nz + k];
          v[q][1] = phi[((long)im * Ny + j) * Nz + k];
          v[q][2] = phi[((long)i * Ny + jp) * Nz + k];
          v[q][3] = phi[((long)i * Ny + jm) * Nz + k];
          v[q][4] = row[kp];
          v[q][5] = row[km];
        }
      }
#pragma unroll
      for (int q = 0; q < GE_U; ++q) {
        if (r + q < r1) {
          const double gx = (v[q][0] - v[q][1]) * hx;
          const double gy = (v[q][2] - v[q][3]) * hy;
          const double gz = (v[q][4] - v[q][5]) * hz;
          acc[0] += 0.5 * kappa * (gx * gx + gy * gy + gz * gz);
        }
      }
    }
  }
  int kinds[1] = {RED_SUM};
  block_reduce_store<1>(acc, kinds, partials + blockIdx.x);
}
__global__ void __launch_bounds__(NT) grad_energy_kernel(const double* phi, int Nx, int Ny, int Nz, double dx, double dy, double dz, double kappa, double* partials) {
  grad_energy_kernel_body(phi, Nx, Ny, Nz, dx, dy, dz, kappa, partials, PF_BIDX, PF_GDIM);
}
struct KGradEnergy {
  struct Args {
    const double* phi;
    int Nx;
    int Ny;
    int Nz;
    double dx;
    double dy;
    double dz;
    double kappa;
    double* partials;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { grad_energy_kernel_body(a.phi, a.Nx, a.Ny, a.Nz, a.dx, a.dy, a.dz, a.kappa, a.partials, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL(grad_energy_kernel, dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.phi, a.Nx, a.Ny, a.Nz, a.dx, a.dy, a.dz, a.kappa, a.partials);
  }
};


// Zero-mode evaluation (compute_pf_Y_zero_mode_terms_kernel, cuda_kernels.cu:3409-3420):
// mass(lambda) = sum (1-h) sigma(Y*+lambda) + h v_B, its derivative, and the
// count of shifted values outside [Y_lo, Y_hi] (validate_pf_Y_zero_mode_bounds_kernel 3422-3432).
//
// The zero-mode root find (run_pf_conserved_y_zero_mode_host, main_cuda.cu:4269-4405)
// with its decisions taken on the device, so the host waits once per step
// instead of once per evaluation.  The three evaluations that open the search
// (the bracket ends lo, hi and the start clamp(0, lo, hi)) share one pass over
// the field; each later Newton/bisection step is an evaluation, its partial
// sums and a one-thread controller that applies the host loop's rules in the
// same FP64 operations.  Steps enqueued beyond convergence see ZS_ACTIVE == 0
// and return at once.
enum ZmSlot {
  ZS_LO = 0, ZS_HI, ZS_LAM, ZS_R, ZS_DERIV, ZS_INVALID, ZS_ACTIVE, ZS_STATUS, ZS_ITERS, ZS_NEXT, ZS_NEWTON,
  ZS_BISECT,
  ZS_SUMS = 16,   // 9 sums of the opening pass, then 3 of each step
  ZS_COUNT = 32
};
enum ZmStatus { ZM_OK = 0, ZM_EMPTY_BRACKET = 1, ZM_NONFINITE = 2, ZM_OUTSIDE = 3 };
// Y* + lambda is bounded to [-Y_clip, Y_cap]: Y_cap = Y_clip in the dynamics,
// the minimize cap logit(xB_max_safe) under the minimize mass constraint
struct ZmConst {
  double Y_clip, Y_cap, vB, target, tol;
  int max_it;
};
// an accepted (or the last rejected) search: shift, residual, d mass/d lambda, steps
struct ZmResultT {
  double lam = NAN, r = NAN, deriv = NAN;
  int iters = 0, newton = 0, bisect = 0;
  double ymin = NAN, ymax = NAN, lo = NAN, hi = NAN;   // Y* extremes, final bracket
};

constexpr int ZM_U = 4;   // grid strides per iteration of the opening zero-mode evaluation

__device__ __forceinline__ void zm_terms(double ys, double ph, double lam, double vB, double ylo, double yhi,
                                         double* acc) {
  const double h = h_of_phi(ph);
  const double alpha = 1.0 - h;
  const double y = ys + lam;
  const double x = sigmoid_unclipped(y);
  acc[0] += alpha * x + h * vB;
  acc[1] += alpha * x * (1.0 - x);
  if (!isfinite(y) || y < ylo || y > yhi) acc[2] += 1.0;
}

// opening pass: lambda = lo, hi, clamp(0, lo, hi) from the Y* extremes ymm[0..1]
__device__ __forceinline__ void zm_eval3_kernel_body(const double* Ystar, const double* phi, const double* ymm, double Y_clip, double Y_cap, double vB, long n, double* partials, const B3 blockIdx, const B3 gridDim) {
  const double lo = -Y_clip - ymm[0];
  const double hi = Y_cap - ymm[1];
  const double lam0 = fmin(fmax(0.0, lo), hi);
  double acc[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  // ZM_U grid strides' loads issued together; the terms are added in the
  // order of the plain grid-stride loop
  const long stride = (long)gridDim.x * NT;
  for (long g0 = blockIdx.x * (long)NT + threadIdx.x; g0 < n; g0 += ZM_U * stride) {
    double ys[ZM_U], ph[ZM_U];
#pragma unroll
    for (int u = 0; u < ZM_U; ++u) {
      const long g = g0 + u * stride;
      ys[u] = g < n ? Ystar[g] : 0.0;
      ph[u] = g < n ? phi[g] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < ZM_U; ++u) {
      if (g0 + u * stride >= n) break;
      zm_terms(ys[u], ph[u], lo, vB, -Y_clip, Y_cap, acc);
      zm_terms(ys[u], ph[u], hi, vB, -Y_clip, Y_cap, acc + 3);
      zm_terms(ys[u], ph[u], lam0, vB, -Y_clip, Y_cap, acc + 6);
    }
  }
  int kinds[9];
#pragma unroll
  for (int r = 0; r < 9; ++r) kinds[r] = RED_SUM;
  block_reduce_store<9>(acc, kinds, partials + (size_t)blockIdx.x * 9);
}
__global__ void __launch_bounds__(NT) zm_eval3_kernel(const double* Ystar, const double* phi, const double* ymm, double Y_clip, double Y_cap, double vB, long n, double* partials) {
  zm_eval3_kernel_body(Ystar, phi, ymm, Y_clip, Y_cap, vB, n, partials, PF_BIDX, PF_GDIM);
}
struct KZmEval3 {
  struct Args {
    const double* Ystar;
    const double* phi;
    const double* ymm;
    double Y_clip;
    double Y_cap;
    double vB;
    long n;
    double* partials;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { zm_eval3_kernel_body(a.Ystar, a.phi, a.ymm, a.Y_clip, a.Y_cap, a.vB, a.n, a.partials, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL(zm_eval3_kernel, dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.Ystar, a.phi, a.ymm, a.Y_clip, a.Y_cap, a.vB, a.n, a.partials);
  }
};


// one search step at st[ZS_NEXT]
__device__ __forceinline__ void zm_eval_kernel_body(const double* Ystar, const double* phi, const double* st, double Y_clip, double Y_cap, double vB, long n, double* partials, const B3 blockIdx, const B3 gridDim) {
  if (st[ZS_ACTIVE] == 0.0) return;
  const double lam = st[ZS_NEXT];
  double acc[3] = {0.0, 0.0, 0.0};
  for (long g = blockIdx.x * (long)NT + threadIdx.x; g < n; g += (long)gridDim.x * NT)
    zm_terms(Ystar[g], phi[g], lam, vB, -Y_clip, Y_cap, acc);
  int kinds[3] = {RED_SUM, RED_SUM, RED_SUM};
  block_reduce_store<3>(acc, kinds, partials + (size_t)blockIdx.x * 3);
}
__global__ void __launch_bounds__(NT) zm_eval_kernel(const double* Ystar, const double* phi, const double* st, double Y_clip, double Y_cap, double vB, long n, double* partials) {
  zm_eval_kernel_body(Ystar, phi, st, Y_clip, Y_cap, vB, n, partials, PF_BIDX, PF_GDIM);
}
struct KZmEval {
  struct Args {
    const double* Ystar;
    const double* phi;
    const double* st;
    double Y_clip;
    double Y_cap;
    double vB;
    long n;
    double* partials;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { zm_eval_kernel_body(a.Ystar, a.phi, a.st, a.Y_clip, a.Y_cap, a.vB, a.n, a.partials, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL(zm_eval_kernel, dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.Ystar, a.phi, a.st, a.Y_clip, a.Y_cap, a.vB, a.n, a.partials);
  }
};


// the loop condition and the next candidate (main_cuda.cu:4370-4390)
__device__ void zm_propose(double* st, const ZmConst& c) {
  const double r = st[ZS_R];
  if (!(st[ZS_ITERS] < (double)c.max_it && fabs(r) > c.tol)) {
    st[ZS_ACTIVE] = 0.0;
    return;
  }
  const double lam = st[ZS_LAM];
  double lo = st[ZS_LO], hi = st[ZS_HI];
  if (r < 0.0)
    lo = lam;
  else
    hi = lam;
  const double deriv = st[ZS_DERIV];
  double cand = NAN;
  if (deriv > 0.0 && isfinite(deriv)) cand = lam - r / deriv;
  if (isfinite(cand) && cand > lo && cand < hi) {
    st[ZS_NEWTON] += 1.0;
  } else {
    cand = 0.5 * (lo + hi);
    st[ZS_BISECT] += 1.0;
  }
  st[ZS_LO] = lo;
  st[ZS_HI] = hi;
  st[ZS_NEXT] = cand;
  st[ZS_ACTIVE] = 1.0;
}

__device__ void zm_ctrl_open(double* st, const double* ymm, ZmConst c) {
  const double* s = st + ZS_SUMS;
  const double lo = -c.Y_clip - ymm[0];
  const double hi = c.Y_cap - ymm[1];
  st[ZS_ACTIVE] = 0.0;
  st[ZS_ITERS] = 0.0;
  st[ZS_NEWTON] = 0.0;
  st[ZS_BISECT] = 0.0;
  st[ZS_LO] = lo;
  st[ZS_HI] = hi;
  st[ZS_LAM] = NAN;
  st[ZS_R] = NAN;
  st[ZS_INVALID] = 0.0;
  const double roundoff = 128.0 * 2.220446049250313e-16 * fmax(fmax(fabs(lo), fabs(hi)), 1.0);
  if (!isfinite(lo) || !isfinite(hi) || lo > hi + roundoff) {
    st[ZS_STATUS] = ZM_EMPTY_BRACKET;
    return;
  }
  if (!isfinite(s[0]) || !isfinite(s[1]) || !isfinite(s[3]) || !isfinite(s[4])) {
    st[ZS_STATUS] = ZM_NONFINITE;
    return;
  }
  const double F_lo = s[0] - c.target, F_hi = s[3] - c.target;
  if (F_lo > c.tol || F_hi < -c.tol) {
    st[ZS_STATUS] = ZM_OUTSIDE;
    return;
  }
  if (!isfinite(s[6]) || !isfinite(s[7])) {
    st[ZS_STATUS] = ZM_NONFINITE;
    return;
  }
  st[ZS_STATUS] = ZM_OK;
  st[ZS_LO] = lo;
  st[ZS_HI] = hi;
  st[ZS_LAM] = fmin(fmax(0.0, lo), hi);
  st[ZS_R] = s[6] - c.target;
  st[ZS_DERIV] = s[7];
  st[ZS_INVALID] = s[8];
  zm_propose(st, c);
}

__device__ void zm_ctrl_step(double* st, ZmConst c) {
  const double* s = st + ZS_SUMS;
  st[ZS_LAM] = st[ZS_NEXT];
  if (!isfinite(s[0]) || !isfinite(s[1])) {
    st[ZS_STATUS] = ZM_NONFINITE;
    st[ZS_ACTIVE] = 0.0;
    return;
  }
  st[ZS_R] = s[0] - c.target;
  st[ZS_DERIV] = s[1];
  st[ZS_INVALID] = s[2];
  st[ZS_ITERS] += 1.0;
  zm_propose(st, c);
}

// The evaluation's sums into the search state, then the bracket set-up
// (open) or one safeguarded step (step); one workgroup.  A step launched
// after the search has ended does nothing.
__device__ __forceinline__ void zm_open_end_kernel_body(const double* partials, int nblocks, double* st, const double* ymm, ZmConst c, const B3 blockIdx, const B3 gridDim) {
  __shared__ double outv[16];
  __shared__ double ss[ZS_COUNT];   // the search state, worked on in LDS by one thread
  for (int i = threadIdx.x; i < ZS_COUNT; i += NT) ss[i] = st[i];
  sum_partials16(partials, nblocks, 9, outv);
  if (threadIdx.x == 0) {
    for (int r = 0; r < 9; ++r) ss[ZS_SUMS + r] = outv[r];
    zm_ctrl_open(ss, ymm, c);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < ZS_COUNT; i += NT) st[i] = ss[i];
}
__global__ void __launch_bounds__(NT) zm_open_end_kernel(const double* partials, int nblocks, double* st, const double* ymm, ZmConst c) {
  zm_open_end_kernel_body(partials, nblocks, st, ymm, c, PF_BIDX, PF_GDIM);
}
struct KZmOpenEnd {
  struct Args {
    const double* partials;
    int nblocks;
    double* st;
    const double* ymm;
    ZmConst c;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { zm_open_end_kernel_body(a.partials, a.nblocks, a.st, a.ymm, a.c, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL(zm_open_end_kernel, dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.partials, a.nblocks, a.st, a.ymm, a.c);
  }
};

__device__ __forceinline__ void zm_step_end_kernel_body(const double* partials, int nblocks, double* st, ZmConst c, const B3 blockIdx, const B3 gridDim) {
  if (st[ZS_ACTIVE] == 0.0) return;
  __shared__ double outv[16];
  __shared__ double ss[ZS_COUNT];
  for (int i = threadIdx.x; i < ZS_COUNT; i += NT) ss[i] = st[i];
  sum_partials16(partials, nblocks, 3, outv);
  if (threadIdx.x == 0) {
    for (int r = 0; r < 3; ++r) ss[ZS_SUMS + r] = outv[r];
    zm_ctrl_step(ss, c);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < ZS_COUNT; i += NT) st[i] = ss[i];
}
__global__ void __launch_bounds__(NT) zm_step_end_kernel(const double* partials, int nblocks, double* st, ZmConst c) {
  zm_step_end_kernel_body(partials, nblocks, st, c, PF_BIDX, PF_GDIM);
}
struct KZmStepEnd {
  struct Args {
    const double* partials;
    int nblocks;
    double* st;
    ZmConst c;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { zm_step_end_kernel_body(a.partials, a.nblocks, a.st, a.c, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL(zm_step_end_kernel, dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.partials, a.nblocks, a.st, a.c);
  }
};


// ---- accelerated elastic solver, decided on the device (control block AaCtl)
struct AaConst {
  int m, iter_min, iter_max;
  double tol, floor;
};

__device__ __forceinline__ void aa_init_kernel_body(double* ctl, const B3 blockIdx, const B3 gridDim) {
  for (int i = threadIdx.x; i < AC_COUNT; i += blockDim.x) ctl[i] = 0.0;
  if (threadIdx.x == 0) {
    ctl[AC_ACTIVE] = 1.0;
    ctl[AC_RES] = NAN;
  }
}
__global__ void __launch_bounds__(64) aa_init_kernel(double* ctl) {
  aa_init_kernel_body(ctl, PF_BIDX, PF_GDIM);
}
struct KAaInit {
  struct Args {
    double* ctl;
  };
  static constexpr int BLOCK = 64;
  __device__ static void body(const Args& a, B3 b, B3 g) { aa_init_kernel_body(a.ctl, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL(aa_init_kernel, dim3(g.x, g.y, g.z), dim3(64), 0, s, a.ctl);
  }
};


// mixing coefficients: min |g - sum gamma_j dg_j| by the normal equations
// with a 1e-12 relative Tikhonov shift, Gaussian elimination with partial
// pivoting
__device__ void aa_coef(double* ctl) {
  if (ctl[AC_ACTIVE] == 0.0) return;
  const int nh = (int)ctl[AC_NH];
  double gam[AA_MAX] = {0.0, 0.0, 0.0, 0.0};
  if (nh > 0) {
    // Every index is a compile-time constant (loops over AA_MAX with nh
    // guards, the pivot row found by value), so the system stays in registers
    // instead of private memory; the operations are those of the plain
    // elimination over nh rows, in the same order.
    double A[AA_MAX][AA_MAX], b[AA_MAX];
    double tr = 0.0;
#pragma unroll
    for (int i = 0; i < AA_MAX; ++i)
      if (i < nh) tr += ctl[AC_H + i * AA_MAX + i];
#pragma unroll
    for (int i = 0; i < AA_MAX; ++i) {
#pragma unroll
      for (int j = 0; j < AA_MAX; ++j)
        A[i][j] = (i < nh && j < nh) ? ctl[AC_H + i * AA_MAX + j] + (i == j ? 1e-12 * tr / nh : 0.0) : 0.0;
      b[i] = i < nh ? ctl[AC_B + i] : 0.0;
    }
#pragma unroll
    for (int col = 0; col < AA_MAX; ++col) {
      if (col < nh) {
        int piv = col;
        double pv = fabs(A[col][col]);
#pragma unroll
        for (int r = col + 1; r < AA_MAX; ++r)
          if (r < nh && fabs(A[r][col]) > pv) {
            piv = r;
            pv = fabs(A[r][col]);
          }
#pragma unroll
        for (int r = col + 1; r < AA_MAX; ++r)
          if (r == piv) {
#pragma unroll
            for (int q = 0; q < AA_MAX; ++q) {
              const double t = A[col][q];
              A[col][q] = A[r][q];
              A[r][q] = t;
            }
            const double t = b[col];
            b[col] = b[r];
            b[r] = t;
          }
        if (A[col][col] != 0.0) {
#pragma unroll
          for (int r = col + 1; r < AA_MAX; ++r)
            if (r < nh) {
              const double f = A[r][col] / A[col][col];
#pragma unroll
              for (int q = col; q < AA_MAX; ++q)
                if (q < nh) A[r][q] -= f * A[col][q];
              b[r] -= f * b[col];
            }
        }
      }
    }
#pragma unroll
    for (int i = AA_MAX - 1; i >= 0; --i)
      if (i < nh) {
        double v = b[i];
#pragma unroll
        for (int j = i + 1; j < AA_MAX; ++j)
          if (j < nh) v -= A[i][j] * gam[j];
        gam[i] = (A[i][i] != 0.0) ? v / A[i][i] : 0.0;
      }
  }
  for (int j = 0; j < AA_MAX; ++j) ctl[AC_GAM + j] = gam[j];
}

// after a Green application: residual, history Gram matrix, stopping rule
__device__ void aa_ctrl(double* ctl, int mode, AaConst c) {
  if (ctl[AC_ACTIVE] == 0.0) return;
  const double* s = ctl + AC_SUMS;
  const int it = (int)ctl[AC_IT] + 1;
  ctl[AC_IT] = it;
  const double gg = s[0], ff = s[1];
  if (!isfinite(gg) || !isfinite(ff) || gg < 0.0 || ff < 0.0) {
    ctl[AC_STATUS] = 1.0;
    ctl[AC_ACTIVE] = 0.0;
    return;
  }
  const double res = sqrt(gg / fmax(ff, c.floor));
  ctl[AC_RES] = res;
  int nh = (int)ctl[AC_NH];
  if (mode == 1 && c.m > 0) {
    const int next = (int)ctl[AC_NEXT];
    const int nh_after = nh + 1 < c.m ? nh + 1 : c.m;
    for (int j = 0; j < nh_after; ++j) {
      ctl[AC_H + next * AA_MAX + j] = s[2 + j];
      ctl[AC_H + j * AA_MAX + next] = s[2 + j];
    }
    nh = nh_after;
    ctl[AC_NH] = nh;
    ctl[AC_NEXT] = (next + 1) % c.m;
  }
  for (int j = 0; j < nh; ++j) ctl[AC_B + j] = s[2 + AA_MAX + j];
  if (it >= c.iter_min && res <= c.tol) {
    ctl[AC_CONV] = 1.0;
    ctl[AC_ACTIVE] = 0.0;
  } else if (it >= c.iter_max) {
    ctl[AC_ACTIVE] = 0.0;
  }
}

// End of an iteration, one workgroup: the Green update's sums into the
// control block, the residual / history / stopping decision, and the next
// iteration's mixing coefficients (unused when the solve has stopped).
__device__ __forceinline__ void aa_end_kernel_body(const double* partials, int nblocks, double* ctl, int mode, AaConst c, const B3 blockIdx, const B3 gridDim) {
  if (ctl[AC_ACTIVE] == 0.0) return;
  __shared__ double outv[16];
  __shared__ double sc[AC_COUNT];   // the control block, worked on in LDS by one thread
  for (int i = threadIdx.x; i < AC_COUNT; i += NT) sc[i] = ctl[i];
  sum_partials16(partials, nblocks, AA_NRED, outv);
  if (threadIdx.x == 0) {
    for (int r = 0; r < AA_NRED; ++r) sc[AC_SUMS + r] = outv[r];
    aa_ctrl(sc, mode, c);
    aa_coef(sc);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < AC_COUNT; i += NT) ctl[i] = sc[i];
}
__global__ void __launch_bounds__(NT) aa_end_kernel(const double* partials, int nblocks, double* ctl, int mode, AaConst c) {
  aa_end_kernel_body(partials, nblocks, ctl, mode, c, PF_BIDX, PF_GDIM);
}
struct KAaEnd {
  struct Args {
    const double* partials;
    int nblocks;
    double* ctl;
    int mode;
    AaConst c;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { aa_end_kernel_body(a.partials, a.nblocks, a.ctl, a.mode, a.c, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL(aa_end_kernel, dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.partials, a.nblocks, a.ctl, a.mode, a.c);
  }
};


// Y = Y* + lambda, x_B = sigma(Y) unclipped (apply_pf_Y_zero_mode_shift_kernel
// 3434-3443), dY/dt = (Y - Y^n)/dt (update_dY_dt_prev_kernel 3545-3556) when
// dYdt is given; the minimize mass constraint keeps the dY/dt of its clamp
// (main_cuda.cu:31622 runs before the constraint at 32297).
__device__ __forceinline__ void zm_apply_kernel_body(const double* Ystar, double lam, double* Y, double* xB, double* dYdt, double dt, long n, const B3 blockIdx, const B3 gridDim) {
  for (long g = blockIdx.x * (long)NT + threadIdx.x; g < n; g += (long)gridDim.x * NT) {
    const double y = Ystar[g] + lam;
    if (dYdt) dYdt[g] = (y - Y[g]) / dt;
    Y[g] = y;
    xB[g] = sigmoid_unclipped(y);
  }
}
__global__ void __launch_bounds__(NT) zm_apply_kernel(const double* Ystar, double lam, double* Y, double* xB, double* dYdt, double dt, long n) {
  zm_apply_kernel_body(Ystar, lam, Y, xB, dYdt, dt, n, PF_BIDX, PF_GDIM);
}
struct KZmApply {
  struct Args {
    const double* Ystar;
    double lam;
    double* Y;
    double* xB;
    double* dYdt;
    double dt;
    long n;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 b, B3 g) { zm_apply_kernel_body(a.Ystar, a.lam, a.Y, a.xB, a.dYdt, a.dt, a.n, b, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL(zm_apply_kernel, dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.Ystar, a.lam, a.Y, a.xB, a.dYdt, a.dt, a.n);
  }
};


}  // namespace

// identity operators used by the transform test entry points
template <typename T>
struct OpLoadReal {
  struct Args {
    const double* f;
  };
  static constexpr int NRED = 0;
  static constexpr int kinds[1] = {0};
  static constexpr int NPRE = 1;
  __device__ static void load(const Args& a, size_t g, double* pre) { pre[0] = a.f[g]; }
  __device__ static void apply(const Args&, size_t, const double* pre, const T*, T* out, double*) { out[0] = T(pre[0]); }
};
template <typename T>
struct OpStoreReal {
  struct Args {
    double* f;
  };
  static constexpr int NRED = 0;
  static constexpr int kinds[1] = {0};
  static constexpr int NPRE = 0;
  __device__ static void load(const Args&, size_t, double*) {}
  __device__ static void apply(const Args& a, size_t g, const double*, const T* vin, T*, double*) { a.f[g] = double(vin[0]); }
};

// ------------------------------------------------------------ learned warm start
// Network input at each voxel: phi s_phi and the eigenstress
// (C + h dC):eps0 s_tau, BF16 channels-last [S][16].
namespace {
__global__ void __launch_bounds__(NT) nn_features_kernel(PhysParams P, const double* phi, const double* xB, long n,
                                                         double s_phi, double s_tau, uint16_t* X) {
  for (long e = (long)blockIdx.x * NT + threadIdx.x; e < n; e += (long)gridDim.x * NT) {
    const double ph = phi[e];
    const double h = h_of_phi(ph);
    double e0[6], tau[6];
    eigenstrain(P, h, clamp01(xB[e]), e0);
    voigt_apply_h<double>(P, h, e0, tau);
    uint16_t o[16];
    o[0] = nn::f2bf((float)(ph * s_phi));
#pragma unroll
    for (int c = 0; c < 6; ++c) o[1 + c] = nn::f2bf((float)(tau[c] * s_tau));
#pragma unroll
    for (int c = 7; c < 16; ++c) o[c] = 0;
    uint4* dst = reinterpret_cast<uint4*>(X + (size_t)e * 16);
    dst[0] = make_uint4(nn::pk(o[0], o[1]), nn::pk(o[2], o[3]), nn::pk(o[4], o[5]), nn::pk(o[6], o[7]));
    dst[1] = make_uint4(nn::pk(o[8], o[9]), nn::pk(o[10], o[11]), nn::pk(o[12], o[13]), nn::pk(o[14], o[15]));
  }
}
// Features of a pooled network: each coarse voxel (x, y, z) of an (Nx/2,
// Ny/2, Nz/2) grid holds the mean of the FP64 features of the 8 voxels
// (2x + i, 2y + j, 2z + k), in the BF16 layout of nn_features_kernel.
__global__ void __launch_bounds__(NT) nn_features_pool_kernel(PhysParams P, const double* phi, const double* xB,
                                                              int Nx, int Ny, int Nz, double s_phi, double s_tau,
                                                              uint16_t* X) {
  const int cy = Ny / 2, cz = Nz / 2;
  const long nc = (long)(Nx / 2) * cy * cz;
  for (long e = (long)blockIdx.x * NT + threadIdx.x; e < nc; e += (long)gridDim.x * NT) {
    const int z = (int)(e % cz), y = (int)((e / cz) % cy), x = (int)(e / ((long)cz * cy));
    double f[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < 8; ++q) {
      const long g = ((long)(2 * x + (q >> 2)) * Ny + (2 * y + ((q >> 1) & 1))) * Nz + 2 * z + (q & 1);
      const double ph = phi[g];
      const double h = h_of_phi(ph);
      double e0[6], tau[6];
      eigenstrain(P, h, clamp01(xB[g]), e0);
      voigt_apply_h<double>(P, h, e0, tau);
      f[0] += ph;
#pragma unroll
      for (int c = 0; c < 6; ++c) f[1 + c] += tau[c];
    }
    uint16_t o[16];
    o[0] = nn::f2bf((float)(0.125 * f[0] * s_phi));
#pragma unroll
    for (int c = 0; c < 6; ++c) o[1 + c] = nn::f2bf((float)(0.125 * f[1 + c] * s_tau));
#pragma unroll
    for (int c = 7; c < 16; ++c) o[c] = 0;
    uint4* dst = reinterpret_cast<uint4*>(X + (size_t)e * 16);
    dst[0] = make_uint4(nn::pk(o[0], o[1]), nn::pk(o[2], o[3]), nn::pk(o[4], o[5]), nn::pk(o[6], o[7]));
    dst[1] = make_uint4(nn::pk(o[8], o[9]), nn::pk(o[10], o[11]), nn::pk(o[12], o[13]), nn::pk(o[14], o[15]));
  }
}

// FP32 planar [3][Nx*Ny*Nz] -> the 2x2x2 means [3][(Nx/2)(Ny/2)(Nz/2)] (a
// pooled network's training target)
__global__ void __launch_bounds__(NT) pool3_kernel(const float* in, int Nx, int Ny, int Nz, float* out) {
  const int cy = Ny / 2, cz = Nz / 2;
  const long nc = (long)(Nx / 2) * cy * cz, n = (long)Nx * Ny * Nz;
  for (long e = (long)blockIdx.x * NT + threadIdx.x; e < 3 * nc; e += (long)gridDim.x * NT) {
    const int c = (int)(e / nc);
    const long v = e % nc;
    const int z = (int)(v % cz), y = (int)((v / cz) % cy), x = (int)(v / ((long)cz * cy));
    float acc = 0.f;
    for (int q = 0; q < 8; ++q)
      acc += in[c * n + ((long)(2 * x + (q >> 2)) * Ny + (2 * y + ((q >> 1) & 1))) * Nz + 2 * z + (q & 1)];
    out[c * nc + v] = 0.125f * acc;
  }
}
}  // namespace

// three FP32 planar real fields [3][R] (times scale) -> forward z of 3 components
template <typename T>
struct OpLoadF3 {
  struct Args {
    const float* y3;
    size_t R;        // voxels of y3 per component
    double scale;
    int pool;        // 2: y3 is on the (Nx/2, Ny/2, Nz/2) grid, read at each voxel's coarse cell
    int Ny, Nz;
  };
  static constexpr int NRED = 0;
  static constexpr int kinds[1] = {0};
  static constexpr int NPRE = 3;
  __device__ static void load(const Args& a, size_t g, double* pre) {
    if (a.pool == 2) {
      const size_t z = g % a.Nz, y = (g / a.Nz) % a.Ny, x = g / ((size_t)a.Nz * a.Ny);
      g = ((x >> 1) * (a.Ny >> 1) + (y >> 1)) * (a.Nz >> 1) + (z >> 1);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) pre[c] = (double)a.y3[c * a.R + g];
  }
  __device__ static void apply(const Args& a, size_t, const double* pre, const T*, T* out, double*) {
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = T(pre[c] * a.scale);
  }
};
// inverse z of 3 components -> FP32 planar real fields [3][R]
template <typename T>
struct OpStoreF3 {
  struct Args {
    float* y3;
    size_t R;
    T invN;
  };
  static constexpr int NRED = 0;
  static constexpr int kinds[1] = {0};
  static constexpr int NPRE = 0;
  __device__ static void load(const Args&, size_t, double*) {}
  __device__ static void apply(const Args& a, size_t g, const double*, const T* vin, T*, double*) {
#pragma unroll
    for (int c = 0; c < 3; ++c) a.y3[c * a.R + g] = (float)(double(vin[c]) * double(a.invN));
  }
};

// u_i += d_i on the compact spectrum, k = 0 excluded (the Green operator's
// zero mode stays 0); kop_add3_state adds into solver state (retained-mode
// index) from k-space slots
template <typename T, typename S>
__global__ void __launch_bounds__(NT) kop_add3(long n, const cx<T>* d0, const cx<T>* d1, const cx<T>* d2, S* u0,
                                               S* u1, S* u2) {
  for (long c = (long)blockIdx.x * NT + threadIdx.x; c < n; c += (long)gridDim.x * NT) {
    if (c == 0) continue;
    u0[c].re += d0[c].re;
    u0[c].im += d0[c].im;
    u1[c].re += d1[c].re;
    u1[c].im += d1[c].im;
    u2[c].re += d2[c].re;
    u2[c].im += d2[c].im;
  }
}
template <typename T>
__global__ void __launch_bounds__(NT) kop_add3_state(KDims d, long n, const cx<T>* d0, const cx<T>* d1,
                                                     const cx<T>* d2, cx<double>* u0, cx<double>* u1,
                                                     cx<double>* u2) {
  const long nact_ = (n / d.Az) * d.Azr;
  for (long j_ = (long)blockIdx.x * NT + threadIdx.x; j_ < nact_; j_ += (long)gridDim.x * NT) {
    if (j_ == 0) continue;
    const long c = kslot(d, j_);
    u0[j_].re += d0[c].re;
    u0[j_].im += d0[c].im;
    u1[j_].re += d1[c].re;
    u1[j_].im += d1[c].im;
    u2[j_].re += d2[c].re;
    u2[j_].im += d2[c].im;
  }
}
// The network's 3 channels are a force density f whose Green response is the
// displacement correction: u^ = Omega(k) f^ with Omega = K^-1 the Green
// tensor of the homogeneous stiffness (the operator of the reference's
// Green update, cuda_kernels.cu:6484-6633).  TO_FORCE applies K (target
// side), otherwise Omega (prediction side); k = 0 maps to 0 either way.
template <typename T>
struct ForceMapArgs {
  KDims d;
  long n;
  cx<T>* a[3];
  double C[21];
  int to_force;
};
template <typename T>
__global__ void __launch_bounds__(NT) kop_force_map(ForceMapArgs<T> fa) {
  const long nact_ = (fa.n / fa.d.Az) * fa.d.Azr;   // retained modes only (no kz padding)
  for (long j_ = (long)blockIdx.x * NT + threadIdx.x; j_ < nact_; j_ += (long)gridDim.x * NT) {
    const long c = kslot(fa.d, j_);
    double kx, ky, kz;
    int iz;
    kvec(fa.d, c, kx, ky, kz, iz);
    double M[6];
    bool ok = c != 0;
    if (ok) {
      if (fa.to_force)
        acoustic_tensor(fa.C, kx, ky, kz, M);
      else
        ok = green_tensor(fa.C, kx, ky, kz, M);
    }
    cx<double> v[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) v[i] = cx<double>{double(fa.a[i][c].re), double(fa.a[i][c].im)};
    const double m[3][3] = {{M[0], M[1], M[2]}, {M[1], M[3], M[4]}, {M[2], M[4], M[5]}};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double re = 0.0, im = 0.0;
      if (ok) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          re += m[i][j] * v[j].re;
          im += m[i][j] * v[j].im;
        }
      }
      fa.a[i][c] = cx<T>{T(re), T(im)};
    }
  }
}

// d_i = u_i - v_i
template <typename T, typename S>
__global__ void __launch_bounds__(NT) kop_diff3(long n, const S* u0, const S* u1, const S* u2, const cx<T>* v0,
                                                const cx<T>* v1, const cx<T>* v2, cx<T>* d0, cx<T>* d1, cx<T>* d2) {
  for (long c = (long)blockIdx.x * NT + threadIdx.x; c < n; c += (long)gridDim.x * NT) {
    d0[c] = cx<T>{T(double(u0[c].re) - double(v0[c].re)), T(double(u0[c].im) - double(v0[c].im))};
    d1[c] = cx<T>{T(double(u1[c].re) - double(v1[c].re)), T(double(u1[c].im) - double(v1[c].im))};
    d2[c] = cx<T>{T(double(u2[c].re) - double(v2[c].re)), T(double(u2[c].im) - double(v2[c].im))};
  }
}

// d = u - v with u solver state (retained-mode index), v and d at k-space slots
template <typename T>
__global__ void __launch_bounds__(NT) kop_diff3_state(KDims dm, long n, const cx<double>* u0, const cx<double>* u1,
                                                      const cx<double>* u2, const cx<T>* v0, const cx<T>* v1,
                                                      const cx<T>* v2, cx<T>* d0, cx<T>* d1, cx<T>* d2) {
  const long nact_ = (n / dm.Az) * dm.Azr;
  for (long j_ = (long)blockIdx.x * NT + threadIdx.x; j_ < nact_; j_ += (long)gridDim.x * NT) {
    const long c = kslot(dm, j_);
    d0[c] = cx<T>{T(u0[j_].re - double(v0[c].re)), T(u0[j_].im - double(v0[c].im))};
    d1[c] = cx<T>{T(u1[j_].re - double(v1[c].re)), T(u1[j_].im - double(v1[c].im))};
    d2[c] = cx<T>{T(u2[j_].re - double(v2[c].re)), T(u2[j_].im - double(v2[c].im))};
  }
}

// ------------------------------------------------------------ dispatch
template <typename T, int N, int NCI, int NCO, int NPRE, int BUDGET = ZPASS_LDS_BUDGET>
constexpr int zpass_rows_for() {
  constexpr int NCM = NCI > NCO ? NCI : NCO;
  return zpass_rows<T, N, (NCM > 0 ? NCM : 1), NPRE, BUDGET>();
}

// z-pass implementation: the LDS-staged kernel (default), or the
// register-resident transform for power-of-two lengths up to 256 (PF_ZPASS=reg;
// measured slower at 256^3: its 256-VGPR footprint allows one wave per SIMD).
inline bool zpass_use_reg() {
  static const bool use = [] {
    const char* e = std::getenv("PF_ZPASS");
    return e && std::strcmp(e, "reg") == 0;
  }();
  return use;
}

template <typename T, int NCI, int NCO, int NPRE, int BUDGET = ZPASS_LDS_BUDGET>
int zpass_grid(int N, int rows) {
  switch (N) {
#define PF_CASE_ZG(v)                                                                  \
  case v:                                                                              \
    if constexpr (RegPlan<v>::ok)                                                      \
      if (zpass_use_reg()) return (rows + 64 / RegPlan<v>::L - 1) / (64 / RegPlan<v>::L); \
    return (rows + zpass_rows_for<T, v, NCI, NCO, NPRE, BUDGET>() - 1) /                        \
           zpass_rows_for<T, v, NCI, NCO, NPRE, BUDGET>();
    PF_SIZES(PF_CASE_ZG)
#undef PF_CASE_ZG
    default:
      return -1;
  }
}

// Returns the grid size (the number of per-block partials the operator's
// reductions wrote).
template <typename T, int NCI, int NCO, class Op, int BUDGET = ZBudget<Op>::value>
int launch_zpass(int N, hipStream_t s, const ZPassArgs<T>& za, const typename Op::Args& oa) {
  const int grid = zpass_grid<T, NCI, NCO, Op::NPRE, BUDGET>(N, za.rows);
  switch (N) {
#define PF_CASE_Z(v)                                                                                \
  case v:                                                                                           \
    if constexpr (RegPlan<v>::ok) {                                                                 \
      if (zpass_use_reg()) {                                                                        \
        hipLaunchKernelGGL((zpass_reg_kernel<T, v, NCI, NCO, Op>), dim3(grid), dim3(64), 0, s, za, oa); \
        break;                                                                                      \
      }                                                                                             \
    }                                                                                               \
    launch<KZPass<T, v, NCI, NCO, Op, BUDGET>>(s, dim3(grid), {za, oa});                          \
    break;
    PF_SIZES(PF_CASE_Z)
#undef PF_CASE_Z
    default:
      break;
  }
  return grid;
}

// The z pass with the LDS budget ZPASS_BUDGETS[bsel] (the solver's z pass,
// whose budget the plan tuner picks at pf_create).
template <typename T, int NCI, int NCO, class Op>
int launch_zpass_b(int bsel, int N, hipStream_t s, const ZPassArgs<T>& za, const typename Op::Args& oa) {
  switch (bsel) {
    case 0: return launch_zpass<T, NCI, NCO, Op, ZPASS_BUDGETS[0]>(N, s, za, oa);
    case 2: return launch_zpass<T, NCI, NCO, Op, ZPASS_BUDGETS[2]>(N, s, za, oa);
    case 3: return launch_zpass<T, NCI, NCO, Op, ZPASS_BUDGETS[3]>(N, s, za, oa);
    default: return launch_zpass<T, NCI, NCO, Op, ZPASS_BUDGETS[1]>(N, s, za, oa);
  }
}
// rows per workgroup of that z pass (the tuner skips budgets that give the same tile)
template <typename T, int NCI, int NCO, class Op>
int zpass_rows_of(int bsel, int N) {
  switch (N) {
#define PF_CASE_ZR(v)                                                                                      \
  case v:                                                                                                  \
    switch (bsel) {                                                                                        \
      case 0: return zpass_rows_for<T, v, NCI, NCO, Op::NPRE, ZPASS_BUDGETS[0]>();                          \
      case 2: return zpass_rows_for<T, v, NCI, NCO, Op::NPRE, ZPASS_BUDGETS[2]>();                          \
      case 3: return zpass_rows_for<T, v, NCI, NCO, Op::NPRE, ZPASS_BUDGETS[3]>();                          \
      default: return zpass_rows_for<T, v, NCI, NCO, Op::NPRE, ZPASS_BUDGETS[1]>();                         \
    }
    PF_SIZES(PF_CASE_ZR)
#undef PF_CASE_ZR
    default:
      return 0;
  }
}

// Column pass along an axis of length N with the tile width CPASS_TILES[tsel]
// (0: cpass_cols' width; widths that do not fit the LDS fall back to it).
template <typename T, bool INV, int TCW, int v>
void launch_cpass_w(hipStream_t s, const CPassArgs<T>& ca, int planes, int comps) {
  constexpr int TC = (TCW == 0 || !cpass_tile_fits<T, v, TCW>()) ? cpass_cols<T, v>() : TCW;
  dim3 grid((ca.cols + TC - 1) / TC, planes, comps);
  launch<KCPass<T, v, INV, TC>>(s, grid, ca);
}
template <typename T, bool INV>
void launch_cpass(int N, hipStream_t s, const CPassArgs<T>& ca, int planes, int comps, int tsel = 0) {
  switch (N) {
#define PF_CASE_C(v)                                                    \
  case v:                                                               \
    switch (tsel) {                                                     \
      case 1: launch_cpass_w<T, INV, CPASS_TILES[1], v>(s, ca, planes, comps); break; \
      case 2: launch_cpass_w<T, INV, CPASS_TILES[2], v>(s, ca, planes, comps); break; \
      case 3: launch_cpass_w<T, INV, CPASS_TILES[3], v>(s, ca, planes, comps); break; \
      default: launch_cpass_w<T, INV, 0, v>(s, ca, planes, comps); break;            \
    }                                                                   \
    break;
    PF_SIZES(PF_CASE_C)
#undef PF_CASE_C
    default:
      break;
  }
}
// the tile width launch_cpass uses for (N, tsel)
template <typename T>
int cpass_cols_of(int tsel, int N) {
  switch (N) {
#define PF_CASE_CW(v)                                                                                     \
  case v:                                                                                                 \
    switch (tsel) {                                                                                       \
      case 1: return cpass_tile_fits<T, v, CPASS_TILES[1]>() ? CPASS_TILES[1] : cpass_cols<T, v>();      \
      case 2: return cpass_tile_fits<T, v, CPASS_TILES[2]>() ? CPASS_TILES[2] : cpass_cols<T, v>();      \
      case 3: return cpass_tile_fits<T, v, CPASS_TILES[3]>() ? CPASS_TILES[3] : cpass_cols<T, v>();      \
      default: return cpass_cols<T, v>();                                                                 \
    }
    PF_SIZES(PF_CASE_CW)
#undef PF_CASE_CW
    default:
      return 0;
  }
}

}  // namespace pf
#include "pf_pass_inst.hpp"
namespace pf {

inline int grid_for(long n) {
  long g = (n + NT - 1) / NT;
  if (g > 2048) g = 2048;
  if (g < 1) g = 1;
  return (int)g;
}

// ------------------------------------------------------------ engine
template <typename T>
struct Engine final : EngineBase {
  using C = cx<T>;
  pf_params prm{};
  PhysParams P{};
  hipStream_t stream = nullptr;
  int Nx = 0, Ny = 0, Nz = 0;
  AxisPlan ax, ay, az;
  long R = 0, nSz = 0, nSy = 0, nSc = 0;   // spectral array sizes (kz stride AzS)
  long nSzA = 0, nSyA = 0, nScA = 0;       // retained modes only (Az), for byte counts
  int AzS = 0;
  // device memory
  double *phi[2] = {nullptr, nullptr}, *Y = nullptr, *Ystar = nullptr, *xB = nullptr, *dYdt = nullptr;
  // mixed precision: the unclamped band-limited phi* (real, FP64) and its
  // retained spectrum (compact, FP64), so that the FP32 transforms of the
  // phi update carry its increment only (kop_phi_state)
  double* phis = nullptr;
  cx<double>* Phik = nullptr;
  T* trs = nullptr;
  C *Sz = nullptr, *Sy = nullptr, *Sc = nullptr, *u = nullptr, *un = nullptr;
  cx<double>* Ud = nullptr;   // FP64 accumulated displacement (accelerated solver)
  cx<double>* Gd = nullptr;   // FP64 fixed-point residual g = F(U) - U
  T* hcache = nullptr;         // h(phi) per voxel for the strain-increment passes of a solve
  cx<double>* Hist = nullptr; // last three converged U (warm-start extrapolation)
  C* dinc = nullptr;          // current increment
  C* dU = nullptr;            // Anderson history of increments [AA_MAX][3]
  C* dG = nullptr;            // Anderson history of residual differences [AA_MAX][3]
  int hist_count = 0, hist_head = 0;
  bool extended = false;      // accelerated solver with U/g state in use
  static constexpr bool kMixed = sizeof(T) == 4;
  C *twx = nullptr, *twy = nullptr, *twz = nullptr;
  double *kx = nullptr, *ky = nullptr, *kz = nullptr;
  double* partials = nullptr;
  double* dscal = nullptr;   // device scalars
  double* hscal = nullptr;   // pinned host mirror
  size_t npartials = 0;
  std::vector<void*> allocations;
  int cur = 0;
  bool warm_valid = false;
  // The device fields form a consistent state: set by pf_set_fields and an
  // accepted restart, cleared while a step, minimize iteration or restart is
  // in progress (kept cleared when one fails part-way, so the handle refuses
  // further steps until it is restored) and by the measurement entry points
  // that overwrite the buffers.
  bool state_ok = false;
  // Y is band-limited: it is the inverse transform of a dealiased spectrum
  // plus the zero-mode shift (false for fields from pf_set_fields)
  bool y_banded = false;
  // phis / Phik hold the unclamped phi* of the current phi (mixed precision;
  // false for fields from pf_set_fields, after minimize and after a restart
  // without the exact-resume companion's phi* block)
  bool phi_banded = false;
  int require_state(const char* what) const {
    if (state_ok) return PF_OK;
    return fail(PF_E_ARG, std::string(what) +
                              ": the handle holds no valid state (no fields yet, or a step, minimize iteration, "
                              "restart or kernel timing left it incomplete); call pf_set_fields or "
                              "pf_checkpoint_read");
  }
  double target_mass = 0.0;
  // statistics
  long long steps_done = 0, elastic_total = 0;
  int last_iters = 0, zm_iters = 0;
  double last_res = NAN, zm_lambda = 0.0, mean_DY = 0.0;
  // zero-mode runtime of the last accepted step (main_cuda.cu:30880-30888)
  double zm_res = 0.0, zm_deriv = 0.0;
  unsigned long long zm_accepted = 0;
  double e_el = 0.0, e_dw = 0.0, e_chem = 0.0, e_grad = 0.0;
  // The phi right-hand-side pass leaves its energy sums as block partials in
  // epartials; they are summed when the scalars are read (nothing in the step
  // uses them), eng_g blocks of 3 values.
  double* epartials = nullptr;
  int eng_g = 0;
  bool eng_pending = false;
  bool egrad_pending = false;   // e_grad of the last step not evaluated yet
  T* sig6 = nullptr;            // stress components of the last step (keep_stress)
  bool keep_sig6 = false, sig6_valid = false;
  // minimize state (phi-only and full model)
  bool minimizing = false;
  pf_minimize_opts mopt{};
  double* m_gexp = nullptr;   // g_exp of the current phi
  double* m_gfull = nullptr;  // g_exp - kappa Lap phi
  double* m_s00 = nullptr;    // sigma : eps00 of the last elastic solve
  double* m_xzero = nullptr;  // composition absent: x_B = 0 for the eigenstrain
  T* m_eps = nullptr;         // total strain of the last elastic solve, 6 x R
  double m_lambda = 0.0, m_V0 = 0.0;
  int m_iter = 0, m_conv = 0;
  bool m_converged = false;
  // full model: chemical baseline x_ref and g_bulk0 = mu_mix(x_ref)
  // (main_cuda.cu:27400-27432), the Y cap of minimize_xB_max_safe
  // (main_cuda.cu:30639-30646) and the previous iteration's excess energy
  double m_xref = 0.0, m_gbulk0 = 0.0, m_ycap = 0.0, m_Fprev = NAN;
  double m_mass_target = NAN;   // minimize mass constraint: the conserved sum of x_Btot
  ZmResultT m_mass{};           // its last accepted search
  // V5 component constraints: labels, device multipliers [0, 64), KKT
  // multipliers / projection corrections [64, 128), per-label sums [128,
  // 128 + 3*64), then the per-block partials
  int m_ncomp = 0;
  int* m_labels = nullptr;
  double* m_cbuf = nullptr;
  std::vector<double> m_ctarget, m_clambda, m_csum;
  double m_cerr = NAN;
  const double* xB_el() const { return (minimizing && !mopt.full_model) ? m_xzero : xB; }
  // cold-start hooks: learned correction and caller callback
  nn::Net* nn_net = nullptr;
  uint16_t *nn_x0 = nullptr, *nn_a = nullptr, *nn_b = nullptr;
  float* nn_y3 = nullptr;
  pf_warm_start_fn ws_fn = nullptr;
  void* ws_user = nullptr;
  pf_handle* ws_handle = nullptr;
  bool in_cold_hook = false;

  ~Engine() override {
    probe_release();
    for (void* q : allocations) (void)hipFree(q);
    if (hscal) (void)hipHostFree(hscal);
    if (stream) (void)hipStreamDestroy(stream);
  }

  template <typename U>
  int dalloc(U** ptr, size_t count) {
    void* q = nullptr;
    PF_HIP(hipMalloc(&q, count * sizeof(U) + 256));
    allocations.push_back(q);
    *ptr = static_cast<U*>(q);
    return PF_OK;
  }

  // Grid of the k-space kernels without reductions: one retained mode per
  // thread (measured faster than a 2048-workgroup grid-stride loop at 256^3;
  // the kernels with block partials keep grid_for so their sums keep one order)
  unsigned gmodes() const { return (unsigned)((nScA + NT - 1) / NT); }

  KDims kd() const {
    return KDims{ax.A, ay.A, AzS, kx, ky, kz, (Nz % 2) == 0, az.A};
  }

  int init(const pf_params& p) override {
    prm = p;
    Nx = p.Nx;
    Ny = p.Ny;
    Nz = p.Nz;
    if (!size_supported(Nx) || !size_supported(Ny) || !size_supported(Nz))
      return fail(PF_E_ARG, "unsupported grid size " + std::to_string(Nx) + "x" + std::to_string(Ny) +
                                "x" + std::to_string(Nz));
    P = make_phys(p);
    ax = make_axis(Nx, p.dx, true);
    ay = make_axis(Ny, p.dy, true);
    az = make_axis(Nz, p.dz, false);
    if (az.A < 1 || ax.A < 1 || ay.A < 1 || 2 * az.A > Nz + 1)
      return fail(PF_E_ARG, "degenerate retained-mode set");
    // the pruned transforms (Pruned<N>, pf_fft.hpp) and the z pass's lane
    // blocks rely on at most N/3 + 1 retained indices per axis half
    if (3 * (az.A - 1) > Nz || 3 * (ax.a - 1) > Nx || 3 * (ay.a - 1) > Ny)
      return fail(PF_E_ARG, "retained mode count exceeds the 2/3 rule bound");
    R = (long)Nx * Ny * Nz;
    {
      // kz stride candidates: whole 128- or 64-byte lines on large grids
      // (aligned column tiles in the y and x passes, partial lines in the
      // mode-by-mode k-space kernels) or none; the plan tuner picks one at
      // the end of init, PF_KZ_PAD (elements, 0: none) fixes it
      auto pad = [&](int line) { return (az.A >= 40 && line > 0) ? (az.A + line - 1) / line * line : az.A; };
      const int line = 128 / (int)sizeof(C);
      kz_cands.clear();
      if (const char* e = std::getenv("PF_KZ_PAD")) {
        kz_cands.push_back(pad(std::atoi(e)));
      } else {
        for (int c : {pad(line), pad(line / 2), az.A})
          if (std::find(kz_cands.begin(), kz_cands.end(), c) == kz_cands.end()) kz_cands.push_back(c);
      }
      AzS = *std::max_element(kz_cands.begin(), kz_cands.end());   // allocations hold the widest
    }
    nSz = (long)Nx * Ny * AzS;
    nSy = (long)Nx * ay.A * AzS;
    nSc = (long)ax.A * ay.A * AzS;
    nSzA = (long)Nx * Ny * az.A;
    nSyA = (long)Nx * ay.A * az.A;
    nScA = (long)ax.A * ay.A * az.A;
    PF_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    int rc;
    if ((rc = dalloc(&phi[0], R)) || (rc = dalloc(&phi[1], R)) || (rc = dalloc(&Y, R)) ||
        (rc = dalloc(&Ystar, R)) || (rc = dalloc(&xB, R)) || (rc = dalloc(&dYdt, R)) ||
        (rc = dalloc(&trs, R)) || (rc = dalloc(&Sz, 6 * nSz)) || (rc = dalloc(&Sy, 6 * nSy)) ||
        (rc = dalloc(&Sc, 6 * nSc)) || (rc = dalloc(&u, 3 * nSc)) || (rc = dalloc(&un, 3 * nSc)))
      return rc;
    {
      const bool accel = p.elastic_enabled && (p.elastic_warm_start_enabled || p.elastic_residual_control_enabled);
      extended = accel && (kMixed || p.elastic_warm_start_order > 1 || p.elastic_anderson_depth > 0);
      // solver state by retained-mode index (nScA entries per component)
      if ((kMixed || extended) && (rc = dalloc(&Ud, 3 * nScA))) return rc;
      if (kMixed && ((rc = dalloc(&phis, R)) || (rc = dalloc(&Phik, nScA)))) return rc;
      if (extended) {
        if ((rc = dalloc(&Gd, 3 * nScA)) || (rc = dalloc(&Hist, 9 * nScA)) || (rc = dalloc(&dinc, 3 * nScA)) ||
            (rc = dalloc(&dU, (size_t)AA_MAX * 3 * nScA)) || (rc = dalloc(&dG, (size_t)AA_MAX * 3 * nScA)) ||
            (rc = dalloc(&hcache, R)))
          return rc;
        for (int k = 0; k < 3; ++k) hbuf[k] = Hist + (size_t)k * 3 * nScA;
        PF_HIP(hipMemset(hcache, 0, R * sizeof(T)));
      }
    }
    auto wx = twiddles<T>(Nx), wy = twiddles<T>(Ny), wz = twiddles<T>(Nz);
    if ((rc = dalloc(&twx, Nx)) || (rc = dalloc(&twy, Ny)) || (rc = dalloc(&twz, Nz)) ||
        (rc = dalloc(&kx, ax.A)) || (rc = dalloc(&ky, ay.A)) || (rc = dalloc(&kz, AzS)))
      return rc;
    PF_HIP(hipMemcpy(twx, wx.data(), Nx * sizeof(C), hipMemcpyHostToDevice));
    PF_HIP(hipMemcpy(twy, wy.data(), Ny * sizeof(C), hipMemcpyHostToDevice));
    PF_HIP(hipMemcpy(twz, wz.data(), Nz * sizeof(C), hipMemcpyHostToDevice));
    PF_HIP(hipMemcpy(kx, ax.k.data(), ax.A * sizeof(double), hipMemcpyHostToDevice));
    PF_HIP(hipMemcpy(ky, ay.k.data(), ay.A * sizeof(double), hipMemcpyHostToDevice));
    {
      std::vector<double> kzp(AzS, 0.0);   // padding entries: k = 0
      std::copy(az.k.begin(), az.k.begin() + az.A, kzp.begin());
      PF_HIP(hipMemcpy(kz, kzp.data(), AzS * sizeof(double), hipMemcpyHostToDevice));
    }
    // spectral arrays start at zero so the kz padding holds zeros
    {
      const std::pair<void*, size_t> zero[] = {
          {Sz, 6 * nSz * sizeof(C)}, {Sy, 6 * nSy * sizeof(C)}, {Sc, 6 * nSc * sizeof(C)},
          {u, 3 * nSc * sizeof(C)},  {un, 3 * nSc * sizeof(C)}};
      for (const auto& z : zero) PF_HIP(hipMemset(z.first, 0, z.second));
      if (Ud) PF_HIP(hipMemset(Ud, 0, 3 * nScA * sizeof(cx<double>)));
      if (Phik) PF_HIP(hipMemset(Phik, 0, nScA * sizeof(cx<double>)));
      if (extended) {
        PF_HIP(hipMemset(Gd, 0, 3 * nScA * sizeof(cx<double>)));
        PF_HIP(hipMemset(Hist, 0, 9 * nScA * sizeof(cx<double>)));
        PF_HIP(hipMemset(dinc, 0, 3 * nScA * sizeof(C)));
        PF_HIP(hipMemset(dU, 0, (size_t)AA_MAX * 3 * nScA * sizeof(C)));
        PF_HIP(hipMemset(dG, 0, (size_t)AA_MAX * 3 * nScA * sizeof(C)));
      }
    }
    // partial-sum scratch: largest grid times 4 reductions
    // (a z pass takes at least one row pair per workgroup)
    long maxg = 2048;
    const int rows = Nx * Ny;
    maxg = std::max<long>(maxg, (rows + 1) / 2);
    npartials = (size_t)maxg * 16;   // up to 16 reductions per block (Anderson dots use AA_NRED)
    if ((rc = dalloc(&partials, npartials)) || (rc = dalloc(&dscal, NSCAL)) ||
        (rc = dalloc(&epartials, (size_t)maxg * 3)))
      return rc;
    PF_HIP(hipHostMalloc((void**)&hscal, NSCAL * sizeof(double), hipHostMallocDefault));
    if (kz_cands.size() > 1 && (rc = tune_kz_stride())) return rc;
    if ((rc = tune_plans())) return rc;
    return PF_OK;
  }

  // ---------------------------------------------------------------- plan tuner
  // The kz stride changes only where spectra sit in memory, not any
  // arithmetic or summation order, so every candidate gives bitwise the same
  // results; the tuner times one elastic update's kernels (z pass, y and x
  // passes both ways, the Green update and the Anderson update) on each and
  // keeps the fastest.  Buffers are cleared afterwards.
  std::vector<int> kz_cands;
  std::vector<double> kz_times;   // microseconds per candidate, in kz_cands order
  void set_kz_stride(int azs) {
    AzS = azs;
    nSz = (long)Nx * Ny * AzS;
    nSy = (long)Nx * ay.A * AzS;
    nSc = (long)ax.A * ay.A * AzS;
  }
  int clear_spectra() {
    const size_t cs = sizeof(C);
    const long wz = (long)Nx * Ny * AzS, wy = (long)Nx * ay.A * AzS, wc = (long)ax.A * ay.A * AzS;
    PF_HIP(hipMemset(Sz, 0, 6 * wz * cs));
    PF_HIP(hipMemset(Sy, 0, 6 * wy * cs));
    PF_HIP(hipMemset(Sc, 0, 6 * wc * cs));
    PF_HIP(hipMemset(u, 0, 3 * wc * cs));
    PF_HIP(hipMemset(un, 0, 3 * wc * cs));
    if (Ud) PF_HIP(hipMemset(Ud, 0, 3 * nScA * sizeof(cx<double>)));
    if (Phik) PF_HIP(hipMemset(Phik, 0, nScA * sizeof(cx<double>)));
    if (extended) {
      PF_HIP(hipMemset(Gd, 0, 3 * nScA * sizeof(cx<double>)));
      PF_HIP(hipMemset(Hist, 0, 9 * nScA * sizeof(cx<double>)));
      PF_HIP(hipMemset(dinc, 0, 3 * nScA * cs));
      PF_HIP(hipMemset(dU, 0, (size_t)AA_MAX * 3 * nScA * cs));
      PF_HIP(hipMemset(dG, 0, (size_t)AA_MAX * 3 * nScA * cs));
      PF_HIP(hipMemset(hcache, 0, R * sizeof(T)));
    }
    return PF_OK;
  }
  // One elastic update's launches in the solve's order (Anderson update, x
  // and y passes back, z pass, y and x passes forward, Green update), or the
  // four 6-component passes without elasticity.
  int update_chain(int* ids) const {
    if (!P.elastic) {
      const int c[] = {PF_K_XPASS_INV6, PF_K_YPASS_INV6, PF_K_YPASS_FWD6, PF_K_XPASS_FWD6};
      std::copy(c, c + 4, ids);
      return 4;
    }
    const int c[] = {PF_K_AA_UPDATE, PF_K_XPASS_INV6, PF_K_YPASS_INV6, PF_K_ZPASS_HIJ,
                     PF_K_YPASS_FWD6, PF_K_XPASS_FWD6, PF_K_GREEN};
    std::copy(c, c + 7, ids);
    return 7;
  }
  // Times the update chain as a step runs it: one untimed chain, then `reps`
  // chains with HIP events around each launch, so every kernel reads what
  // the one before it wrote rather than its own input of a moment ago
  // (repeated launches of one kernel find much of a 256^3 operand still in
  // the device's last-level cache, which a step never does).  us[PF_K_*]
  // receives each kernel's mean microseconds; *total their sum per chain.
  int chain_time(int reps, double* us, double* total) {
    int ids[8];
    const int n = update_chain(ids);
    int rc = aa_ctl_steady();
    if (rc) return rc;
    for (int i = 0; i < n; ++i) launch_one(ids[i]);
    std::vector<hipEvent_t> ev((size_t)2 * n * reps, nullptr);
    auto destroy = [&]() {
      for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    };
    for (auto& e : ev)
      if (hipEventCreate(&e) != hipSuccess) {
        destroy();
        return fail(PF_E_HIP, "hipEventCreate failed in the plan tuner");
      }
    bool ok = true;
    for (int r = 0; r < reps && ok; ++r)
      for (int i = 0; i < n && ok; ++i) {
        ok = hipEventRecord(ev[2 * (r * n + i)], stream) == hipSuccess;
        launch_one(ids[i]);
        ok = ok && hipEventRecord(ev[2 * (r * n + i) + 1], stream) == hipSuccess;
      }
    ok = ok && hipEventSynchronize(ev.back()) == hipSuccess;
    for (int k = 0; k < PF_K_COUNT; ++k) us[k] = 0.0;
    *total = 0.0;
    for (int r = 0; r < reps && ok; ++r)
      for (int i = 0; i < n && ok; ++i) {
        float ms = 0.f;
        ok = hipEventElapsedTime(&ms, ev[2 * (r * n + i)], ev[2 * (r * n + i) + 1]) == hipSuccess;
        us[ids[i]] += 1000.0 * (double)ms / reps;
        *total += 1000.0 * (double)ms / reps;
      }
    destroy();
    if (!ok) return fail(PF_E_HIP, "event timing failed in the plan tuner");
    PF_HIP(hipGetLastError());
    return PF_OK;
  }
  int tune_kz_stride() {
    const int widest = AzS;
    PF_HIP(hipMemset(phi[0], 0, R * sizeof(double)));
    PF_HIP(hipMemset(phi[1], 0, R * sizeof(double)));
    PF_HIP(hipMemset(xB, 0, R * sizeof(double)));
    kz_times.assign(kz_cands.size(), 0.0);
    size_t best = 0;
    for (size_t c = 0; c < kz_cands.size(); ++c) {
      set_kz_stride(kz_cands[c]);
      double us[PF_K_COUNT];
      int rc = chain_time(3, us, &kz_times[c]);
      if (rc) return rc;
      if (kz_times[c] < kz_times[best]) best = c;
    }
    set_kz_stride(widest);
    int rc = clear_spectra();
    if (rc) return rc;
    set_kz_stride(kz_cands[best]);
    PF_HIP(hipStreamSynchronize(stream));
    return PF_OK;
  }

  // Tile plans: the LDS budget of the solver's z pass (rows per workgroup,
  // ZPASS_BUDGETS) and the column-tile widths of the y and x passes
  // (CPASS_TILES).  Each choice only changes which rows or columns a
  // workgroup transforms together (the solver's z-pass operators reduce
  // nothing), so results are bitwise the same for every choice; the tuner
  // times each distinct tile on the handle's own grid and keeps the fastest.
  // PF_PLAN=<z>,<y>,<x> fixes the three indices.
  // The first handle of a grid shape and solver kind in a process tunes; the
  // others take its plan, so handles of one shape compute identically (the
  // tiles change the compiled code of the transforms, and with it the
  // rounding of some of their operations), and the checkpoint companion
  // carries the plan to a restart.
  int zb_sel = 1, ty_sel = 0, tx_sel = 0;
  double plan_us[3] = {0.0, 0.0, 0.0};   // the chosen plans' timed microseconds
  struct PlanKey {
    int nx, ny, nz, el, ext;
    bool operator<(const PlanKey& o) const {
      return std::tie(nx, ny, nz, el, ext) < std::tie(o.nx, o.ny, o.nz, o.el, o.ext);
    }
  };
  static std::mutex& plan_mu() {
    static std::mutex m;
    return m;
  }
  static std::map<PlanKey, std::array<int, 3>>& plan_cache() {
    static std::map<PlanKey, std::array<int, 3>> c;
    return c;
  }
  int tune_plans() {
    const PlanKey key{Nx, Ny, Nz, P.elastic, extended ? 1 : 0};
    std::lock_guard<std::mutex> lk(plan_mu());
    const auto it = plan_cache().find(key);
    if (it != plan_cache().end() && !std::getenv("PF_PLAN")) {
      zb_sel = it->second[0];
      ty_sel = it->second[1];
      tx_sel = it->second[2];
      return PF_OK;
    }
    const int rc = tune_plans_timed();
    if (rc == PF_OK && !std::getenv("PF_PLAN")) plan_cache()[key] = {zb_sel, ty_sel, tx_sel};
    return rc;
  }
  int tune_plans_timed() {
    if (const char* e = std::getenv("PF_PLAN")) {
      int z = 1, y = 0, x = 0;
      if (std::sscanf(e, "%d,%d,%d", &z, &y, &x) != 3 || z < 0 || z > 3 || y < 0 || y > 3 || x < 0 || x > 3)
        return fail(PF_E_ARG, "PF_PLAN must be <z>,<y>,<x> with indices 0..3");
      zb_sel = z;
      ty_sel = y;
      tx_sel = x;
      return PF_OK;
    }
    if (R < 32768) return PF_OK;   // small grids keep the default tiles (launch latency dominates)
    // each candidate's kernels timed inside the update chain
    auto timed = [&](std::initializer_list<int> ids, double* out) -> int {
      double us[PF_K_COUNT], chain = 0.0;
      int rc = chain_time(3, us, &chain);
      if (rc) return rc;
      double tot = 0.0;
      for (int id : ids) tot += us[id];
      *out = tot;
      return PF_OK;
    };
    int rc;
    if (P.elastic) {
      int seen[4], ns = 0, best = 1;
      double bt = 0.0;
      for (int b = 0; b < 4; ++b) {
        const int rows = zpass_rows_of<T, 6, 6, OpHijInc<T>>(b, Nz);
        bool dup = false;
        for (int k = 0; k < ns; ++k) dup |= seen[k] == rows;
        if (dup) continue;
        seen[ns++] = rows;
        zb_sel = b;
        double us;
        if ((rc = timed({PF_K_ZPASS_HIJ}, &us))) return rc;
        if (ns == 1 || us < bt) {
          bt = us;
          best = b;
        }
      }
      zb_sel = best;
      plan_us[0] = bt;
    }
    for (int axis = 0; axis < 2; ++axis) {
      const int N = axis == 0 ? Ny : Nx;
      int* sel = axis == 0 ? &ty_sel : &tx_sel;
      int seen[4], ns = 0, best = 0;
      double bt = 0.0;
      for (int t = 0; t < 4; ++t) {
        const int cols = cpass_cols_of<T>(t, N);
        bool dup = false;
        for (int k = 0; k < ns; ++k) dup |= seen[k] == cols;
        if (dup) continue;
        seen[ns++] = cols;
        *sel = t;
        double us;
        if ((rc = axis == 0 ? timed({PF_K_YPASS_FWD6, PF_K_YPASS_INV6}, &us)
                            : timed({PF_K_XPASS_FWD6, PF_K_XPASS_INV6}, &us)))
          return rc;
        if (ns == 1 || us < bt) {
          bt = us;
          best = t;
        }
      }
      *sel = best;
      plan_us[1 + axis] = bt;
    }
    if ((rc = clear_spectra())) return rc;
    PF_HIP(hipStreamSynchronize(stream));
    return PF_OK;
  }

  // ---------------------------------------------------------------- helpers
  C* sz(int c) { return Sz + (size_t)c * nSz; }
  C* sy(int c) { return Sy + (size_t)c * nSy; }
  C* sc(int c) { return Sc + (size_t)c * nSc; }
  C* uu(C* base, int c) { return base + (size_t)c * nSc; }

  ZPassArgs<T> zargs(int c_in0, int nci, int c_out0, int nco) {
    ZPassArgs<T> za{};
    za.rows = Nx * Ny;
    za.Az = az.A;
    za.AzS = AzS;
    for (int c = 0; c < nci; ++c) za.in[c] = sz(c_in0 + c);
    for (int c = 0; c < nco; ++c) za.out[c] = sz(c_out0 + c);
    za.tw = twz;
    za.partials = partials;
    za.gate = gate_;
    return za;
  }

  // y pass over comps [c0, c0+nc): Sz <-> Sy
  void ypass(bool inv, int c0, int nc) {
    CPassArgs<T> ca{};
    ca.cols = AzS;
    ca.A = ay.A;
    ca.a = ay.a;
    ca.tw = twy;
    ca.gate = gate_;
    if (!inv) {
      ca.in = sz(c0);
      ca.out = sy(c0);
      ca.in_pstride = (long)Ny * AzS;
      ca.out_pstride = (long)ay.A * AzS;
      ca.in_cstride = nSz;
      ca.out_cstride = nSy;
      launch_cpass<T, false>(Ny, stream, ca, Nx, nc, ty_sel);
    } else {
      ca.in = sy(c0);
      ca.out = sz(c0);
      ca.in_pstride = (long)ay.A * AzS;
      ca.out_pstride = (long)Ny * AzS;
      ca.in_cstride = nSy;
      ca.out_cstride = nSz;
      launch_cpass<T, true>(Ny, stream, ca, Nx, nc, ty_sel);
    }
  }
  // x pass: Sy <-> compact (dst/src compact base with component stride nSc)
  void xpass(bool inv, int c0, int nc, C* compact) {
    CPassArgs<T> ca{};
    ca.cols = ay.A * AzS;
    ca.A = ax.A;
    ca.a = ax.a;
    ca.tw = twx;
    ca.gate = gate_;
    ca.in_pstride = ca.out_pstride = 0;
    if (!inv) {
      ca.in = sy(c0);
      ca.out = compact;
      ca.in_cstride = nSy;
      ca.out_cstride = nSc;
      launch_cpass<T, false>(Nx, stream, ca, 1, nc, tx_sel);
    } else {
      ca.in = compact;
      ca.out = sy(c0);
      ca.in_cstride = nSc;
      ca.out_cstride = nSy;
      launch_cpass<T, true>(Nx, stream, ca, 1, nc, tx_sel);
    }
  }

  void finalize(int nblocks, int nred, RedKinds k, double* out, double scale) {
    launch<KFinalize>(stream, dim3(1), {partials, nblocks, nred, k, out, scale});
  }

  int fetch(int count) {
    PF_HIP(hipMemcpyAsync(hscal, dscal, count * sizeof(double), hipMemcpyDeviceToHost, stream));
    PF_HIP(hipStreamSynchronize(stream));
    return PF_OK;
  }

  // ---------------------------------------------------------------- fields
  int set_fields(const double* phi0, const double* xB0, const double* dYdt0, double xB_cap) override {
    std::vector<double> ph(R), xb(R), yy(R), dd(R, 0.0);
    for (long i = 0; i < R; ++i) {
      double p = phi0[i], x = xB0[i];
      if (!std::isfinite(p) || !std::isfinite(x) || (dYdt0 && !std::isfinite(dYdt0[i])))
        return fail(PF_E_NONFINITE, "raw init contains NaN/Inf at idx=" + std::to_string(i));
      p = clamp01(p);
      x = clamp_eps(x, P.xB_eps);
      if (std::isfinite(xB_cap) && x > xB_cap) x = xB_cap;
      ph[i] = p;
      xb[i] = x;
      yy[i] = logit_from_fraction(x, P.xB_eps, P.Y_clip);
      if (dYdt0) dd[i] = dYdt0[i];
    }
    state_ok = false;
    y_banded = false;
    phi_banded = false;
    cur = 0;
    PF_HIP(hipMemcpy(phi[0], ph.data(), R * sizeof(double), hipMemcpyHostToDevice));
    PF_HIP(hipMemcpy(xB, xb.data(), R * sizeof(double), hipMemcpyHostToDevice));
    PF_HIP(hipMemcpy(Y, yy.data(), R * sizeof(double), hipMemcpyHostToDevice));
    PF_HIP(hipMemcpy(dYdt, dd.data(), R * sizeof(double), hipMemcpyHostToDevice));
    warm_valid = false;
    minimizing = false;
    hist_count = 0;
    hist_head = 0;
    steps_done = 0;
    elastic_total = 0;
    double m, vf;
    int rc = mass_vf(&m, &vf);
    if (rc) return rc;
    target_mass = m;
    egrad_pending = false;
    eng_pending = false;
    sig6_valid = false;
    e_grad = 0.0;
    zm_lambda = zm_res = zm_deriv = 0.0;
    zm_iters = 0;
    zm_accepted = 0;
    state_ok = true;
    return PF_OK;
  }

  int mass_vf(double* m, double* vf, double* nin = nullptr, double* nif = nullptr) {
    const int g = grid_for(R);
    hipLaunchKernelGGL(mass_vf_kernel, dim3(g), dim3(NT), 0, stream, phi[cur], xB, P.v_B, R, partials);
    finalize(g, 4, RedKinds{{RED_SUM, RED_SUM, RED_SUM, RED_SUM}}, dscal + 24, 1.0);
    PF_HIP(hipMemcpyAsync(hscal + 24, dscal + 24, 4 * sizeof(double), hipMemcpyDeviceToHost, stream));
    PF_HIP(hipStreamSynchronize(stream));
    *m = hscal[24];
    *vf = hscal[25] / (double)R;
    if (nin) *nin = hscal[26];
    if (nif) *nif = hscal[27];
    return PF_OK;
  }

  // ---------------------------------------------------------------- elastic
  int elastic_solve() {
    if (extended) return elastic_solve_accel();
    return kMixed ? elastic_solve_mixed() : elastic_solve_plain();
  }

  // History ring of the last converged states: three buffers, rotated with
  // Ud by pointer (the solve leaves its result in Ud and swaps it into the
  // ring; the buffer it gets back is overwritten by the next warm start).
  cx<double>* hbuf[3] = {nullptr, nullptr, nullptr};   // set with the Hist allocation
  cx<double>* hist_slot(int i) { return hbuf[i]; }
  // the last converged displacement (FP64 state)
  const cx<double>* U_last() {
    return (extended && hist_count > 0) ? hist_slot((hist_head + 2) % 3) : Ud;
  }
  C* du_slot(int j, int i) { return dU + ((size_t)j * 3 + i) * nScA; }
  C* dg_slot(int j, int i) { return dG + ((size_t)j * 3 + i) * nScA; }

  // Accelerated fixed point (same map F, residual |F(U)-U|/|F(U)| and test as
  // main_cuda.cu:28887-29046): warm start by polynomial extrapolation of the
  // last converged displacements, FP64 accumulation of U and g, increments in
  // T through the linear part of the map, and Anderson mixing of depth m.
  // Accelerated fixed point (same map F, residual |F(U)-U|/|F(U)| and test as
  // main_cuda.cu:28887-29046): warm start by polynomial extrapolation of the
  // last converged displacements, FP64 accumulation of U and g, increments in
  // T through the linear part of the map, and Anderson mixing of depth m.
  // The iteration's decisions (mixing coefficients, history bookkeeping,
  // stopping rule) are taken on the device in the control block; the host
  // enqueues as many iterations as the previous solve took and waits once,
  // enqueuing more only if the solve is still running.  The solve is split
  // into accel_begin (start and the first iterations), accel_iters and
  // accel_finish so that a batch runner can drive many handles' solves
  // through the same waits.
  int acc_k = 0, acc_kmax = 0;   // iterations enqueued in the running solve
  int elastic_solve_accel() {
    int rc = accel_begin();
    if (rc) return rc;
    for (;;) {
      PF_HIP(hipMemcpyAsync(hscal + AA_SLOT, dscal + AA_SLOT, AC_GAM * sizeof(double), hipMemcpyDeviceToHost,
                            stream));
      PF_HIP(hipStreamSynchronize(stream));
      if (!accel_more()) break;
      accel_iters(2);
    }
    return accel_finish();
  }
  // the start of the solve (extrapolated warm start or cold start) and its
  // first aa_spec iterations
  int accel_begin() {
    if (tl_rec && cold_hooks() && !(warm_valid && hist_count > 0))
      return fail(PF_E_ARG, "a recorded solve cannot run a cold-start hook");
    const KDims d = kd();
    const int gk = grid_for(nSc);
    if (warm_valid && hist_count > 0) {
      const int order = std::min(prm.elastic_warm_start_order, hist_count);
      const double coef[3][3] = {{1.0, 0.0, 0.0}, {2.0, -1.0, 0.0}, {3.0, -3.0, 1.0}};
      auto h_at = [&](int back) -> const cx<double>* {   // back = 0: latest
        return back < order ? hist_slot((hist_head - 1 - back + 3) % 3) : nullptr;
      };
      // extrapolated start and its strain in one pass over the retained modes
      // (the kz padding of Ud stays zero)
      StrainArgs<T, double> sa{};
      sa.ustate = 1;
      sa.d = d;
      sa.n = nSc;
      for (int i = 0; i < 3; ++i) {
        sa.u[i] = Ud + (size_t)i * nScA;
        sa.Uout[i] = Ud + (size_t)i * nScA;
        for (int k = 0; k < order; ++k) sa.H[k][i] = h_at(k) + (size_t)i * nScA;
      }
      for (int k = 0; k < order; ++k) sa.cf[k] = coef[order - 1][k];
      sa.nh = order;
      for (int i = 0; i < 6; ++i) sa.eps[i] = sc(i);
      launch<KStrain<T, double>>(stream, dim3(gmodes()), {sa});
    } else {
      typename OpElasticCold<T>::Args oa{P, phi[cur], xB_el()};
      launch_zpass<T, 0, 6, OpElasticCold<T>>(Nz, stream, zargs(0, 0, 0, 6), oa);
      ypass(false, 0, 6);
      xpass(false, 0, 6, Sc);
      GreenAccArgs<T> ga{};
      ga.d = d;
      ga.n = nSc;
      for (int i = 0; i < 6; ++i) {
        ga.h[i] = sc(i);
        ga.eps[i] = sc(i);
      }
      for (int i = 0; i < 3; ++i) ga.U[i] = Ud + (size_t)i * nScA;
      ga.partials = nullptr;
      for (int i = 0; i < 21; ++i) ga.C[i] = P.C[i];
      ga.mode = 0;
      launch<KGreenAcc<T>>(stream, dim3(gk), {ga});
      if (cold_hooks()) {
        int rc = cold_hook();
        if (rc) return rc;
        StrainArgs<T, double> sa{};
        sa.ustate = 1;
        sa.d = d;
        sa.n = nSc;
        for (int i = 0; i < 3; ++i) sa.u[i] = Ud + (size_t)i * nScA;
        for (int i = 0; i < 6; ++i) sa.eps[i] = sc(i);
        launch<KStrain<T, double>>(stream, dim3(gmodes()), {sa});
      }
    }
    double* ctl = dscal + AA_SLOT;
    launch<KAaInit>(stream, dim3(1), {ctl});
    acc_kmax = std::max(prm.elastic_iter_max, 1);
    acc_k = 0;
    gate_ = ctl + AC_ACTIVE;
    accel_iters(std::max(aa_spec, 1));
    return PF_OK;
  }
  // true while the solve runs and may take more iterations (control block
  // mirrored into hscal)
  bool accel_more() const { return hscal[AA_SLOT + AC_ACTIVE] != 0.0 && acc_k < acc_kmax; }
  // enqueue up to `batch` more iterations (each does nothing once the solve has stopped)
  void accel_iters(int batch) {
    const int m = prm.elastic_anderson_depth;
    const KDims d = kd();
    const int gk = grid_for(nSc);
    const T invN = T(1.0 / (double)R);
    double* ctl = dscal + AA_SLOT;
    const AaConst ac{m, prm.elastic_iter_min, prm.elastic_iter_max, prm.elastic_residual_tolerance,
                     prm.elastic_residual_absolute_floor};
    for (int b = 0; b < batch && acc_k < acc_kmax; ++b, ++acc_k) {
      const int mode = (acc_k == 0) ? 0 : 1;
      if (mode == 1) {
        // mixing coefficients: from the previous iteration's aa_end_kernel
        AAUpdateArgs<T> ua{};
        ua.d = d;
        ua.n = nSc;
        for (int i = 0; i < 3; ++i) {
          ua.g[i] = Gd + (size_t)i * nScA;
          ua.U[i] = Ud + (size_t)i * nScA;
          ua.dinc[i] = dinc + (size_t)i * nScA;
          for (int j = 0; j < AA_MAX; ++j) {
            ua.du[j][i] = du_slot(j, i);
            ua.dg[j][i] = dg_slot(j, i);
          }
        }
        for (int i = 0; i < 6; ++i) ua.eps[i] = sc(i);
        ua.ctl = ctl;
        ua.m = m;
        int pr = probe_begin(PF_K_AA_UPDATE);
        launch<KAaUpdate<T>>(stream, dim3(gmodes()), {ua});
        probe_end(pr, acc_k, probe_bytes_of(PF_K_AA_UPDATE, acc_k));
      }
      int pr = probe_begin(PF_K_XPASS_INV6);
      xpass(true, 0, 6, Sc);
      probe_end(pr, acc_k, probe_bytes_of(PF_K_XPASS_INV6, acc_k));
      pr = probe_begin(PF_K_YPASS_INV6);
      ypass(true, 0, 6);
      probe_end(pr, acc_k, probe_bytes_of(PF_K_YPASS_INV6, acc_k));
      pr = probe_begin(PF_K_ZPASS_HIJ);
      if (mode == 0) {
        typename OpHij<T>::Args oa{P, phi[cur], xB_el(), invN, 0, hcache};
        launch_zpass_b<T, 6, 6, OpHij<T>>(zb_sel, Nz, stream, zargs(0, 6, 0, 6), oa);
      } else {
        typename OpHijInc<T>::Args oa{P, hcache, invN};
        launch_zpass_b<T, 6, 6, OpHijInc<T>>(zb_sel, Nz, stream, zargs(0, 6, 0, 6), oa);
      }
      probe_end(pr, acc_k, probe_bytes_of(PF_K_ZPASS_HIJ, acc_k));
      pr = probe_begin(PF_K_YPASS_FWD6);
      ypass(false, 0, 6);
      probe_end(pr, acc_k, probe_bytes_of(PF_K_YPASS_FWD6, acc_k));
      pr = probe_begin(PF_K_XPASS_FWD6);
      xpass(false, 0, 6, Sc);
      probe_end(pr, acc_k, probe_bytes_of(PF_K_XPASS_FWD6, acc_k));
      GreenAAArgs<T> ga{};
      ga.d = d;
      ga.n = nSc;
      for (int i = 0; i < 6; ++i) ga.h[i] = sc(i);
      for (int i = 0; i < 3; ++i) {
        ga.U[i] = Ud + (size_t)i * nScA;
        ga.g[i] = Gd + (size_t)i * nScA;
        ga.dinc[i] = dinc + (size_t)i * nScA;
        for (int j = 0; j < AA_MAX; ++j) {
          ga.dg[j][i] = dg_slot(j, i);
          ga.du[j][i] = du_slot(j, i);
        }
      }
      ga.mode = mode;
      ga.partials = partials;
      for (int i = 0; i < 21; ++i) ga.C[i] = P.C[i];
      ga.ctl = ctl;
      ga.m = m;
      pr = probe_begin(PF_K_GREEN);
      launch<KGreenAA<T>>(stream, dim3(gk), {ga});
      probe_end(pr, acc_k, probe_bytes_of(PF_K_GREEN, acc_k));
      launch<KAaEnd>(stream, dim3(1), {partials, gk, ctl, mode, ac});
    }
  }
  // the accepted result: host checks, U <- U + g and its strain, history rotation
  int accel_finish() {
    const KDims d = kd();
    gate_ = nullptr;
    const double* hc = hscal + AA_SLOT;
    const int it = (int)hc[AC_IT];
    const double res = hc[AC_RES];
    const bool converged = hc[AC_CONV] != 0.0;
    aa_spec = std::max(it, 1);
    if (hc[AC_STATUS] != 0.0) {
      last_iters = it;
      probe_settle(it);
      return fail(PF_E_NONFINITE, "elastic residual is non-finite at step " + std::to_string(steps_done + 1));
    }
    launch<KAaFinal<T>>(stream, dim3(gmodes()), {d, nSc, Ud, Ud + nScA, Ud + 2 * nScA, Gd, Gd + nScA,
                                                  Gd + 2 * nScA, sc(0), sc(1), sc(2), sc(3), sc(4), sc(5)});
    int pr = probe_begin(PF_K_XPASS_INV6);
    xpass(true, 0, 6, Sc);
    probe_end(pr, -1, probe_bytes_of(PF_K_XPASS_INV6, 0));
    pr = probe_begin(PF_K_YPASS_INV6);
    ypass(true, 0, 6);
    probe_end(pr, -1, probe_bytes_of(PF_K_YPASS_INV6, 0));
    probe_settle(it);
    std::swap(hbuf[hist_head], Ud);
    hist_head = (hist_head + 1) % 3;
    hist_count = std::min(hist_count + 1, 3);
    last_iters = it;
    last_res = res;
    elastic_total += it;
    warm_valid = true;
    if (!converged && prm.elastic_fail_on_nonconvergence) {
      char buf[256];
      std::snprintf(buf, sizeof buf,
                    "accelerated elastic solver did not converge at step %lld: iterations=%d residual=%.17e "
                    "tolerance=%.17e",
                    steps_done + 1, it, res, prm.elastic_residual_tolerance);
      return fail(PF_E_NONCONVERGED, buf);
    }
    return PF_OK;
  }

  // Mixed precision (main_cuda.cu:28760-29095 semantics):
  // iterate on FP32 increments d_{n+1} = A d_n with the FP64 sum U += d and
  // the FP64 residual |d|/|U|.  The first update of a step evaluates the full
  // affine map, later ones only its linear part -h dC:eps(d).
  int elastic_solve_mixed() {
    const bool accel = prm.elastic_warm_start_enabled || prm.elastic_residual_control_enabled;
    const bool warm = accel && warm_valid;
    const KDims d = kd();
    const int gk = grid_for(nSc);
    auto green = [&](int mode, bool resid) {
      GreenAccArgs<T> ga{};
      ga.d = d;
      ga.n = nSc;
      for (int i = 0; i < 6; ++i) {
        ga.h[i] = sc(i);
        ga.eps[i] = sc(i);
      }
      for (int i = 0; i < 3; ++i) ga.U[i] = Ud + (size_t)i * nScA;
      ga.partials = resid ? partials : nullptr;
      for (int i = 0; i < 21; ++i) ga.C[i] = P.C[i];
      ga.mode = mode;
      hipLaunchKernelGGL((kop_green_acc<T>), dim3(gk), dim3(NT), 0, stream, ga);
    };
    auto strain_of_U = [&]() {
      StrainArgs<T, double> sa{};
      sa.ustate = 1;
      sa.d = d;
      sa.n = nSc;
      for (int i = 0; i < 3; ++i) sa.u[i] = Ud + (size_t)i * nScA;
      for (int i = 0; i < 6; ++i) sa.eps[i] = sc(i);
      hipLaunchKernelGGL((kop_strain<T, double>), dim3(gmodes()), dim3(NT), 0, stream, sa);
    };
    if (!warm) {
      typename OpElasticCold<T>::Args oa{P, phi[cur], xB_el()};
      launch_zpass<T, 0, 6, OpElasticCold<T>>(Nz, stream, zargs(0, 0, 0, 6), oa);
      ypass(false, 0, 6);
      xpass(false, 0, 6, Sc);
      green(0, false);
      if (cold_hooks()) {
        int rc = cold_hook();
        if (rc) return rc;
        strain_of_U();
      }
    } else {
      strain_of_U();
    }
    xpass(true, 0, 6, Sc);
    ypass(true, 0, 6);
    int it = accel ? 0 : 1;
    bool converged = !accel;
    double res = NAN;
    const T invN = T(1.0 / (double)R);
    while (true) {
      const int mode = !accel ? 3 : (it == 0 ? 1 : 2);
      typename OpHij<T>::Args oa{P, phi[cur], xB_el(), invN, mode == 2 ? 1 : 0};
      launch_zpass_b<T, 6, 6, OpHij<T>>(zb_sel, Nz, stream, zargs(0, 6, 0, 6), oa);
      ypass(false, 0, 6);
      xpass(false, 0, 6, Sc);
      green(mode, accel);
      if (accel) {
        finalize(gk, 2, RedKinds{{RED_SUM, RED_SUM, 0, 0}}, dscal, 1.0);
        PF_HIP(hipMemcpyAsync(hscal, dscal, 2 * sizeof(double), hipMemcpyDeviceToHost, stream));
      }
      xpass(true, 0, 6, Sc);
      ypass(true, 0, 6);
      ++it;
      if (accel) {
        PF_HIP(hipStreamSynchronize(stream));
        const double ds = hscal[0], rs = hscal[1];
        if (!std::isfinite(ds) || !std::isfinite(rs) || ds < 0.0 || rs < 0.0) {
          last_iters = it;
          return fail(PF_E_NONFINITE, "elastic residual is non-finite at step " +
                                          std::to_string(steps_done + 1));
        }
        res = std::sqrt(ds / std::fmax(rs, prm.elastic_residual_absolute_floor));
        if (it >= prm.elastic_iter_min && res <= prm.elastic_residual_tolerance) {
          converged = true;
          break;
        }
        if (it >= prm.elastic_iter_max) break;
      } else {
        if (it >= prm.elastic_iter_max) break;
      }
    }
    if (accel) {
      // the loop leaves strain(increment) behind; the stress needs strain(U)
      strain_of_U();
      xpass(true, 0, 6, Sc);
      ypass(true, 0, 6);
    }
    last_iters = it;
    last_res = res;
    elastic_total += it;
    if (accel) {
      warm_valid = true;
      if (!converged && prm.elastic_fail_on_nonconvergence) {
        char buf[256];
        std::snprintf(buf, sizeof buf,
                      "accelerated elastic solver did not converge at step %lld: iterations=%d "
                      "residual=%.17e tolerance=%.17e",
                      steps_done + 1, it, res, prm.elastic_residual_tolerance);
        return fail(PF_E_NONCONVERGED, buf);
      }
    }
    return PF_OK;
  }

  int elastic_solve_plain() {
    const bool accel = prm.elastic_warm_start_enabled || prm.elastic_residual_control_enabled;
    const bool warm = accel && warm_valid;
    const KDims d = kd();
    const int gk = grid_for(nSc);
    if (!warm) {
      typename OpElasticCold<T>::Args oa{P, phi[cur], xB_el()};
      launch_zpass<T, 0, 6, OpElasticCold<T>>(Nz, stream, zargs(0, 0, 0, 6), oa);
      ypass(false, 0, 6);
      xpass(false, 0, 6, Sc);
      GreenArgs<T> ga{};
      ga.d = d;
      ga.n = nSc;
      for (int i = 0; i < 6; ++i) {
        ga.h[i] = sc(i);
        ga.eps[i] = sc(i);
      }
      for (int i = 0; i < 3; ++i) ga.u_new[i] = uu(u, i);
      ga.partials = nullptr;
      for (int i = 0; i < 21; ++i) ga.C[i] = P.C[i];
      ga.have_old = 0;
      hipLaunchKernelGGL((kop_green<T>), dim3(gk), dim3(NT), 0, stream, ga);
      if (cold_hooks()) {
        int rc = cold_hook();
        if (rc) return rc;
        StrainArgs<T> sa{};
        sa.d = d;
        sa.n = nSc;
        for (int i = 0; i < 3; ++i) sa.u[i] = uu(u, i);
        for (int i = 0; i < 6; ++i) sa.eps[i] = sc(i);
        hipLaunchKernelGGL((kop_strain<T>), dim3(gmodes()), dim3(NT), 0, stream, sa);
      }
    } else {
      StrainArgs<T> sa{};
      sa.d = d;
      sa.n = nSc;
      for (int i = 0; i < 3; ++i) sa.u[i] = uu(u, i);
      for (int i = 0; i < 6; ++i) sa.eps[i] = sc(i);
      hipLaunchKernelGGL((kop_strain<T>), dim3(gmodes()), dim3(NT), 0, stream, sa);
    }
    xpass(true, 0, 6, Sc);
    ypass(true, 0, 6);
    int it = accel ? 0 : 1;
    bool converged = !accel;
    double res = NAN;
    const T invN = T(1.0 / (double)R);
    while (true) {
      typename OpHij<T>::Args oa{P, phi[cur], xB_el(), invN, 0};
      launch_zpass_b<T, 6, 6, OpHij<T>>(zb_sel, Nz, stream, zargs(0, 6, 0, 6), oa);
      ypass(false, 0, 6);
      xpass(false, 0, 6, Sc);
      GreenArgs<T> ga{};
      ga.d = d;
      ga.n = nSc;
      for (int i = 0; i < 6; ++i) {
        ga.h[i] = sc(i);
        ga.eps[i] = sc(i);
      }
      for (int i = 0; i < 3; ++i) {
        ga.u_old[i] = uu(u, i);
        ga.u_new[i] = uu(un, i);
      }
      for (int i = 0; i < 21; ++i) ga.C[i] = P.C[i];
      ga.have_old = 1;
      ga.partials = accel ? partials : nullptr;
      hipLaunchKernelGGL((kop_green<T>), dim3(gk), dim3(NT), 0, stream, ga);
      if (accel) finalize(gk, 2, RedKinds{{RED_SUM, RED_SUM, 0, 0}}, dscal, 1.0);
      if (accel)
        PF_HIP(hipMemcpyAsync(hscal, dscal, 2 * sizeof(double), hipMemcpyDeviceToHost, stream));
      // strain of the new iterate back to Sz: needed by the next h_ij or the final stress
      xpass(true, 0, 6, Sc);
      ypass(true, 0, 6);
      std::swap(u, un);
      ++it;
      if (accel) {
        PF_HIP(hipStreamSynchronize(stream));
        const double ds = hscal[0], rs = hscal[1];
        if (!std::isfinite(ds) || !std::isfinite(rs) || ds < 0.0 || rs < 0.0) {
          last_iters = it;
          return fail(PF_E_NONFINITE, "elastic residual is non-finite at step " +
                                          std::to_string(steps_done + 1));
        }
        res = std::sqrt(ds / std::fmax(rs, prm.elastic_residual_absolute_floor));
        if (it >= prm.elastic_iter_min && res <= prm.elastic_residual_tolerance) {
          converged = true;
          break;
        }
        if (it >= prm.elastic_iter_max) break;
      } else {
        if (it >= prm.elastic_iter_max) break;
      }
    }
    last_iters = it;
    last_res = res;
    elastic_total += it;
    if (accel) {
      warm_valid = true;
      if (!converged && prm.elastic_fail_on_nonconvergence) {
        char buf[256];
        std::snprintf(buf, sizeof buf,
                      "accelerated elastic solver did not converge at step %lld: iterations=%d "
                      "residual=%.17e tolerance=%.17e",
                      steps_done + 1, it, res, prm.elastic_residual_tolerance);
        return fail(PF_E_NONCONVERGED, buf);
      }
    }
    return PF_OK;
  }

  // ---------------------------------------------------------------- step
  // One time step.  It runs in three phases around its host waits (the
  // elastic solver's and the zero mode's), so that a batch runner can drive
  // many handles through the same waits:
  //   step_phase_a  the start of the elastic solve (the whole solve for the
  //                 solvers other than the accelerated one);
  //   (the accelerated solve's further iterations: accel_more / accel_iters)
  //   step_phase_b  the solve's result, the phi update and the composition
  //                 update through the opening of the zero-mode search;
  //   (the search's further steps: zm_more / zm_steps)
  //   step_phase_c  the search's result and the shift, the step's bookkeeping.
  static constexpr int STEP_NFETCH = 13;   // leading scalars the step reads back
  int step() override {
    int rc = step_phase_a();
    if (rc) return rc;
    if (P.elastic && extended) {
      for (;;) {
        PF_HIP(hipMemcpyAsync(hscal + AA_SLOT, dscal + AA_SLOT, AC_GAM * sizeof(double), hipMemcpyDeviceToHost,
                              stream));
        PF_HIP(hipStreamSynchronize(stream));
        if (!accel_more()) break;
        accel_iters(2);
      }
    }
    if ((rc = step_phase_b())) return rc;
    const bool zm = prm.zero_mode != 0;
    for (;;) {
      PF_HIP(hipMemcpyAsync(hscal, dscal, STEP_NFETCH * sizeof(double), hipMemcpyDeviceToHost, stream));
      if (zm)
        PF_HIP(hipMemcpyAsync(hscal + ZM_SLOT, dscal + ZM_SLOT, ZS_SUMS * sizeof(double), hipMemcpyDeviceToHost,
                              stream));
      PF_HIP(hipStreamSynchronize(stream));
      if (!zm || !zm_more()) break;
      zm_steps(4);
    }
    return step_phase_c();
  }

  int step_phase_a() {
    int rc = require_state("pf_step");
    if (rc) return rc;
    state_ok = false;
    minimizing = false;
    // The gradient energy of phi^n is a diagnostic (the reference's dynamics
    // loop does not evaluate it, main_cuda.cu:29261-29290 is minimize only):
    // it is computed when the scalars are read, from phi^n, which stays in
    // the other ping-pong buffer until the next step (egrad_pending, set
    // when the step is accepted).
    egrad_pending = false;
    eng_pending = false;
    sig6_valid = false;
    if (P.elastic) {
      if ((rc = extended ? accel_begin() : elastic_solve())) return rc;
    }
    return PF_OK;
  }

  int step_phase_b() {
    int rc;
    const T invN = T(1.0 / (double)R);
    const bool zm = prm.zero_mode != 0;
    double* phi_old = phi[cur];
    double* phi_new = phi[1 - cur];
    // mixed precision: 1 = first update from new fields (direct, sets phi*),
    // 2 = increment form; 0 in FP64
    const int phi_mode = kMixed ? (phi_banded ? 2 : 1) : 0;
    if (P.elastic) {
      if (extended && (rc = accel_finish())) return rc;
      typename OpPhiRhs<T, true>::Args oa{P, phi_old, xB, trs, invN, keep_sig6 ? sig6 : nullptr, (size_t)R,
                                          phi_mode == 2 ? phis : nullptr};
      ZPassArgs<T> za = zargs(0, 6, 0, 1);
      za.partials = epartials;
      eng_g = launch_zpass<T, 6, 1, OpPhiRhs<T, true>>(Nz, stream, za, oa);
    } else {
      typename OpPhiRhs<T, false>::Args oa{P, phi_old, xB, trs, invN, nullptr, 0, phi_mode == 2 ? phis : nullptr};
      ZPassArgs<T> za = zargs(0, 0, 0, 1);
      za.partials = epartials;
      eng_g = launch_zpass<T, 0, 1, OpPhiRhs<T, false>>(Nz, stream, za, oa);
    }
    const KDims d = kd();
    const int gk = grid_for(nSc);
    // phi semi-implicit update
    ypass(false, 0, 1);
    xpass(false, 0, 1, Sc);
    if (phi_mode)
      launch<KPhiState<T>>(stream, dim3(gmodes()),
                           {d, nSc, sc(0), prm.L_phi * prm.dt * prm.kappa_phi, Phik, phi_mode});
    else
      launch<KPhi<T>>(stream, dim3(gmodes()), {d, nSc, sc(0), prm.L_phi * prm.dt * prm.kappa_phi});
    xpass(true, 0, 1, Sc);
    ypass(true, 0, 1);
    {
      const int phiY_old = (zm || prm.enable_Y_rhs_previous_time_level) ? 1 : 0;
      typename OpPhiFinish<T>::Args oa{P, phi_new, phi_old, Y, xB, trs, invN, phiY_old, phis, phi_mode};
      const int g = launch_zpass<T, 1, 1, OpPhiFinish<T>>(Nz, stream, zargs(0, 1, 1, 1), oa);
      finalize(g, 1, RedKinds{{RED_SUM, 0, 0, 0}}, dscal + 4, 1.0 / (double)R);
    }
    // Y: spectral gradient of mu, fluxes, update.  Mixed precision with the
    // zero mode updates Y by its increment once Y^n is band-limited (after
    // the first step): the FP32 transforms then carry dt (div J - s) /
    // (1 + dt D-bar k^2) rather than Y itself, so their rounding scales with
    // the increment (kop_yupdate), and Y^n needs no transform.
    const bool yinc = kMixed && zm && y_banded;
    y_chain_launch(P, phi_old, phi_new, yinc, P.Y_clip);
    return PF_OK;
  }

  int step_phase_c() {
    int rc;
    if ((rc = y_chain_finish())) return rc;
    cur = 1 - cur;
    egrad_pending = true;
    eng_pending = true;
    sig6_valid = keep_sig6 && P.elastic;
    ++steps_done;
    y_banded = true;
    phi_banded = kMixed;
    state_ok = true;
    return PF_OK;
  }

  // The composition update from the mu spectrum in Sz(1) and D-bar in
  // dscal[4] (main_cuda.cu:30283-30911): the spectral gradient of mu, the
  // fluxes, the lagged right-hand side, the semi-implicit update, then the
  // zero mode or, without it, the clamp of Y to [-Y_clip, y_cap].  PY holds
  // the time step of the update; yinc selects the increment form; the first
  // nfetch device scalars come back with the result.
  int y_chain(const PhysParams& PY, const double* phi_old, const double* phi_new, bool yinc, double y_cap,
              int nfetch) {
    y_chain_launch(PY, phi_old, phi_new, yinc, y_cap);
    const bool zm = prm.zero_mode != 0;
    for (;;) {
      PF_HIP(hipMemcpyAsync(hscal, dscal, nfetch * sizeof(double), hipMemcpyDeviceToHost, stream));
      if (zm)
        PF_HIP(hipMemcpyAsync(hscal + ZM_SLOT, dscal + ZM_SLOT, ZS_SUMS * sizeof(double), hipMemcpyDeviceToHost,
                              stream));
      PF_HIP(hipStreamSynchronize(stream));
      if (!zm || !zm_more()) break;
      zm_steps(4);
    }
    return y_chain_finish();
  }
  // the launches of the composition update, through the opening of the
  // zero-mode search (or the clamp without it)
  void y_chain_launch(const PhysParams& PY, const double* phi_old, const double* phi_new, bool yinc,
                      double y_cap) {
    const T invN = T(1.0 / (double)R);
    const KDims d = kd();
    const bool zm = prm.zero_mode != 0;
    ypass(false, 1, 1);
    xpass(false, 1, 1, sc(1));
    launch<KGrad<T>>(stream, dim3(gmodes()), {d, nSc, sc(1), sc(2), sc(3), sc(4)});
    xpass(true, 2, 3, sc(2));
    ypass(true, 2, 3);
    {
      const double* phiY = (zm || prm.enable_Y_rhs_previous_time_level) ? phi_old : phi_new;
      typename OpFlux<T>::Args oa{PY,   phiY, xB,  phi_old, phi_new, Y, dYdt, invN, zm ? 1 : 0,
                                  prm.enable_Y_rhs_previous_time_level, prm.disable_Y_rhs_gamma_term,
                                  prm.Y_rhs_term_h_scale, yinc ? 1 : 0};
      if constexpr (kMixed) {
        if (yinc)
          launch_zpass<T, 3, 4, OpFlux<T>>(Nz, stream, zargs(2, 3, 0, 4), oa);
        else
          launch_zpass<T, 3, 5, OpFlux<T>>(Nz, stream, zargs(2, 3, 0, 5), oa);
      } else {
        launch_zpass<T, 3, 5, OpFlux<T>>(Nz, stream, zargs(2, 3, 0, 5), oa);
      }
    }
    const int nyc = yinc ? 4 : 5;
    ypass(false, 0, nyc);
    xpass(false, 0, nyc, Sc);
    launch<KYUpdate<T>>(stream, dim3(gmodes()),
                        {d, nSc, sc(0), sc(1), sc(2), yinc ? (const C*)nullptr : (const C*)sc(3),
                         yinc ? sc(3) : sc(4), sc(5), (const double*)(dscal + 4), PY.dt});
    xpass(true, 5, 1, sc(5));
    ypass(true, 5, 1);
    if (zm) {
      typename OpYStar<T>::Args oa{Ystar, invN, yinc ? (const double*)Y : nullptr};
      const int g = launch_zpass<T, 1, 0, OpYStar<T>>(Nz, stream, zargs(5, 1, 0, 0), oa);
      finalize(g, 2, RedKinds{{RED_MIN, RED_MAX, 0, 0}}, dscal, 1.0);
      zm_begin(phi_new, zm_step_const());
    } else {
      typename OpYFinal<T>::Args oa{PY, Y, xB, dYdt, invN, y_cap};
      launch_zpass<T, 1, 0, OpYFinal<T>>(Nz, stream, zargs(5, 1, 0, 0), oa);
    }
  }
  // after the scalars (and the search state) are mirrored: the search's
  // result and the shift; mean D-bar
  int y_chain_finish() {
    int rc = PF_OK;
    if (prm.zero_mode) {
      ZmResult res;
      rc = zero_mode_apply(zm_end(zm_step_const(), "PF zero-mode", &res), res);
    }
    mean_DY = hscal[4];
    return rc;
  }

  // run_pf_conserved_y_zero_mode_host (main_cuda.cu:4269-4405), decided on the
  // device (zm_open_end_kernel, zm_step_end_kernel): the Y* extremes are in
  // dscal[0..1]; `nfetch` more leading scalars come back with the search
  // state in the one wait.
  static constexpr int ZM_SLOT = 64;   // dscal offset of the search state
  int zm_spec = 3;                     // search steps enqueued before the first wait
  static constexpr int AA_SLOT = 128;  // dscal offset of the accelerated solver's control block
  static constexpr int MIN_SLOT = 256; // dscal offset of the phi-only minimize control block (MinCtlSlot)
  static constexpr int NSCAL = 512;    // device scalars (and their pinned host mirror)
  int aa_spec = 4;                     // iterations enqueued before the solver's first wait
  const double* gate_ = nullptr;       // passes do nothing while *gate_ == 0 (set during the solve)
  using ZmResult = ZmResultT;
  ZmConst zm_step_const() const {
    return ZmConst{P.Y_clip, P.Y_clip, P.v_B, target_mass,
                   prm.zero_mode_tolerance_relative * std::fmax(std::fabs(target_mass), 1.0),
                   prm.zero_mode_max_iterations};
  }
  int zero_mode(const double* phi_np1, int nfetch) {
    ZmResult res;
    const int rc = zero_mode_solve(phi_np1, nfetch, zm_step_const(), "PF zero-mode", &res);
    return zero_mode_apply(rc, res);
  }
  // the time step's use of a finished search: its statistics and the shift
  int zero_mode_apply(int rc, const ZmResult& res) {
    zm_iters = res.iters;
    if (rc) return rc;
    zm_lambda = res.lam;
    zm_res = res.r;
    zm_deriv = res.deriv;
    ++zm_accepted;
    launch<KZmApply>(stream, dim3((unsigned)((R + NT - 1) / NT)), {Ystar, res.lam, Y, xB, dYdt, prm.dt, R});
    return PF_OK;
  }
  // the search on Y* (member Ystar) with the extremes in dscal[0..1]; on
  // success res holds the accepted shift, nothing is applied yet
  // The search runs in three parts so that a batch runner can drive it:
  // zm_begin (the opening evaluation of both bracket ends and the start, and
  // the first zm_spec Newton/bisection steps), zm_steps (more steps while the
  // mirrored search state says it runs, zm_more) and zm_end (the host checks
  // on the mirrored state).
  ZmConst zm_c{};
  const double* zm_phi = nullptr;
  int zm_left = 0;
  void zm_begin(const double* phi_np1, const ZmConst& c) {
    double* st = dscal + ZM_SLOT;
    const int g = grid_for(R);
    zm_c = c;
    zm_phi = phi_np1;
    launch<KZmEval3>(stream, dim3(g), {Ystar, phi_np1, (const double*)dscal, c.Y_clip, c.Y_cap, c.vB, R, partials});
    launch<KZmOpenEnd>(stream, dim3(1), {partials, g, st, (const double*)dscal, c});
    zm_left = std::max(c.max_it, 0);
    zm_steps(std::min(zm_spec, zm_left));
  }
  void zm_steps(int batch) {
    double* st = dscal + ZM_SLOT;
    const int g = grid_for(R);
    const ZmConst& c = zm_c;
    batch = std::min(batch, zm_left);
    for (int k = 0; k < batch; ++k) {
      launch<KZmEval>(stream, dim3(g), {Ystar, zm_phi, (const double*)st, c.Y_clip, c.Y_cap, c.vB, R, partials});
      launch<KZmStepEnd>(stream, dim3(1), {partials, g, st, c});
    }
    zm_left -= batch;
  }
  bool zm_more() const { return hscal[ZM_SLOT + ZS_ACTIVE] != 0.0 && zm_left > 0; }
  int zero_mode_solve(const double* phi_np1, int nfetch, const ZmConst& c, const char* what, ZmResult* res) {
    zm_begin(phi_np1, c);
    for (;;) {
      PF_HIP(hipMemcpyAsync(hscal, dscal, nfetch * sizeof(double), hipMemcpyDeviceToHost, stream));
      PF_HIP(hipMemcpyAsync(hscal + ZM_SLOT, dscal + ZM_SLOT, ZS_SUMS * sizeof(double), hipMemcpyDeviceToHost,
                            stream));
      PF_HIP(hipStreamSynchronize(stream));
      if (!zm_more()) break;
      zm_steps(4);
    }
    return zm_end(c, what, res);
  }
  // the accepted shift of a finished search (host mirror), or the rejection
  int zm_end(const ZmConst& c, const char* what, ZmResult* res) {
    const double* hs = hscal + ZM_SLOT;
    const int status = (int)hs[ZS_STATUS];
    res->ymin = hscal[0];
    res->ymax = hscal[1];
    res->lo = hs[ZS_LO];
    res->hi = hs[ZS_HI];
    char buf[256];
    if (status != ZM_OK) {
      std::snprintf(buf, sizeof buf, "%s rejected: %s", what,
                    status == ZM_EMPTY_BRACKET ? "empty bracket"
                    : status == ZM_NONFINITE   ? "evaluation non-finite"
                                               : "target outside bracket");
      return fail(PF_E_ZERO_MODE, buf);
    }
    const int iters = (int)hs[ZS_ITERS];
    // enqueue one step more than this search took, next time
    zm_spec = std::min(std::max(iters + 1, 1), 8);
    const double r = hs[ZS_R], lam = hs[ZS_LAM], invalid = hs[ZS_INVALID];
    res->iters = iters;
    res->newton = (int)hs[ZS_NEWTON];
    res->bisect = (int)hs[ZS_BISECT];
    res->lam = lam;
    res->r = r;
    res->deriv = hs[ZS_DERIV];
    if (std::fabs(r) > c.tol || invalid > 0.0) {
      std::snprintf(buf, sizeof buf, "%s rejected: residual=%.17e lambda=%.17e invalid=%.0f", what, r, lam,
                    invalid);
      return fail(PF_E_ZERO_MODE, buf);
    }
    return PF_OK;
  }

  // ---------------------------------------------------------------- cold-start hooks
  bool state_fp64() const { return kMixed || extended; }   // warm state in Ud (FP64) or u (T)
  bool cold_hooks() const { return nn_net != nullptr || ws_fn != nullptr; }

  int nn_buffers() {
    int rc;
    if (!nn_x0 && ((rc = dalloc(&nn_x0, 16 * (size_t)R)) || (rc = dalloc(&nn_a, 16 * (size_t)R)) ||
                   (rc = dalloc(&nn_b, 16 * (size_t)R)) || (rc = dalloc(&nn_y3, 3 * (size_t)R))))
      return rc;
    return PF_OK;
  }

  // network prediction (force density), transformed to sc(0..2)
  int nn_predict_to_sc() {
    int rc;
    if ((rc = nn_buffers())) return rc;
    double sp, st, so;
    nn::scales(nn_net, &sp, &st, &so);
    const int pool = nn::pool(nn_net);
    if (pool == 2) {
      if (Nx % 2 || Ny % 2 || Nz % 2) return fail(PF_E_ARG, "a pooled network needs even grid dimensions");
      hipLaunchKernelGGL(nn_features_pool_kernel, dim3(grid_for(R / 8)), dim3(NT), 0, stream, P,
                         (const double*)phi[cur], xB_el(), Nx, Ny, Nz, sp, st, nn_x0);
      if ((rc = nn::infer(nn_net, stream, Nx / 2, Ny / 2, Nz / 2, nn_x0, nn_a, nn_b, nn_y3))) return rc;
    } else {
      hipLaunchKernelGGL(nn_features_kernel, dim3(grid_for(R)), dim3(NT), 0, stream, P, (const double*)phi[cur],
                         xB_el(), R, sp, st, nn_x0);
      if ((rc = nn::infer(nn_net, stream, Nx, Ny, Nz, nn_x0, nn_a, nn_b, nn_y3))) return rc;
    }
    typename OpLoadF3<T>::Args oa{nn_y3, (size_t)(pool == 2 ? R / 8 : R), so, pool, Ny, Nz};
    launch_zpass<T, 0, 3, OpLoadF3<T>>(Nz, stream, zargs(0, 0, 0, 3), oa);
    ypass(false, 0, 3);
    xpass(false, 0, 3, Sc);
    return PF_OK;
  }

  void force_map(bool to_force) {
    ForceMapArgs<T> fa{};
    fa.d = kd();
    fa.n = nSc;
    for (int i = 0; i < 3; ++i) fa.a[i] = sc(i);
    for (int i = 0; i < 21; ++i) fa.C[i] = P.C[i];
    fa.to_force = to_force ? 1 : 0;
    hipLaunchKernelGGL((kop_force_map<T>), dim3(grid_for(nSc)), dim3(NT), 0, stream, fa);
  }

  int cold_hook() {
    const int gk = grid_for(nSc);
    if (nn_net) {
      int rc = nn_predict_to_sc();
      if (rc) return rc;
      force_map(false);
      if (state_fp64())
        hipLaunchKernelGGL((kop_add3_state<T>), dim3(gmodes()), dim3(NT), 0, stream, kd(), nSc, (const C*)sc(0),
                           (const C*)sc(1), (const C*)sc(2), Ud, Ud + nScA, Ud + 2 * nScA);
      else
        hipLaunchKernelGGL((kop_add3<T, C>), dim3(gk), dim3(NT), 0, stream, nSc, (const C*)sc(0), (const C*)sc(1),
                           (const C*)sc(2), uu(u, 0), uu(u, 1), uu(u, 2));
    }
    if (ws_fn) {
      PF_HIP(hipStreamSynchronize(stream));
      in_cold_hook = true;
      const int r = ws_fn(ws_handle, ws_user);
      in_cold_hook = false;
      if (r) return fail(PF_E_ARG, "warm-start callback returned " + std::to_string(r));
    }
    return PF_OK;
  }

  int set_warm_start(const double* in) override {
    if (!in) {
      if (in_cold_hook) return PF_OK;
      warm_valid = false;
      hist_count = 0;
      hist_head = 0;
      return PF_OK;
    }
    if (!in_cold_hook) return set_elastic_state(in);
    // inside a cold start: replace the start of this solve only
    if (state_fp64()) {
      PF_HIP(hipMemcpyAsync(Ud, in, 3 * nScA * sizeof(cx<double>), hipMemcpyHostToDevice, stream));
    } else {
      std::vector<C> t;
      state_to_slots(in, &t);
      PF_HIP(hipMemcpyAsync(u, t.data(), 3 * nSc * sizeof(C), hipMemcpyHostToDevice, stream));
      PF_HIP(hipStreamSynchronize(stream));
    }
    return PF_OK;
  }

  int set_warm_start_callback(pf_warm_start_fn fn, void* user, pf_handle* h) override {
    ws_fn = fn;
    ws_user = user;
    ws_handle = h;
    return PF_OK;
  }

  int attach_nn(nn::Net* net) override {
    nn_net = net;
    return net ? nn_buffers() : PF_OK;
  }

  // Converge the solve for the current phi, then store the network input
  // and the target: the force density K(k) (u_converged - u_cold)^ in real
  // space, u_cold the homogeneous cold start.
  int nn_collect(nn::Net* net) override {
    uint16_t* feat;
    float* tgt;
    int rc;
    const int pool = nn::pool(net);
    if (pool == 2 && (Nx % 2 || Ny % 2 || Nz % 2))
      return fail(PF_E_ARG, "pf_nn_collect: a pooled network needs even grid dimensions");
    if ((rc = nn::dataset_slot(net, Nx / pool, Ny / pool, Nz / pool, &feat, &tgt))) return rc;
    if (!P.elastic) return fail(PF_E_ARG, "pf_nn_collect: elasticity is disabled");
    if ((rc = require_state("pf_nn_collect"))) return rc;
    if (pool == 2 && (rc = nn_buffers())) return rc;
    if ((rc = elastic_solve())) return rc;
    double sp, st, so;
    nn::scales(net, &sp, &st, &so);
    if (pool == 2)
      hipLaunchKernelGGL(nn_features_pool_kernel, dim3(grid_for(R / 8)), dim3(NT), 0, stream, P,
                         (const double*)phi[cur], xB_el(), Nx, Ny, Nz, sp, st, feat);
    else
      hipLaunchKernelGGL(nn_features_kernel, dim3(grid_for(R)), dim3(NT), 0, stream, P, (const double*)phi[cur],
                         xB_el(), R, sp, st, feat);
    // homogeneous cold start into un
    typename OpElasticCold<T>::Args oa{P, phi[cur], xB_el()};
    launch_zpass<T, 0, 6, OpElasticCold<T>>(Nz, stream, zargs(0, 0, 0, 6), oa);
    ypass(false, 0, 6);
    xpass(false, 0, 6, Sc);
    const KDims d = kd();
    const int gk = grid_for(nSc);
    GreenArgs<T> ga{};
    ga.d = d;
    ga.n = nSc;
    for (int i = 0; i < 6; ++i) {
      ga.h[i] = sc(i);
      ga.eps[i] = sc(i);
    }
    for (int i = 0; i < 3; ++i) ga.u_new[i] = uu(un, i);
    ga.partials = nullptr;
    for (int i = 0; i < 21; ++i) ga.C[i] = P.C[i];
    ga.have_old = 0;
    hipLaunchKernelGGL((kop_green<T>), dim3(gk), dim3(NT), 0, stream, ga);
    if (state_fp64())
      hipLaunchKernelGGL((kop_diff3_state<T>), dim3(gmodes()), dim3(NT), 0, stream, d, nSc, U_last(),
                         U_last() + nScA, U_last() + 2 * nScA, (const C*)uu(un, 0), (const C*)uu(un, 1),
                         (const C*)uu(un, 2), sc(0), sc(1), sc(2));
    else
      hipLaunchKernelGGL((kop_diff3<T, C>), dim3(gk), dim3(NT), 0, stream, nSc, (const C*)uu(u, 0),
                         (const C*)uu(u, 1), (const C*)uu(u, 2), (const C*)uu(un, 0), (const C*)uu(un, 1),
                         (const C*)uu(un, 2), sc(0), sc(1), sc(2));
    force_map(true);
    xpass(true, 0, 3, Sc);
    ypass(true, 0, 3);
    // the target at full resolution, then (pooled network) its 2x2x2 means
    typename OpStoreF3<T>::Args sa{pool == 2 ? nn_y3 : tgt, (size_t)R, T(1.0 / (double)R)};
    launch_zpass<T, 3, 0, OpStoreF3<T>>(Nz, stream, zargs(0, 3, 0, 0), sa);
    if (pool == 2)
      hipLaunchKernelGGL(pool3_kernel, dim3(grid_for(3 * (R / 8))), dim3(NT), 0, stream, (const float*)nn_y3, Nx,
                         Ny, Nz, tgt);
    PF_HIP(hipStreamSynchronize(stream));
    rc = nn::dataset_commit(net);
    return rc < 0 ? rc : PF_OK;
  }

  // ---------------------------------------------------------------- minimize
  int minimize_begin(const pf_minimize_opts& o) override {
    int rc = require_state("pf_minimize_begin");
    if (rc) return rc;
    if (!m_gexp) {
      if ((rc = dalloc(&m_gexp, R)) || (rc = dalloc(&m_gfull, R)) || (rc = dalloc(&m_s00, R)) ||
          (rc = dalloc(&m_xzero, R)) || (rc = dalloc(&m_eps, 6 * (size_t)R)))
        return rc;
      PF_HIP(hipMemsetAsync(m_xzero, 0, R * sizeof(double), stream));
    }
    mopt = o;
    minimizing = true;
    phi_banded = false;   // minimize moves phi outside the dynamics update
    m_lambda = 0.0;
    m_V0 = o.V0;
    m_iter = 0;
    min_ctl(MCS_INIT);
    m_conv = 0;
    m_converged = false;
    warm_valid = false;   // the eigenstrain differs from the dynamics one
    hist_count = 0;
    hist_head = 0;
    if (o.full_model) {
      if (prm.zero_mode)
        return fail(PF_E_ARG,
                    "full-model minimize with the dynamics zero mode in its composition update is not built; "
                    "the minimize mass constraint conserves the composition instead");
      y_banded = false;
      // chemical baseline (main_cuda.cu:27400-27432): ic_23d_xB_out, else <x_B>
      double xr = prm.ic_23d_xB_out;
      if (!(xr > 0.0)) {
        const int g = grid_for(R);
        hipLaunchKernelGGL(sum_sq_kernel, dim3(g), dim3(NT), 0, stream, (const double*)xB, 0.0, R, partials);
        finalize(g, 2, RedKinds{{RED_SUM, RED_SUM, 0, 0}}, dscal + 58, 1.0);
        if ((rc = fetch(60))) return rc;
        xr = hscal[58] / (double)R;
      }
      m_xref = xr;
      m_gbulk0 = (1.0 - xr) * mu_A(P, xr) + xr * mu_B(P, xr);
      double c = o.xB_max_safe;
      if (c < prm.xB_eps) c = prm.xB_eps;
      if (c > 1.0 - prm.xB_eps) c = 1.0 - prm.xB_eps;
      m_ycap = std::fmin(prm.Y_clip, std::log(c / (1.0 - c)));
      m_Fprev = NAN;
      m_mass = ZmResultT{};
      m_mass_target = NAN;
      if (o.mass_constraint) {
        // main_cuda.cu:24599-24614: the constraint owns the Y zero mode
        if (prm.enable_Y_rhs_previous_time_level || prm.disable_Y_rhs_gamma_term || prm.Y_rhs_term_h_scale != 1.0)
          return fail(PF_E_ARG,
                      "the minimize mass constraint requires the legacy Y update: previous-time-level rhs off, "
                      "gamma term on, Y_rhs_term_h_scale 1");
        // the exact initial ledger, or an explicit target that must match it
        // (main_cuda.cu:28117-28146)
        double m, vf;
        if ((rc = mass_vf(&m, &vf))) return rc;
        const double t = o.target_mass_code > 0.0 ? o.target_mass_code : m;
        const double tol = o.mass_tolerance_relative * std::fmax(std::fabs(t), 1.0);
        if (!std::isfinite(m) || std::fabs(m - t) > tol) {
          char buf[256];
          std::snprintf(buf, sizeof buf,
                        "explicit minimize target mass does not match the initial ledger: current=%.17e "
                        "target=%.17e residual=%.17e tolerance=%.17e",
                        m, t, m - t, tol);
          return fail(PF_E_ARG, buf);
        }
        m_mass_target = t;
      }
      m_ncomp = 0;
      if (o.component_count > 0) {
        // ownership labels (read_raw_component_labels_i32, main_cuda.cu:2805-2833,
        // 27262-27290): every label in range and owning a cell
        const int K = o.component_count;
        std::vector<long> owned(K, 0);
        for (long i = 0; i < R; ++i) {
          const int c = o.component_labels[i];
          if (c < 0 || c >= K) {
            char buf[160];
            std::snprintf(buf, sizeof buf, "component label out of range at idx=%ld: %d not in [0,%d)", i, c, K);
            return fail(PF_E_ARG, buf);
          }
          ++owned[c];
        }
        for (int c = 0; c < K; ++c)
          if (owned[c] == 0) {
            char buf[96];
            std::snprintf(buf, sizeof buf, "V5 component %d owns no grid cells", c);
            return fail(PF_E_ARG, buf);
          }
        const int nb = comp_blocks();
        if (!m_labels && ((rc = dalloc(&m_labels, R)) ||
                          (rc = dalloc(&m_cbuf, 128 + 3 * COMP_MAX + (size_t)nb * 3 * COMP_MAX))))
          return rc;
        PF_HIP(hipMemcpy(m_labels, o.component_labels, R * sizeof(int), hipMemcpyHostToDevice));
        m_ctarget.assign(o.component_target_h_sums, o.component_target_h_sums + K);
        m_clambda.assign(K, 0.0);
        m_csum.assign(3 * K, 0.0);
        m_cerr = NAN;
        m_ncomp = K;
      }
    } else if (o.mass_constraint || o.freeze_phi || o.component_count) {
      return fail(PF_E_ARG, "the mass constraint, freeze_phi and the component constraints belong to the "
                            "full-model minimize");
    }
    return PF_OK;
  }

  int comp_blocks() const { return (int)((R + (long)NT * COMP_VPT - 1) / ((long)NT * COMP_VPT)); }
  double* comp_lambda_dev() { return m_cbuf; }
  double* comp_aux_dev() { return m_cbuf + 64; }
  // per-label sums of h' g (g may be null), h'^2 and h(clamp01 phi) into m_csum
  int comp_moments(const double* ph, const double* g) {
    const int nb = comp_blocks();
    double* sums = m_cbuf + 128;
    double* part = sums + 3 * COMP_MAX;
    hipLaunchKernelGGL(comp_moments_kernel, dim3(nb), dim3(NT), 0, stream, ph, g, (const int*)m_labels, m_ncomp,
                       R, part);
    hipLaunchKernelGGL(comp_sum_kernel, dim3(1), dim3(NT), 0, stream, (const double*)part, nb, 3 * m_ncomp, sums);
    PF_HIP(hipMemcpyAsync(m_csum.data(), sums, 3 * m_ncomp * sizeof(double), hipMemcpyDeviceToHost, stream));
    PF_HIP(hipStreamSynchronize(stream));
    return PF_OK;
  }
  // -sum h' g / sum h'^2 per label (main_cuda.cu:29468-29478, 29694-29700)
  void comp_multipliers(std::vector<double>* lam) {
    lam->resize(m_ncomp);
    for (int c = 0; c < m_ncomp; ++c) {
      const double d = m_csum[3 * c + 1];
      (*lam)[c] = std::fabs(d) > 1.0e-30 ? -m_csum[3 * c] / d : 0.0;
    }
  }
  int comp_upload(double* dst, const std::vector<double>& v) {
    PF_HIP(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    PF_HIP(hipStreamSynchronize(stream));   // v is pageable host memory reused by the caller
    return PF_OK;
  }
  // the per-label projection: up to three first-order corrections, then the
  // largest relative error of the label sums (main_cuda.cu:32207-32275)
  int comp_project(double* ph) {
    int rc;
    std::vector<double> corr(m_ncomp);
    for (int k = 0; k < 3; ++k) {
      if ((rc = comp_moments(ph, nullptr))) return rc;
      bool need = false;
      for (int c = 0; c < m_ncomp; ++c) {
        const double delta = m_ctarget[c] - m_csum[3 * c + 2];
        corr[c] = m_csum[3 * c + 1] > 1.0e-30 ? delta / m_csum[3 * c + 1] : 0.0;
        if (std::fabs(delta) > 1.0e-14) need = true;
      }
      if (!need) break;
      if ((rc = comp_upload(comp_aux_dev(), corr))) return rc;
      hipLaunchKernelGGL(comp_project_kernel, dim3(grid_for(R)), dim3(NT), 0, stream, ph, (const int*)m_labels,
                         (const double*)comp_aux_dev(), R);
    }
    if ((rc = comp_moments(ph, nullptr))) return rc;
    m_cerr = 0.0;
    for (int c = 0; c < m_ncomp; ++c)
      m_cerr = std::fmax(m_cerr, std::fabs(m_csum[3 * c + 2] - m_ctarget[c]) / std::fmax(m_ctarget[c], 1.0e-30));
    return PF_OK;
  }

  // The minimize mass constraint (main_cuda.cu:32297-32366): the zero-mode
  // search on the clamped Y of this iteration at the final phi, with Y bounded
  // to [-Y_clip, logit(xB_max_safe)]; Y and x_B take the shift, dY/dt keeps
  // the clamp's.  m_mass holds the search.
  int min_mass_constraint(const double* ph) {
    const int gR = grid_for(R);
    hipLaunchKernelGGL(copy_minmax_kernel, dim3(gR), dim3(NT), 0, stream, (const double*)Y, Ystar, R, partials);
    finalize(gR, 2, RedKinds{{RED_MIN, RED_MAX, 0, 0}}, dscal, 1.0);
    const ZmConst c{P.Y_clip, m_ycap, P.v_B, m_mass_target,
                    mopt.mass_tolerance_relative * std::fmax(std::fabs(m_mass_target), 1.0),
                    mopt.mass_max_iterations};
    char what[96];
    std::snprintf(what, sizeof what, "minimize mass constraint (iteration %d)", m_iter);
    m_mass = ZmResultT{};
    int rc = zero_mode_solve(ph, 2, c, what, &m_mass);
    if (rc) return rc;
    hipLaunchKernelGGL(zm_apply_kernel, dim3((unsigned)((R + NT - 1) / NT)), dim3(NT), 0, stream, Ystar,
                       m_mass.lam, Y, xB, (double*)nullptr, mopt.dt, R);
    return PF_OK;
  }

  // phi (real, FP64) -> forward z via Op, then y and x passes into sc(0)
  void min_forward_to_sc0() {
    ypass(false, 0, 1);
    xpass(false, 0, 1, Sc);
  }
  // Lap^ of sc(0) -> inverse x, y passes into Sz(1)
  void min_lap_to_sz1() {
    const KDims d = kd();
    launch<KLap<T>>(stream, dim3(grid_for(nSc)), {d, nSc, (const C*)sc(0), sc(1)});
    xpass(true, 1, 1, sc(1));
    ypass(true, 1, 1);
  }
  // g_full of phi from Sz(1); sums of h' g_full and h'^2 into dscal[o], dscal[o+1]
  void min_gfull(const double* ph, int o) {
    const T invN = T(1.0 / (double)R);
    typename OpMinGfull<T>::Args oa{ph, m_gexp, m_gfull, prm.kappa_phi, invN};
    const int gz = launch_zpass<T, 1, 0, OpMinGfull<T>>(Nz, stream, zargs(1, 1, 0, 0), oa);
    finalize(gz, 2, RedKinds{{RED_SUM, RED_SUM, 0, 0}}, dscal + o, 1.0);
  }
  // semi-implicit phi update from the field phi - L dt (g_exp + lambda h') of ph
  void min_update(const double* ph, double lambda, const int* labels = nullptr, const double* lambdas = nullptr,
                  const double* lambda_p = nullptr) {
    const double Ldt = prm.L_phi * mopt.dt;
    typename OpMinUpd<T>::Args oa{ph, m_gexp, lambda, Ldt, labels, lambdas, lambda_p};
    launch_zpass<T, 0, 1, OpMinUpd<T>>(Nz, stream, zargs(0, 0, 0, 1), oa);
    min_forward_to_sc0();
    launch<KPhi<T>>(stream, dim3(grid_for(nSc)), {kd(), nSc, sc(0), Ldt * prm.kappa_phi});
    xpass(true, 0, 1, Sc);
    ypass(true, 0, 1);
  }
  static double min_multiplier(double s_hg, double s_h2) { return std::fabs(s_h2) > 1e-30 ? -s_hg / s_h2 : 0.0; }

  // the phi-only iteration's control decisions (min_ctl_kernel)
  void min_ctl(int stage) {
    const MinCtlConst c{(double)R, std::fmin(std::fmax(mopt.eta_lambda_vol, 0.0), 1.0), m_V0, m_iter == 1 ? 1 : 0};
    launch<KMinCtl>(stream, dim3(1), {(const double*)dscal, dscal + MIN_SLOT, stage, c});
  }

  // One phi-only iteration.  Every decision between the passes (the
  // multipliers, V0 from the first iteration, whether and how far to
  // project) is taken on the device (min_ctl_kernel), so the iteration waits
  // on the host only inside the elastic solve and once at its end, for the
  // scalars of the stopping rule.
  // One phi-only iteration.  Every decision between the passes (the
  // multipliers, V0 from the first iteration, whether and how far to
  // project) is taken on the device (min_ctl_kernel), so the iteration waits
  // on the host only inside the elastic solve and once at its end, for the
  // scalars of the stopping rule.  It runs in three phases so that a batch
  // runner can drive many handles through the same waits:
  //   min_phase_a  energies and <h> of phi^n, the start of the elastic solve;
  //   (the solve's remaining iterations: accel_more / accel_iters)
  //   min_phase_b  the solve's result and everything up to the final scalars;
  //   min_phase_c  the stopping rule and the statistics, from the scalars.
  int minimize_iterate(pf_minimize_stats* st) override {
    if (!minimizing) return fail(PF_E_ARG, "pf_minimize_begin has not been called on this handle");
    if (int rs = require_state("pf_minimize_iterate")) return rs;
    state_ok = false;
    if (mopt.full_model) return minimize_iterate_full(st);
    int rc = min_phase_a();
    if (rc) return rc;
    if (P.elastic && extended) {
      for (;;) {
        PF_HIP(hipMemcpyAsync(hscal + AA_SLOT, dscal + AA_SLOT, AC_GAM * sizeof(double), hipMemcpyDeviceToHost,
                              stream));
        PF_HIP(hipStreamSynchronize(stream));
        if (!accel_more()) break;
        accel_iters(2);
      }
    }
    if ((rc = min_phase_b())) return rc;
    PF_HIP(hipMemcpyAsync(hscal, dscal, 64 * sizeof(double), hipMemcpyDeviceToHost, stream));
    PF_HIP(hipMemcpyAsync(hscal + MIN_SLOT, dscal + MIN_SLOT, MC_COUNT * sizeof(double), hipMemcpyDeviceToHost,
                          stream));
    PF_HIP(hipStreamSynchronize(stream));
    return min_phase_c(st);
  }

  int min_phase_a() {
    const double* phi_old = phi[cur];
    const int gR = grid_for(R);
    ++m_iter;
    last_iters = 0;
    // energies and <h> of phi^n (main_cuda.cu:29264-29346)
    launch<KGradEnergy>(stream, dim3(gR), {phi_old, Nx, Ny, Nz, prm.dx, prm.dy, prm.dz, prm.kappa_phi, partials});
    finalize(gR, 1, RedKinds{{RED_SUM, 0, 0, 0}}, dscal + 8, 1.0 / (double)R);
    launch<KMinVolume>(stream, dim3(gR), {phi_old, R, partials});
    finalize(gR, 2, RedKinds{{RED_SUM, RED_SUM, 0, 0}}, dscal + MS_VOL, 1.0);
    if (P.elastic) {
      // the accelerated solve is driven by the caller; the others run here
      const int rc = extended ? accel_begin() : elastic_solve();
      if (rc) return rc;
    }
    return PF_OK;
  }

  int min_phase_b() {
    const T invN = T(1.0 / (double)R);
    const bool el = P.elastic != 0;
    double* phi_old = phi[cur];
    double* phi_new = phi[1 - cur];
    const int gR = grid_for(R);
    double* mc = dscal + MIN_SLOT;
    if (el && extended) {
      const int rc = accel_finish();
      if (rc) return rc;
    }
    if (el) {
      typename OpMinRhs<T, true>::Args oa{P, phi_old, m_gexp, m_eps, m_s00, (size_t)R, invN};
      const int gz = launch_zpass<T, 6, 1, OpMinRhs<T, true>>(Nz, stream, zargs(0, 6, 0, 1), oa);
      finalize(gz, 3, RedKinds{{RED_SUM, RED_SUM, RED_SUM, 0}}, dscal + 9, 1.0);
    } else {
      typename OpMinRhs<T, false>::Args oa{P, phi_old, m_gexp, m_eps, m_s00, (size_t)R, invN};
      const int gz = launch_zpass<T, 0, 1, OpMinRhs<T, false>>(Nz, stream, zargs(0, 0, 0, 1), oa);
      finalize(gz, 3, RedKinds{{RED_SUM, RED_SUM, RED_SUM, 0}}, dscal + 9, 1.0);
    }
    // g_full(phi^n) and the under-relaxed multiplier (main_cuda.cu:29436-29500)
    min_forward_to_sc0();
    min_lap_to_sz1();
    min_gfull(phi_old, MS_RAW);
    min_ctl(MCS_RAW);
    // phi^{n+1}
    min_update(phi_old, 0.0, nullptr, nullptr, mc + MC_LAMBDA);
    int gz;
    if (el) {
      typename OpMinFinish<T, true, true>::Args oa{P, phi_old, phi_new, m_gexp, m_eps, m_s00, (size_t)R, invN};
      gz = launch_zpass<T, 1, 1, OpMinFinish<T, true, true>>(Nz, stream, zargs(0, 1, 0, 1), oa);
    } else {
      typename OpMinFinish<T, false, true>::Args oa{P, phi_old, phi_new, m_gexp, m_eps, m_s00, (size_t)R, invN};
      gz = launch_zpass<T, 1, 1, OpMinFinish<T, false, true>>(Nz, stream, zargs(0, 1, 0, 1), oa);
    }
    finalize(gz, 1, RedKinds{{RED_SUM, 0, 0, 0}}, dscal + 52, 1.0);
    // KKT residual at phi^{n+1} with the exact multiplier (main_cuda.cu:29596-29728)
    min_forward_to_sc0();
    min_lap_to_sz1();
    min_gfull(phi_new, MS_KKT);
    min_ctl(MCS_KKT);
    launch<KMinKkt>(stream, dim3(gR), {phi_new, m_gfull, 0.0, (const int*)nullptr, (const double*)nullptr, R, partials, (const double*)(mc + MC_LAM_KKT)});
    finalize(gR, 1, RedKinds{{RED_SUM, 0, 0, 0}}, dscal + 55, 1.0);
    // post-projection sub-steps, eta = 1, stress frozen (main_cuda.cu:32104-32197)
    for (int k = 0; k < mopt.post_projection_iters; ++k) {
      if (k > 0) {
        if (el) {
          typename OpMinEval<T, true>::Args oa{P, phi_new, m_gexp, m_eps, m_s00, (size_t)R};
          launch_zpass<T, 0, 1, OpMinEval<T, true>>(Nz, stream, zargs(0, 0, 0, 1), oa);
        } else {
          typename OpMinEval<T, false>::Args oa{P, phi_new, m_gexp, m_eps, m_s00, (size_t)R};
          launch_zpass<T, 0, 1, OpMinEval<T, false>>(Nz, stream, zargs(0, 0, 0, 1), oa);
        }
        min_forward_to_sc0();
        min_lap_to_sz1();
        min_gfull(phi_new, MS_PP);
        min_ctl(MCS_PP);
      }
      min_ctl(MCS_PP_COMMIT);
      min_update(phi_new, 0.0, nullptr, nullptr, mc + MC_LAMBDA);
      typename OpMinFinish<T, false, false>::Args oa{P, phi_new, phi_new, m_gexp, m_eps, m_s00, (size_t)R, invN};
      launch_zpass<T, 1, 0, OpMinFinish<T, false, false>>(Nz, stream, zargs(0, 1, 0, 0), oa);
    }
    // first-order volume projection (main_cuda.cu:32274-32290)
    launch<KMinVolume>(stream, dim3(gR), {phi_new, R, partials});
    finalize(gR, 2, RedKinds{{RED_SUM, RED_SUM, 0, 0}}, dscal + MS_VOL, 1.0);
    min_ctl(MCS_VOLUME);
    launch<KMinProject>(stream, dim3(gR), {phi_new, 0.0, R, partials, (const double*)mc});
    finalize(gR, 1, RedKinds{{RED_SUM, 0, 0, 0}}, dscal + MS_PROJ, 1.0);
    min_ctl(MCS_FINAL);
    return PF_OK;
  }

  int min_phase_c(pf_minimize_stats* st) {
    const bool el = P.elastic != 0;
    const double* hm = hscal + MIN_SLOT;
    const double grad_mean = hscal[8];
    const double gel_sum = hscal[9], dw_sum = hscal[10], core_sum = hscal[11];
    const double rms_dphi = std::sqrt(hscal[52] / (double)R) / mopt.dt;
    const double rms_res = std::sqrt(hscal[55] / (double)R);
    const double lam_raw = hm[MC_LAM_RAW], lam_kkt = hm[MC_LAM_KKT];
    const double mean_h = hm[MC_MEAN_H];
    m_lambda = hm[MC_LAMBDA];
    m_V0 = hm[MC_V0];
    cur = 1 - cur;
    // convergence bookkeeping (main_cuda.cu:32400-32560); condB needs the
    // excess free energy, which phi-only minimize leaves undefined
    const double vol_err_rel = m_V0 > 1e-30 ? std::fabs(mean_h - m_V0) / (m_V0 + 1e-30) : 0.0;
    const bool condA = std::isfinite(rms_dphi) && rms_dphi < mopt.rms_dphi_threshold && std::isfinite(rms_res) &&
                       rms_res < mopt.rms_res_threshold && std::isfinite(vol_err_rel) &&
                       vol_err_rel < mopt.vol_err_rel_threshold;
    const bool t_ok = (double)m_iter * mopt.dt + 1e-14 >= mopt.min_pseudo_time;
    m_conv = (condA && t_ok) ? m_conv + 1 : 0;
    if (m_conv >= mopt.convergence_steps) m_converged = true;
    e_grad = grad_mean;
    e_dw = dw_sum / (double)R;
    e_el = gel_sum / (double)R;
    e_chem = 0.0;
    eng_pending = false;
    ++steps_done;
    if (st) {
      st->iteration = m_iter;
      st->converged = m_converged ? 1 : 0;
      st->convergence_count = m_conv;
      st->elastic_iterations = el ? last_iters : 0;
      st->mean_h = mean_h;
      st->V0 = m_V0;
      st->lambda_vol = m_lambda;
      st->F_surf_hat = grad_mean + dw_sum / (double)R;
      st->F_el_hat = gel_sum / (double)R;
      st->interface_sum = grad_mean * (double)R + dw_sum;
      st->el_core_sum = core_sum;
      st->rms_res = rms_res;
      st->rms_dphi = rms_dphi;
      st->vol_err_rel = vol_err_rel;
      st->lambda_raw = lam_raw;
      st->lambda_kkt = lam_kkt;
      st->F_chem_excess_hat = st->F_total_excess_hat = st->F_chem_CNT_hat = st->F_total_CNT_hat = NAN;
      st->rms_dY = st->energy_diff_rel = NAN;
      st->mass_target = st->mass_current = st->mass_err_rel = st->mass_lambda = NAN;
      st->mass_residual = st->mass_derivative = st->mass_Ystar_min = st->mass_Ystar_max = NAN;
      st->mass_lambda_lower = st->mass_lambda_upper = NAN;
      st->mass_iterations = st->mass_newton_steps = st->mass_bisection_steps = 0;
    }
    if (!std::isfinite(mean_h) || !std::isfinite(m_lambda))
      return fail(PF_E_NONFINITE, "minimize produced a non-finite phi");
    state_ok = true;
    return PF_OK;
  }

  // One full-model iteration (main_cuda.cu:28427-32600 with P.mode == 1 and
  // minimize_full_model == 1): the solve with the composition eigenstrain;
  // energies of phi^n with the chemical excess; x_B refreshed from Y and the
  // dynamics phi right-hand side as g_exp; the multiplier step; the KKT
  // residual at phi^{n+1}; the composition update at the pseudo-time step
  // with Y <= logit(minimize_xB_max_safe); the post-projection sub-steps with
  // the new x_B; the volume projection; rms_dY and the full-model condA/condB.
  int minimize_iterate_full(pf_minimize_stats* st) {
    const T invN = T(1.0 / (double)R);
    const bool el = P.elastic != 0;
    double* phi_old = phi[cur];
    double* phi_new = phi[1 - cur];
    const int gR = grid_for(R);
    int rc;
    ++m_iter;
    last_iters = 0;
    hipLaunchKernelGGL(grad_energy_kernel, dim3(gR), dim3(NT), 0, stream, phi_old, Nx, Ny, Nz, prm.dx, prm.dy,
                       prm.dz, prm.kappa_phi, partials);
    finalize(gR, 1, RedKinds{{RED_SUM, 0, 0, 0}}, dscal + 8, 1.0 / (double)R);
    hipLaunchKernelGGL(min_volume_kernel, dim3(gR), dim3(NT), 0, stream, phi_old, R, partials);
    finalize(gR, 2, RedKinds{{RED_SUM, RED_SUM, 0, 0}}, dscal + 48, 1.0);
    if (el) {
      if ((rc = elastic_solve())) return rc;
      typename OpMinRhsFull<T, true>::Args oa{P, phi_old, Y, xB, m_gexp, m_eps, m_s00, trs, (size_t)R, invN};
      const int gz = launch_zpass<T, 6, 1, OpMinRhsFull<T, true>>(Nz, stream, zargs(0, 6, 0, 1), oa);
      finalize(gz, 4, RedKinds{{RED_SUM, RED_SUM, RED_SUM, RED_SUM}}, dscal + 9, 1.0);
    } else {
      typename OpMinRhsFull<T, false>::Args oa{P, phi_old, Y, xB, m_gexp, m_eps, m_s00, trs, (size_t)R, invN};
      const int gz = launch_zpass<T, 0, 1, OpMinRhsFull<T, false>>(Nz, stream, zargs(0, 0, 0, 1), oa);
      finalize(gz, 4, RedKinds{{RED_SUM, RED_SUM, RED_SUM, RED_SUM}}, dscal + 9, 1.0);
    }
    min_forward_to_sc0();
    min_lap_to_sz1();
    min_gfull(phi_old, 50);
    if ((rc = fetch(58))) return rc;
    const double grad_mean = hscal[8];
    const double gel_sum = hscal[9], dw_sum = hscal[10], core_sum = hscal[11], chem_sum = hscal[12];
    const double mean_h0 = hscal[48] / (double)R;
    if (m_iter == 1 && m_V0 <= 0.0) m_V0 = mean_h0;
    const double F_surf = grad_mean + dw_sum / (double)R;
    const double F_el = gel_sum / (double)R;
    const double F_chem_excess = chem_sum / (double)R - m_gbulk0;
    const double F_total_excess = F_chem_excess + F_surf + F_el;
    double F_chem_cnt = NAN, F_total_cnt = NAN;
    if (prm.ic_23d_xB_out > 0.0 && mean_h0 > 0.0) {
      F_chem_cnt = mean_h0 * (P.mu0 - mu_B(P, prm.ic_23d_xB_out));
      F_total_cnt = F_surf + F_el + F_chem_cnt;
    }
    const double F_prev = m_Fprev;
    m_Fprev = F_total_excess;
    // with freeze_phi, phi stays as it is (main_cuda.cu:29357, 32108, 32192, 32276)
    const bool frozen = mopt.freeze_phi != 0;
    const double* phi_end = frozen ? phi_old : phi_new;
    double lam_raw = NAN, lam_kkt = NAN, rms_dphi = NAN, rms_res = NAN;
    const bool comp = m_ncomp > 0;
    if (!frozen) {
      const double eta = std::fmin(std::fmax(mopt.eta_lambda_vol, 0.0), 1.0);
      if (comp) {
        // one multiplier per label (main_cuda.cu:29446-29510)
        std::vector<double> raw;
        if ((rc = comp_moments(phi_old, m_gfull))) return rc;
        comp_multipliers(&raw);
        for (int c = 0; c < m_ncomp; ++c) m_clambda[c] = (1.0 - eta) * m_clambda[c] + eta * raw[c];
        if ((rc = comp_upload(comp_lambda_dev(), m_clambda))) return rc;
        min_update(phi_old, 0.0, m_labels, comp_lambda_dev());
      } else {
        lam_raw = min_multiplier(hscal[50], hscal[51]);
        m_lambda = (1.0 - eta) * m_lambda + eta * lam_raw;
        // phi^{n+1}, and g_exp there for the KKT residual (x_B unchanged: Y is)
        min_update(phi_old, m_lambda);
      }
      int gz;
      if (el) {
        typename OpMinFinishFull<T, true>::Args oa{P, phi_old, phi_new, m_gexp, m_eps, m_s00, trs, xB, (size_t)R,
                                                   invN};
        gz = launch_zpass<T, 1, 1, OpMinFinishFull<T, true>>(Nz, stream, zargs(0, 1, 0, 1), oa);
      } else {
        typename OpMinFinishFull<T, false>::Args oa{P, phi_old, phi_new, m_gexp, m_eps, m_s00, trs, xB, (size_t)R,
                                                    invN};
        gz = launch_zpass<T, 1, 1, OpMinFinishFull<T, false>>(Nz, stream, zargs(0, 1, 0, 1), oa);
      }
      finalize(gz, 1, RedKinds{{RED_SUM, 0, 0, 0}}, dscal + 52, 1.0);
      min_forward_to_sc0();
      min_lap_to_sz1();
      min_gfull(phi_new, 53);
      if ((rc = fetch(58))) return rc;
      rms_dphi = std::sqrt(hscal[52] / (double)R) / mopt.dt;
      if (comp) {
        // the exact multiplier of every label at phi^{n+1} (main_cuda.cu:29660-29728)
        std::vector<double> kkt;
        if ((rc = comp_moments(phi_new, m_gfull))) return rc;
        comp_multipliers(&kkt);
        if ((rc = comp_upload(comp_aux_dev(), kkt))) return rc;
        hipLaunchKernelGGL(min_kkt_kernel, dim3(gR), dim3(NT), 0, stream, phi_new, m_gfull, 0.0,
                           (const int*)m_labels, (const double*)comp_aux_dev(), R, partials,
                           (const double*)nullptr);
      } else {
        lam_kkt = min_multiplier(hscal[53], hscal[54]);
        hipLaunchKernelGGL(min_kkt_kernel, dim3(gR), dim3(NT), 0, stream, phi_new, m_gfull, lam_kkt,
                           (const int*)nullptr, (const double*)nullptr, R, partials, (const double*)nullptr);
      }
      finalize(gR, 1, RedKinds{{RED_SUM, 0, 0, 0}}, dscal + 55, 1.0);
    }
    // the composition update at the pseudo-time step (P.dt is the minimize dt)
    {
      const double* phiY = prm.enable_Y_rhs_previous_time_level ? phi_old : phi_end;
      typename OpMuX<T>::Args oa{P, Y, phiY, xB, trs};
      const int g = launch_zpass<T, 0, 1, OpMuX<T>>(Nz, stream, zargs(0, 0, 1, 1), oa);
      finalize(g, 1, RedKinds{{RED_SUM, 0, 0, 0}}, dscal + 4, 1.0 / (double)R);
      PhysParams PY = P;
      PY.dt = mopt.dt;
      if ((rc = y_chain(PY, phi_old, phi_end, false, m_ycap, 13))) return rc;
    }
    // post-projection sub-steps with the updated x_B, elastic state frozen
    for (int k = 0; !frozen && k < mopt.post_projection_iters; ++k) {
      if (el) {
        typename OpMinEvalFull<T, true>::Args oa{P, phi_new, m_gexp, m_eps, m_s00, trs, xB, (size_t)R};
        launch_zpass<T, 0, 1, OpMinEvalFull<T, true>>(Nz, stream, zargs(0, 0, 0, 1), oa);
      } else {
        typename OpMinEvalFull<T, false>::Args oa{P, phi_new, m_gexp, m_eps, m_s00, trs, xB, (size_t)R};
        launch_zpass<T, 0, 1, OpMinEvalFull<T, false>>(Nz, stream, zargs(0, 0, 0, 1), oa);
      }
      min_forward_to_sc0();
      min_lap_to_sz1();
      min_gfull(phi_new, 56);
      if ((rc = fetch(58))) return rc;
      m_lambda = min_multiplier(hscal[56], hscal[57]);
      min_update(phi_new, m_lambda);
      typename OpMinFinish<T, false, false>::Args oa{P, phi_new, phi_new, m_gexp, m_eps, m_s00, (size_t)R, invN};
      launch_zpass<T, 1, 0, OpMinFinish<T, false, false>>(Nz, stream, zargs(0, 1, 0, 0), oa);
    }
    // first-order volume projection
    hipLaunchKernelGGL(min_volume_kernel, dim3(gR), dim3(NT), 0, stream, phi_end, R, partials);
    finalize(gR, 2, RedKinds{{RED_SUM, RED_SUM, 0, 0}}, dscal + 48, 1.0);
    if ((rc = fetch(58))) return rc;
    if (!frozen) rms_res = std::sqrt(hscal[55] / (double)R);
    double mean_h = hscal[48] / (double)R;
    if (comp) {
      if ((rc = comp_project(phi_new))) return rc;
      hipLaunchKernelGGL(min_volume_kernel, dim3(gR), dim3(NT), 0, stream, (const double*)phi_new, R, partials);
      finalize(gR, 2, RedKinds{{RED_SUM, RED_SUM, 0, 0}}, dscal + 48, 1.0);
      if ((rc = fetch(58))) return rc;
      mean_h = hscal[48] / (double)R;
    } else if (!frozen && m_V0 > 1e-30 && std::isfinite(mean_h)) {
      const double vol_err = m_V0 - mean_h;
      if (std::fabs(vol_err) > 1e-14) {
        const double mean_hp2 = hscal[49] / (double)R;
        const double lc = mean_hp2 > 1e-30 ? vol_err / mean_hp2 : 0.0;
        hipLaunchKernelGGL(min_project_kernel, dim3(gR), dim3(NT), 0, stream, phi_new, lc, R, partials,
                           (const double*)nullptr);
        finalize(gR, 1, RedKinds{{RED_SUM, 0, 0, 0}}, dscal + 48, 1.0);
        if ((rc = fetch(58))) return rc;
        mean_h = hscal[48] / (double)R;
      }
    }
    if (!frozen) cur = 1 - cur;
    // the exact mass after every phi correction (main_cuda.cu:32291-32296),
    // then rms_dY = rms(Y^{n+1} - Y^n)/dt with the shift included
    double mass_cur = NAN, mass_err_rel = NAN, lam_dt = 0.0;
    if (mopt.mass_constraint) {
      if ((rc = min_mass_constraint(phi[cur]))) return rc;
      double vf;
      if ((rc = mass_vf(&mass_cur, &vf))) return rc;
      mass_err_rel = std::fabs(mass_cur - m_mass_target) / std::fmax(std::fabs(m_mass_target), 1.0);
      lam_dt = m_mass.lam / mopt.dt;
    }
    hipLaunchKernelGGL(sum_sq_kernel, dim3(gR), dim3(NT), 0, stream, (const double*)dYdt, lam_dt, R, partials);
    finalize(gR, 2, RedKinds{{RED_SUM, RED_SUM, 0, 0}}, dscal + 58, 1.0);
    if ((rc = fetch(60))) return rc;
    const double rms_dY = std::sqrt(hscal[59] / (double)R);
    const double energy_diff_rel =
        std::isfinite(F_prev) ? std::fabs(F_total_excess - F_prev) / std::fmax(std::fabs(F_prev), 1.0e-20) : NAN;
    const double vol_err_rel =
        comp ? m_cerr : (m_V0 > 1e-30 ? std::fabs(mean_h - m_V0) / (m_V0 + 1e-30) : 0.0);
    auto fin = [](double v) { return std::isfinite(v); };
    const bool vol_ok = fin(vol_err_rel) && vol_err_rel < mopt.vol_err_rel_threshold;
    const bool res_ok = fin(rms_res) && rms_res < mopt.rms_res_threshold;
    const bool mass_ok =
        !mopt.mass_constraint || (fin(mass_err_rel) && mass_err_rel < mopt.mass_tolerance_relative);
    bool condA, condB;
    if (frozen) {
      // main_cuda.cu:32508-32522: no phi residual, no rms_dphi
      condA = fin(rms_dY) && rms_dY < mopt.rms_dY_threshold && vol_ok && mass_ok;
      condB = fin(energy_diff_rel) && energy_diff_rel < mopt.energy_diff_rel_threshold && fin(rms_dY) &&
              rms_dY < 5.0 * mopt.rms_dY_threshold && vol_ok && mass_ok;
    } else {
      condA = fin(rms_dphi) && fin(rms_dY) && rms_dphi < mopt.rms_dphi_threshold &&
              rms_dY < mopt.rms_dY_threshold && res_ok && vol_ok && mass_ok;
      condB = fin(energy_diff_rel) && energy_diff_rel < mopt.energy_diff_rel_threshold && fin(rms_dphi) &&
              fin(rms_dY) && rms_dphi < 5.0 * mopt.rms_dphi_threshold && rms_dY < 5.0 * mopt.rms_dY_threshold &&
              res_ok && vol_ok && mass_ok;
    }
    const bool t_ok = (double)m_iter * mopt.dt + 1e-14 >= mopt.min_pseudo_time;
    m_conv = ((condA || condB) && t_ok) ? m_conv + 1 : 0;
    if (m_conv >= mopt.convergence_steps) m_converged = true;
    e_grad = grad_mean;
    e_dw = dw_sum / (double)R;
    e_el = gel_sum / (double)R;
    e_chem = chem_sum / (double)R;
    eng_pending = false;
    ++steps_done;
    if (st) {
      st->iteration = m_iter;
      st->converged = m_converged ? 1 : 0;
      st->convergence_count = m_conv;
      st->elastic_iterations = el ? last_iters : 0;
      st->mean_h = mean_h;
      st->V0 = m_V0;
      st->lambda_vol = m_lambda;
      st->F_surf_hat = F_surf;
      st->F_el_hat = F_el;
      st->interface_sum = grad_mean * (double)R + dw_sum;
      st->el_core_sum = core_sum;
      st->rms_res = rms_res;
      st->rms_dphi = rms_dphi;
      st->vol_err_rel = vol_err_rel;
      st->lambda_raw = lam_raw;
      st->lambda_kkt = lam_kkt;
      st->F_chem_excess_hat = F_chem_excess;
      st->F_total_excess_hat = F_total_excess;
      st->F_chem_CNT_hat = F_chem_cnt;
      st->F_total_CNT_hat = F_total_cnt;
      st->rms_dY = rms_dY;
      st->energy_diff_rel = energy_diff_rel;
      st->mass_target = m_mass_target;
      st->mass_current = mass_cur;
      st->mass_err_rel = mass_err_rel;
      st->mass_lambda = mopt.mass_constraint ? m_mass.lam : NAN;
      st->mass_iterations = mopt.mass_constraint ? m_mass.iters : 0;
      st->mass_residual = mopt.mass_constraint ? m_mass.r : NAN;
      st->mass_derivative = mopt.mass_constraint ? m_mass.deriv : NAN;
      st->mass_Ystar_min = mopt.mass_constraint ? m_mass.ymin : NAN;
      st->mass_Ystar_max = mopt.mass_constraint ? m_mass.ymax : NAN;
      st->mass_lambda_lower = mopt.mass_constraint ? m_mass.lo : NAN;
      st->mass_lambda_upper = mopt.mass_constraint ? m_mass.hi : NAN;
      st->mass_newton_steps = mopt.mass_constraint ? m_mass.newton : 0;
      st->mass_bisection_steps = mopt.mass_constraint ? m_mass.bisect : 0;
    }
    if (!std::isfinite(mean_h) || !std::isfinite(m_lambda) || !std::isfinite(rms_dY))
      return fail(PF_E_NONFINITE, "minimize produced a non-finite phi or Y");
    state_ok = true;
    return PF_OK;
  }

  // ---------------------------------------------------------------- GP eta
  double* eta_buf = nullptr;   // eta, phi, x_B(alpha), tr sigma, result: 5 x R
  int eta_update(const pf_eta_params& q, const double* eta, const double* ph, const double* xa, const double* trs,
                 double* out) override {
    int rc;
    if (!eta_buf && (rc = dalloc(&eta_buf, 5 * (size_t)R))) return rc;
    double* b[5];
    for (int i = 0; i < 5; ++i) b[i] = eta_buf + (size_t)i * R;
    const size_t bytes = R * sizeof(double);
    PF_HIP(hipMemcpyAsync(b[0], eta, bytes, hipMemcpyHostToDevice, stream));
    PF_HIP(hipMemcpyAsync(b[1], ph, bytes, hipMemcpyHostToDevice, stream));
    PF_HIP(hipMemcpyAsync(b[2], xa, bytes, hipMemcpyHostToDevice, stream));
    if (trs) PF_HIP(hipMemcpyAsync(b[3], trs, bytes, hipMemcpyHostToDevice, stream));
    EtaConst c{};
    c.xB_gp = clamp_eps(q.gp_xB_fixed, 1.0e-12);
    c.delta_g0 = q.gp_delta_g0;
    c.W_eta = q.gp_W_eta;
    c.eps_gp = q.gp_eps_iso;
    c.el_scale = q.gp_elastic_derivative_scale;
    c.nu_A = q.gp_reaction_nu_A;
    c.nu_B = q.gp_reaction_nu_B;
    c.mu_ref_raw = q.gp_mu_reference_raw;
    c.Ldt = q.gp_L_eta * prm.dt;
    c.raw_drive = q.gp_raw_reaction_drive_only;
    c.raw_units = q.gp_raw_reaction_drive_use_raw_units_debug;
    c.elastic = q.gp_elastic;
    typename OpEtaRhs<T>::Args oa{P, c, b[0], b[1], b[2], trs ? b[3] : nullptr};
    launch_zpass<T, 0, 1, OpEtaRhs<T>>(Nz, stream, zargs(0, 0, 0, 1), oa);
    ypass(false, 0, 1);
    xpass(false, 0, 1, Sc);
    hipLaunchKernelGGL((kop_phi<T>), dim3(grid_for(nSc)), dim3(NT), 0, stream, kd(), nSc, sc(0),
                       c.Ldt * q.gp_kappa_eta);
    xpass(true, 0, 1, Sc);
    ypass(true, 0, 1);
    typename OpEtaFinish<T>::Args fa{b[4], T(1.0 / (double)R)};
    launch_zpass<T, 1, 0, OpEtaFinish<T>>(Nz, stream, zargs(0, 1, 0, 0), fa);
    PF_HIP(hipMemcpyAsync(out, b[4], bytes, hipMemcpyDeviceToHost, stream));
    PF_HIP(hipStreamSynchronize(stream));
    return PF_OK;
  }

  // ---------------------------------------------------------------- queries
  // The 6 stress components of the step's elastic solution at the
  // start-of-step fields (the reference's sigma arrays) are written by the
  // phi right-hand-side pass while enabled: 6 R values of the transform
  // precision, allocated on the first enable and kept until the handle goes.
  int keep_stress(int enable) override {
    if (enable && !sig6) {
      int rc = dalloc(&sig6, (size_t)6 * R);
      if (rc) return rc;
    }
    keep_sig6 = enable != 0;
    if (!keep_sig6) sig6_valid = false;
    return PF_OK;
  }

  int get_field(int which, double* out) override {
    PF_HIP(hipStreamSynchronize(stream));
    const double* src = nullptr;
    switch (which) {
      case PF_PHI: src = phi[cur]; break;
      case PF_XB: src = xB; break;
      case PF_Y: src = Y; break;
      case PF_DYDT: src = dYdt; break;
      case PF_XBTOT: {
        std::vector<double> ph(R), xb(R);
        PF_HIP(hipMemcpy(ph.data(), phi[cur], R * sizeof(double), hipMemcpyDeviceToHost));
        PF_HIP(hipMemcpy(xb.data(), xB, R * sizeof(double), hipMemcpyDeviceToHost));
        for (long i = 0; i < R; ++i) {
          const double h = h_of_phi(ph[i]);
          out[i] = (1.0 - h) * xb[i] + P.v_B * h;
        }
        return PF_OK;
      }
      case PF_TRSIGMA: {
        std::vector<T> t(R);
        PF_HIP(hipMemcpy(t.data(), trs, R * sizeof(T), hipMemcpyDeviceToHost));
        for (long i = 0; i < R; ++i) out[i] = double(t[i]);
        return PF_OK;
      }
      case PF_SIGMA_XX: case PF_SIGMA_YY: case PF_SIGMA_ZZ:
      case PF_SIGMA_YZ: case PF_SIGMA_XZ: case PF_SIGMA_XY: {
        if (!sig6_valid)
          return fail(PF_E_ARG, "stress components are kept only for steps taken after pf_keep_stress(h, 1)");
        std::vector<T> t(R);
        PF_HIP(hipMemcpy(t.data(), sig6 + (size_t)(which - PF_SIGMA_XX) * R, R * sizeof(T),
                         hipMemcpyDeviceToHost));
        for (long i = 0; i < R; ++i) out[i] = double(t[i]);
        return PF_OK;
      }
      default:
        return fail(PF_E_ARG, "unknown field id");
    }
    PF_HIP(hipMemcpy(out, src, R * sizeof(double), hipMemcpyDeviceToHost));
    return PF_OK;
  }

  int scalars(pf_scalars* s) override {
    std::memset(s, 0, sizeof(*s));
    if (egrad_pending) {
      // phi^n of the last step is in the buffer the step did not end in
      const int g = grid_for(R);
      hipLaunchKernelGGL(grad_energy_kernel, dim3(g), dim3(NT), 0, stream, (const double*)phi[1 - cur], Nx, Ny,
                         Nz, prm.dx, prm.dy, prm.dz, prm.kappa_phi, partials);
      finalize(g, 1, RedKinds{{RED_SUM, 0, 0, 0}}, dscal + 8, 1.0 / (double)R);
      PF_HIP(hipMemcpyAsync(hscal + 8, dscal + 8, sizeof(double), hipMemcpyDeviceToHost, stream));
      PF_HIP(hipStreamSynchronize(stream));
      e_grad = hscal[8];
      egrad_pending = false;
    }
    if (eng_pending) {
      hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(NT), 0, stream, (const double*)epartials, eng_g, 3,
                         RedKinds{{RED_SUM, RED_SUM, RED_SUM, 0}}, dscal + 9, 1.0 / (double)R);
      PF_HIP(hipMemcpyAsync(hscal + 9, dscal + 9, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
      PF_HIP(hipStreamSynchronize(stream));
      e_el = hscal[9];
      e_dw = hscal[10];
      e_chem = hscal[11];
      eng_pending = false;
    }
    double m, vf, nin, nif;
    int rc = mass_vf(&m, &vf, &nin, &nif);
    if (rc) return rc;
    s->n_in = nin;
    s->n_if = nif;
    s->mass_xBtot = m;
    s->mean_xBtot = m / (double)R;
    s->volume_fraction = vf;
    s->f_grad = e_grad;
    s->f_dw = e_dw;
    s->f_elastic = e_el;
    s->f_chem = e_chem;
    s->f_total = e_grad + e_dw + e_el + e_chem;
    s->mean_DY = mean_DY;
    s->last_elastic_residual = last_res;
    s->zero_mode_lambda = zm_lambda;
    s->last_elastic_iterations = last_iters;
    s->zero_mode_iterations = zm_iters;
    s->steps_done = steps_done;
    s->elastic_iterations_total = elastic_total;
    s->zero_mode_target_mass = prm.zero_mode ? target_mass : NAN;
    s->kz_stride = AzS;
    s->zpass_rows = zpass_rows_of<T, 6, 6, OpHijInc<T>>(zb_sel, Nz);
    s->ypass_cols = cpass_cols_of<T>(ty_sel, Ny);
    s->xpass_cols = cpass_cols_of<T>(tx_sel, Nx);
    return PF_OK;
  }

  // The external warm state is 3 components over the retained modes,
  // [Ax][Ay][Az] with no kz padding, whatever stride the spectra use; the
  // plain solver keeps its state at k-space slots and is converted here.
  long slot_of(long j) const { return (j / az.A) * AzS + j % az.A; }
  void state_to_slots(const double* in, std::vector<C>* t) const {
    t->assign(3 * nSc, C{T(0), T(0)});
    for (int i = 0; i < 3; ++i)
      for (long j = 0; j < nScA; ++j) {
        const long q = (long)i * nScA + j;
        (*t)[(size_t)i * nSc + slot_of(j)] = C{T(in[2 * q]), T(in[2 * q + 1])};
      }
  }
  void slots_to_state(const std::vector<C>& t, cx<double>* out) const {
    for (int i = 0; i < 3; ++i)
      for (long j = 0; j < nScA; ++j) {
        const C v = t[(size_t)i * nSc + slot_of(j)];
        out[(size_t)i * nScA + j] = cx<double>{double(v.re), double(v.im)};
      }
  }
  long long elastic_state_size() override { return nScA; }
  int get_elastic_state(double* out) override {
    PF_HIP(hipStreamSynchronize(stream));
    if (kMixed || extended) {
      PF_HIP(hipMemcpy(out, U_last(), 3 * nScA * sizeof(cx<double>), hipMemcpyDeviceToHost));
      return PF_OK;
    }
    std::vector<C> t(3 * nSc);
    PF_HIP(hipMemcpy(t.data(), u, 3 * nSc * sizeof(C), hipMemcpyDeviceToHost));
    slots_to_state(t, reinterpret_cast<cx<double>*>(out));
    return PF_OK;
  }
  int set_elastic_state(const double* in) override {
    if (kMixed || extended) {
      PF_HIP(hipMemcpy(Ud, in, 3 * nScA * sizeof(cx<double>), hipMemcpyHostToDevice));
      if (extended) {
        PF_HIP(hipMemcpy(hist_slot(0), Ud, 3 * nScA * sizeof(cx<double>), hipMemcpyDeviceToDevice));
        hist_head = 1;
        hist_count = 1;
      }
      warm_valid = true;
      return PF_OK;
    }
    std::vector<C> t;
    state_to_slots(in, &t);
    PF_HIP(hipMemcpy(u, t.data(), 3 * nSc * sizeof(C), hipMemcpyHostToDevice));
    warm_valid = true;
    return PF_OK;
  }

  // ---------------------------------------------------------------- checkpoint / restart
  // The warm state after the last accepted step, FP64, retained modes (3 x nScA).
  int warm_state_host(std::vector<cx<double>>& v) {
    v.resize(3 * nScA);
    if (state_fp64()) {
      PF_HIP(hipMemcpy(v.data(), U_last(), 3 * nScA * sizeof(cx<double>), hipMemcpyDeviceToHost));
    } else {
      std::vector<C> t(3 * nSc);
      PF_HIP(hipMemcpy(t.data(), u, 3 * nSc * sizeof(C), hipMemcpyDeviceToHost));
      slots_to_state(t, v.data());
    }
    return PF_OK;
  }
  // retained-mode index <-> full half-spectrum index [Nx][Ny][Nz/2+1]
  template <typename F>
  void for_retained(F f) const {
    const long kfull = (long)Nx * Ny * (Nz / 2 + 1);
    (void)kfull;
    for (int cx_ = 0; cx_ < ax.A; ++cx_) {
      const int ix = cx_ < ax.a ? cx_ : cx_ + Nx - ax.A;
      for (int cy = 0; cy < ay.A; ++cy) {
        const int iy = cy < ay.a ? cy : cy + Ny - ay.A;
        for (int cz = 0; cz < az.A; ++cz)
          f(((long)cx_ * ay.A + cy) * az.A + cz, ((long)ix * Ny + iy) * (Nz / 2 + 1) + cz);
      }
    }
  }

  int checkpoint_write(const char* path, const pf_checkpoint_provenance* q) override {
    if (!prm.zero_mode)
      return fail(PF_E_ARG, "PF checkpoint/restart provenance is available only with PF_CONSERVED_Y_ZERO_MODE_V1");
    if (minimizing) return fail(PF_E_ARG, "pf_checkpoint_write: the handle is in minimize mode");
    pf_checkpoint_provenance def;
    if (!q) {
      pf_checkpoint_provenance_default(&prm, &def);
      q = &def;
    }
    PF_HIP(hipStreamSynchronize(stream));
    ckpt::Checkpoint c;
    c.accepted_step = (std::uint64_t)steps_done;
    c.nx = Nx;
    c.ny = Ny;
    c.nz = Nz;
    c.dt_code = prm.dt;
    c.temperature_K = prm.temperature_C + 273.15;
    c.prov = provenance_from_abi(q);
    c.target_mass_code = target_mass;
    c.last_lambda = zm_lambda;
    c.last_residual_code = zm_res;
    c.last_derivative_code = zm_deriv;
    c.last_iterations = (std::uint64_t)zm_iters;
    c.accepted_zero_mode_steps = zm_accepted;
    c.phi.resize(R);
    c.Y.resize(R);
    c.xB.resize(R);
    c.dY_dt_prev.resize(R);
    PF_HIP(hipMemcpy(c.phi.data(), phi[cur], R * sizeof(double), hipMemcpyDeviceToHost));
    PF_HIP(hipMemcpy(c.Y.data(), Y, R * sizeof(double), hipMemcpyDeviceToHost));
    PF_HIP(hipMemcpy(c.xB.data(), xB, R * sizeof(double), hipMemcpyDeviceToHost));
    PF_HIP(hipMemcpy(c.dY_dt_prev.data(), dYdt, R * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<cx<double>> w;
    if (elastic_state_required(&prm)) {
      // main_cuda.cu:33509-33533: the converged state of the last accepted step
      if (!warm_valid || steps_done < 1)
        return fail(PF_E_ARG, "refusing to checkpoint an invalid elastic warm state at step " +
                                  std::to_string(steps_done));
      c.elastic_present = true;
      c.elastic_source_field_step = (std::uint64_t)steps_done - 1;
      c.elastic_last_iterations = (std::uint64_t)last_iters;
      c.elastic_last_relative_residual = last_res;
      int rc = warm_state_host(w);
      if (rc) return rc;
      const long kfull = (long)Nx * Ny * (Nz / 2 + 1);
      c.displacement_k.assign(6 * kfull, 0.0f);
      for (int i = 0; i < 3; ++i) {
        float* d = c.displacement_k.data() + 2 * kfull * i;
        const cx<double>* src = w.data() + (size_t)i * nScA;
        for_retained([&](long cc, long fi) {
          d[2 * fi] = (float)src[cc].re;
          d[2 * fi + 1] = (float)src[cc].im;
        });
      }
    }
    std::string err;
    if (!ckpt::write(path, c, &err))
      return fail(PF_E_IO, "PF zero-mode checkpoint write failed at step " + std::to_string(steps_done) + ": " + err);
    const bool phi_block = kMixed && phi_banded;
    if (c.elastic_present || phi_block) {
      // exact-resume companion: the FP64 warm state and the extrapolation
      // history, and in mixed precision the unclamped phi* with its spectrum
      ckpt::ExactState x;
      x.bound_checksum = c.payload_checksum;
      x.accepted_step = c.accepted_step;
      x.precision = kMixed ? PF_MIXED : PF_FP64;
      x.n = (std::uint64_t)nScA;   // retained modes: independent of the tuned kz stride
      const int nh = extended ? hist_count : 0;
      x.nstates = c.elastic_present ? 1 + nh : 0;
      x.states.resize((size_t)x.nstates * 6 * nScA);
      if (c.elastic_present) {
        std::memcpy(x.states.data(), w.data(), 3 * nScA * sizeof(cx<double>));
        for (int k = 0; k < nh; ++k)
          PF_HIP(hipMemcpy(x.states.data() + (size_t)(1 + k) * 6 * nScA, hist_slot((hist_head + 2 - k) % 3),
                           3 * nScA * sizeof(cx<double>), hipMemcpyDeviceToHost));
      }
      x.plan[0] = zb_sel;
      x.plan[1] = ty_sel;
      x.plan[2] = tx_sel;
      if (phi_block) {
        x.phi_star.resize(R);
        x.phi_spec.resize(2 * (size_t)nScA);
        PF_HIP(hipMemcpy(x.phi_star.data(), phis, R * sizeof(double), hipMemcpyDeviceToHost));
        PF_HIP(hipMemcpy(x.phi_spec.data(), Phik, nScA * sizeof(cx<double>), hipMemcpyDeviceToHost));
      }
      if (!ckpt::write_exact(std::string(path) + ".state", x, &err)) return fail(PF_E_IO, err);
    }
    return PF_OK;
  }

  int checkpoint_read(const char* path, const pf_checkpoint_provenance* q) override {
    if (!prm.zero_mode)
      return fail(PF_E_ARG, "PF checkpoint/restart provenance is available only with PF_CONSERVED_Y_ZERO_MODE_V1");
    pf_checkpoint_provenance def;
    if (!q) {
      pf_checkpoint_provenance_default(&prm, &def);
      q = &def;
    }
    ckpt::Checkpoint c;
    std::string err;
    if (!ckpt::read(path, provenance_from_abi(q), &c, &err))
      return fail(PF_E_IO, "PF zero-mode restart rejected: " + err);
    // main_cuda.cu:27350-27358 (exact comparisons, as there)
    if (c.nx != Nx || c.ny != Ny || c.nz != Nz || c.dt_code != prm.dt || c.temperature_K != prm.temperature_C + 273.15)
      return fail(PF_E_IO, "PF zero-mode restart grid/dt/temperature mismatch");
    const bool el = elastic_state_required(&prm);
    if (el && !c.elastic_present) return fail(PF_E_IO, "accelerated elastic restart lacks the registered V4 warm state");
    // the exact-resume companion is read (and a broken one refused) before any
    // device state changes
    ckpt::ExactState x;
    bool have_x = false;
    if (el || kMixed) {
      have_x = ckpt::read_exact(std::string(path) + ".state", &x, &err);
      if (!have_x && !err.empty()) return fail(PF_E_IO, err);
    }
    // the companion belongs to this file and this handle
    const bool x_ok = have_x && x.bound_checksum == c.payload_checksum && x.accepted_step == c.accepted_step &&
                      x.n == (std::uint64_t)nScA && x.precision == (kMixed ? PF_MIXED : PF_FP64);
    if (x_ok && x.plan[0] >= 0) {   // the run continues with the tiles it was computed with
      zb_sel = x.plan[0];
      ty_sel = x.plan[1];
      tx_sel = x.plan[2];
    }
    PF_HIP(hipStreamSynchronize(stream));
    // from here on a rejection leaves the handle without a valid state
    state_ok = false;
    cur = 0;
    PF_HIP(hipMemcpy(phi[0], c.phi.data(), R * sizeof(double), hipMemcpyHostToDevice));
    PF_HIP(hipMemcpy(Y, c.Y.data(), R * sizeof(double), hipMemcpyHostToDevice));
    PF_HIP(hipMemcpy(xB, c.xB.data(), R * sizeof(double), hipMemcpyHostToDevice));
    PF_HIP(hipMemcpy(dYdt, c.dY_dt_prev.data(), R * sizeof(double), hipMemcpyHostToDevice));
    // main_cuda.cu:28081-28105: the field ledger must match the frozen target
    double m, vf;
    int rc = mass_vf(&m, &vf);
    if (rc) return rc;
    const double tolm = 1.0e-12 * std::fmax(std::fabs(c.target_mass_code), 1.0);
    if (!std::isfinite(m) || std::fabs(m - c.target_mass_code) > tolm) {
      char buf[256];
      std::snprintf(buf, sizeof buf,
                    "checkpoint field ledger does not match frozen zero-mode target: current=%.17e target=%.17e "
                    "residual=%.17e",
                    m, c.target_mass_code, m - c.target_mass_code);
      return fail(PF_E_ZERO_MODE, buf);
    }
    target_mass = c.target_mass_code;
    zm_lambda = c.last_lambda;
    zm_res = c.last_residual_code;
    zm_deriv = c.last_derivative_code;
    zm_iters = (int)c.last_iterations;
    zm_accepted = c.accepted_zero_mode_steps;
    steps_done = (long long)c.accepted_step;
    y_banded = c.accepted_step > 0;   // the state of an accepted step
    elastic_total = 0;
    egrad_pending = false;
    eng_pending = false;
    sig6_valid = false;
    minimizing = false;
    warm_valid = false;
    hist_count = 0;
    hist_head = 0;
    // mixed precision: phi* and its spectrum from the companion continue the
    // increment form bit for bit; without them the next update starts it anew
    phi_banded = false;
    if (kMixed && x_ok && x.phi_star.size() == (size_t)R && x.phi_spec.size() == 2 * (size_t)nScA) {
      PF_HIP(hipMemcpy(phis, x.phi_star.data(), R * sizeof(double), hipMemcpyHostToDevice));
      PF_HIP(hipMemcpy(Phik, x.phi_spec.data(), nScA * sizeof(cx<double>), hipMemcpyHostToDevice));
      phi_banded = true;
    }
    if (!el) {
      state_ok = true;
      return PF_OK;
    }
    last_iters = (int)c.elastic_last_iterations;
    last_res = c.elastic_last_relative_residual;
    std::vector<cx<double>> cur_state(3 * nScA);
    int nh = 0;
    if (x_ok && x.nstates >= 1) {
      std::memcpy(cur_state.data(), x.states.data(), 3 * nScA * sizeof(cx<double>));
      nh = extended ? std::min(x.nstates - 1, 3) : 0;
      for (int k = 0; k < nh; ++k)   // newest first -> ring slots nh-1, ..., 0
        PF_HIP(hipMemcpy(hist_slot(nh - 1 - k), x.states.data() + (size_t)(1 + k) * 6 * nScA,
                         3 * nScA * sizeof(cx<double>), hipMemcpyHostToDevice));
    } else {
      // the reference's restart: the FP32 warm state of the V4 file
      std::fill(cur_state.begin(), cur_state.end(), cx<double>{0.0, 0.0});
      const long kfull = (long)Nx * Ny * (Nz / 2 + 1);
      for (int i = 0; i < 3; ++i) {
        const float* s = c.displacement_k.data() + 2 * kfull * i;
        cx<double>* d = cur_state.data() + (size_t)i * nScA;
        for_retained([&](long cc, long fi) { d[cc] = cx<double>{(double)s[2 * fi], (double)s[2 * fi + 1]}; });
      }
      if (extended) {
        PF_HIP(hipMemcpy(hist_slot(0), cur_state.data(), 3 * nScA * sizeof(cx<double>), hipMemcpyHostToDevice));
        nh = 1;
      }
    }
    if (state_fp64()) {
      PF_HIP(hipMemcpy(Ud, cur_state.data(), 3 * nScA * sizeof(cx<double>), hipMemcpyHostToDevice));
    } else {
      std::vector<C> t;
      state_to_slots(reinterpret_cast<const double*>(cur_state.data()), &t);
      PF_HIP(hipMemcpy(u, t.data(), 3 * nSc * sizeof(C), hipMemcpyHostToDevice));
    }
    hist_count = nh;
    hist_head = nh % 3;
    warm_valid = true;
    state_ok = true;
    return PF_OK;
  }

  // ---------------------------------------------------------------- transform test
  int fft_forward(const double* in, double* out_half) override {
    double* f = phi[0];
    PF_HIP(hipMemcpy(f, in, R * sizeof(double), hipMemcpyHostToDevice));
    typename OpLoadReal<T>::Args oa{f};
    launch_zpass<T, 0, 1, OpLoadReal<T>>(Nz, stream, zargs(0, 0, 0, 1), oa);
    ypass(false, 0, 1);
    xpass(false, 0, 1, Sc);
    PF_HIP(hipStreamSynchronize(stream));
    std::vector<C> c(nSc);
    PF_HIP(hipMemcpy(c.data(), Sc, nSc * sizeof(C), hipMemcpyDeviceToHost));
    const int nzc = Nz / 2 + 1;
    std::memset(out_half, 0, sizeof(double) * 2 * (size_t)Nx * Ny * nzc);
    for (int cx_ = 0; cx_ < ax.A; ++cx_) {
      const int ix = cx_ < ax.a ? cx_ : cx_ + Nx - ax.A;
      for (int cy = 0; cy < ay.A; ++cy) {
        const int iy = cy < ay.a ? cy : cy + Ny - ay.A;
        for (int cz = 0; cz < az.A; ++cz) {
          const C v = c[((size_t)cx_ * ay.A + cy) * AzS + cz];
          const size_t o = ((size_t)ix * Ny + iy) * nzc + cz;
          out_half[2 * o] = double(v.re);
          out_half[2 * o + 1] = double(v.im);
        }
      }
    }
    return PF_OK;
  }
  int fft_inverse(const double* in_half, double* out) override {
    const int nzc = Nz / 2 + 1;
    std::vector<C> c(nSc);
    for (int cx_ = 0; cx_ < ax.A; ++cx_) {
      const int ix = cx_ < ax.a ? cx_ : cx_ + Nx - ax.A;
      for (int cy = 0; cy < ay.A; ++cy) {
        const int iy = cy < ay.a ? cy : cy + Ny - ay.A;
        for (int cz = 0; cz < az.A; ++cz) {
          const size_t o = ((size_t)ix * Ny + iy) * nzc + cz;
          c[((size_t)cx_ * ay.A + cy) * AzS + cz] = C{T(in_half[2 * o]), T(in_half[2 * o + 1])};
        }
      }
    }
    PF_HIP(hipMemcpy(Sc, c.data(), nSc * sizeof(C), hipMemcpyHostToDevice));
    xpass(true, 0, 1, Sc);
    ypass(true, 0, 1);
    double* f = phi[0];
    typename OpStoreReal<T>::Args oa{f};
    launch_zpass<T, 1, 0, OpStoreReal<T>>(Nz, stream, zargs(0, 1, 0, 0), oa);
    PF_HIP(hipStreamSynchronize(stream));
    PF_HIP(hipMemcpy(out, f, R * sizeof(double), hipMemcpyDeviceToHost));
    return PF_OK;
  }

  int sync() override {
    PF_HIP(hipStreamSynchronize(stream));
    return PF_OK;
  }

  // ---------------------------------------------------------------- fused batch
  // The phases of a phi-only minimize iteration with the accelerated solver,
  // recorded by the fused batch runner (pf_batch_minimize_fused).  Between
  // phases the runner mirrors every handle's device scalars into its host
  // mirror (mb_gather / mb_scatter) with one copy for all handles.
  int mb_check() override {
    if (!minimizing) return fail(PF_E_ARG, "pf_minimize_begin has not been called on this handle");
    if (mopt.full_model) return fail(PF_E_ARG, "fused batches run phi-only minimize");
    if (P.elastic && !extended) return fail(PF_E_ARG, "fused batches need the accelerated elastic solver");
    if (cold_hooks()) return fail(PF_E_ARG, "fused batches run no cold-start hooks");
    if (zpass_use_reg()) return fail(PF_E_ARG, "fused batches use the LDS z pass");
    return PF_OK;
  }
  int mb_spec() override { return std::max(aa_spec, 1); }
  int mb_phase_a(int spec) override {
    if (int rs = require_state("pf_minimize_iterate")) return rs;
    state_ok = false;
    aa_spec = std::max(spec, 1);
    return min_phase_a();
  }
  int mb_solve_more(int* more) override {
    *more = 0;
    if (!(P.elastic && extended) || !accel_more()) return PF_OK;
    accel_iters(2);
    *more = 1;
    return PF_OK;
  }
  int mb_phase_b() override { return min_phase_b(); }
  int mb_phase_c(pf_minimize_stats* st) override { return min_phase_c(st); }
  int mb_converged() override { return m_converged ? 1 : 0; }
  int sb_check() override {
    if (P.elastic && !extended) return fail(PF_E_ARG, "fused batches need the accelerated elastic solver");
    if (cold_hooks()) return fail(PF_E_ARG, "fused batches run no cold-start hooks");
    if (zpass_use_reg()) return fail(PF_E_ARG, "fused batches use the LDS z pass");
    return require_state("pf_batch_step_fused");
  }
  int sb_phase_a(int spec) override {
    aa_spec = std::max(spec, 1);
    return step_phase_a();
  }
  int sb_zspec() override { return std::max(zm_spec, 1); }
  int sb_phase_b(int zspec) override {
    zm_spec = std::max(zspec, 1);
    return step_phase_b();
  }
  int sb_zm_more(int* more) override {
    *more = 0;
    if (!prm.zero_mode || !zm_more()) return PF_OK;
    zm_steps(4);
    *more = 1;
    return PF_OK;
  }
  int sb_phase_c() override { return step_phase_c(); }
  int mb_nscal() override { return NSCAL; }
  void mb_gather(double* dst) override { launch<KGather>(stream, dim3(1), {dscal, dst, NSCAL}); }
  void mb_scatter(const double* src) override { std::memcpy(hscal, src, NSCAL * sizeof(double)); }

  // One launch of kernel `id` with the current buffers; returns its bytes.
  double launch_one(int id) {
    const double cs = (double)sizeof(C);
    const T invN = T(1.0 / (double)R);
    const KDims d = kd();
    const int gk = grid_for(nSc);
    switch (id) {
      case PF_K_ZPASS_HIJ: {
        // the form the solve repeats: increments through the linear part of
        // the map in the mixed and accelerated solvers (h(phi) from the
        // per-solve cache in the accelerated one, phi in the mixed one), the
        // full map (phi and x_B) in the reference iteration
        const int inc = (kMixed || extended) ? 1 : 0;
        if (extended) {
          typename OpHijInc<T>::Args oa{P, hcache, invN};
          launch_zpass_b<T, 6, 6, OpHijInc<T>>(zb_sel, Nz, stream, zargs(0, 6, 0, 6), oa);
        } else {
          typename OpHij<T>::Args oa{P, phi[cur], xB, invN, inc, nullptr};
          launch_zpass_b<T, 6, 6, OpHij<T>>(zb_sel, Nz, stream, zargs(0, 6, 0, 6), oa);
        }
        return 12.0 * nSzA * cs + (inc ? (extended ? (double)sizeof(T) : 8.0) : 16.0) * R;
      }
      case PF_K_YPASS_FWD6: ypass(false, 0, 6); return 6.0 * (nSzA + nSyA) * cs;
      case PF_K_YPASS_INV6: ypass(true, 0, 6); return 6.0 * (nSzA + nSyA) * cs;
      case PF_K_XPASS_FWD6: xpass(false, 0, 6, Sc); return 6.0 * (nSyA + nScA) * cs;
      case PF_K_XPASS_INV6: xpass(true, 0, 6, Sc); return 6.0 * (nSyA + nScA) * cs;
      case PF_K_GREEN: {
        if (extended) {
          // the solve's form: control block with a full history (set by time_kernel)
          const int m = prm.elastic_anderson_depth;
          GreenAAArgs<T> ga{};
          ga.d = d;
          ga.n = nSc;
          for (int i = 0; i < 6; ++i) ga.h[i] = sc(i);
          for (int i = 0; i < 3; ++i) {
            ga.U[i] = Ud + (size_t)i * nScA;
            ga.g[i] = Gd + (size_t)i * nScA;
            ga.dinc[i] = dinc + (size_t)i * nScA;
            for (int j = 0; j < AA_MAX; ++j) {
              ga.dg[j][i] = dg_slot(j, i);
              ga.du[j][i] = du_slot(j, i);
            }
          }
          ga.mode = 1;
          ga.partials = partials;
          for (int i = 0; i < 21; ++i) ga.C[i] = P.C[i];
          ga.ctl = dscal + AA_SLOT;
          ga.m = m;
          hipLaunchKernelGGL((kop_green_aa<T>), dim3(gk), dim3(NT), 0, stream, ga);
          // h read; U, g read and written; with a history: du[slot] read and
          // written, dg[slot] written, the other m-1 dg slots read
          return (m > 0 ? 12.0 + 3.0 * m : 9.0) * nScA * cs + 12.0 * nScA * 16.0;
        }
        GreenArgs<T> ga{};
        ga.d = d;
        ga.n = nSc;
        for (int i = 0; i < 6; ++i) {
          ga.h[i] = sc(i);
          ga.eps[i] = sc(i);
        }
        for (int i = 0; i < 3; ++i) {
          ga.u_old[i] = uu(u, i);
          ga.u_new[i] = uu(un, i);
        }
        for (int i = 0; i < 21; ++i) ga.C[i] = P.C[i];
        ga.have_old = 1;
        ga.partials = partials;
        hipLaunchKernelGGL((kop_green<T>), dim3(gk), dim3(NT), 0, stream, ga);
        return (12.0 + 6.0) * nScA * cs;
      }
      case PF_K_AA_UPDATE: {
        const int m = prm.elastic_anderson_depth;
        if (!extended || m == 0) return 0.0;
        AAUpdateArgs<T> ua{};
        ua.d = d;
        ua.n = nSc;
        for (int i = 0; i < 3; ++i) {
          ua.g[i] = Gd + (size_t)i * nScA;
          ua.U[i] = Ud + (size_t)i * nScA;
          ua.dinc[i] = dinc + (size_t)i * nScA;
          for (int j = 0; j < AA_MAX; ++j) {
            ua.du[j][i] = du_slot(j, i);
            ua.dg[j][i] = dg_slot(j, i);
          }
        }
        for (int i = 0; i < 6; ++i) ua.eps[i] = sc(i);
        ua.ctl = dscal + AA_SLOT;
        ua.m = m;
        hipLaunchKernelGGL((kop_aa_update<T>), dim3(gmodes()), dim3(NT), 0, stream, ua);
        // g read; the m history slots read; du[slot] and the 6 strains written
        return (3.0 * m + 9.0) * nScA * cs + 3.0 * nScA * 16.0;
      }
      case PF_K_ZPASS_PHIRHS: {
        // the form the step runs: in mixed precision phi* is read as well
        typename OpPhiRhs<T, true>::Args oa{P, phi[cur], xB, trs, invN, nullptr, 0, kMixed ? phis : nullptr};
        launch_zpass<T, 6, 1, OpPhiRhs<T, true>>(Nz, stream, zargs(0, 6, 0, 1), oa);
        return 7.0 * nSzA * cs + (kMixed ? 3.0 : 2.0) * R * 8.0 + R * (double)sizeof(T);
      }
      case PF_K_ZPASS_FLUX: {
        const double* ph = phi[cur];
        // the form the step runs: (J, s) in mixed precision with the zero mode
        // after the first step, (J, Y^n, s) otherwise
        const bool inc = kMixed && prm.zero_mode;
        typename OpFlux<T>::Args oa{P, ph, xB, ph, phi[1 - cur], Y, dYdt, invN, 1, 0, 0, 1.0, inc ? 1 : 0};
        if constexpr (kMixed) {
          if (inc) {
            launch_zpass<T, 3, 4, OpFlux<T>>(Nz, stream, zargs(0, 3, 0, 4), oa);
            return 7.0 * nSzA * cs + 6.0 * R * 8.0;
          }
        }
        launch_zpass<T, 3, 5, OpFlux<T>>(Nz, stream, zargs(0, 3, 0, 5), oa);
        return 8.0 * nSzA * cs + 6.0 * R * 8.0;
      }
      default:
        return 0.0;
    }
  }

  // ------------------------------------------------------------ launch probe
  // HIP events around the accelerated solve's launches of the kernel
  // classes in probe_mask (bit PF_K_*), on the handle's stream, for the
  // classes' in-step durations and bytes.  A launch of an iteration after
  // the one the device stopped at does nothing (its gate is off): it counts
  // its time and no bytes.  Launches recorded for a fused batch are not
  // probed.
  struct ProbeRec {
    hipEvent_t e0, e1;
    int id, k;          // class; iteration of the solve (-1: the solve's end, always runs)
    double bytes;       // algorithmic bytes when the launch does its work
  };
  unsigned probe_mask = 0;
  bool probe_err = false;
  std::vector<ProbeRec> probe_open, probe_done;
  std::vector<hipEvent_t> probe_pool;
  long long probe_n[PF_K_COUNT] = {}, probe_worked[PF_K_COUNT] = {};
  double probe_us[PF_K_COUNT] = {}, probe_bytes[PF_K_COUNT] = {};

  void probe_release() {
    for (auto* v : {&probe_open, &probe_done})
      for (auto& r : *v) {
        (void)hipEventDestroy(r.e0);
        (void)hipEventDestroy(r.e1);
      }
    probe_open.clear();
    probe_done.clear();
    for (hipEvent_t e : probe_pool) (void)hipEventDestroy(e);
    probe_pool.clear();
  }
  hipEvent_t probe_event() {
    hipEvent_t e = nullptr;
    if (!probe_pool.empty()) {
      e = probe_pool.back();
      probe_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
    }
    return e;
  }
  // before a launch of class `id`: the record's index, or -1 when not probed
  int probe_begin(int id) {
    if (!(probe_mask >> id & 1u) || tl_rec || probe_err) return -1;
    ProbeRec r{probe_event(), probe_event(), id, -1, 0.0};
    if (!r.e0 || !r.e1 || hipEventRecord(r.e0, stream) != hipSuccess) {
      probe_err = true;
      if (r.e0) probe_pool.push_back(r.e0);
      if (r.e1) probe_pool.push_back(r.e1);
      return -1;
    }
    probe_open.push_back(r);
    return (int)probe_open.size() - 1;
  }
  void probe_end(int idx, int k, double bytes) {
    if (idx < 0) return;
    ProbeRec& r = probe_open[idx];
    if (hipEventRecord(r.e1, stream) != hipSuccess) probe_err = true;
    r.k = k;
    r.bytes = bytes;
  }
  // the solve took `it` iterations: launches of later ones moved nothing
  void probe_settle(int it) {
    for (ProbeRec& r : probe_open) {
      if (r.k >= it) r.bytes = 0.0;
      probe_done.push_back(r);
    }
    probe_open.clear();
  }
  // algorithmic bytes of iteration k's launch of class id in the accelerated
  // solve (the counts of launch_one, with the history the iteration has:
  // min(k, m) slots after the Green update of iteration k >= 1)
  double probe_bytes_of(int id, int k) const {
    const double cs = (double)sizeof(C);
    const int m = prm.elastic_anderson_depth;
    switch (id) {
      case PF_K_ZPASS_HIJ:   // iteration 0 maps phi and x_B and fills the h(phi) cache
        return 12.0 * nSzA * cs + (k == 0 ? 16.0 * R + (double)sizeof(T) * R : (double)sizeof(T) * R);
      case PF_K_YPASS_FWD6:
      case PF_K_YPASS_INV6: return 6.0 * (nSzA + nSyA) * cs;
      case PF_K_XPASS_FWD6:
      case PF_K_XPASS_INV6: return 6.0 * (nSyA + nScA) * cs;
      case PF_K_GREEN:
        if (k == 0) return 6.0 * nScA * cs + 6.0 * nScA * 16.0;   // h, U read; g written
        return (m > 0 ? 12.0 + 3.0 * std::min(k, m) : 9.0) * nScA * cs + 12.0 * nScA * 16.0;
      case PF_K_AA_UPDATE: return (3.0 * std::min(k - 1, m) + 9.0) * nScA * cs + 3.0 * nScA * 16.0;
      default: return 0.0;
    }
  }
  int probe_kernels(unsigned mask) override {
    PF_HIP(hipStreamSynchronize(stream));
    probe_release();
    probe_mask = mask;
    probe_err = false;
    for (int i = 0; i < PF_K_COUNT; ++i) {
      probe_n[i] = probe_worked[i] = 0;
      probe_us[i] = probe_bytes[i] = 0.0;
    }
    return PF_OK;
  }
  int probe_read(int id, long long* launches, long long* worked, double* us, double* bytes) override {
    PF_HIP(hipStreamSynchronize(stream));
    if (probe_err) return fail(PF_E_HIP, "launch probe: event record failed");
    for (ProbeRec& r : probe_done) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, r.e0, r.e1) != hipSuccess) probe_err = true;
      probe_n[r.id] += 1;
      probe_us[r.id] += 1000.0 * (double)ms;
      if (r.bytes > 0.0) {
        probe_worked[r.id] += 1;
        probe_bytes[r.id] += r.bytes;
      }
      probe_pool.push_back(r.e0);
      probe_pool.push_back(r.e1);
    }
    probe_done.clear();
    if (probe_err) return fail(PF_E_HIP, "launch probe: event timing failed");
    *launches = probe_n[id];
    *worked = probe_worked[id];
    *us = probe_us[id];
    *bytes = probe_bytes[id];
    return PF_OK;
  }

  // the accelerated solver's control block in its steady state (running, a
  // full history, slot 0 next, fixed mixing coefficients) for launch_one
  int aa_ctl_steady() {
    if (!extended) return PF_OK;
    double* hc = hscal + AA_SLOT;
    for (int i = 0; i < AC_COUNT; ++i) hc[i] = 0.0;
    hc[AC_ACTIVE] = 1.0;
    hc[AC_NH] = (double)prm.elastic_anderson_depth;
    for (int j = 0; j < AA_MAX; ++j) hc[AC_GAM + j] = 0.25;
    PF_HIP(hipMemcpyAsync(dscal + AA_SLOT, hc, AC_COUNT * sizeof(double), hipMemcpyHostToDevice, stream));
    PF_HIP(hipStreamSynchronize(stream));
    return PF_OK;
  }

  int time_kernel(int id, int reps, double* us, double* bytes) override {
    state_ok = false;   // the timed launches overwrite the simulation buffers
    if (id == PF_K_GREEN || id == PF_K_AA_UPDATE) {
      const int rc = aa_ctl_steady();
      if (rc) return rc;
    }
    hipEvent_t e0, e1;
    PF_HIP(hipEventCreate(&e0));
    PF_HIP(hipEventCreate(&e1));
    launch_one(id);   // warm-up launch
    PF_HIP(hipEventRecord(e0, stream));
    double b = 0.0;
    for (int r = 0; r < reps; ++r) b = launch_one(id);
    PF_HIP(hipEventRecord(e1, stream));
    PF_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    PF_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *us = 1000.0 * (double)ms / reps;
    *bytes = b;
    warm_valid = false;
    return PF_OK;
  }
};

}  // namespace pf
