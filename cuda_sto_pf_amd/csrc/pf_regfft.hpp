// Register-resident z pass for power-of-two lengths N <= 256.
//
// A transform of length N = E * L is held by a group of L lanes of one wave64,
// lane l owning positions n = l + L e (e < E) in registers.  The transform is
// the two-step (four-step) factorisation
//     X[k1 + E k2] = sum_l W_L^{l k2} W_N^{l k1} sum_e x[l + L e] W_E^{e k1}
// with the E-point and L-point DFTs done in registers (compile-time radix-4/2
// Stockham, constant twiddles) and one LDS transpose between them.  With
// L | E the output positions owned by a lane are again {l + L e}, so the
// inverse transform, the point-wise operator and the forward transform run
// without any further data movement; a workgroup is one wave, so no block
// barrier waits on another wave.  Used for the same fused z-pass operators as
// zpass_kernel (pf_passes.hpp) and with the same spectral layouts.
#pragma once
#include "pf_passes.hpp"

namespace pf {

// Length-N transform over the L-lane group; v[e] holds position l + L e on
// entry and on exit.  `scr` is this wave's LDS scratch of 64 * (E + 1) values.
template <typename T, int N, int L, bool INV>
__device__ __forceinline__ void fft_reg(cx<T>* v, cx<T>* scr, const cx<T>* tw) {
  constexpr int E = N / L;
  constexpr int C = E / L;
  constexpr int P = E + 1;
  static_assert(C * L == E, "L must divide E");
  const int lane = threadIdx.x;
  const int l = lane % L;
  const int g0 = lane - l;
  dft_reg<T, E, INV>(v);
#pragma unroll
  for (int k1 = 1; k1 < E; ++k1) {
    cx<T> w = tw[l * k1];
    if (INV) w.im = -w.im;
    v[k1] = cmul(v[k1], w);
  }
#pragma unroll
  for (int k1 = 0; k1 < E; ++k1) scr[lane * P + k1] = v[k1];
  __syncthreads();
  cx<T> w2[C][L];
#pragma unroll
  for (int t = 0; t < C; ++t)
#pragma unroll
    for (int lp = 0; lp < L; ++lp) w2[t][lp] = scr[(g0 + lp) * P + l + L * t];
  __syncthreads();
#pragma unroll
  for (int t = 0; t < C; ++t) dft_reg<T, L, INV>(w2[t]);
#pragma unroll
  for (int t = 0; t < C; ++t)
#pragma unroll
    for (int k2 = 0; k2 < L; ++k2) v[t + C * k2] = w2[t][k2];
}

template <int N>
struct RegPlan {
  static constexpr int L = (N >= 256) ? 16 : (N >= 64 ? 8 : (N >= 16 ? 4 : 2));
  static constexpr int E = N / L;
  static constexpr bool ok = (N & (N - 1)) == 0 && N >= 8 && N <= 256 && (E % L == 0);
};

template <typename T, int N, int NCI, int NCO, class Op>
__global__ void __launch_bounds__(64) zpass_reg_kernel(ZPassArgs<T> za, typename Op::Args oa) {
  if (za.gate && *za.gate == 0.0) return;
  constexpr int L = RegPlan<N>::L;
  constexpr int E = N / L;
  constexpr int G = 64 / L;
  constexpr int PI = (NCI + 1) / 2;
  constexpr int PO = (NCO + 1) / 2;
  constexpr int PM = (PI > PO ? PI : PO) > 0 ? (PI > PO ? PI : PO) : 1;
  __shared__ cx<T> scr[64 * (E + 1)];
  const int lane = threadIdx.x;
  const int l = lane % L;
  const int grp = lane / L;
  const int row = blockIdx.x * G + grp;
  const bool live = row < za.rows;
  const int Az = za.Az;
  const size_t rb = (size_t)row * za.AzS;
  cx<T> v[PM][E];

  // ---- inverse: Hermitian extension of the retained kz, two channels per transform
#pragma unroll
  for (int p = 0; p < PI; ++p) {
    const int ca = 2 * p, cbi = 2 * p + 1;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = l + L * e;
      cx<T> z = {T(0), T(0)};
      if (live) {
        if (k < Az) {
          cx<T> A = za.in[ca][rb + k];
          cx<T> B = (cbi < NCI) ? za.in[cbi][rb + k] : cx<T>{T(0), T(0)};
          if (k == 0) {
            A.im = T(0);
            B.im = T(0);
          }
          z = cx<T>{A.re - B.im, A.im + B.re};
        } else if (N - k < Az) {
          cx<T> A = za.in[ca][rb + (N - k)];
          cx<T> B = (cbi < NCI) ? za.in[cbi][rb + (N - k)] : cx<T>{T(0), T(0)};
          z = cx<T>{A.re + B.im, B.re - A.im};
        }
      }
      v[p][e] = z;
    }
    fft_reg<T, N, L, true>(v[p], scr, za.tw);
  }

  // ---- point-wise operator at positions z = l + L e
  double red[Op::NRED > 0 ? Op::NRED : 1];
#pragma unroll
  for (int r = 0; r < Op::NRED; ++r) red[r] = red_identity(Op::kinds[r]);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    T xin[NCI > 0 ? NCI : 1];
    T yout[NCO > 0 ? NCO : 1];
#pragma unroll
    for (int c = 0; c < NCI; ++c) xin[c] = (c & 1) ? v[c >> 1][e].im : v[c >> 1][e].re;
    if (live) {
      double pre[Op::NPRE > 0 ? Op::NPRE : 1];
      Op::load(oa, (size_t)row * N + (l + L * e), pre);
      Op::apply(oa, (size_t)row * N + (l + L * e), pre, xin, yout, red);
    } else {
#pragma unroll
      for (int c = 0; c < NCO; ++c) yout[c] = T(0);
    }
#pragma unroll
    for (int q = 0; q < PO; ++q)
      v[q][e] = cx<T>{yout[2 * q], (2 * q + 1 < NCO) ? yout[2 * q + 1] : T(0)};
  }
  if constexpr (Op::NRED > 0) {
    // one wave per block: xor butterfly, lane 0 writes the block partial
#pragma unroll
    for (int r = 0; r < Op::NRED; ++r) {
      double x = red[r];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) x = red_combine(Op::kinds[r], x, __shfl_xor(x, off, 64));
      if (lane == 0) za.partials[(size_t)blockIdx.x * Op::NRED + r] = x;
    }
  }

  // ---- forward, then split the two real channels and keep kz < Az
#pragma unroll
  for (int q = 0; q < PO; ++q) {
    fft_reg<T, N, L, false>(v[q], scr, za.tw);
    // partner of position k is N - k: lane (L - l) % L, register (l == 0 ? (E - e) % E : E - 1 - e)
    constexpr int P = E + 1;
#pragma unroll
    for (int e = 0; e < E; ++e) scr[lane * P + e] = v[q][e];
    __syncthreads();
    const int pl = (L - l) % L;
    const int ca = 2 * q, cbi = 2 * q + 1;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = l + L * e;
      if (live && k < Az) {
        const int pe = (l == 0) ? ((E - e) % E) : (E - 1 - e);
        const cx<T> Zk = v[q][e];
        const cx<T> Zm = scr[(lane - l + pl) * P + pe];
        const cx<T> A = {T(0.5) * (Zk.re + Zm.re), T(0.5) * (Zk.im - Zm.im)};
        za.out[ca][rb + k] = A;
        if (cbi < NCO) za.out[cbi][rb + k] = cx<T>{T(0.5) * (Zk.im + Zm.im), T(0.5) * (Zm.re - Zk.re)};
      }
    }
    __syncthreads();
  }
}

}  // namespace pf
