// LDS-staged mixed-radix FFT core for gfx950 (wave64).
//
// Replaces the cuFFT D2Z/Z2D (FP64) and R2C/C2R (FP32) plans the reference
// creates at main_cuda.cu:28158-28214.  A workgroup holds BATCH transforms of
// compile-time length N in LDS and runs a Stockham auto-sort sequence of
// radix-4/2/3/5 stages (in place, butterflies staged through registers);
// any other prime factor p (e.g. the 41 of 246 = 2*3*41) runs as a symmetric
// p-point DFT stage out of place through a second LDS buffer.
//
// Element (b, j) of the batch lives at  COLMAJ ? j*BATCH + b : b*N + j.
// Twiddles W[t] = exp(-2 pi i t / N) come from a per-length table staged in
// LDS; INV uses conj(W).  Transforms are unnormalised, as cuFFT's are.
#pragma once
#include <hip/hip_runtime.h>

namespace pf {

template <typename T>
struct cx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> cconj(cx<T> a) { return {a.re, -a.im}; }

// ----------------------------------------------------------- radix selection
constexpr int smallest_prime_factor(int n) {
  for (int p = 2; p * p <= n; ++p)
    if (n % p == 0) return p;
  return n;
}
// Largest power-of-two radix up to 16 first (fewer LDS round trips), then
// 3, 5 and any other prime.
constexpr int pick_radix(int n) {
  return (n % 16 == 0) ? 16 : (n % 8 == 0) ? 8 : (n % 4 == 0) ? 4 : smallest_prime_factor(n);
}
constexpr bool is_pow2_radix(int r) { return r == 2 || r == 4 || r == 8 || r == 16; }
constexpr bool plan_has_big_prime(int n) {
  while (n > 1) {
    int r = pick_radix(n);
    if (r > 5 && !is_pow2_radix(r)) return true;
    n /= r;
  }
  return false;
}
// A prime stage needs the second LDS buffer unless it is the last stage
// (NS * R == N): there each butterfly's outputs land on its own inputs' slots
// and the stage runs in place.
constexpr bool plan_needs_alt(int n) {
  int ns = 1;
  while (ns < n) {
    const int r = pick_radix(n / ns);
    if (r > 5 && !is_pow2_radix(r) && ns * r != n) return true;
    ns *= r;
  }
  return false;
}

// cos(2 pi k / 32), k = 0..8, as literals (constant-folded after unrolling)
__host__ __device__ constexpr double cos32_quarter(int k) {
  return k == 0 ? 1.0
       : k == 1 ? 0.980785280403230449126182236134
       : k == 2 ? 0.923879532511286756128183189397
       : k == 3 ? 0.831469612302545237078788377618
       : k == 4 ? 0.707106781186547524400844362105
       : k == 5 ? 0.555570233019602224742830813949
       : k == 6 ? 0.382683432365089771728459984030
       : k == 7 ? 0.195090322016128267848284868477
       : 0.0;
}
__host__ __device__ constexpr double cos32(int m) {
  m = ((m % 32) + 32) % 32;
  return m <= 8 ? cos32_quarter(m)
                : (m <= 16 ? -cos32_quarter(16 - m) : (m <= 24 ? -cos32_quarter(m - 16) : cos32_quarter(32 - m)));
}
__host__ __device__ constexpr double sin32(int m) { return cos32(m - 8); }

// ----------------------------------------------------------- small DFTs
template <typename T, int R, bool INV>
struct Dft;

template <typename T, bool INV>
struct Dft<T, 2, INV> {
  static __device__ __forceinline__ void run(cx<T>* v) {
    cx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <typename T, bool INV>
struct Dft<T, 4, INV> {
  static __device__ __forceinline__ void run(cx<T>* v) {
    cx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    cx<T> a2 = cadd(v[1], v[3]), a3 = csub(v[1], v[3]);
    // multiply a3 by -i (forward) or +i (inverse)
    cx<T> a3r = INV ? cx<T>{-a3.im, a3.re} : cx<T>{a3.im, -a3.re};
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3r);
    v[3] = csub(a1, a3r);
  }
};

template <typename T, bool INV>
struct Dft<T, 3, INV> {
  static __device__ __forceinline__ void run(cx<T>* v) {
    const T c = T(-0.5);
    const T s = INV ? T(0.86602540378443864676372317075294) : T(-0.86602540378443864676372317075294);
    cx<T> t = cadd(v[1], v[2]);
    cx<T> d = csub(v[1], v[2]);
    cx<T> m = {v[0].re + c * t.re, v[0].im + c * t.im};
    cx<T> r = {-s * d.im, s * d.re};   // i*s*d
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
};

template <typename T, bool INV>
struct Dft<T, 5, INV> {
  static __device__ __forceinline__ void run(cx<T>* v) {
    const T c1 = T(0.30901699437494742410229341718282);    // cos(2pi/5)
    const T c2 = T(-0.80901699437494742410229341718282);   // cos(4pi/5)
    const T s1a = T(0.95105651629515357211643933337938);   // sin(2pi/5)
    const T s2a = T(0.58778525229247312916870595463907);   // sin(4pi/5)
    const T s1 = INV ? s1a : -s1a;
    const T s2 = INV ? s2a : -s2a;
    cx<T> t1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
    cx<T> t2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
    cx<T> m1 = {v[0].re + c1 * t1.re + c2 * t2.re, v[0].im + c1 * t1.im + c2 * t2.im};
    cx<T> m2 = {v[0].re + c2 * t1.re + c1 * t2.re, v[0].im + c2 * t1.im + c1 * t2.im};
    // n1 = i*(s1 d1 + s2 d2), n2 = i*(s2 d1 - s1 d2)
    T a1r = s1 * d1.re + s2 * d2.re, a1i = s1 * d1.im + s2 * d2.im;
    T a2r = s2 * d1.re - s1 * d2.re, a2i = s2 * d1.im - s1 * d2.im;
    cx<T> n1 = {-a1i, a1r};
    cx<T> n2 = {-a2i, a2r};
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
  }
};

// W_E^m = exp(-2 pi i m / E) for E | 32 (conjugated for the inverse)
template <typename T, int E, bool INV>
__device__ __forceinline__ cx<T> wconst(int m) {
  const int q = (m * (32 / E)) % 32;
  return cx<T>{T(cos32(q)), INV ? T(sin32(q)) : T(-sin32(q))};
}

// In-register E-point DFT (E | 32), natural order in and out: Stockham stages
// of radix 4/2 with compile-time twiddles.
template <typename T, int E, int NS, bool INV>
__device__ __forceinline__ void dft_reg_stages(cx<T>* a) {
  if constexpr (NS < E) {
    constexpr int R = ((E / NS) % 4 == 0) ? 4 : 2;
    constexpr int NB = E / R;
    cx<T> b[E];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      cx<T> v[R];
      const int jm = j % NS;
#pragma unroll
      for (int s = 0; s < R; ++s) {
        v[s] = a[j + s * NB];
        if (NS > 1 && s > 0 && jm != 0) v[s] = cmul(v[s], wconst<T, E, INV>(jm * s * (E / (NS * R))));
      }
      Dft<T, R, INV>::run(v);
      const int d = (j / NS) * NS * R + jm;
#pragma unroll
      for (int s = 0; s < R; ++s) b[d + s * NS] = v[s];
    }
#pragma unroll
    for (int i = 0; i < E; ++i) a[i] = b[i];
    dft_reg_stages<T, E, NS * R, INV>(a);
  }
}

template <typename T, int E, bool INV>
__device__ __forceinline__ void dft_reg(cx<T>* a) {
  if constexpr (E > 1) dft_reg_stages<T, E, 1, INV>(a);
}

template <typename T, bool INV>
struct Dft<T, 8, INV> {
  static __device__ __forceinline__ void run(cx<T>* v) { dft_reg<T, 8, INV>(v); }
};
template <typename T, bool INV>
struct Dft<T, 16, INV> {
  static __device__ __forceinline__ void run(cx<T>* v) { dft_reg<T, 16, INV>(v); }
};

// ----------------------------------------------------------- LDS layout
// Padded to keep the strided Stockham writes free of bank conflicts: one
// extra element after every 16 along the transform axis, and (column-major)
// an odd row pitch.
template <int N, int BATCH, bool COLMAJ>
struct Lds {
  static constexpr int NPAD = N + N / 16;
  static constexpr int PITCH = BATCH + 1;
  static constexpr int size = COLMAJ ? NPAD * PITCH : BATCH * NPAD;
  static __device__ __forceinline__ int idx(int b, int j) {
    return COLMAJ ? (j + (j >> 4)) * PITCH + b : b * NPAD + j + (j >> 4);
  }
};

// ----------------------------------------------------------- one stage
template <int N, int BATCH, bool COLMAJ>
__device__ __forceinline__ int lds_index(int b, int j) {
  return Lds<N, BATCH, COLMAJ>::idx(b, j);
}

// Positions a pruned transform never needs.  The 2/3 rule keeps the
// retained wrapped indices of an axis of length N inside [0, N/3 + 1) and
// (N - N/3 - 1, N) (|k| < (2/3) pi / d with k = 2 pi i / (d N) is i < N/3),
// so [N/3 + 1, N - N/3 - 1] holds zeros on the input of a pruned inverse
// transform and outputs nobody reads on a pruned forward transform.
template <int N>
struct Pruned {
  static constexpr int LO = N / 3 + 1;
  static constexpr int HI = N - N / 3 - 1;
  // all of [a, a + len) inside [LO, HI]
  static constexpr bool covers(int a, int len) { return LO <= a && a + len - 1 <= HI; }
};

// The inputs a pruned inverse transform's first stage never reads: the
// union [lo, hi] of its fully pruned input blocks (empty, lo > hi, when the
// first stage is a prime stage or no block is fully pruned).  A caller that
// zero-fills the unretained middle before such a transform may skip it.
template <int N>
struct PrunedSkip {
  static constexpr int R = pick_radix(N);
  static constexpr bool INPLACE = R <= 5 || is_pow2_radix(R);
  static constexpr int NB = N / R;
  static constexpr int first() {
    for (int s = 0; s < R; ++s)
      if (Pruned<N>::covers(s * NB, NB)) return s * NB;
    return N;
  }
  static constexpr int last() {
    for (int s = R - 1; s >= 0; --s)
      if (Pruned<N>::covers(s * NB, NB)) return s * NB + NB - 1;
    return -1;
  }
  static constexpr int lo = INPLACE ? first() : N;
  static constexpr int hi = INPLACE ? last() : -1;
};

// Positions of [a, b] outside [PrunedSkip<N>::lo, hi]: count() of them, and
// the i-th in increasing order (two intervals at most).
template <int N>
struct ZeroRange {
  int a, b;
  __device__ __forceinline__ int n1() const {
    constexpr int lo = PrunedSkip<N>::lo;
    const int e = (lo <= PrunedSkip<N>::hi && lo - 1 < b) ? lo - 1 : b;
    return e >= a ? e - a + 1 : 0;
  }
  __device__ __forceinline__ int s2() const {
    constexpr int hi = PrunedSkip<N>::hi;
    return (PrunedSkip<N>::lo <= hi && hi + 1 > a) ? hi + 1 : b + 1;
  }
  __device__ __forceinline__ int count() const {
    const int st = s2();
    return n1() + (b >= st ? b - st + 1 : 0);
  }
  __device__ __forceinline__ int at(int i) const {
    const int m = n1();
    return i < m ? a + i : s2() + (i - m);
  }
};

// In-place Stockham stage for radix 2/3/4/5/8/16 (butterflies staged in registers).
// PIN: the first stage (NS = 1) takes the inputs Pruned<N> marks as zero as
// compile-time zeros (neither read nor multiplied); POUT: the last stage
// (NS * R = N) stores no output Pruned<N> marks as unread, and the arithmetic
// that feeds only those outputs is dropped by the compiler.
template <typename T, int N, int R, int NS, int BATCH, bool COLMAJ, bool INV, int NT, bool PIN = false,
          bool POUT = false>
__device__ __forceinline__ void stage_inplace(cx<T>* buf, const cx<T>* tw) {
  constexpr int NB = N / R;
  constexpr int TOT = NB * BATCH;
  constexpr int PER = (TOT + NT - 1) / NT;
  constexpr int TSTEP = N / (NS * R);
  constexpr bool ZIN = PIN && NS == 1;        // input j + s NB, j < NB
  constexpr bool ZOUT = POUT && NS * R == N;  // output j + s NS, j < NS
  cx<T> v[PER][R];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int t = threadIdx.x + q * NT;
    if (TOT % NT == 0 || t < TOT) {
      int b, j;
      if (COLMAJ) {
        b = t % BATCH;
        j = t / BATCH;
      } else {
        j = t % NB;
        b = t / NB;
      }
      const int jm = j % NS;
#pragma unroll
      for (int s = 0; s < R; ++s)
        v[q][s] = (ZIN && Pruned<N>::covers(s * NB, NB)) ? cx<T>{T(0), T(0)}
                                                         : buf[lds_index<N, BATCH, COLMAJ>(b, j + s * NB)];
      if (NS > 1) {
        // twiddles W^(jm s TSTEP) read from the table (jm s TSTEP < N): one
        // LDS read each instead of a chain of complex products
#pragma unroll
        for (int s = 1; s < R; ++s) {
          cx<T> w = tw[jm * s * TSTEP];
          if (INV) w.im = -w.im;
          v[q][s] = cmul(v[q][s], w);
        }
      }
      Dft<T, R, INV>::run(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int t = threadIdx.x + q * NT;
    if (TOT % NT == 0 || t < TOT) {
      int b, j;
      if (COLMAJ) {
        b = t % BATCH;
        j = t / BATCH;
      } else {
        j = t % NB;
        b = t / NB;
      }
      const int jm = j % NS;
      const int d = (j / NS) * NS * R + jm;
#pragma unroll
      for (int s = 0; s < R; ++s)
        if (!(ZOUT && Pruned<N>::covers(s * NS, NS))) buf[lds_index<N, BATCH, COLMAJ>(b, d + s * NS)] = v[q][s];
    }
  }
  __syncthreads();
}

// DFT stage for an odd prime radix R > 5 (the 41 of 246), out of place
// (src -> dst) or, as the last stage of a plan, in place (dst == src: a
// butterfly's R outputs occupy its R input slots, and its lanes are one
// wave's, which read all inputs before any output is written).
//  1. The inter-stage twiddles W_N^(jm s TSTEP) are applied to the R-1
//     non-zero inputs of every butterfly, in place in `src`.
//  2. Each output pair (X_k, X_{R-k}) comes from the symmetric sums
//        A = x_0 + sum_s (x_s + x_{R-s}) cos(2 pi s k / R)
//        B =       sum_s (x_s - x_{R-s}) (-/+) sin(2 pi s k / R),  s = 1..(R-1)/2,
//     X_k = A + iB, X_{R-k} = A - iB (a quarter of the multiplies of the
//     direct sum).  Two pairs per lane, k and k + G, G = (R+3)/4 lanes per
//     butterfly (lane 0 of the group gives X_0 as its first); the group's
//     lanes read x_s and x_{R-s} at the same LDS addresses in the same
//     iteration (broadcast reads) and form x_s +/- x_{R-s} once for both
//     pairs, and each pair's cos/sin of s*k mod R comes from the length-N
//     twiddle table.
template <typename T, int N, int R, int NS, int BATCH, bool COLMAJ, bool INV, int NT>
__device__ __forceinline__ void stage_prime(cx<T>* src, cx<T>* dst, const cx<T>* tw) {
  constexpr int NB = N / R;
  constexpr int TSTEP = N / (NS * R);
  constexpr int WSTEP = N / R;   // W_R^m = W_N^(m*N/R)
  constexpr int NBF = NB * BATCH;   // butterflies in the workgroup
  constexpr int H = (R - 1) / 2;
  constexpr int G = (H + 2) / 2;    // lanes per butterfly: lane k gives pairs k and k + G
  static_assert(R % 2 == 1 && G <= 64, "odd prime radix up to 255");
  constexpr int GPW = 64 / G;       // butterflies per wave and pass
  constexpr int GROUPS = (NT / 64) * GPW;
  auto split = [](int bt, int& b, int& j) {
    if (COLMAJ) {
      b = bt % BATCH;
      j = bt / BATCH;
    } else {
      j = bt % NB;
      b = bt / NB;
    }
  };
  if constexpr (NS > 1) {
    constexpr int TOT = NBF * (R - 1);
    for (int t = threadIdx.x; t < TOT; t += NT) {
      const int s = 1 + t % (R - 1);
      int b, j;
      split(t / (R - 1), b, j);
      const int jm = j % NS;
      if (jm != 0) {
        const int at = lds_index<N, BATCH, COLMAJ>(b, j + s * NB);
        cx<T> w = tw[jm * s * TSTEP];   // jm s TSTEP < NS R TSTEP = N
        if (INV) w.im = -w.im;
        src[at] = cmul(src[at], w);
      }
    }
    __syncthreads();
  }
  const int lane = threadIdx.x & 63;
  const int g = lane / G;
  const int k = lane - g * G;
  const int k2 = k + G;          // the lane's second pair (none when k2 > H)
  const int slot = (threadIdx.x >> 6) * GPW + g;
  for (int base = 0; base < NBF; base += GROUPS) {
    const int bt = base + slot;
    if (g < GPW && bt < NBF) {
      int b, j;
      split(bt, b, j);
      const cx<T> x0 = src[lds_index<N, BATCH, COLMAJ>(b, j)];
      cx<T> A = x0, B = {T(0), T(0)}, A2 = x0, B2 = {T(0), T(0)};
      int m = 0, m2 = 0;   // s k mod R, s k2 mod R
#pragma unroll 4
      for (int s = 1; s <= H; ++s) {
        m += k;
        if (m >= R) m -= R;
        m2 += k2;
        if (m2 >= R) m2 -= R;
        const cx<T> a = src[lds_index<N, BATCH, COLMAJ>(b, j + s * NB)];
        const cx<T> c = src[lds_index<N, BATCH, COLMAJ>(b, j + (R - s) * NB)];
        const cx<T> sp = {a.re + c.re, a.im + c.im}, sm = {a.re - c.re, a.im - c.im};
        const cx<T> w = tw[m * WSTEP];   // (cos, -sin) of 2 pi m / R
        const T wi = INV ? -w.im : w.im;
        A.re += sp.re * w.re;
        A.im += sp.im * w.re;
        B.re += sm.re * wi;
        B.im += sm.im * wi;
        const cx<T> w2 = tw[m2 * WSTEP];
        const T wi2 = INV ? -w2.im : w2.im;
        A2.re += sp.re * w2.re;
        A2.im += sp.im * w2.re;
        B2.re += sm.re * wi2;
        B2.im += sm.im * wi2;
      }
      const int jm = j % NS;
      const int d = (j / NS) * NS * R + jm;
      dst[lds_index<N, BATCH, COLMAJ>(b, d + k * NS)] = cx<T>{A.re - B.im, A.im + B.re};
      if (k > 0) dst[lds_index<N, BATCH, COLMAJ>(b, d + (R - k) * NS)] = cx<T>{A.re + B.im, A.im - B.re};
      if (k2 <= H) {
        dst[lds_index<N, BATCH, COLMAJ>(b, d + k2 * NS)] = cx<T>{A2.re - B2.im, A2.im + B2.re};
        dst[lds_index<N, BATCH, COLMAJ>(b, d + (R - k2) * NS)] = cx<T>{A2.re + B2.im, A2.im - B2.re};
      }
    }
  }
  __syncthreads();
}

// Runs all stages.  `buf` holds the data on entry; the result is returned
// (either `buf` or `alt`, the second buffer needed only by big-prime plans).
template <typename T, int N, int NS, int BATCH, bool COLMAJ, bool INV, int NT, bool PIN, bool POUT>
__device__ __forceinline__ cx<T>* fft_stages(cx<T>* buf, cx<T>* alt, const cx<T>* tw) {
  if constexpr (NS >= N) {
    return buf;
  } else {
    constexpr int R = pick_radix(N / NS);
    if constexpr (R <= 5 || is_pow2_radix(R)) {
      stage_inplace<T, N, R, NS, BATCH, COLMAJ, INV, NT, PIN, POUT>(buf, tw);
      return fft_stages<T, N, NS * R, BATCH, COLMAJ, INV, NT, PIN, POUT>(buf, alt, tw);
    } else if constexpr (NS * R == N) {
      stage_prime<T, N, R, NS, BATCH, COLMAJ, INV, NT>(buf, buf, tw);
      return buf;
    } else {
      stage_prime<T, N, R, NS, BATCH, COLMAJ, INV, NT>(buf, alt, tw);
      return fft_stages<T, N, NS * R, BATCH, COLMAJ, INV, NT, PIN, POUT>(alt, buf, tw);
    }
  }
}

// PIN: the caller guarantees zeros at the inputs Pruned<N> marks; POUT: the
// caller reads no output Pruned<N> marks (those positions are left stale).
template <typename T, int N, int BATCH, bool COLMAJ, bool INV, int NT, bool PIN = false, bool POUT = false>
__device__ __forceinline__ cx<T>* fft_lds(cx<T>* buf, cx<T>* alt, const cx<T>* tw) {
  return fft_stages<T, N, 1, BATCH, COLMAJ, INV, NT, PIN, POUT>(buf, alt, tw);
}

}  // namespace pf
