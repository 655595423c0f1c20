// FFT pass kernels of the pruned 3-D transform.
//
// Spectral data never exists in the reference's full half-spectrum form: the
// reference zeroes every mode with |k_i| >= (2/3) pi / d_i after every
// transform (dealias_23_kernel, cuda_kernels.cu:103-161, applied at every
// cufftExec* site of main_cuda.cu:28846-30806), so only the retained modes are
// ever computed or stored.  Per axis the retained wrapped indices are
// |ii| < a (x, y) and kz < az (z); they are computed on the host with the
// reference's own floating-point expression.
//
// Layouts (complex, last index fastest):
//   Sz  [Nx][Ny][AzS]  after the z pass (real rows -> Az retained kz)
//   Sy  [Nx][Ay][AzS]  after the y pass
//   Sc  [Ax][Ay][AzS]  compact spectrum after the x pass
// The kz stride AzS >= Az is rounded up to whole 128-byte lines on large
// grids so the column tiles of the y and x passes are line-aligned; the
// padding entries hold zeros, and every operator maps zeros to zeros.
// Compact index c of a wrapped axis maps to the full index c (c < a) or
// c + N - A (c >= a), A = 2a - 1.
#pragma once
#include "pf_launch.hpp"
#include <type_traits>

#include "pf_fft.hpp"

namespace pf {

constexpr int NT = 256;                 // threads per workgroup (4 wave64)
constexpr int LDS_BUDGET = 48 * 1024;   // bytes of FFT data per workgroup

constexpr int ZPASS_LDS_BUDGET = 32 * 1024;   // smaller tiles: more resident workgroups

// Register budget (doubles per thread) for the operator's real-space operands
// loaded ahead of the inverse transform.
constexpr int ZPASS_PREFETCH_DOUBLES = 48;

// Rows per workgroup: as many row pairs as the LDS budget holds; an operator
// with NPRE real-space operands per point whose operands would not fit the
// prefetch budget at that tile gets a smaller tile where they fit in half of
// it (so their loads overlap the inverse transform instead of being issued
// a chunk at a time after it).
template <typename T, int N, int NC, int NPRE = 0, int BUDGET = ZPASS_LDS_BUDGET>
constexpr int zpass_rows() {
  constexpr int per_pair = NC * (N + N / 16) * (int)sizeof(cx<T>) * (plan_needs_alt(N) ? 2 : 1);
  constexpr int pairs = BUDGET / per_pair;
  int p = pairs < 1 ? 1 : (pairs > 32 ? 32 : pairs);
  if (NPRE > 0 && ((2 * p * N + NT - 1) / NT) * NPRE > ZPASS_PREFETCH_DOUBLES)
    while (p > 1 && ((2 * p * N + NT - 1) / NT) * NPRE > ZPASS_PREFETCH_DOUBLES / 2) --p;
  return 2 * p;
}

// Upper bound (a multiple of 64) on the retained kz count: the 2/3 rule keeps
// kz < N/3, so Az <= N/3 + 1.  The engine rejects grids that exceed it.
template <int N>
constexpr int zpass_kmax() {
  return ((N / 3 + 1 + 63) / 64) * 64;
}

template <typename T, int N>
constexpr int cpass_cols() {
  constexpr int per_col = (N + N / 16) * (int)sizeof(cx<T>) * (plan_needs_alt(N) ? 2 : 1);
  constexpr int c = (LDS_BUDGET / per_col) - 1;
  return c >= 64 ? 64 : c >= 32 ? 32 : c >= 16 ? 16 : c >= 8 ? 8 : c >= 4 ? 4 : c >= 2 ? 2 : 1;
}

// --------------------------------------------------------------- reductions
enum RedKind { RED_SUM = 0, RED_MIN = 1, RED_MAX = 2 };

__device__ __forceinline__ double red_identity(int kind) {
  return kind == RED_SUM ? 0.0 : (kind == RED_MIN ? 1.0e300 : -1.0e300);
}
__device__ __forceinline__ double red_combine(int kind, double a, double b) {
  return kind == RED_SUM ? a + b : (kind == RED_MIN ? fmin(a, b) : fmax(a, b));
}

// Deterministic block reduction: fixed xor-butterfly inside each wave64,
// then wave partials combined in wave order by thread 0.
template <int NRED>
__device__ __forceinline__ void block_reduce_store(double* vals, const int* kinds, double* out) {
  __shared__ double wsum[NRED > 0 ? NRED : 1][NT / 64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < NRED; ++r) {
    double v = vals[r];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = red_combine(kinds[r], v, __shfl_xor(v, off, 64));
    if (lane == 0) wsum[r][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NRED; ++r) {
      double v = wsum[r][0];
      for (int w = 1; w < NT / 64; ++w) v = red_combine(kinds[r], v, wsum[r][w]);
      out[r] = v;
    }
  }
}

// --------------------------------------------------------------- z pass
template <typename T>
struct ZPassArgs {
  int rows;          // Nx*Ny
  int Az;            // retained kz count
  int AzS;           // kz stride of the spectral rows (>= Az)
  const cx<T>* in[6];
  cx<T>* out[6];
  const cx<T>* tw;   // twiddles of length Nz
  double* partials;  // [gridDim.x][Op::NRED]
  const double* gate;   // when set, the launch does nothing while *gate == 0
};

// One workgroup: RT consecutive rows, all components.
//  1. NCI>0: Hermitian-extend the Az retained kz of row pairs, inverse complex
//     FFT (two real rows per complex row), giving N * (real rows).
//  2. Op::apply at every (row, z): consumes NCI reals, produces NCO reals and
//     may read/write real fields and accumulate reductions.
//  3. NCO>0: forward complex FFT of row pairs, split into two real-row
//     spectra, store the Az retained kz.
// An operator may set its own LDS budget (static constexpr int LDS_BUDGET).
template <class Op, class = void>
struct ZBudget {
  static constexpr int value = ZPASS_LDS_BUDGET;
};
template <class Op>
struct ZBudget<Op, std::void_t<decltype(Op::LDS_BUDGET)>> {
  static constexpr int value = Op::LDS_BUDGET;
};

template <typename T, int N, int NCI, int NCO, class Op, int BUDGET = ZBudget<Op>::value>
__device__ __forceinline__ void zpass_body(const ZPassArgs<T>& za, const typename Op::Args& oa, const B3 blockIdx) {
  if (za.gate && *za.gate == 0.0) return;
  constexpr int NCM = NCI > NCO ? NCI : NCO;
  constexpr int NCB = NCM > 0 ? NCM : 1;
  constexpr int RT = zpass_rows<T, N, NCB, Op::NPRE, BUDGET>();
  constexpr int NP = RT / 2;
  constexpr bool BIG = plan_needs_alt(N);
  using LY = Lds<N, NCB * NP, false>;   // row-major: layout independent of the batch
  __shared__ cx<T> buf[LY::size];
  __shared__ cx<T> alt[BIG ? LY::size : 1];
  __shared__ cx<T> tw[N];

  const int row0 = blockIdx.x * RT;
  const int Az = za.Az;
  const int AzS = za.AzS;
  // wave index as a wave-uniform (scalar) value, so per-wave component
  // pointers come from scalar loads of the kernel arguments
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  for (int i = threadIdx.x; i < N; i += NT) tw[i] = za.tw[i];
  // real-space operands of the operator, issued before the inverse transform
  // so their latency overlaps it
  // (only while the operands of all of a thread's points fit a modest
  // register budget; otherwise each point loads its operands just before use)
  constexpr int PTS = (RT * N + NT - 1) / NT;
  constexpr int NPRE = Op::NPRE > 0 ? Op::NPRE : 1;
  constexpr bool PREFETCH = PTS * NPRE <= ZPASS_PREFETCH_DOUBLES;
  double pre[PREFETCH ? PTS : 1][NPRE];
  if constexpr (PREFETCH) {
#pragma unroll
    for (int i = 0; i < PTS; ++i) {
      const int e = threadIdx.x + i * NT;
      const int row = row0 + e / N;
      if (e < RT * N && row < za.rows) Op::load(oa, (size_t)row * N + (e % N), pre[i]);
    }
  }

  if constexpr (NCI > 0) {
    // zero the unretained middle, then scatter the retained modes (one wave
    // per row); positions the pruned inverse transform never reads stay as
    // they are
    {
      const ZeroRange<N> zr{Az, N - Az};
      const int nz = zr.count();
      for (int e = threadIdx.x; e < NCI * NP * nz; e += NT) buf[LY::idx(e / nz, zr.at(e % nz))] = cx<T>{T(0), T(0)};
    }
    // all retained-mode loads of a wave are issued before the LDS scatter
    // (when they fit a register budget; small grids use the loop below)
    constexpr int CPW0 = (NCI * NP + NT / 64 - 1) / (NT / 64);
    constexpr int KI = zpass_kmax<N>() / 64;
    constexpr bool BATCHED = CPW0 * KI * 2 <= 24;
    constexpr int CPW = BATCHED ? CPW0 : 0;
    cx<T> va[CPW > 0 ? CPW : 1][KI], vb[CPW > 0 ? CPW : 1][KI];
    const cx<T>* pin[CPW > 0 ? CPW : 1];
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
      const int cp = wave + q * (NT / 64);
      pin[q] = za.in[cp < NCI * NP ? cp / NP : 0];
    }
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
      const int cp = wave + q * (NT / 64);
      const int p = cp % NP;
      const int ra = row0 + 2 * p, rb = ra + 1;
      const bool live = cp < NCI * NP;
      const cx<T>* ia = pin[q] + (size_t)ra * AzS;
      const cx<T>* ib = pin[q] + (size_t)rb * AzS;
#pragma unroll
      for (int i = 0; i < KI; ++i) {
        const int k = lane + 64 * i;
        va[q][i] = (live && k < Az && ra < za.rows) ? ia[k] : cx<T>{T(0), T(0)};
        vb[q][i] = (live && k < Az && rb < za.rows) ? ib[k] : cx<T>{T(0), T(0)};
      }
    }
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
      const int cp = wave + q * (NT / 64);
      if (cp >= NCI * NP) continue;
#pragma unroll
      for (int i = 0; i < KI; ++i) {
        const int k = lane + 64 * i;
        if (k >= Az) continue;
        cx<T> A = va[q][i], B = vb[q][i];
        if (k == 0) {
          A.im = T(0);
          B.im = T(0);
        }
        buf[LY::idx(cp, k)] = cx<T>{A.re - B.im, A.im + B.re};
        if (k > 0) buf[LY::idx(cp, N - k)] = cx<T>{A.re + B.im, B.re - A.im};
      }
    }
    if constexpr (!BATCHED) {
      for (int cp = wave; cp < NCI * NP; cp += NT / 64) {
        const int p = cp % NP;
        const int c = cp / NP;
        const int ra = row0 + 2 * p, rb = ra + 1;
        const cx<T>* ia = za.in[c] + (size_t)ra * AzS;
        const cx<T>* ib = za.in[c] + (size_t)rb * AzS;
        for (int k = lane; k < Az; k += 64) {
          cx<T> A = (ra < za.rows) ? ia[k] : cx<T>{T(0), T(0)};
          cx<T> B = (rb < za.rows) ? ib[k] : cx<T>{T(0), T(0)};
          if (k == 0) {
            A.im = T(0);
            B.im = T(0);
          }
          buf[LY::idx(cp, k)] = cx<T>{A.re - B.im, A.im + B.re};
          if (k > 0) buf[LY::idx(cp, N - k)] = cx<T>{A.re + B.im, B.re - A.im};
        }
      }
    }
    __syncthreads();
    cx<T>* res = fft_lds<T, N, NCI * NP, false, true, NT, true, false>(buf, alt, tw);
    if (BIG && res != buf) {
      for (int e = threadIdx.x; e < LY::size; e += NT) buf[e] = res[e];
      __syncthreads();
    }
  }

  // point-wise operator
  double red[Op::NRED > 0 ? Op::NRED : 1];
#pragma unroll
  for (int r = 0; r < Op::NRED; ++r) red[r] = red_identity(Op::kinds[r]);
  auto point = [&](int e, const double* pp) {
    const int z = e % N;
    const int r = e / N;
    const int row = row0 + r;
    const int p = r >> 1;
    const int hi = r & 1;
    T vin[NCI > 0 ? NCI : 1];
    T vout[NCO > 0 ? NCO : 1];
#pragma unroll
    for (int c = 0; c < NCI; ++c) vin[c] = reinterpret_cast<const T*>(&buf[LY::idx(c * NP + p, z)])[hi];
    if (row < za.rows) {
      double lp[NPRE];
      if (pp == nullptr) {
        Op::load(oa, (size_t)row * N + z, lp);
        pp = lp;
      }
      Op::apply(oa, (size_t)row * N + z, pp, vin, vout, red);
    } else {
#pragma unroll
      for (int c = 0; c < NCO; ++c) vout[c] = T(0);
    }
    if constexpr (NCO > 0) {
      // the two rows of a pair are handled by different threads: write halves
#pragma unroll
      for (int c = 0; c < NCO; ++c) reinterpret_cast<T*>(&buf[LY::idx(c * NP + p, z)])[hi] = vout[c];
    }
  };
  if constexpr (PREFETCH) {
#pragma unroll
    for (int i = 0; i < PTS; ++i) {
      const int e = threadIdx.x + i * NT;
      if (e < RT * N) point(e, pre[i]);
    }
  } else {
    // operands loaded a chunk of points at a time, all in flight together
    constexpr int CHK = 12 / NPRE > 0 ? 12 / NPRE : 1;
    for (int base = 0; base < PTS; base += CHK) {
      double lp[CHK][NPRE];
#pragma unroll
      for (int i = 0; i < CHK; ++i) {
        const int e = threadIdx.x + (base + i) * NT;
        const int row = row0 + e / N;
        if (base + i < PTS && e < RT * N && row < za.rows) Op::load(oa, (size_t)row * N + (e % N), lp[i]);
      }
#pragma unroll
      for (int i = 0; i < CHK; ++i) {
        const int e = threadIdx.x + (base + i) * NT;
        if (base + i < PTS && e < RT * N) point(e, lp[i]);
      }
    }
  }
  if constexpr (Op::NRED > 0) {
    int kinds[Op::NRED > 0 ? Op::NRED : 1];
#pragma unroll
    for (int r = 0; r < Op::NRED; ++r) kinds[r] = Op::kinds[r];
    block_reduce_store<Op::NRED>(red, kinds, za.partials + (size_t)blockIdx.x * Op::NRED);
  }

  if constexpr (NCO > 0) {
    __syncthreads();
    cx<T>* res = fft_lds<T, N, NCO * NP, false, false, NT, false, true>(buf, alt, tw);
    for (int cp = wave; cp < NCO * NP; cp += NT / 64) {
      const int p = cp % NP;
      const int c = cp / NP;
      const int ra = row0 + 2 * p, rb = ra + 1;
      cx<T>* oa_ = za.out[c] + (size_t)ra * AzS;
      cx<T>* ob_ = za.out[c] + (size_t)rb * AzS;
      for (int k = lane; k < Az; k += 64) {
        const cx<T> Zk = res[LY::idx(cp, k)];
        const cx<T> Zm = res[LY::idx(cp, k == 0 ? 0 : N - k)];
        if (ra < za.rows) oa_[k] = cx<T>{T(0.5) * (Zk.re + Zm.re), T(0.5) * (Zk.im - Zm.im)};
        if (rb < za.rows) ob_[k] = cx<T>{T(0.5) * (Zk.im + Zm.im), T(0.5) * (Zm.re - Zk.re)};
      }
    }
  }
}

template <typename T, int N, int NCI, int NCO, class Op, int BUDGET = ZBudget<Op>::value>
__global__ void __launch_bounds__(NT) zpass_kernel(ZPassArgs<T> za, typename Op::Args oa) {
  zpass_body<T, N, NCI, NCO, Op, BUDGET>(za, oa, PF_BIDX);
}
// launch descriptor (pf_launch.hpp)
template <typename T, int N, int NCI, int NCO, class Op, int BUDGET = ZBudget<Op>::value>
struct KZPass {
  struct Args {
    ZPassArgs<T> za;
    typename Op::Args oa;
  };
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 bid, B3) {
    zpass_body<T, N, NCI, NCO, Op, BUDGET>(a.za, a.oa, bid);
  }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL((zpass_kernel<T, N, NCI, NCO, Op, BUDGET>), dim3(g.x, g.y, g.z), dim3(NT), 0, s, a.za,
                       a.oa);
  }
};

// --------------------------------------------------------------- column pass
// Strided complex FFT along y (planes = Nx) or x (planes = 1) with pruning.
// Forward: N full inputs -> A retained outputs.  Inverse: A -> N.
template <typename T>
struct CPassArgs {
  const cx<T>* in;
  cx<T>* out;
  int cols;                 // contiguous columns per j (AzS for y, Ay*AzS for x)
  long in_pstride, out_pstride;   // plane strides (elements)
  long in_cstride, out_cstride;   // component strides (elements)
  int A, a;                 // compact count and retained half-width
  const cx<T>* tw;
  const double* gate;       // when set, the launch does nothing while *gate == 0
};

// Work item of a workgroup in XCD-aware order.  The dispatcher deals
// workgroup ids round-robin to the 8 XCDs, each with its own L2; column tiles
// next to each other share 128-byte lines wherever a row is not a whole
// number of lines (kz strides of 64-byte lines), so each XCD takes one
// contiguous range of (tile, plane, component) items and neighbouring tiles
// meet in the same L2 close together in time.  A bijection on the grid.
__device__ __forceinline__ B3 xcd_item(const B3 b, const B3 g) {
  constexpr unsigned NXCD = 8;
  const unsigned total = g.x * g.y * g.z;
  const unsigned i = b.x + g.x * (b.y + g.y * b.z);
  const unsigned per = (total + NXCD - 1) / NXCD;
  const unsigned tall = total - (per - 1) * NXCD;   // XCDs that take `per` items
  const unsigned xcd = i % NXCD, loc = i / NXCD;
  const unsigned w = xcd < tall ? xcd * per + loc : tall * per + (xcd - tall) * (per - 1) + loc;
  return B3{w % g.x, (w / g.x) % g.y, w / (g.x * g.y)};
}

template <typename T, int N, bool INV, int TC = cpass_cols<T, N>()>
__device__ __forceinline__ void cpass_body(const CPassArgs<T>& ca, const B3 bid, const B3 gdim) {
  if (ca.gate && *ca.gate == 0.0) return;
  constexpr bool BIG = plan_needs_alt(N);
  using LY = Lds<N, TC, true>;
  __shared__ cx<T> buf[LY::size];
  __shared__ cx<T> alt[BIG ? LY::size : 1];
  __shared__ cx<T> tw[N];
  const B3 item = xcd_item(bid, gdim);
  const int col0 = item.x * TC;
  const int plane = item.y;
  const int comp = item.z;
  const int cols = ca.cols;
  const int A = ca.A, a = ca.a;
  const cx<T>* in = ca.in + comp * ca.in_cstride + plane * ca.in_pstride;
  cx<T>* out = ca.out + comp * ca.out_cstride + plane * ca.out_pstride;
  for (int i = threadIdx.x; i < N; i += NT) tw[i] = ca.tw[i];
  // All of a thread's global loads are issued before the first LDS write, so
  // they are in flight together instead of one round trip each.
  constexpr int PER = (N * TC + NT - 1) / NT;
  cx<T> v[PER];
  if (!INV) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = threadIdx.x + i * NT;
      const int b = e % TC, j = e / TC;
      const int col = col0 + b;
      v[i] = (e < N * TC && col < cols) ? in[(size_t)j * cols + col] : cx<T>{T(0), T(0)};
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = threadIdx.x + i * NT;
      if (e < N * TC) buf[LY::idx(e % TC, e / TC)] = v[i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = threadIdx.x + i * NT;
      const int b = e % TC, jc = e / TC;
      const int col = col0 + b;
      v[i] = (e < A * TC && col < cols) ? in[(size_t)jc * cols + col] : cx<T>{T(0), T(0)};
    }
    {
      // rows the pruned inverse transform never reads stay as they are
      const ZeroRange<N> zr{a, N - a};
      const int nz = zr.count();
      for (int e = threadIdx.x; e < nz * TC; e += NT) buf[LY::idx(e % TC, zr.at(e / TC))] = cx<T>{T(0), T(0)};
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = threadIdx.x + i * NT;
      if (e < A * TC) {
        const int b = e % TC, jc = e / TC;
        const int jf = jc < a ? jc : jc + N - A;
        buf[LY::idx(b, jf)] = v[i];
      }
    }
  }
  __syncthreads();
  cx<T>* res = fft_lds<T, N, TC, true, INV, NT, INV, !INV>(buf, alt, tw);
  if (!INV) {
    for (int e = threadIdx.x; e < A * TC; e += NT) {
      const int b = e % TC, jc = e / TC;
      const int col = col0 + b;
      const int jf = jc < a ? jc : jc + N - A;
      if (col < cols) out[(size_t)jc * cols + col] = res[LY::idx(b, jf)];
    }
  } else {
    for (int e = threadIdx.x; e < N * TC; e += NT) {
      const int b = e % TC, j = e / TC;
      const int col = col0 + b;
      if (col < cols) out[(size_t)j * cols + col] = res[LY::idx(b, j)];
    }
  }
}

template <typename T, int N, bool INV, int TC = cpass_cols<T, N>()>
__global__ void __launch_bounds__(NT) cpass_kernel(CPassArgs<T> ca) {
  cpass_body<T, N, INV, TC>(ca, PF_BIDX, PF_GDIM);
}
template <typename T, int N, bool INV, int TC = cpass_cols<T, N>()>
struct KCPass {
  using Args = CPassArgs<T>;
  static constexpr int BLOCK = NT;
  __device__ static void body(const Args& a, B3 bid, B3 g) { cpass_body<T, N, INV, TC>(a, bid, g); }
  static void direct(hipStream_t s, B3 g, const Args& a) {
    hipLaunchKernelGGL((cpass_kernel<T, N, INV, TC>), dim3(g.x, g.y, g.z), dim3(NT), 0, s, a);
  }
};

// Column-tile widths the plan tuner tries (0: the width cpass_cols picks),
// and whether a width fits the workgroup's LDS (at most 96 KB with the
// twiddles).
constexpr int CPASS_TILES[4] = {0, 8, 16, 32};
template <typename T, int N, int TC>
constexpr bool cpass_tile_fits() {
  return TC == 0 || (Lds<N, TC, true>::size * (plan_needs_alt(N) ? 2 : 1) + N) * (int)sizeof(cx<T>) <= 96 * 1024;
}

// z-pass LDS budgets (bytes of FFT data per workgroup) the plan tuner tries
// for the solver's z pass; index 1 is the default
constexpr int ZPASS_BUDGETS[4] = {16 * 1024, ZPASS_LDS_BUDGET, 64 * 1024, 96 * 1024};

}  // namespace pf
