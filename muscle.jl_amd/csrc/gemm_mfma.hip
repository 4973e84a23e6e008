// Gather-GEMM on CDNA4 matrix cores.
//
// C[bC + rowC[m] + colC[n]] = sum_k A[bA + rowA[m] + kA[k]] * B[bB + kB[k] + colB[n]]
//
// This replaces BackendBase's permutedims(a) / permutedims(b) / BLAS gemm /
// permutedims(c) sequence (src/Operations/binary_einsum.jl:89-95) with one
// kernel: operand tiles are gathered from their strided layout straight into
// LDS through per-mode offset tables, multiplied on MFMA, and scattered to C's
// own layout in the epilogue.
//
// Complex products run on real MFMAs either as 4M
//   re += ar*br - ai*bi,  im += ar*bi + ai*br               (4 MFMAs)
// or as 3M (Gauss):
//   P1 += ar*br, P2 += ai*bi, P3 += (ar+ai)*(br+bi);
//   re = P1 - P2, im = P3 - P1 - P2                         (3 MFMAs)
//
// Tile: 64 (M) x 64 (N) x BK (K) per 256-thread workgroup, four waves in a
// 2 x 2 arrangement, each wave a 32 x 32 block of 2 x 2 MFMA 16x16x4 tiles.
// Operands are staged in LDS as separate real and imaginary planes, double
// buffered, so one barrier per K-tile suffices.  The global gather of K-tile
// t+1 and the offset-table reads of K-tile t+2 are in flight while K-tile t
// is multiplied.  Rows of 64 + 16 elements put the four 16-lane groups of an
// MFMA operand read in disjoint LDS banks, and the column of element (k, m)
// is m ^ (k & 15) (a permutation inside each aligned 16-block): the 16 lanes
// that store 16 consecutive k of one row (operands contiguous along K) then
// hit 16 different banks instead of one.  Workgroups are renumbered so each
// XCD (the dispatcher deals workgroups round-robin over 8 XCDs) gets a
// contiguous, grouped range of tiles and reuses A and B panels in its own L2.
//
// v_mfma_{f64_16x16x4f64, f32_16x16x4f32} lane layout on gfx950 (64 lanes),
// established with exact small-integer operands on the device:
//   A operand: lane l holds A[i = l % 16][k = l / 16]
//   B operand: lane l holds B[k = l / 16][j = l % 16]
//   D (4 regs), f32: lane l, reg r holds D[i = 4 * (l / 16) + r][j = l % 16]
//   D (4 regs), f64: lane l, reg r holds D[i = l / 16 + 4 * r][j = l % 16]
#include "device_types.hpp"
#include "kernels.hpp"

namespace mh {
namespace {

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ inline acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ inline acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int BM = 64, BN = 64, NT = 256;
constexpr int LDP = 16;  // LDS row padding (elements)


// T: arithmetic type (MFMA precision); S: storage type of A, B and C.  S = T
// is the plain path; S = double with T = float is the FP32 rung of the
// precision ladder (operands rounded to FP32 as they are gathered).
// MAJ: bit 1 set if A runs contiguous along K (else along M), bit 0 the same
// for B (else along N), or -1 to read both from the arguments.  Fixed at
// compile time, the K-offset (or row-offset) registers that every element of
// a thread shares collapse into one (FP64: 233 -> 227 VGPRs, config 2 from
// 68 to 70.5 TF; ComplexF32: 160 -> 142..154 VGPRs, config 2 from 128 to
// 131 TF and the dims-2 4096-cube from 119 to 126 TF).  The FP32 rung of the
// ladder keeps the run-time form, which measured the same there.
template <typename T, typename S, bool CA, bool CB, bool G3, int BK, int MAJ>
__global__ __launch_bounds__(NT, 2) void gemm_tile_kernel(GemmArgs p) {
  const bool AK = MAJ < 0 ? p.a_kmajor != 0 : (MAJ & 2) != 0;
  const bool BKM = MAJ < 0 ? p.b_kmajor != 0 : (MAJ & 1) != 0;
  using M_ = Mfma<T>;
  using acc_t = typename M_::acc_t;
  constexpr int NE = BK / 4;  // elements of A (and of B) each thread gathers per K-tile
  constexpr int PA = CA ? 2 : 1, PB = CB ? 2 : 1;
  constexpr bool CC = CA || CB;
  static_assert(!G3 || (CA && CB), "3M needs two complex operands");
  __shared__ T As[2][PA][BK][BM + LDP];
  __shared__ T Bs[2][PB][BK][BN + LDP];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  int tm, tn;
  gemm_tile_coords(blockIdx.x, gridDim.x, p.tiles_m, p.tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int bz = blockIdx.y;
  const int64_t offA0 = p.bA[bz], offB0 = p.bB[bz], offC0 = p.bC[bz];
  // this workgroup's K range (split-K slice)
  const int kbeg = blockIdx.z * p.kchunk;
  const int kend = min(p.K, kbeg + p.kchunk);

  // which (m|n, k) of the tile each thread gathers
  int a_mm[NE], a_kk[NE], b_nn[NE], b_kk[NE];
#pragma unroll
  for (int i = 0; i < NE; ++i) {
    const int e = tid + NT * i;
    if (AK) {
      a_kk[i] = e % BK;
      a_mm[i] = e / BK;
    } else {
      a_mm[i] = e & (BM - 1);
      a_kk[i] = e >> 6;
    }
    if (BKM) {
      b_kk[i] = e % BK;
      b_nn[i] = e / BK;
    } else {
      b_nn[i] = e & (BN - 1);
      b_kk[i] = e >> 6;
    }
  }
  int64_t rowA_r[NE], colB_r[NE];
  bool a_ok[NE], b_ok[NE];
#pragma unroll
  for (int i = 0; i < NE; ++i) {
    const int m = m0 + a_mm[i], n = n0 + b_nn[i];
    a_ok[i] = m < p.M;
    b_ok[i] = n < p.N;
    rowA_r[i] = p.rowA[a_ok[i] ? m : p.M - 1] + offA0;
    colB_r[i] = p.colB[b_ok[i] ? n : p.N - 1] + offB0;
    a_mm[i] ^= a_kk[i] & 15;  // from here on: the swizzled LDS columns
    b_nn[i] ^= b_kk[i] & 15;
  }
  const int nk = (kend - kbeg + BK - 1) / BK;

  // K-offset table entries of the tile gathered next (indices clamped in range)
  int64_t tA[NE], tB[NE];
  auto load_tables = [&](int kt) {
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      tA[i] = p.kA[min(kbeg + kt * BK + a_kk[i], kend - 1)];
      tB[i] = p.kB[min(kbeg + kt * BK + b_kk[i], kend - 1)];
    }
  };
  Cplx<T> ra[NE], rb[NE];
  // a tile whose rows and columns all lie inside M and N needs no per-element
  // checks on the K-tiles that lie inside its K range (all but the last);
  // FP64 data on FP32 MFMA keeps the checked form, which measured faster there
  const bool full_mn = sizeof(T) == sizeof(S) && m0 + BM <= p.M && n0 + BN <= p.N;
  auto gather = [&](int kt) {
    if (full_mn && kbeg + (kt + 1) * BK <= kend) {
#pragma unroll
      for (int i = 0; i < NE; ++i) {
        const Cplx<S> xa = load_elem<S, CA>(p.A, rowA_r[i] + tA[i]);
        const Cplx<S> xb = load_elem<S, CB>(p.B, colB_r[i] + tB[i]);
        ra[i] = Cplx<T>{(T)xa.re, (T)xa.im};
        rb[i] = Cplx<T>{(T)xb.re, (T)xb.im};
      }
      return;
    }
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      const Cplx<S> xa = load_elem<S, CA>(p.A, rowA_r[i] + tA[i]);
      const Cplx<S> xb = load_elem<S, CB>(p.B, colB_r[i] + tB[i]);
      ra[i] = Cplx<T>{(T)xa.re, (T)xa.im};
      rb[i] = Cplx<T>{(T)xb.re, (T)xb.im};
      const bool ka = kbeg + kt * BK + a_kk[i] < kend, kb = kbeg + kt * BK + b_kk[i] < kend;
      if (!(a_ok[i] && ka)) ra[i] = Cplx<T>{T(0), T(0)};
      if (!(b_ok[i] && kb)) rb[i] = Cplx<T>{T(0), T(0)};
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      // a_mm / b_nn hold the swizzled LDS columns by now
      As[buf][0][a_kk[i]][a_mm[i]] = ra[i].re;
      if constexpr (CA) As[buf][PA - 1][a_kk[i]][a_mm[i]] = ra[i].im;
      Bs[buf][0][b_kk[i]][b_nn[i]] = rb[i].re;
      if constexpr (CB) Bs[buf][PB - 1][b_kk[i]][b_nn[i]] = rb[i].im;
    }
  };

  acc_t acc0[2][2], acc1[2][2], acc2[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        acc0[i][j][r] = T(0);
        acc1[i][j][r] = T(0);
        acc2[i][j][r] = T(0);
      }

  load_tables(0);
  gather(0);
  stash(0);
  if (nk > 1) load_tables(1);
  __syncthreads();

  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gather(kt + 1);
    if (kt + 2 < nk) load_tables(kt + 2);
    // FP64: the MFMA section issues at raised wave priority, so a wave that is
    // multiplying is not held back by the other workgroup's waves issuing
    // their gathers and LDS stores (config 2: 71.1 -> 76.2 TF)
    if constexpr (sizeof(T) == 8) __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      T ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int mm = (wm * 32 + i * 16 + (lane & 15)) ^ (kr & 15);
        ar[i] = As[cur][0][kr][mm];
        ai[i] = CA ? As[cur][PA - 1][kr][mm] : T(0);
        const int nn = (wn * 32 + i * 16 + (lane & 15)) ^ (kr & 15);
        br[i] = Bs[cur][0][kr][nn];
        bi[i] = CB ? Bs[cur][PB - 1][kr][nn] : T(0);
      }
      if constexpr (G3) {
        T as[2], bs[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          as[i] = ar[i] + ai[i];
          bs[i] = br[i] + bi[i];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            acc0[i][j] = M_::mma(ar[i], br[j], acc0[i][j]);
            acc1[i][j] = M_::mma(ai[i], bi[j], acc1[i][j]);
            acc2[i][j] = M_::mma(as[i], bs[j], acc2[i][j]);
          }
      } else {
        // acc0: real part, acc1: imaginary part
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            acc0[i][j] = M_::mma(ar[i], br[j], acc0[i][j]);
            if constexpr (CA && CB) acc0[i][j] = M_::mma(-ai[i], bi[j], acc0[i][j]);
            if constexpr (CB) acc1[i][j] = M_::mma(ar[i], bi[j], acc1[i][j]);
            if constexpr (CA) acc1[i][j] = M_::mma(ai[i], br[j], acc1[i][j]);
          }
      }
    }
    if constexpr (sizeof(T) == 8) __builtin_amdgcn_s_setprio(0);
    if (kt + 1 < nk) stash(cur ^ 1);
    __syncthreads();
  }

  // partial sums of a split-K slice go to the workspace, [split][batch][m][n]
  const bool to_partial = p.ksplit > 1;
  const int64_t pbase = ((int64_t)blockIdx.z * gridDim.y + bz) * (int64_t)p.M * p.N;
  // C's offsets for this thread's rows and columns, read before any store
  // (a table load after a store to C could not be hoisted: they may alias)
  int64_t crow[2][4], ccol[2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m0 + wm * 32 + i * 16 + M_::row(lane, r);
      crow[i][r] = (!to_partial && m < p.M) ? p.rowC[m] : 0;
    }
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int n = n0 + wn * 32 + j * 16 + (lane & 15);
    ccol[j] = n < p.N ? (to_partial ? pbase + n : offC0 + p.colC[n]) : 0;
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn * 32 + j * 16 + (lane & 15);
      if (n >= p.N) continue;
      const int64_t cn = ccol[j];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * 32 + i * 16 + M_::row(lane, r);
        if (m >= p.M) continue;
        T re, im;
        if constexpr (G3) {
          re = acc0[i][j][r] - acc1[i][j][r];
          im = acc2[i][j][r] - acc0[i][j][r] - acc1[i][j][r];
        } else {
          re = acc0[i][j][r];
          im = acc1[i][j][r];
        }
        if (to_partial) store_elem<T, CC>(p.partial, cn + (int64_t)m * p.N, Cplx<T>{re, im});
        else if (p.c_stream) store_elem_nt<S, CC>(p.C, cn + crow[i][r], Cplx<S>{(S)re, (S)im});
        else store_elem<S, CC>(p.C, cn + crow[i][r], Cplx<S>{(S)re, (S)im});
      }
    }
}

// C[bz, m, n] = sum over slices s of partial[s][bz][m][n], in slice order.
// Small (64-thread) workgroups spread the few outputs of a split-K GEMM over
// many CUs, and the unrolled slice loop keeps 8 loads of each thread in flight.
constexpr int NTF = 64;
template <typename T, typename S, bool CC>
__global__ __launch_bounds__(NTF) void splitk_finish_kernel(GemmArgs p) {
  const int64_t mn = (int64_t)p.M * p.N;
  const int64_t total = mn * p.batch;
  const int64_t o = (int64_t)blockIdx.x * NTF + threadIdx.x;
  if (o >= total) return;
  T re = 0, im = 0;
#pragma unroll 8
  for (int s = 0; s < p.ksplit; ++s) {
    const Cplx<T> v = load_elem<T, CC>(p.partial, (int64_t)s * total + o);
    re += v.re;
    im += v.im;
  }
  const int bz = (int)(o / mn);
  const int64_t r = o - bz * mn;
  const int m = (int)(r / p.N), n = (int)(r - (int64_t)m * p.N);
  store_elem<S, CC>(p.C, p.bC[bz] + p.rowC[m] + p.colC[n], Cplx<S>{(S)re, (S)im});
}

template <typename T, typename S, bool CA, bool CB, bool G3>
void launch_t(const GemmArgs& g, hipStream_t s) {
  dim3 grid((unsigned)(g.tiles_m * g.tiles_n), (unsigned)g.batch, (unsigned)g.ksplit);
  if constexpr (sizeof(T) != sizeof(S)) {
    hipLaunchKernelGGL((gemm_tile_kernel<T, S, CA, CB, G3, 16, -1>), grid, dim3(NT), 0, s, g);
  } else {
    switch ((g.a_kmajor ? 2 : 0) | (g.b_kmajor ? 1 : 0)) {
      case 0: hipLaunchKernelGGL((gemm_tile_kernel<T, S, CA, CB, G3, 16, 0>), grid, dim3(NT), 0, s, g); break;
      case 1: hipLaunchKernelGGL((gemm_tile_kernel<T, S, CA, CB, G3, 16, 1>), grid, dim3(NT), 0, s, g); break;
      case 2: hipLaunchKernelGGL((gemm_tile_kernel<T, S, CA, CB, G3, 16, 2>), grid, dim3(NT), 0, s, g); break;
      default: hipLaunchKernelGGL((gemm_tile_kernel<T, S, CA, CB, G3, 16, 3>), grid, dim3(NT), 0, s, g); break;
    }
  }
  if (g.ksplit > 1) {
    const int64_t total = (int64_t)g.M * g.N * g.batch;
    hipLaunchKernelGGL((splitk_finish_kernel<T, S, CA || CB>), dim3((unsigned)((total + NTF - 1) / NTF)), dim3(NTF),
                       0, s, g);
  }
}

}  // namespace

void launch_gemm_mfma(bool dbl, bool fp32_math, bool ca, bool cb, const GemmArgs& g0, hipStream_t s) {
  GemmArgs g = g0;
  g.tiles_m = (g.M + BM - 1) / BM;
  g.tiles_n = (g.N + BN - 1) / BN;
  if (g.ksplit < 1) {
    g.ksplit = 1;
    g.kchunk = g.K;
  }
  const bool g3 = ca && cb && g.complex3m;
  const int key = (dbl ? 8 : 0) | (ca ? 4 : 0) | (cb ? 2 : 0) | (g3 ? 1 : 0);
  if (dbl && fp32_math) {
    switch (key) {
      case 8: launch_t<float, double, false, false, false>(g, s); break;
      case 10: launch_t<float, double, false, true, false>(g, s); break;
      case 12: launch_t<float, double, true, false, false>(g, s); break;
      case 14: launch_t<float, double, true, true, false>(g, s); break;
      case 15: launch_t<float, double, true, true, true>(g, s); break;
    }
    return;
  }
  switch (key) {
    case 0: launch_t<float, float, false, false, false>(g, s); break;
    case 2: launch_t<float, float, false, true, false>(g, s); break;
    case 4: launch_t<float, float, true, false, false>(g, s); break;
    case 6: launch_t<float, float, true, true, false>(g, s); break;
    case 7: launch_t<float, float, true, true, true>(g, s); break;
    case 8: launch_t<double, double, false, false, false>(g, s); break;
    case 10: launch_t<double, double, false, true, false>(g, s); break;
    case 12: launch_t<double, double, true, false, false>(g, s); break;
    case 14: launch_t<double, double, true, true, false>(g, s); break;
    case 15: launch_t<double, double, true, true, true>(g, s); break;
  }
}

}  // namespace mh
