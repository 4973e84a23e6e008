// Gather-GEMM on CDNA4 matrix cores.
//
// C[bC + rowC[m] + colC[n]] = sum_k A[bA + rowA[m] + kA[k]] * B[bB + kB[k] + colB[n]]
//
// This replaces BackendBase's permutedims(a) / permutedims(b) / BLAS gemm /
// permutedims(c) sequence (src/Operations/binary_einsum.jl:89-95) with one
// kernel: operand tiles are gathered from their strided layout straight into
// LDS through per-mode offset tables, multiplied on MFMA, and scattered to C's
// own layout in the epilogue.  Complex products use four real MFMAs (4M):
//   re += ar*br - ai*bi,  im += ar*bi + ai*br.
//
// Tile: 64 (M) x 64 (N) x 16 (K) per 256-thread workgroup, four waves in a
// 2 x 2 arrangement, each wave a 32 x 32 block of 2 x 2 MFMA 16x16x4 tiles.
// Operands are staged in LDS as separate real and imaginary planes so each
// MFMA operand fetch is one ds_read of a T.  The next K-tile is fetched from
// global memory into registers while the current one is multiplied.
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
  // row of D held by (lane, reg)
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

constexpr int BM = 64, BN = 64, BK = 16, NT = 256;
constexpr int LDP = 4;  // LDS row padding (elements)

template <typename T, bool CA, bool CB>
__global__ __launch_bounds__(NT) void gemm_tile_kernel(GemmArgs p) {
  using M_ = Mfma<T>;
  using acc_t = typename M_::acc_t;
  __shared__ T As[CA ? 2 : 1][BK][BM + LDP];
  __shared__ T Bs[CB ? 2 : 1][BK][BN + LDP];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  // tile coordinates; consecutive workgroups walk N first so neighbours share A rows
  const int tile = blockIdx.x;
  const int tm = tile / p.tiles_n;
  const int tn = tile - tm * p.tiles_n;
  const int m0 = tm * BM, n0 = tn * BN;
  const int bz = blockIdx.y;
  const int64_t offA0 = p.bA[bz], offB0 = p.bB[bz], offC0 = p.bC[bz];

  // global -> register staging: 4 elements of A and 4 of B per thread
  int a_mm[4], a_kk[4], b_nn[4], b_kk[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int e = tid + NT * i;
    if (p.a_kmajor) {
      a_kk[i] = e & (BK - 1);
      a_mm[i] = e >> 4;
    } else {
      a_mm[i] = e & (BM - 1);
      a_kk[i] = e >> 6;
    }
    if (p.b_kmajor) {
      b_kk[i] = e & (BK - 1);
      b_nn[i] = e >> 4;
    } else {
      b_nn[i] = e & (BN - 1);
      b_kk[i] = e >> 6;
    }
  }
  int64_t rowA_r[4], colB_r[4];
  bool a_ok[4], b_ok[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + a_mm[i], n = n0 + b_nn[i];
    a_ok[i] = m < p.M;
    b_ok[i] = n < p.N;
    rowA_r[i] = a_ok[i] ? p.rowA[m] : 0;
    colB_r[i] = b_ok[i] ? p.colB[n] : 0;
  }

  Cplx<T> ra[4], rb[4];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int ka = k0 + a_kk[i];
      ra[i] = Cplx<T>{T(0), T(0)};
      if (a_ok[i] && ka < p.K) ra[i] = load_elem<T, CA>(p.A, offA0 + rowA_r[i] + p.kA[ka]);
      const int kb = k0 + b_kk[i];
      rb[i] = Cplx<T>{T(0), T(0)};
      if (b_ok[i] && kb < p.K) rb[i] = load_elem<T, CB>(p.B, offB0 + colB_r[i] + p.kB[kb]);
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      As[0][a_kk[i]][a_mm[i]] = ra[i].re;
      if constexpr (CA) As[CA ? 1 : 0][a_kk[i]][a_mm[i]] = ra[i].im;
      Bs[0][b_kk[i]][b_nn[i]] = rb[i].re;
      if constexpr (CB) Bs[CB ? 1 : 0][b_kk[i]][b_nn[i]] = rb[i].im;
    }
  };

  acc_t accr[2][2], acci[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        accr[i][j][r] = T(0);
        acci[i][j][r] = T(0);
      }

  const int nk = (p.K + BK - 1) / BK;
  fetch(0);
  stash();
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    if (kt + 1 < nk) fetch((kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      T ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int mm = wm * 32 + i * 16 + (lane & 15);
        ar[i] = As[0][kr][mm];
        ai[i] = CA ? -As[CA ? 1 : 0][kr][mm] : T(0);  // negated: feeds re -= ai * bi
        const int nn = wn * 32 + i * 16 + (lane & 15);
        br[i] = Bs[0][kr][nn];
        bi[i] = CB ? Bs[CB ? 1 : 0][kr][nn] : T(0);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          accr[i][j] = M_::mma(ar[i], br[j], accr[i][j]);
          if constexpr (CA && CB) accr[i][j] = M_::mma(ai[i], bi[j], accr[i][j]);
          if constexpr (CB) acci[i][j] = M_::mma(ar[i], bi[j], acci[i][j]);
          if constexpr (CA) acci[i][j] = M_::mma(-ai[i], br[j], acci[i][j]);
        }
    }
    __syncthreads();
    if (kt + 1 < nk) {
      stash();
      __syncthreads();
    }
  }

  constexpr bool CC = CA || CB;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn * 32 + j * 16 + (lane & 15);
      if (n >= p.N) continue;
      const int64_t cn = offC0 + p.colC[n];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * 32 + i * 16 + M_::row(lane, r);
        if (m < p.M) store_elem<T, CC>(p.C, cn + p.rowC[m], Cplx<T>{accr[i][j][r], acci[i][j][r]});
      }
    }
}

template <typename T, bool CA, bool CB>
void launch_t(const GemmArgs& g, hipStream_t s) {
  dim3 grid((unsigned)(g.tiles_m * g.tiles_n), (unsigned)g.batch);
  hipLaunchKernelGGL((gemm_tile_kernel<T, CA, CB>), grid, dim3(NT), 0, s, g);
}

}  // namespace

void launch_gemm_mfma(bool dbl, bool ca, bool cb, const GemmArgs& g0, hipStream_t s) {
  GemmArgs g = g0;
  g.tiles_m = (g.M + BM - 1) / BM;
  g.tiles_n = (g.N + BN - 1) / BN;
  const int key = (dbl ? 4 : 0) | (ca ? 2 : 0) | (cb ? 1 : 0);
  switch (key) {
    case 0: launch_t<float, false, false>(g, s); break;
    case 1: launch_t<float, false, true>(g, s); break;
    case 2: launch_t<float, true, false>(g, s); break;
    case 3: launch_t<float, true, true>(g, s); break;
    case 4: launch_t<double, false, false>(g, s); break;
    case 5: launch_t<double, false, true>(g, s); break;
    case 6: launch_t<double, true, false>(g, s); break;
    case 7: launch_t<double, true, true>(g, s); break;
  }
}

}  // namespace mh
