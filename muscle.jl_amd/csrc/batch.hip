// Persistent batched contraction kernel.
//
// Thousands of uneven pairwise contractions (the pairs of a contraction
// tree) run in one launch: a fixed grid of workgroups (2 per CU) pulls
// output tiles from a global work queue (one atomic per tile), largest pairs
// first, until the queue is drained -- every workgroup reaches the exit test
// with the same item number.  Each tile is a gather-GEMM tile like the plain
// kernel's (offset tables into LDS, MFMA 16x16x4, scatter to C), with a tile
// shape chosen per pair (64x64, 32x128 or 128x32) so skinny contractions do
// not waste the workgroup.  The 64 x 64 shape double-buffers its LDS tiles,
// the skinny shapes use one buffer plus a register prefetch of the next
// K-tile; all shapes use 4 waves x 4 MFMA tiles, so register use is the same
// for every shape.
#include "batch.hpp"
#include "device_types.hpp"

namespace mh {
namespace {

constexpr int NT = 256, BK = 16, LDP = 16;

template <typename T>
struct Mma;
template <>
struct Mma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ inline acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ inline acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// DB: double-buffered LDS (one barrier per K-tile, the gather of K-tile
// t + 1 in flight while t is multiplied); the 64 x 64 shape uses it, the
// skinny shapes keep one buffer so every shape fits the same LDS budget.
template <typename T, bool CP, bool G3, int BM, int BN, int WM, int WN, bool DB>
__device__ __forceinline__ void batch_tile(const BatchPair& pr, int bz, int tm, int tn, T* lds) {
  using MM = Mma<T>;
  using acc_t = typename MM::acc_t;
  constexpr int PL = CP ? 2 : 1;
  constexpr int NEA = BM * BK / NT, NEB = BN * BK / NT;
  constexpr int TM = BM / WM / 16, TN = BN / WN / 16;
  constexpr int LA = BM + LDP, LB = BN + LDP;
  constexpr int STAGE = PL * BK * (LA + LB);  // elements of one LDS buffer
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int m0 = tm * BM, n0 = tn * BN;
  const int M = pr.M, N = pr.N, K = pr.K;
  const int64_t offA0 = pr.bA[bz], offB0 = pr.bB[bz], offC0 = pr.bC[bz];

  int a_mm[NEA], a_kk[NEA], b_nn[NEB], b_kk[NEB];
  bool a_ok[NEA], b_ok[NEB];
  int64_t a_row[NEA], b_col[NEB];
#pragma unroll
  for (int i = 0; i < NEA; ++i) {
    const int e = tid + NT * i;
    if (pr.a_kmajor) {
      a_kk[i] = e % BK;
      a_mm[i] = e / BK;
    } else {
      a_mm[i] = e % BM;
      a_kk[i] = e / BM;
    }
    const int m = m0 + a_mm[i];
    a_ok[i] = m < M;
    a_row[i] = pr.rowA[a_ok[i] ? m : M - 1] + offA0;
  }
#pragma unroll
  for (int i = 0; i < NEB; ++i) {
    const int e = tid + NT * i;
    if (pr.b_kmajor) {
      b_kk[i] = e % BK;
      b_nn[i] = e / BK;
    } else {
      b_nn[i] = e % BN;
      b_kk[i] = e / BN;
    }
    const int n = n0 + b_nn[i];
    b_ok[i] = n < N;
    b_col[i] = pr.colB[b_ok[i] ? n : N - 1] + offB0;
  }
  Cplx<T> ra[NEA], rb[NEB];
  acc_t acc0[TM][TN], acc1[TM][TN], acc2[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc0[i][j][r] = acc1[i][j][r] = acc2[i][j][r] = T(0);

#define MH_GATHER(kt)                                                            \
  do {                                                                           \
    _Pragma("unroll") for (int i = 0; i < NEA; ++i) {                            \
      const int ka = (kt) * BK + a_kk[i];                                        \
      const Cplx<T> v = load_elem<T, CP>(pr.A, a_row[i] + pr.kA[min(ka, K - 1)]); \
      ra[i] = (a_ok[i] && ka < K) ? v : Cplx<T>{T(0), T(0)};                     \
    }                                                                            \
    _Pragma("unroll") for (int i = 0; i < NEB; ++i) {                            \
      const int kb = (kt) * BK + b_kk[i];                                        \
      const Cplx<T> v = load_elem<T, CP>(pr.B, b_col[i] + pr.kB[min(kb, K - 1)]); \
      rb[i] = (b_ok[i] && kb < K) ? v : Cplx<T>{T(0), T(0)};                     \
    }                                                                            \
  } while (0)
#define MH_STASH(buf)                                                            \
  do {                                                                           \
    T* As = lds + (buf) * STAGE;                                                 \
    T* Bs = As + PL * BK * LA;                                                   \
    _Pragma("unroll") for (int i = 0; i < NEA; ++i) {                            \
      As[a_kk[i] * LA + a_mm[i]] = ra[i].re;                                     \
      if constexpr (CP) As[(BK + a_kk[i]) * LA + a_mm[i]] = ra[i].im;            \
    }                                                                            \
    _Pragma("unroll") for (int i = 0; i < NEB; ++i) {                            \
      Bs[b_kk[i] * LB + b_nn[i]] = rb[i].re;                                     \
      if constexpr (CP) Bs[(BK + b_kk[i]) * LB + b_nn[i]] = rb[i].im;            \
    }                                                                            \
  } while (0)
#define MH_COMPUTE(buf)                                                          \
  do {                                                                           \
    const T* As = lds + (buf) * STAGE;                                           \
    const T* Bs = As + PL * BK * LA;                                             \
    _Pragma("unroll") for (int ks = 0; ks < BK / 4; ++ks) {                      \
      const int kr = ks * 4 + (lane >> 4);                                       \
      T ar[TM], ai[TM], br[TN], bi[TN];                                          \
      _Pragma("unroll") for (int i = 0; i < TM; ++i) {                           \
        const int mm = wm * (BM / WM) + i * 16 + (lane & 15);                    \
        ar[i] = As[kr * LA + mm];                                                \
        ai[i] = CP ? As[(BK + kr) * LA + mm] : T(0);                             \
      }                                                                          \
      _Pragma("unroll") for (int j = 0; j < TN; ++j) {                           \
        const int nn = wn * (BN / WN) + j * 16 + (lane & 15);                    \
        br[j] = Bs[kr * LB + nn];                                                \
        bi[j] = CP ? Bs[(BK + kr) * LB + nn] : T(0);                             \
      }                                                                          \
      _Pragma("unroll") for (int i = 0; i < TM; ++i)                             \
        _Pragma("unroll") for (int j = 0; j < TN; ++j) {                         \
          if constexpr (G3) {                                                    \
            acc0[i][j] = MM::mma(ar[i], br[j], acc0[i][j]);                      \
            acc1[i][j] = MM::mma(ai[i], bi[j], acc1[i][j]);                      \
            acc2[i][j] = MM::mma(ar[i] + ai[i], br[j] + bi[j], acc2[i][j]);      \
          } else {                                                               \
            acc0[i][j] = MM::mma(ar[i], br[j], acc0[i][j]);                      \
            if constexpr (CP) {                                                  \
              acc0[i][j] = MM::mma(-ai[i], bi[j], acc0[i][j]);                   \
              acc1[i][j] = MM::mma(ar[i], bi[j], acc1[i][j]);                    \
              acc1[i][j] = MM::mma(ai[i], br[j], acc1[i][j]);                    \
            }                                                                    \
          }                                                                      \
        }                                                                        \
    }                                                                            \
  } while (0)

  const int nk = (K + BK - 1) / BK;
  if (nk > 0) MH_GATHER(0);
  if constexpr (DB) {
    __syncthreads();  // the previous tile's readers of buffer 0 are done
    if (nk > 0) MH_STASH(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
      const int cur = kt & 1;
      if (kt + 1 < nk) MH_GATHER(kt + 1);
      MH_COMPUTE(cur);
      if (kt + 1 < nk) MH_STASH(cur ^ 1);
      __syncthreads();
    }
  } else {
    for (int kt = 0; kt < nk; ++kt) {
      __syncthreads();  // earlier readers of the LDS tile are done
      MH_STASH(0);
      __syncthreads();
      if (kt + 1 < nk) MH_GATHER(kt + 1);
      MH_COMPUTE(0);
    }
  }
#undef MH_GATHER
#undef MH_STASH
#undef MH_COMPUTE
  // C's offsets for this thread's rows and columns, read before any store
  // (a table load after a store to C could not be hoisted: they may alias)
  int64_t crow[TM][4], ccol[TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m0 + wm * (BM / WM) + i * 16 + MM::row(lane, r);
      crow[i][r] = m < M ? pr.rowC[m] : 0;
    }
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int n = n0 + wn * (BN / WN) + j * 16 + (lane & 15);
    ccol[j] = n < N ? offC0 + pr.colC[n] : 0;
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int n = n0 + wn * (BN / WN) + j * 16 + (lane & 15);
      if (n >= N) continue;
      const int64_t cn = ccol[j];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * (BM / WM) + i * 16 + MM::row(lane, r);
        if (m >= M) continue;
        T re, im;
        if constexpr (G3) {
          re = acc0[i][j][r] - acc1[i][j][r];
          im = acc2[i][j][r] - acc0[i][j][r] - acc1[i][j][r];
        } else {
          re = acc0[i][j][r];
          im = acc1[i][j][r];
        }
        store_elem<T, CP>(pr.C, cn + crow[i][r], Cplx<T>{re, im});
      }
    }
}

__global__ void reset_counter_kernel(int32_t* c) { *c = 0; }

template <typename T, bool CP, bool G3>
__global__ __launch_bounds__(NT, kBatchWgPerCu) void batch_kernel(BatchArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  // the broadcast word lives in dynamic LDS no tile shape of the launch
  // touches (see batch_item_offset)
  int* s_item = reinterpret_cast<int*>(smem_raw + a.item_off);
  for (;;) {
    if (threadIdx.x == 0) *s_item = atomicAdd(a.counter, 1);
    __syncthreads();
    const int64_t it = *s_item;
    __syncthreads();  // every thread holds the item before thread 0 fetches the next
    if (it >= a.nitems) break;
    const int2 w = a.items[it];
    const BatchPair& pr = a.pairs[w.x];
    const int per_batch = pr.tiles_m * pr.tiles_n;
    const int bz = w.y / per_batch;
    const int r = w.y - bz * per_batch;
    const int tm = r / pr.tiles_n, tn = r - tm * pr.tiles_n;
    switch (pr.tile) {
      case 0: batch_tile<T, CP, G3, 64, 64, 2, 2, true>(pr, bz, tm, tn, lds); break;
      case 1: batch_tile<T, CP, G3, 32, 128, 1, 4, false>(pr, bz, tm, tn, lds); break;
      default: batch_tile<T, CP, G3, 128, 32, 4, 1, false>(pr, bz, tm, tn, lds); break;
    }
  }
}

template <typename T, bool CP, bool G3>
void launch_t(const BatchArgs& a, int wg, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, a.counter);
  hipLaunchKernelGGL((batch_kernel<T, CP, G3>), dim3(wg), dim3(NT), lds, s, a);
}

}  // namespace

size_t batch_item_offset(bool dbl, bool cplx, size_t& lds_bytes) {
  const size_t budget = (size_t(160) << 10) / kBatchWgPerCu;
  if (lds_bytes + 16 <= budget) {
    const size_t off = (lds_bytes + 15) & ~size_t(15);
    lds_bytes = off + 16;
    return off;
  }
  // only the double-buffered 64 x 64 complex FP64 tile fills the budget: use
  // a padding element of its second buffer's B row 0 (column 64), beyond
  // the end of the single-buffered shapes' tiles
  const size_t el = dbl ? 8 : 4, pl = cplx ? 2 : 1;
  const size_t stage = pl * BK * ((size_t)(64 + LDP) + (64 + LDP));
  return (stage + pl * BK * (64 + LDP) + 64) * el;
}

size_t batch_lds_bytes(bool dbl, bool cplx, int tile) {
  const TileShape& t = kTileShapes[tile];
  const size_t el = dbl ? 8 : 4, pl = cplx ? 2 : 1;
  const size_t bufs = tile == MUSCLE_HIP_TILE_64x64 ? 2 : 1;  // the double-buffered shape
  return bufs * pl * BK * ((size_t)(t.bm + LDP) + (t.bn + LDP)) * el;
}

void launch_batch(bool dbl, bool cplx, const BatchArgs& a, int wg, size_t lds, hipStream_t s) {
  const bool g3 = cplx && a.complex3m;
  if (dbl) {
    if (g3) launch_t<double, true, true>(a, wg, lds, s);
    else if (cplx) launch_t<double, true, false>(a, wg, lds, s);
    else launch_t<double, false, false>(a, wg, lds, s);
  } else {
    if (g3) launch_t<float, true, true>(a, wg, lds, s);
    else if (cplx) launch_t<float, true, false>(a, wg, lds, s);
    else launch_t<float, false, false>(a, wg, lds, s);
  }
}

}  // namespace mh
