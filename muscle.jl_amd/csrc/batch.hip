// Batched contraction kernels: the persistent MFMA kernel (pairs with K > 16),
// the shallow vector-FMA kernel (pairs with K <= 16) and the skinny vector-FMA
// kernel (pairs with an output side of at most 16), the last two further below.
//
// Thousands of uneven pairwise contractions (the pairs of a contraction
// tree) run in one launch: a fixed grid of workgroups (2 per CU) pulls
// output tiles from work queues (one per XCD in a large launch, one atomic per
// tile), largest pairs first, until every queue is drained -- every workgroup
// reaches the exit test with no item left.  Each tile is a gather-GEMM tile like the plain
// kernel's (offset tables into LDS, MFMA 16x16x4, scatter to C), with a tile
// shape chosen per pair (64x64, 32x128 or 128x32) so skinny contractions do
// not waste the workgroup.  The 64 x 64 shape double-buffers its LDS tiles,
// the skinny shapes use one buffer plus a register prefetch of the next
// K-tile; all shapes use 4 waves x 4 MFMA tiles, so register use is the same
// for every shape.
#include "batch.hpp"
#include "device_types.hpp"

#include <algorithm>

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

// The last slice of a split tile to arrive (the partial tiles hold pm x pn
// elements per tile and slice, pm = min(M, BM), pn = min(N, BN)) sums the
// slices in slice order and writes C, so the result does not depend on which
// slice finishes last.  Every thread of the workgroup calls it after writing
// its part of this slice's partial tile.
template <typename T, bool CP, int BM, int BN>
__device__ __forceinline__ void split_tile_fixup(const BatchPair& pr, int bz, int tm, int tn, int* s_flag) {
  const int tid = threadIdx.x;
  const int m0 = tm * BM, n0 = tn * BN;
  const int pm = min(pr.M, BM), pn = min(pr.N, BN);
  const int mr = min(BM, pr.M - m0), nr = min(BN, pr.N - n0);
  const int64_t tile = (int64_t)bz * pr.tiles_m * pr.tiles_n + (int64_t)tm * pr.tiles_n + tn;
  const int64_t pstride = (int64_t)pm * pn, pbase = tile * pr.ksplit * pstride;
  __threadfence();  // this slice's partial is visible device-wide before it is counted
  __syncthreads();
  if (tid == 0) *s_flag = atomicAdd(pr.tilecnt + tile, 1) == pr.ksplit - 1;
  __syncthreads();
  const bool last = *s_flag != 0;
  __syncthreads();  // every thread has read the flag before its word is reused
  if (!last) return;
  __threadfence();  // the other slices' partials are read after their counts
  const int64_t offC = pr.bC[bz];
  for (int e = tid; e < mr * nr; e += NT) {
    const int m = e / nr, n = e - m * nr;
    Cplx<T> acc = {T(0), T(0)};
    for (int q = 0; q < pr.ksplit; ++q) {
      const Cplx<T> v = load_elem<T, CP>(pr.partial, pbase + q * pstride + (int64_t)m * pn + n);
      acc.re += v.re;
      acc.im += v.im;
    }
    if (pr.c_stream) store_elem_nt<T, CP>(pr.C, offC + pr.rowC[m0 + m] + pr.colC[n0 + n], acc);
    else store_elem<T, CP>(pr.C, offC + pr.rowC[m0 + m] + pr.colC[n0 + n], acc);
  }
}

// DB: double-buffered LDS (one barrier per K-tile, the gather of K-tile
// t + 1 in flight while t is multiplied); the 64 x 64 shape uses it, the
// skinny shapes keep one buffer so every shape fits the same LDS budget.
template <typename T, bool CP, bool G3, int BM, int BN, int WM, int WN, bool DB, bool SPLIT>
__device__ __forceinline__ void batch_tile(const BatchPair& pr, int bz, int tm, int tn, int slice, T* lds,
                                           int* s_flag) {
  using MM = Mma<T>;
  using acc_t = typename MM::acc_t;
  constexpr int PL = CP ? 2 : 1;
  constexpr int NEA = BM * BK / NT, NEB = BN * BK / NT;
  constexpr int TM = BM / WM / 16, TN = BN / WN / 16;
  constexpr int LA = BM + LDP, LB = BN + LDP;
  constexpr int STAGE = PL * BK * (LA + LB);  // elements of one LDS buffer
  // LDS column of row (operand index) m at k is m ^ k, a permutation inside
  // each aligned 16-block: a K-contiguous operand's 16 lanes of one m then
  // store to 16 different banks (rows LA apart all start on one bank), and
  // the MFMA reads of one k stay conflict-free.
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int m0 = tm * BM, n0 = tn * BN;
  const int M = pr.M, N = pr.N, K = pr.K;
  // this item's K range: the whole of K, or one slice of a split pair
  const int kbeg = slice * pr.kchunk, kend = min(K, kbeg + pr.kchunk);
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
    a_mm[i] ^= a_kk[i];  // from here on: the swizzled LDS column
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
    b_nn[i] ^= b_kk[i];
  }
  // this lane's MFMA reads at K step ks (rows k = 4 ks + (lane >> 4)): the
  // swizzled column (c ^ k) with c = lane & 15 is (c ^ (lane >> 4)) ^ 4 ks, so
  // row and column together are one lane offset per ks plus constants
  int xa[BK / 4], xb[BK / 4];
#pragma unroll
  for (int ks = 0; ks < BK / 4; ++ks) {
    const int x = ((lane & 15) ^ (lane >> 4)) ^ (ks << 2);
    xa[ks] = (lane >> 4) * LA + x;
    xb[ks] = (lane >> 4) * LB + x;
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
      const Cplx<T> v = load_elem<T, CP>(pr.A, a_row[i] + pr.kA[min(ka, kend - 1)]); \
      ra[i] = (a_ok[i] && ka < kend) ? v : Cplx<T>{T(0), T(0)};                  \
    }                                                                            \
    _Pragma("unroll") for (int i = 0; i < NEB; ++i) {                            \
      const int kb = (kt) * BK + b_kk[i];                                        \
      const Cplx<T> v = load_elem<T, CP>(pr.B, b_col[i] + pr.kB[min(kb, kend - 1)]); \
      rb[i] = (b_ok[i] && kb < kend) ? v : Cplx<T>{T(0), T(0)};                  \
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
      T ar[TM], ai[TM], br[TN], bi[TN];                                          \
      _Pragma("unroll") for (int i = 0; i < TM; ++i) {                           \
        const int o = xa[ks] + ks * 4 * LA + wm * (BM / WM) + i * 16;            \
        ar[i] = As[o];                                                           \
        ai[i] = CP ? As[BK * LA + o] : T(0);                                     \
      }                                                                          \
      _Pragma("unroll") for (int j = 0; j < TN; ++j) {                           \
        const int o = xb[ks] + ks * 4 * LB + wn * (BN / WN) + j * 16;            \
        br[j] = Bs[o];                                                           \
        bi[j] = CP ? Bs[BK * LB + o] : T(0);                                     \
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

  // K-tiles kt0 .. kt1 - 1 (kbeg is a multiple of BK)
  const int kt0 = kbeg / BK, kt1 = (kend + BK - 1) / BK;
  if (kt0 < kt1) MH_GATHER(kt0);
  if constexpr (DB) {
    __syncthreads();  // the previous tile's readers of buffer 0 are done
    if (kt0 < kt1) MH_STASH(0);
    __syncthreads();
    for (int kt = kt0; kt < kt1; ++kt) {
      const int cur = (kt - kt0) & 1;
      if (kt + 1 < kt1) MH_GATHER(kt + 1);
      // FP64: the MFMA section at raised wave priority, as in the plain tiles
      if constexpr (sizeof(T) == 8) __builtin_amdgcn_s_setprio(1);
      MH_COMPUTE(cur);
      if constexpr (sizeof(T) == 8) __builtin_amdgcn_s_setprio(0);
      if (kt + 1 < kt1) MH_STASH(cur ^ 1);
      __syncthreads();
    }
  } else {
    for (int kt = kt0; kt < kt1; ++kt) {
      __syncthreads();  // earlier readers of the LDS tile are done
      MH_STASH(0);
      __syncthreads();
      if (kt + 1 < kt1) MH_GATHER(kt + 1);
      if constexpr (sizeof(T) == 8) __builtin_amdgcn_s_setprio(1);
      MH_COMPUTE(0);
      if constexpr (sizeof(T) == 8) __builtin_amdgcn_s_setprio(0);
    }
  }
#undef MH_GATHER
#undef MH_STASH
#undef MH_COMPUTE
  // C's offsets for this thread's rows and columns, read before any store
  // (a table load after a store to C could not be hoisted: they may alias).
  // One slice of a split tile stores to its partial tile instead (pm x pn,
  // row-major), with the same loop.
  constexpr bool split = SPLIT;
  void* const dst = split ? pr.partial : pr.C;
  const int pn = min(N, BN);
  const int64_t pb = split ? (((int64_t)bz * pr.tiles_m * pr.tiles_n + (int64_t)tm * pr.tiles_n + tn) * pr.ksplit +
                              slice) * ((int64_t)min(M, BM) * pn)
                           : 0;
  int64_t crow[TM][4], ccol[TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int ml = wm * (BM / WM) + i * 16 + MM::row(lane, r), m = m0 + ml;
      crow[i][r] = m < M ? (split ? pb + (int64_t)ml * pn : pr.rowC[m]) : 0;
    }
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int nl = wn * (BN / WN) + j * 16 + (lane & 15), n = n0 + nl;
    ccol[j] = n < N ? (split ? (int64_t)nl : offC0 + pr.colC[n]) : 0;
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
        if (split || !pr.c_stream) store_elem<T, CP>(dst, cn + crow[i][r], Cplx<T>{re, im});  // partials: read back soon
        else store_elem_nt<T, CP>(dst, cn + crow[i][r], Cplx<T>{re, im});
      }
    }
  if constexpr (split) split_tile_fixup<T, CP, BM, BN>(pr, bz, tm, tn, s_flag);
}

// Clears the work counter and the split tiles' arrival counters (a kernel,
// not a memset, so that it can be graph captured).
__global__ void reset_counter_kernel(int32_t* c, int32_t* tilecnt, int64_t n) {
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i0 < kXcds) c[i0] = 0;
  for (int64_t i = i0; i < n; i += (int64_t)gridDim.x * blockDim.x) tilecnt[i] = 0;
}

template <typename T, bool CP, bool G3>
__global__ __launch_bounds__(NT, kBatchWgPerCu) void batch_kernel(BatchArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  // the broadcast word lives in dynamic LDS no tile shape of the launch
  // touches (see batch_item_offset)
  int* s_item = reinterpret_cast<int*>(smem_raw + a.item_off);
  // this workgroup's XCD queue first; once it is drained, the others in turn
  // (a small launch has one queue)
  const int x0 = a.nq > 1 ? blockIdx.x % kXcds : 0;
  int q = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      int item = -1;
      for (; q < a.nq; ++q) {
        const int x = (x0 + q) % a.nq;
        const int i = a.qoff[x] + atomicAdd(&a.counter[x], 1);
        if (i < a.qoff[x + 1]) {
          item = i;
          break;
        }
      }
      *s_item = item;
    }
    __syncthreads();
    const int64_t it = *s_item;
    __syncthreads();  // every thread holds the item before thread 0 fetches the next
    if (it < 0) break;
    const int2 w = a.items[it];
    const BatchPair& pr = a.pairs[w.x];
    const int per_batch = pr.tiles_m * pr.tiles_n;
    // w.y = tile * ksplit + slice
    const int t = w.y / pr.ksplit, slice = w.y - t * pr.ksplit;
    const int bz = t / per_batch;
    const int r = t - bz * per_batch;
    const int tm = r / pr.tiles_n, tn = r - tm * pr.tiles_n;
    switch (pr.tile) {
      case 0: batch_tile<T, CP, G3, 64, 64, 2, 2, true, false>(pr, bz, tm, tn, slice, lds, s_item); break;
      case 1: batch_tile<T, CP, G3, 32, 128, 1, 4, false, false>(pr, bz, tm, tn, slice, lds, s_item); break;
      default: batch_tile<T, CP, G3, 128, 32, 4, 1, false, false>(pr, bz, tm, tn, slice, lds, s_item); break;
    }
  }
}

// K slices of the split pairs: one (tile, slice) item per workgroup pass,
// in a kernel of its own so that the persistent kernel above keeps its
// register allocation.  The partial-tile stores and the last arriver's sum
// are in batch_tile's SPLIT form.
template <typename T, bool CP, bool G3>
__global__ __launch_bounds__(NT, kBatchWgPerCu) void split_kernel(BatchArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  int* s_flag = reinterpret_cast<int*>(smem_raw + a.item_off);
  for (int64_t it = blockIdx.x; it < a.nitems; it += gridDim.x) {
    const int2 w = a.items[it];
    const BatchPair& pr = a.pairs[w.x];
    const int per_batch = pr.tiles_m * pr.tiles_n;
    const int t = w.y / pr.ksplit, slice = w.y - t * pr.ksplit;
    const int bz = t / per_batch;
    const int r = t - bz * per_batch;
    const int tm = r / pr.tiles_n, tn = r - tm * pr.tiles_n;
    switch (pr.tile) {
      case 0: batch_tile<T, CP, G3, 64, 64, 2, 2, true, true>(pr, bz, tm, tn, slice, lds, s_flag); break;
      case 1: batch_tile<T, CP, G3, 32, 128, 1, 4, false, true>(pr, bz, tm, tn, slice, lds, s_flag); break;
      default: batch_tile<T, CP, G3, 128, 32, 4, 1, false, true>(pr, bz, tm, tn, slice, lds, s_flag); break;
    }
  }
}

// ---- shallow pairs --------------------------------------------------------
// A pair whose contracted extent K is at most kShallowK does at most
// 2 * kShallowK real flops per complex output element it writes, so it is
// bound by writing C; the MFMA tile above would spend its time waiting on
// its chain of dependent table and data loads at 2 workgroups per CU.  Here
// one workgroup of 512 threads writes one output tile (the pair's tile shape,
// BM x BN = 4096 elements) with vector FMAs and few registers, so several
// workgroups share a CU and hide each other's latency:
//   1. every offset the tile needs (rows and columns of A, B and C, the K
//      offsets, the batch offsets) is looked up once, one lookup per thread,
//      into LDS;
//   2. per pass of SKC K elements, A's BM x SKC and B's SKC x BN elements are
//      gathered into LDS (k-major, so either operand is read along its row
//      or column without bank conflicts);
//   3. each thread owns 8 outputs along one line of C: along a row of C
//      (fixed m, 8 columns) when C is contiguous along M, else along a
//      column; its own operand element is read once per k and the other
//      operand's 8 elements are wave-wide LDS broadcasts;
//   4. C is stored with consecutive lanes on consecutive addresses of C's
//      contiguous direction.
// Each output sums its K products in increasing k with the complex product
// a * b evaluated the same way in both orientations, so a pair's result does
// not depend on the orientation or on what else is in the batch.
constexpr int SKC = 8;                // K elements staged per pass
constexpr int SMAXE = 160;            // BM + BN of every tile shape
constexpr int SNT = 512;              // threads per workgroup
// Workgroups of a shallow launch (a grid-stride loop over the items).  A
// launch of up to 2^20 items gives each workgroup one; resident grids of 512
// to 2048 workgroups that loop over items measured no faster for complex
// FP64 and 2-6 % slower for complex FP32.
constexpr int64_t kShallowGrid = int64_t(1) << 20;
constexpr int SOUT = 4096 / SNT;      // outputs per thread
constexpr int SGA = (128 * SKC + SNT - 1) / SNT;  // gathered elements per operand and thread
constexpr int SOFF_ROWA = 0, SOFF_ROWC = 128, SOFF_COLB = 256, SOFF_COLC = 384, SOFF_KA = 512,
              SOFF_KB = SOFF_KA + kShallowK, SOFF_BATCH = SOFF_KB + kShallowK, SOFF_N = SOFF_BATCH + 3;

template <typename T, bool CP>
__device__ inline void cmac(Cplx<T>& acc, const Cplx<T>& a, const Cplx<T>& b) {
  acc.re = fma(a.re, b.re, acc.re);
  if constexpr (CP) {
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
  }
}

// The scalars of one shallow item that outlive its table phase.
struct ShallowItem {
  const void* A;
  const void* B;
  void* C;
  int BM, BN, mr, nr, K, a_kmajor, b_kmajor, c_mfast;
};

// Table phase of item `it`: one lookup per thread (rows, columns, K offsets,
// batch offsets) into soff.  The caller synchronizes before soff is read.
__device__ __forceinline__ ShallowItem shallow_tables(const ShallowArgs& a, int64_t it, int64_t* soff) {
  const int tid = threadIdx.x;
  const int2 w = a.items[it];
  const BatchPair& pr = a.pairs[w.x];
  ShallowItem s;
  s.BM = pr.tile == MUSCLE_HIP_TILE_32x128 ? 32 : pr.tile == MUSCLE_HIP_TILE_128x32 ? 128 : 64;
  s.BN = 4096 / s.BM;
  const int per_batch = pr.tiles_m * pr.tiles_n;
  const int bz = w.y / per_batch;
  const int r = w.y - bz * per_batch;
  const int tm = r / pr.tiles_n, tn = r - tm * pr.tiles_n;
  const int m0 = tm * s.BM, n0 = tn * s.BN;
  s.mr = min(s.BM, pr.M - m0);
  s.nr = min(s.BN, pr.N - n0);
  s.K = pr.K;
  s.A = pr.A;
  s.B = pr.B;
  s.C = pr.C;
  s.a_kmajor = pr.a_kmajor;
  s.b_kmajor = pr.b_kmajor;
  s.c_mfast = pr.c_mfast;
  const int e = tid, BM = s.BM, BN = s.BN, K = s.K;
  if (e < BM) {
    soff[SOFF_ROWA + e] = e < s.mr ? pr.rowA[m0 + e] : 0;
    soff[SOFF_ROWC + e] = e < s.mr ? pr.rowC[m0 + e] : 0;
  } else if (e < BM + BN) {
    const int n = e - BM;
    soff[SOFF_COLB + n] = n < s.nr ? pr.colB[n0 + n] : 0;
    soff[SOFF_COLC + n] = n < s.nr ? pr.colC[n0 + n] : 0;
  } else if (e < BM + BN + K) {
    const int k = e - BM - BN;
    soff[SOFF_KA + k] = pr.kA[k];
    soff[SOFF_KB + k] = pr.kB[k];
  } else if (e < BM + BN + K + 3) {
    const int j = e - BM - BN - K;
    soff[SOFF_BATCH + j] = j == 0 ? pr.bA[bz] : j == 1 ? pr.bB[bz] : pr.bC[bz];
  }
  return s;
}

// Gathers the K pass starting at k0: at most SGA elements of A and SGA of B
// per thread, all loads issued before any is used.
template <typename T, bool CP>
__device__ __forceinline__ void shallow_gather(const ShallowItem& s, const int64_t* soff, int k0, Cplx<T>* ra,
                                               Cplx<T>* rb) {
  const int tid = threadIdx.x;
  const int kn = min(SKC, s.K - k0);
  const int64_t offA = soff[SOFF_BATCH], offB = soff[SOFF_BATCH + 1];
#pragma unroll
  for (int j = 0; j < SGA; ++j) {
    const int e = tid + SNT * j;
    ra[j] = rb[j] = Cplx<T>{T(0), T(0)};
    if (e < s.BM * SKC) {
      const int m = s.a_kmajor ? e / SKC : e & (s.BM - 1), k = s.a_kmajor ? e & (SKC - 1) : e / s.BM;
      if (m < s.mr && k < kn) ra[j] = load_elem<T, CP>(s.A, offA + soff[SOFF_ROWA + m] + soff[SOFF_KA + k0 + k]);
    }
    if (e < s.BN * SKC) {
      const int n = s.b_kmajor ? e / SKC : e & (s.BN - 1), k = s.b_kmajor ? e & (SKC - 1) : e / s.BN;
      if (n < s.nr && k < kn) rb[j] = load_elem<T, CP>(s.B, offB + soff[SOFF_COLB + n] + soff[SOFF_KB + k0 + k]);
    }
  }
}

// Writes a gathered pass into LDS, k-major ([k][m] for A, [k][n] for B).
template <typename T>
__device__ __forceinline__ void shallow_stash(const ShallowItem& s, Cplx<T>* sab, const Cplx<T>* ra,
                                              const Cplx<T>* rb) {
  const int tid = threadIdx.x;
  Cplx<T>* sA = sab;
  Cplx<T>* sB = sab + s.BM * SKC;
#pragma unroll
  for (int j = 0; j < SGA; ++j) {
    const int e = tid + SNT * j;
    if (e < s.BM * SKC) {
      const int m = s.a_kmajor ? e / SKC : e & (s.BM - 1), k = s.a_kmajor ? e & (SKC - 1) : e / s.BM;
      sA[k * s.BM + m] = ra[j];
    }
    if (e < s.BN * SKC) {
      const int n = s.b_kmajor ? e / SKC : e & (s.BN - 1), k = s.b_kmajor ? e & (SKC - 1) : e / s.BN;
      sB[k * s.BN + n] = rb[j];
    }
  }
}

// Accumulates kn staged k of this thread's line of C: fixed index f along C's
// contiguous direction, SOUT indices v0 + vs * i across it.
template <typename T, bool CP, bool CM>
__device__ __forceinline__ void shallow_compute(const ShallowItem& s, const Cplx<T>* sab, int kn, Cplx<T>* acc) {
  const int tid = threadIdx.x;
  const Cplx<T>* sA = sab;
  const Cplx<T>* sB = sab + s.BM * SKC;
  const int LF = CM ? s.BM : s.BN;
  const int f = tid & (LF - 1), v0 = tid / LF, vs = SNT / LF;
  for (int kk = 0; kk < kn; ++kk) {
    if constexpr (CM) {
      const Cplx<T> av = sA[kk * s.BM + f];
#pragma unroll
      for (int i = 0; i < SOUT; ++i) cmac<T, CP>(acc[i], av, sB[kk * s.BN + v0 + vs * i]);
    } else {
      const Cplx<T> bv = sB[kk * s.BN + f];
#pragma unroll
      for (int i = 0; i < SOUT; ++i) cmac<T, CP>(acc[i], sA[kk * s.BM + v0 + vs * i], bv);
    }
  }
}

template <typename T, bool CP, bool CM, bool CS>
__device__ __forceinline__ void shallow_store(const ShallowItem& s, const int64_t* soff, const Cplx<T>* acc) {
  const int tid = threadIdx.x;
  const int LF = CM ? s.BM : s.BN;
  const int f = tid & (LF - 1), v0 = tid / LF, vs = SNT / LF;
  const int64_t offC = soff[SOFF_BATCH + 2];
#pragma unroll
  for (int i = 0; i < SOUT; ++i) {
    const int m = CM ? f : v0 + vs * i, n = CM ? v0 + vs * i : f;
    if (m < s.mr && n < s.nr) {
      const int64_t o = offC + soff[SOFF_ROWC + m] + soff[SOFF_COLC + n];
      if constexpr (CS) store_elem_nt<T, CP>(s.C, o, acc[i]);
      else store_elem<T, CP>(s.C, o, acc[i]);
    }
  }
}

// Two 8-wave workgroups per CU (about 122 registers for complex FP64: the
// next item's gathered pass is held across the stores); one launch per C
// orientation (CM: C runs along M) and store kind (CS: nontemporal).  A
// workgroup that loops over items issues the next item's table phase and
// first gather before the current item's stores, so its loads are not queued
// behind them (vector memory operations complete in order), and stashes the
// gathered pass after the stores.  soff is double-buffered (the stores read
// the current item's C offsets after the next item's table phase).
template <typename T, bool CP, bool CM, bool CS>
__global__ __launch_bounds__(SNT) void shallow_kernel(
    ShallowArgs a) {
  __shared__ Cplx<T> sab[SMAXE * SKC];
  __shared__ int64_t soff[2][SOFF_N];
  int64_t it = blockIdx.x;
  if (it >= a.nitems) return;
  int b = 0;
  ShallowItem cur = shallow_tables(a, it, soff[0]);
  __syncthreads();
  Cplx<T> ra[SGA], rb[SGA];
  shallow_gather<T, CP>(cur, soff[0], 0, ra, rb);
  shallow_stash<T>(cur, sab, ra, rb);
  for (;;) {
    Cplx<T> acc[SOUT];
#pragma unroll
    for (int i = 0; i < SOUT; ++i) acc[i] = Cplx<T>{T(0), T(0)};
    __syncthreads();  // the item's first pass is in sab
    shallow_compute<T, CP, CM>(cur, sab, min(SKC, cur.K), acc);
    for (int k0 = SKC; k0 < cur.K; k0 += SKC) {
      shallow_gather<T, CP>(cur, soff[b], k0, ra, rb);
      __syncthreads();  // the readers of sab are done
      shallow_stash<T>(cur, sab, ra, rb);
      __syncthreads();
      shallow_compute<T, CP, CM>(cur, sab, min(SKC, cur.K - k0), acc);
    }
    const int64_t it2 = it + gridDim.x;
    const bool more = it2 < a.nitems;
    ShallowItem nxt;
    if (more) {
      // soff[b ^ 1] was last read by the previous item's stores, before the
      // barrier at the top of this item
      nxt = shallow_tables(a, it2, soff[b ^ 1]);
      __syncthreads();  // also: the readers of sab are done
      shallow_gather<T, CP>(nxt, soff[b ^ 1], 0, ra, rb);
    }
    shallow_store<T, CP, CM, CS>(cur, soff[b], acc);
    if (!more) break;
    shallow_stash<T>(nxt, sab, ra, rb);
    cur = nxt;
    b ^= 1;
    it = it2;
  }
}

// One item per workgroup (launches of at most kShallowGrid items): nothing is
// held across the stores, so the complex kernels fit in 72 registers and
// three workgroups share a CU, where the looping kernel's 122 allow two
// (real FP64 takes 96: capped at 80 it spilled).
template <typename T, bool CP, bool CM, bool CS>
__global__ __launch_bounds__(SNT, CP ? 6 : 4) void shallow_one_kernel(ShallowArgs a) {
  __shared__ Cplx<T> sab[SMAXE * SKC];
  __shared__ int64_t soff[SOFF_N];
  const ShallowItem cur = shallow_tables(a, blockIdx.x, soff);
  __syncthreads();
  Cplx<T> acc[SOUT];
#pragma unroll
  for (int i = 0; i < SOUT; ++i) acc[i] = Cplx<T>{T(0), T(0)};
  for (int k0 = 0; k0 < cur.K; k0 += SKC) {
    Cplx<T> ra[SGA], rb[SGA];
    shallow_gather<T, CP>(cur, soff, k0, ra, rb);
    if (k0 > 0) __syncthreads();  // the readers of sab are done
    shallow_stash<T>(cur, sab, ra, rb);
    __syncthreads();
    shallow_compute<T, CP, CM>(cur, sab, min(SKC, cur.K - k0), acc);
  }
  shallow_store<T, CP, CM, CS>(cur, soff, acc);
}

// ---- skinny pairs ---------------------------------------------------------
// A pair whose output has a short side of S <= kSkinnyMaxS elements does at
// most S complex products per element of the long side's operand it reads,
// so it is bound by streaming that operand and C, and an MFMA or shallow tile
// (at least 32 wide) would spend most of its work on padding.  Here one
// workgroup of 512 threads covers the whole short side (padded to a power of
// two) and lt positions of the long side (skinny_geom):
//   - each thread owns J long positions and Q = O / J short positions (O = 4
//     outputs for complex FP64, else 8); it reads its long positions' operand
//     straight from memory into registers, once per (position, k), from a
//     ring of O loads refilled D = O / J steps of k ahead;
//   - the short side's operand is staged in LDS in chunks of KSB elements
//     (row stride sp) and read as wave-wide broadcasts;
//   - the item's K offsets (K <= kSkinnyMaxK) and short-side offsets are
//     looked up once, into LDS;
//   - C is stored with consecutive lanes on consecutive long positions.
// Each output sums its K products in increasing k with the shallow kernel's
// complex product (a from A, b from B), so a pair's result does not depend on
// the orientation or on what else is in the batch.
constexpr int KNT = 512;   // threads per workgroup
constexpr int KSB = 2048;  // short-side operand elements staged per chunk
constexpr int KOFF_S = 2 * kSkinnyMaxK, KOFF_C = KOFF_S + kSkinnyMaxS, KOFF_N = KOFF_C + kSkinnyMaxS;

// SM: the short side is M (A is the short operand, B the long one, C's long
// direction is N); else the short side is N.
template <typename T, bool CP, int J, bool SM>
__device__ __forceinline__ void skinny_tile(const BatchPair& pr, int bz, int tl, Cplx<T>* sS, int64_t* sK) {
  constexpr int O = (CP && sizeof(T) == 8) ? 4 : 8;
  constexpr int QM = O / J, D = O / J;
  const int tid = threadIdx.x;
  const int S = SM ? pr.M : pr.N, L = SM ? pr.N : pr.M, K = pr.K;
  const SkinnyGeom g = skinny_geom(S, O);
  const int sps = __builtin_ctz(g.sp);
  const int kc = KSB / g.sp;  // K per staged chunk, a multiple of D
  const int q0 = (tid / g.lw) * QM;
  const int lbase = tl * g.lt + tid % g.lw;
  const void* XL = SM ? pr.B : pr.A;
  const void* XS = SM ? pr.A : pr.B;
  const OffTab& tabL = SM ? pr.colB : pr.rowA;
  const OffTab& tabS = SM ? pr.rowA : pr.colB;
  const OffTab& kL = SM ? pr.kB : pr.kA;
  const OffTab& kS = SM ? pr.kA : pr.kB;
  const OffTab& cL = SM ? pr.colC : pr.rowC;
  const OffTab& cS = SM ? pr.rowC : pr.colC;
  const int64_t bS = SM ? pr.bA[bz] : pr.bB[bz];
  const int64_t bL = SM ? pr.bB[bz] : pr.bA[bz];

  __syncthreads();  // the previous item's readers of LDS are done
  for (int k = tid; k < K; k += KNT) {
    sK[k] = kL[k];
    sK[kSkinnyMaxK + k] = kS[k];
  }
  if (tid < kSkinnyMaxS) {
    const bool ok = tid < S;
    sK[KOFF_S + tid] = ok ? bS + tabS[tid] : 0;
    sK[KOFF_C + tid] = ok ? cS[tid] : 0;
  }
  // this thread's long positions (clamped into range; their outputs are not stored)
  int64_t oL[J];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int l = lbase + g.lw * j;
    oL[j] = bL + tabL[l < L ? l : L - 1];
  }
  __syncthreads();

  // the short operand's chunk of K starting at kb, as sS[kk * sp + q]
  auto stage = [&](int kb) {
    const int kn = min(kc, K - kb);
    Cplx<T> r[KSB / KNT];
#pragma unroll
    for (int i = 0; i < KSB / KNT; ++i) {
      const int e = tid + KNT * i;
      const int kk = e >> sps, q = e & (g.sp - 1);
      r[i] = Cplx<T>{T(0), T(0)};
      if (kk < kn && q < S) r[i] = load_elem<T, CP>(XS, sK[KOFF_S + q] + sK[kSkinnyMaxK + kb + kk]);
    }
#pragma unroll
    for (int i = 0; i < KSB / KNT; ++i) sS[tid + KNT * i] = r[i];
  };

  Cplx<T> acc[J][QM];
#pragma unroll
  for (int j = 0; j < J; ++j)
#pragma unroll
    for (int i = 0; i < QM; ++i) acc[j][i] = Cplx<T>{T(0), T(0)};
  Cplx<T> ring[D][J];
#pragma unroll
  for (int d = 0; d < D; ++d)
#pragma unroll
    for (int j = 0; j < J; ++j) ring[d][j] = d < K ? load_elem<T, CP>(XL, oL[j] + sK[d]) : Cplx<T>{T(0), T(0)};
  if (K > 0) stage(0);
  __syncthreads();
  int kb = 0;
  for (int k0 = 0; k0 < K; k0 += D) {
    if (k0 - kb >= kc) {  // the next chunk (k0 = kb + kc: kc is a multiple of D)
      __syncthreads();
      kb = k0;
      stage(kb);
      __syncthreads();
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int k = k0 + d;
      if (k < K) {
        const Cplx<T>* srow = sS + ((k - kb) << sps) + q0;
#pragma unroll
        for (int i = 0; i < QM; ++i) {
          const Cplx<T> sv = srow[i];
#pragma unroll
          for (int j = 0; j < J; ++j) {
            if constexpr (SM) cmac<T, CP>(acc[j][i], sv, ring[d][j]);
            else cmac<T, CP>(acc[j][i], ring[d][j], sv);
          }
        }
        if (k + D < K) {
#pragma unroll
          for (int j = 0; j < J; ++j) ring[d][j] = load_elem<T, CP>(XL, oL[j] + sK[k + D]);
        }
      }
    }
  }
  const int64_t offC = pr.bC[bz];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int l = lbase + g.lw * j;
    if (l >= L) continue;
    const int64_t cl = offC + cL[l];
#pragma unroll
    for (int i = 0; i < QM; ++i) {
      if (q0 + i >= S) continue;
      // streaming stores only with lanes along C's contiguous direction (SM):
      // otherwise a lane's few consecutive elements leave partial lines
      if (SM && pr.c_stream) store_elem_nt<T, CP>(pr.C, cl + sK[KOFF_C + q0 + i], acc[j][i]);
      else store_elem<T, CP>(pr.C, cl + sK[KOFF_C + q0 + i], acc[j][i]);
    }
  }
}

// One launch for every skinny pair: each item takes the variant of its pair's
// long positions per thread (J = 1, 2 or 4; complex FP64 uses 1 and 2).
template <typename T, bool CP>
__global__ __launch_bounds__(KNT) void skinny_kernel(ShallowArgs a) {
  constexpr int O = (CP && sizeof(T) == 8) ? 4 : 8;
  __shared__ Cplx<T> sS[KSB];
  __shared__ int64_t sK[KOFF_N];
  for (int64_t it = blockIdx.x; it < a.nitems; it += gridDim.x) {
    const int2 w = a.items[it];
    const BatchPair& pr = a.pairs[w.x];
    const bool sm = pr.M <= pr.N;
    const SkinnyGeom g = skinny_geom(sm ? pr.M : pr.N, O);
    const int tiles = ((sm ? pr.N : pr.M) + g.lt - 1) / g.lt;
    const int bz = w.y / tiles, tl = w.y - bz * tiles;
    if (g.j == 1) {
      if (sm) skinny_tile<T, CP, 1, true>(pr, bz, tl, sS, sK);
      else skinny_tile<T, CP, 1, false>(pr, bz, tl, sS, sK);
    } else if (O == 4 || g.j == 2) {
      if (sm) skinny_tile<T, CP, 2, true>(pr, bz, tl, sS, sK);
      else skinny_tile<T, CP, 2, false>(pr, bz, tl, sS, sK);
    } else {
      if (sm) skinny_tile<T, CP, O == 4 ? 2 : 4, true>(pr, bz, tl, sS, sK);
      else skinny_tile<T, CP, O == 4 ? 2 : 4, false>(pr, bz, tl, sS, sK);
    }
  }
}

template <typename T, bool CP, bool G3>
void launch_t(const BatchArgs& a, const BatchArgs& sp, int wg, size_t lds, hipStream_t s) {
  const unsigned rg = (unsigned)std::max<int64_t>(1, std::min<int64_t>((a.ntilecnt + 255) / 256, 1024));
  hipLaunchKernelGGL(reset_counter_kernel, dim3(rg), dim3(256), 0, s, a.counter, a.tilecnt, a.ntilecnt);
  if (a.nitems > 0) hipLaunchKernelGGL((batch_kernel<T, CP, G3>), dim3(wg), dim3(NT), lds, s, a);
  if (sp.nitems > 0) {
    const unsigned g = (unsigned)std::min<int64_t>(sp.nitems, int64_t(wg) * 4);
    hipLaunchKernelGGL((split_kernel<T, CP, G3>), dim3(g), dim3(NT), lds, s, sp);
  }
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

void launch_shallow(bool dbl, bool cplx, bool cm, bool cs, const ShallowArgs& a, hipStream_t s) {
  if (a.nitems <= 0) return;
  const unsigned grid = (unsigned)std::min<int64_t>(a.nitems, kShallowGrid);
#define MH_SHALLOW_CS(T, CP, CM)                                                                     \
  if (a.nitems <= kShallowGrid) {                                                                   \
    if (cs) hipLaunchKernelGGL((shallow_one_kernel<T, CP, CM, true>), dim3(grid), dim3(SNT), 0, s, a); \
    else hipLaunchKernelGGL((shallow_one_kernel<T, CP, CM, false>), dim3(grid), dim3(SNT), 0, s, a);   \
  } else if (cs) {                                                                                  \
    hipLaunchKernelGGL((shallow_kernel<T, CP, CM, true>), dim3(grid), dim3(SNT), 0, s, a);          \
  } else {                                                                                          \
    hipLaunchKernelGGL((shallow_kernel<T, CP, CM, false>), dim3(grid), dim3(SNT), 0, s, a);         \
  }
#define MH_SHALLOW(T, CP)       \
  if (cm) {                     \
    MH_SHALLOW_CS(T, CP, true)  \
  } else {                      \
    MH_SHALLOW_CS(T, CP, false) \
  }
  if (dbl) {
    if (cplx) {
      MH_SHALLOW(double, true)
    } else {
      MH_SHALLOW(double, false)
    }
  } else {
    if (cplx) {
      MH_SHALLOW(float, true)
    } else {
      MH_SHALLOW(float, false)
    }
  }
#undef MH_SHALLOW
#undef MH_SHALLOW_CS
}

void launch_skinny(bool dbl, bool cplx, const ShallowArgs& a, hipStream_t s) {
  if (a.nitems <= 0) return;
  const unsigned grid = (unsigned)std::min<int64_t>(a.nitems, int64_t(1) << 20);
  if (dbl) {
    if (cplx) hipLaunchKernelGGL((skinny_kernel<double, true>), dim3(grid), dim3(KNT), 0, s, a);
    else hipLaunchKernelGGL((skinny_kernel<double, false>), dim3(grid), dim3(KNT), 0, s, a);
  } else {
    if (cplx) hipLaunchKernelGGL((skinny_kernel<float, true>), dim3(grid), dim3(KNT), 0, s, a);
    else hipLaunchKernelGGL((skinny_kernel<float, false>), dim3(grid), dim3(KNT), 0, s, a);
  }
}

void launch_batch(bool dbl, bool cplx, const BatchArgs& a, const BatchArgs& sp, int wg, size_t lds, hipStream_t s) {
  const bool g3 = cplx && a.complex3m;
  if (dbl) {
    if (g3) launch_t<double, true, true>(a, sp, wg, lds, s);
    else if (cplx) launch_t<double, true, false>(a, sp, wg, lds, s);
    else launch_t<double, false, false>(a, sp, wg, lds, s);
  } else {
    if (g3) launch_t<float, true, true>(a, sp, wg, lds, s);
    else if (cplx) launch_t<float, true, false>(a, sp, wg, lds, s);
    else launch_t<float, false, false>(a, sp, wg, lds, s);
  }
}

}  // namespace mh
