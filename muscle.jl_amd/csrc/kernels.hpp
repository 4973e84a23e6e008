// Launch interface between the host layer and the HIP kernels.  Offsets are
// in elements of the operand's own type; every table is a device array.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mh {

// Unsigned 32-bit division by a run-time constant with a multiply-high and a
// shift: the round-up method of Granlund & Montgomery, "Division by invariant
// integers using multiplication" (PLDI 1994).  Valid for any n < 2^32.
struct FastDiv {
  uint32_t d, mul, shift;
  FastDiv() : d(1), mul(1), shift(0) {}
  explicit FastDiv(uint32_t div) : d(div ? div : 1) {
    div = d;  // a divisor of 0 (an unused table) behaves as 1
    uint32_t l = 0;
    while ((uint64_t(1) << l) < div) ++l;
    shift = l;
    mul = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << l) - div)) / div + 1);
  }
  __host__ __device__ inline uint32_t div(uint32_t n) const {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t t = __umulhi(mul, n);
#else
    uint32_t t = (uint32_t)(((uint64_t)mul * n) >> 32);
#endif
    return (uint32_t)(((uint64_t)t + n) >> shift);
  }
};

// Two-level offset table of one index class for one operand:
//   off(i) = lo[i mod P] + hi[i div P]
// where the class's (fused) modes are split into a fast part of product P,
// about the square root of the class size, and a slow part.  Tables stay
// O(sqrt(n)) in size: a 2^28-element output mode class needs 2 x 16384 entries.
// Classes of up to 65536 entries keep a single level (hi = {0}, single = 1),
// read with one load.
struct OffTab {
  const int64_t* lo;
  const int64_t* hi;
  FastDiv div;      // divisor P
  uint32_t base;    // added to every index (batch chunks)
  uint32_t single;  // one-level table: off(i) = lo[i]
  // The tables are device (global) memory: read as such they are global, not
  // flat, loads (a pending flat access makes every later wait a full drain).
  __device__ inline int64_t operator[](uint32_t i) const {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef const __attribute__((address_space(1))) int64_t* gp;
    const gp l = (gp)lo, h = (gp)hi;
#else
    const int64_t *l = lo, *h = hi;
#endif
    i += base;
    if (single) return l[i];
    const uint32_t q = div.div(i);
    return l[i - q * div.d] + h[q];
  }
};

// ---- tile GEMM on MFMA ----------------------------------------------------
struct GemmArgs {
  const void* A;
  const void* B;
  void* C;
  OffTab rowA, kA, kB, colB, rowC, colC, bA, bB, bC;
  int32_t M, N, K, batch;
  int32_t a_kmajor;  // A's contiguous direction is K (else M)
  int32_t b_kmajor;  // B's contiguous direction is K (else N)
  int32_t tiles_m, tiles_n;
  int32_t complex3m;  // complex x complex with 3 real products (Gauss) instead of 4
  // split-K: grid.z = ksplit slices of kchunk (a multiple of the K-tile); with
  // ksplit > 1 every slice writes a partial [ksplit][batch][M][N] array in the
  // compute type and a second kernel adds the slices in order into C.
  int32_t ksplit, kchunk;
  void* partial;
  int32_t c_stream;  // store C with nontemporal (streaming) stores
};
// Complex FP64 outputs of at least this many bytes are written with
// nontemporal stores: they do not stay in the caches anyway (writing C bounds
// small-K products, which this speeds up), while smaller outputs are often
// read again soon (the next contraction of a tree).  Elements of 8 bytes or
// less keep ordinary stores: written nontemporally, the L2 no longer merges
// their partial lines (config 4 in ComplexF32: 138.5 against 149.6 ms).
constexpr double kStreamStoreBytes = double(int64_t(16) << 20);
// dbl: FP64 storage (else FP32); fp32_math: FP64 storage with FP32 MFMA
// arithmetic (the FP32 rung); ca/cb: operand is complex.
void launch_gemm_mfma(bool dbl, bool fp32_math, bool ca, bool cb, const GemmArgs& g, hipStream_t s);

// ---- reduction kernel (skinny contractions, unary sums, traces) -------------
struct ReduceArgs {
  const void* A;
  const void* B;  // null when has_b == false
  void* C;
  void* partial;  // [nsplit][outputs] when nsplit > 1
  OffTab rowA, kA, kB, colB, rowC, colC, bA, bB, bC;
  int64_t M, N, K, batch, outputs;
  int64_t kchunk;
  int32_t gshift;  // lanes per output = 1 << gshift (<= 64)
  int32_t nsplit;
  FastDiv divN, divM;  // valid when outputs < 2^32
  int32_t small_index;  // 1: use FastDiv
  // nonzero when the summed dims fuse into one: offsets k * kstep instead of
  // the k tables
  int64_t kstepA, kstepB;
};
void launch_reduce(bool dbl, bool ca, bool cb, bool has_b, const ReduceArgs& r, hipStream_t s);

// ---- elementwise broadcast product (hadamard) --------------------------------
constexpr int kMaxEwDims = 24;
struct EwArgs {
  void* C;
  const void* A;
  const void* B;
  int64_t total;
  int32_t R;
  int32_t small_index;
  FastDiv div[kMaxEwDims];
  int64_t ext[kMaxEwDims];
  int64_t sA[kMaxEwDims], sB[kMaxEwDims], sC[kMaxEwDims];
  // tiled form (b runs contiguous along other modes than c): dims [0, Rr) are
  // the tile rows (c's fastest), [Rr, Rr + Rc) the tile columns (b's fastest
  // remaining), the rest the outer index
  int32_t tiled, Rr, Rc;
  int64_t nrows, ncols, nouter;
};
// dtA/dtB: muscle_hip_dtype_t of a and b; c has a's type.
void launch_hadamard(int dtA, int dtB, const EwArgs& e, hipStream_t s);

// ---- dtype cast over a strided span (mixed-precision operands) --------------
// dst[i] = (dst type) src[i] for i in [0, n); complexness is preserved.
void launch_cast(int dt_src, int dt_dst, const void* src, void* dst, int64_t n, hipStream_t s);
// Bandwidth probe: streams `bytes` (a multiple of 16) of buf with 16-byte
// stores (write) or loads (else; out receives at most one word).
void launch_probe(bool write, void* buf, int64_t bytes, uint32_t* out, hipStream_t s);

// Copies a GEMM operand from its strided layout into [batch][rows][K] with K
// contiguous (rows = M for A, N for B); source offsets come from the plan's
// row, K and batch tables.  k_fast: read the tile with the K index fastest
// (the source's smaller stride), else the row index fastest.
struct PackOpArgs {
  const void* src;
  void* dst;
  OffTab row, k, bt;
  int32_t R, K, B, k_fast;
};
void launch_pack_operand(int elem_bytes, const PackOpArgs& a, hipStream_t s);

// ---- tiled permutation (unary_einsum without summed modes) -------------------
// Y[rowY[m] + colY[n]] = X[rowX[m] + colX[n]]: m runs along Y's fastest dims,
// n along X's fastest remaining dims; 32 x 32 tiles transposed through LDS so
// that both the reads and the writes are contiguous runs.
struct PermuteArgs {
  const void* X;
  void* Y;
  OffTab rowX, rowY, colX, colY;
  int64_t M, N;
};
void launch_permute(int elem_bytes, const PermuteArgs& a, hipStream_t s);

// dst (k x n, ld ldd) = upper trapezoid of the first k rows of src (m x n, ld m)
void launch_upper_trapezoid(int elem_bytes, const void* src, void* dst, int64_t m, int64_t k, int64_t n, int64_t ldd,
                            hipStream_t s);
// zero `bytes` (a multiple of 4) at a 4-byte aligned dst
void launch_zero(void* dst, int64_t bytes, hipStream_t s);
// `count` elements of dtype with deterministic pseudo-random parts in [-1, 1)
void launch_random_unit(int dtype, void* dst, int64_t count, uint64_t seed, hipStream_t s);
// singular pairs from the top k eigenvectors of the nev x nev augmented matrix
// (eigenvalues w ascending): U (top x k), Vc (bot x k, conjugated), s (k)
void launch_split_pairs(int dtype, const void* H, const void* w, int64_t nev, int64_t top, int64_t bot, int64_t k,
                        void* U, void* Vc, void* s, hipStream_t st);

}  // namespace mh
