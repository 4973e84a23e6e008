// Device helpers of the factorizations: the R factor's upper trapezoid.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "muscle_hip.h"

namespace mh {
namespace {

struct alignas(16) Word16 {
  uint64_t lo, hi;
};

// dst (k x n, leading dimension ldd) = the upper trapezoid of the first k
// rows of src (m x n, leading dimension m), zeros below the diagonal.
template <typename W>
__global__ void __launch_bounds__(256) upper_kernel(const W* __restrict__ src, W* __restrict__ dst, int64_t m,
                                                    int64_t k, int64_t n, int64_t ldd) {
  const int64_t total = k * n;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
    const int64_t j = e / k, i = e - j * k;
    W v{};
    if (i <= j) v = src[i + j * m];
    dst[i + j * ldd] = v;
  }
}

template <typename W>
__global__ void __launch_bounds__(256) zero_kernel(W* __restrict__ dst, int64_t count) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < count; e += (int64_t)gridDim.x * 256) dst[e] = W{};
}

unsigned grid_for(int64_t total) { return (unsigned)std::min<int64_t>((total + 255) / 256, 8192); }

// Splitmix64 finalizer of (seed, index): a deterministic value in [-1, 1).
__device__ inline float hashed_unit(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  return (float)((int64_t)(x >> 11) - (int64_t(1) << 52)) * (1.0f / 4503599627370496.0f);
}

template <typename R, int NC>
__global__ void __launch_bounds__(256) random_kernel(R* __restrict__ dst, int64_t count, uint64_t seed) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < count; e += (int64_t)gridDim.x * 256)
    for (int c = 0; c < NC; ++c) dst[e * NC + c] = (R)hashed_unit(seed * 0x100000001B3ull + (uint64_t)(e * NC + c));
}

// Eigenvector i (ascending eigenvalue order, so column nev - 1 - i) of the
// augmented matrix [[0, M], [M^H, 0]] is [u_i; v_i] / sqrt(2) for the i-th
// largest singular value of M.  One workgroup per i: the two halves are
// normalised separately (a mix of the +sigma and -sigma eigenvectors rescales
// them, it does not turn them), u_i goes to U's column i and conj(v_i) to Vc's;
// s[i] = max(lambda, 0).
template <typename R, int NC>
__global__ void __launch_bounds__(256) split_pairs_kernel(const R* __restrict__ H, const R* __restrict__ w,
                                                          int64_t nev, int64_t top, int64_t bot,
                                                          R* __restrict__ U, R* __restrict__ Vc, R* __restrict__ s) {
  __shared__ R red[2][256 / 64];
  const int64_t i = blockIdx.x;
  const R* e = H + (nev - 1 - i) * nev * NC;
  R su = 0, sv = 0;
  for (int64_t j = threadIdx.x; j < top + bot; j += 256) {
    R a2 = 0;
    for (int c = 0; c < NC; ++c) a2 += e[j * NC + c] * e[j * NC + c];
    if (j < top)
      su += a2;
    else
      sv += a2;
  }
  for (int off = 32; off > 0; off >>= 1) {
    su += __shfl_xor(su, off);
    sv += __shfl_xor(sv, off);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = su;
    red[1][threadIdx.x >> 6] = sv;
  }
  __syncthreads();
  su = red[0][0] + red[0][1] + red[0][2] + red[0][3];
  sv = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  const R fu = su > 0 ? (R)1 / sqrt(su) : (R)0, fv = sv > 0 ? (R)1 / sqrt(sv) : (R)0;
  for (int64_t j = threadIdx.x; j < top; j += 256)
    for (int c = 0; c < NC; ++c) U[(i * top + j) * NC + c] = e[j * NC + c] * fu;
  for (int64_t j = threadIdx.x; j < bot; j += 256)
    for (int c = 0; c < NC; ++c) Vc[(i * bot + j) * NC + c] = (c ? -fv : fv) * e[(top + j) * NC + c];
  if (threadIdx.x == 0) {
    const R l = w[nev - 1 - i];
    s[i] = l > 0 ? l : (R)0;
  }
}

}  // namespace

void launch_upper_trapezoid(int elem_bytes, const void* src, void* dst, int64_t m, int64_t k, int64_t n, int64_t ldd,
                            hipStream_t s) {
  const int64_t total = k * n;
  if (total <= 0) return;
  switch (elem_bytes) {
    case 4:
      upper_kernel<uint32_t><<<grid_for(total), 256, 0, s>>>((const uint32_t*)src, (uint32_t*)dst, m, k, n, ldd);
      break;
    case 8:
      upper_kernel<uint64_t><<<grid_for(total), 256, 0, s>>>((const uint64_t*)src, (uint64_t*)dst, m, k, n, ldd);
      break;
    default:
      upper_kernel<Word16><<<grid_for(total), 256, 0, s>>>((const Word16*)src, (Word16*)dst, m, k, n, ldd);
      break;
  }
}

void launch_zero(void* dst, int64_t bytes, hipStream_t s) {
  if (bytes <= 0) return;
  zero_kernel<uint32_t><<<grid_for(bytes / 4), 256, 0, s>>>((uint32_t*)dst, bytes / 4);  // bytes: a multiple of 4
}

void launch_random_unit(int dtype, void* dst, int64_t count, uint64_t seed, hipStream_t s) {
  if (count <= 0) return;
  switch (dtype) {
    case MUSCLE_HIP_F32:
      random_kernel<float, 1><<<grid_for(count), 256, 0, s>>>((float*)dst, count, seed);
      break;
    case MUSCLE_HIP_F64:
      random_kernel<double, 1><<<grid_for(count), 256, 0, s>>>((double*)dst, count, seed);
      break;
    case MUSCLE_HIP_C32:
      random_kernel<float, 2><<<grid_for(count), 256, 0, s>>>((float*)dst, count, seed);
      break;
    default:
      random_kernel<double, 2><<<grid_for(count), 256, 0, s>>>((double*)dst, count, seed);
      break;
  }
}

void launch_split_pairs(int dtype, const void* H, const void* w, int64_t nev, int64_t top, int64_t bot, int64_t k,
                        void* U, void* Vc, void* s, hipStream_t st) {
  if (k <= 0) return;
  switch (dtype) {
    case MUSCLE_HIP_F32:
      split_pairs_kernel<float, 1><<<(unsigned)k, 256, 0, st>>>((const float*)H, (const float*)w, nev, top, bot,
                                                                 (float*)U, (float*)Vc, (float*)s);
      break;
    case MUSCLE_HIP_F64:
      split_pairs_kernel<double, 1><<<(unsigned)k, 256, 0, st>>>((const double*)H, (const double*)w, nev, top, bot,
                                                                  (double*)U, (double*)Vc, (double*)s);
      break;
    case MUSCLE_HIP_C32:
      split_pairs_kernel<float, 2><<<(unsigned)k, 256, 0, st>>>((const float*)H, (const float*)w, nev, top, bot,
                                                                 (float*)U, (float*)Vc, (float*)s);
      break;
    default:
      split_pairs_kernel<double, 2><<<(unsigned)k, 256, 0, st>>>((const double*)H, (const double*)w, nev, top, bot,
                                                                  (double*)U, (double*)Vc, (double*)s);
      break;
  }
}

}  // namespace mh
