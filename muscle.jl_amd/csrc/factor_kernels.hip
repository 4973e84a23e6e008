// Device helpers of the factorizations: the R factor's upper trapezoid.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace mh {
namespace {

struct alignas(16) Word16 {
  uint64_t lo, hi;
};

// dst (k x n, leading dimension k) = the upper trapezoid of the first k rows
// of src (m x n, leading dimension m), zeros below the diagonal.
template <typename W>
__global__ void __launch_bounds__(256) upper_kernel(const W* __restrict__ src, W* __restrict__ dst, int64_t m,
                                                    int64_t k, int64_t n) {
  const int64_t total = k * n;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
    const int64_t j = e / k, i = e - j * k;
    W v{};
    if (i <= j) v = src[i + j * m];
    dst[e] = v;
  }
}

}  // namespace

void launch_upper_trapezoid(int elem_bytes, const void* src, void* dst, int64_t m, int64_t k, int64_t n,
                            hipStream_t s) {
  const int64_t total = k * n;
  if (total <= 0) return;
  const unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 4096);
  switch (elem_bytes) {
    case 4:
      upper_kernel<uint32_t><<<blocks, 256, 0, s>>>((const uint32_t*)src, (uint32_t*)dst, m, k, n);
      break;
    case 8:
      upper_kernel<uint64_t><<<blocks, 256, 0, s>>>((const uint64_t*)src, (uint64_t*)dst, m, k, n);
      break;
    default:
      upper_kernel<Word16><<<blocks, 256, 0, s>>>((const Word16*)src, (Word16*)dst, m, k, n);
      break;
  }
}

}  // namespace mh
