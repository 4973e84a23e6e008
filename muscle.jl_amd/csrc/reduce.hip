// Reduction-shaped contractions: outputs that are too thin for a 64 x 64
// MFMA tile (matrix-vector, inner products, scalars), outer products (K = 1)
// and every unary_einsum (sums, traces, diagonals; the OMEinsum rule behind
// src/Operations/unary_einsum.jl:20-24 and ext/MuscleOMEinsumExt.jl:31-37).
//
// Each output element (m, n, batch) is owned by a group of G = 2^gshift lanes
// of one wavefront; the lanes stride over the summed index k and combine
// their partial sums with cross-lane shuffles.  Outputs are numbered with n
// fastest, the direction in which C is contiguous after lowering, so stores
// coalesce.  When there are too few outputs to fill the GPU, K is split over
// `nsplit` workgroup columns that write partial sums; a second kernel adds
// the partials in a fixed order, so results do not depend on scheduling.
#include "device_types.hpp"
#include "kernels.hpp"

namespace mh {
namespace {

constexpr int NT = 256;

template <typename T>
__device__ inline T shfl_xor(T v, int mask) {
  return __shfl_xor(v, mask, 64);
}

template <typename T, bool CA, bool CB, bool HAS_B>
__global__ __launch_bounds__(NT) void reduce_kernel(ReduceArgs p) {
  const int64_t gt = (int64_t)blockIdx.x * NT + threadIdx.x;
  const int G = 1 << p.gshift;
  const int64_t o = gt >> p.gshift;
  const int r = (int)(gt & (G - 1));
  const int split = blockIdx.y;
  const bool active = o < p.outputs;

  int64_t m = 0, n = 0, b = 0;
  if (active) {
    if (p.small_index) {
      const uint32_t ou = (uint32_t)o;
      const uint32_t t = p.divN.div(ou);
      n = ou - t * (uint32_t)p.N;
      const uint32_t bb = p.divM.div(t);
      m = t - bb * (uint32_t)p.M;
      b = bb;
    } else {
      const int64_t t = o / p.N;
      n = o - t * p.N;
      b = t / p.M;
      m = t - b * p.M;
    }
  }
  T ar = 0, ai = 0;
  if (active) {
    const int64_t baseA = p.bA[b] + p.rowA[m];
    int64_t baseB = 0;
    if constexpr (HAS_B) baseB = p.bB[b] + p.colB[n];
    const int64_t k0 = (int64_t)split * p.kchunk;
    int64_t k1 = k0 + p.kchunk;
    if (k1 > p.K) k1 = p.K;
    const bool affine = p.kstepA != 0 && (!HAS_B || p.kstepB != 0);
#pragma unroll 4
    for (int64_t k = k0 + r; k < k1; k += G) {
      const Cplx<T> x = load_elem<T, CA>(p.A, baseA + (affine ? k * p.kstepA : p.kA[k]));
      if constexpr (HAS_B) {
        const Cplx<T> y = load_elem<T, CB>(p.B, baseB + (affine ? k * p.kstepB : p.kB[k]));
        if constexpr (CA && CB) {
          ar = fma(x.re, y.re, ar);
          ar = fma(-x.im, y.im, ar);
          ai = fma(x.re, y.im, ai);
          ai = fma(x.im, y.re, ai);
        } else if constexpr (CA) {
          ar = fma(x.re, y.re, ar);
          ai = fma(x.im, y.re, ai);
        } else if constexpr (CB) {
          ar = fma(x.re, y.re, ar);
          ai = fma(x.re, y.im, ai);
        } else {
          ar = fma(x.re, y.re, ar);
        }
      } else {
        ar += x.re;
        if constexpr (CA) ai += x.im;
      }
    }
  }
  // combine the G lanes of each group (groups are aligned inside a wavefront)
  for (int s = G >> 1; s > 0; s >>= 1) {
    ar += shfl_xor(ar, s);
    ai += shfl_xor(ai, s);
  }
  constexpr bool CC = CA || CB;
  if (active && r == 0) {
    if (p.nsplit == 1) {
      store_elem<T, CC>(p.C, p.bC[b] + p.rowC[m] + p.colC[n], Cplx<T>{ar, ai});
    } else {
      store_elem<T, CC>(p.partial, (int64_t)split * p.outputs + o, Cplx<T>{ar, ai});
    }
  }
}

template <typename T, bool CC>
__global__ __launch_bounds__(NT) void reduce_finish_kernel(ReduceArgs p) {
  const int64_t o = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (o >= p.outputs) return;
  T ar = 0, ai = 0;
  for (int s = 0; s < p.nsplit; ++s) {
    const Cplx<T> v = load_elem<T, CC>(p.partial, (int64_t)s * p.outputs + o);
    ar += v.re;
    ai += v.im;
  }
  const int64_t t = o / p.N;
  const int64_t n = o - t * p.N;
  const int64_t b = t / p.M;
  const int64_t m = t - b * p.M;
  store_elem<T, CC>(p.C, p.bC[b] + p.rowC[m] + p.colC[n], Cplx<T>{ar, ai});
}

template <typename T, bool CA, bool CB, bool HAS_B>
void launch_t(const ReduceArgs& r, hipStream_t s) {
  const int64_t threads = r.outputs << r.gshift;
  const int64_t blocks = (threads + NT - 1) / NT;
  hipLaunchKernelGGL((reduce_kernel<T, CA, CB, HAS_B>), dim3((unsigned)blocks, (unsigned)r.nsplit), dim3(NT), 0,
                     s, r);
  if (r.nsplit > 1) {
    const int64_t fb = (r.outputs + NT - 1) / NT;
    hipLaunchKernelGGL((reduce_finish_kernel<T, CA || CB>), dim3((unsigned)fb), dim3(NT), 0, s, r);
  }
}

}  // namespace

void launch_reduce(bool dbl, bool ca, bool cb, bool has_b, const ReduceArgs& r, hipStream_t s) {
  if (!has_b) {
    if (dbl) {
      if (ca) launch_t<double, true, false, false>(r, s);
      else launch_t<double, false, false, false>(r, s);
    } else {
      if (ca) launch_t<float, true, false, false>(r, s);
      else launch_t<float, false, false, false>(r, s);
    }
    return;
  }
  const int key = (dbl ? 4 : 0) | (ca ? 2 : 0) | (cb ? 1 : 0);
  switch (key) {
    case 0: launch_t<float, false, false, true>(r, s); break;
    case 1: launch_t<float, false, true, true>(r, s); break;
    case 2: launch_t<float, true, false, true>(r, s); break;
    case 3: launch_t<float, true, true, true>(r, s); break;
    case 4: launch_t<double, false, false, true>(r, s); break;
    case 5: launch_t<double, false, true, true>(r, s); break;
    case 6: launch_t<double, true, false, true>(r, s); break;
    case 7: launch_t<double, true, true, true>(r, s); break;
  }
}

}  // namespace mh
