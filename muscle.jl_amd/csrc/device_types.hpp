// Device-side element types and arithmetic helpers.
#pragma once

#include <hip/hip_runtime.h>

namespace mh {

template <typename T>
struct Cplx {
  T re, im;
};

// Storage type of one element: T for real, a 2-vector for complex so a
// complex element is one 8- or 16-byte memory access.
template <typename T, bool C>
struct Elem;
template <>
struct Elem<double, false> {
  typedef double type;
};
template <>
struct Elem<float, false> {
  typedef float type;
};
template <>
struct Elem<double, true> {
  typedef double2 type;
};
template <>
struct Elem<float, true> {
  typedef float2 type;
};

// Every operand, output and partial buffer is device (global) memory.
// Addressed as such, its loads and stores are global rather than flat
// instructions: a flat access also counts against the LDS wait counter, so
// each wait for an LDS read or write would drain the outstanding stores of C.
#if defined(__HIP_DEVICE_COMPILE__)
#define MH_GLOBAL __attribute__((address_space(1)))
#else
#define MH_GLOBAL
#endif
template <typename E>
__device__ inline const MH_GLOBAL E* gptr(const void* p) {
  return (const MH_GLOBAL E*)p;
}
template <typename E>
__device__ inline MH_GLOBAL E* gptr(void* p) {
  return (MH_GLOBAL E*)p;
}

template <typename T, bool C>
__device__ inline Cplx<T> load_elem(const void* base, int64_t off) {
  if constexpr (C) {
    auto v = gptr<typename Elem<T, true>::type>(base)[off];
    return Cplx<T>{v.x, v.y};
  } else {
    return Cplx<T>{gptr<T>(base)[off], T(0)};
  }
}

template <typename T, bool C>
__device__ inline void store_elem(void* base, int64_t off, Cplx<T> v) {
  if constexpr (C) {
    typename Elem<T, true>::type w;
    w.x = v.re;
    w.y = v.im;
    gptr<typename Elem<T, true>::type>(base)[off] = w;
  } else {
    gptr<T>(base)[off] = v.re;
  }
}

// The same with a nontemporal (streaming) store.
template <typename T, bool C>
__device__ inline void store_elem_nt(void* base, int64_t off, Cplx<T> v) {
  if constexpr (C) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    v2 w = {v.re, v.im};
    __builtin_nontemporal_store(w, gptr<v2>(base) + off);
  } else {
    __builtin_nontemporal_store(v.re, gptr<T>(base) + off);
  }
}

// Tile order of the GEMM kernels: the dispatcher deals workgroups
// round-robin over the 8 XCDs, so workgroup L is renumbered to give each XCD
// a contiguous range; the range is walked in groups of 8 tile rows, row
// fastest, so workgroups running together share A and B panels in their L2.
__device__ inline void gemm_tile_coords(int L, int nwg, int tiles_m, int tiles_n, int& tm, int& tn) {
  constexpr int GROUP_M = 8;
  if ((nwg & 7) == 0) L = (L & 7) * (nwg >> 3) + (L >> 3);
  const int per_group = GROUP_M * tiles_n;
  const int g = L / per_group;
  const int first_m = g * GROUP_M;
  const int gsize = min(tiles_m - first_m, GROUP_M);
  const int r = L - g * per_group;
  tm = first_m + r % gsize;
  tn = r / gsize;
}

}  // namespace mh
