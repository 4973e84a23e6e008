// Index classification and lowering of a contraction to a fused
// batch x M x N x K description with per-operand strides.  Pure host code:
// nothing here touches the GPU, so it is exercised by CPU tests through
// muscle_hip_describe.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace mh {

enum ModeClass { CLS_BATCH = 0, CLS_M = 1, CLS_N = 2, CLS_K = 3, CLS_SUM_A = 4, CLS_SUM_B = 5 };

// One (possibly fused) loop dimension; stride[0..2] are the strides of the
// A, B and C operands (0 where the operand does not carry the mode).
struct Dim {
  int64_t extent;
  int64_t s[3];
};

struct ModeInfo {
  int32_t label;
  int64_t extent;
  int64_t s[3];  // summed over repeated occurrences (diagonal)
  bool in[3];
  int cls;
};

// batch x (M x K) * (K x N) -> (M x N); dims lists are fastest-first.
struct ContractionDesc {
  int dt[3] = {0, 0, 0};  // dtype of A, B, C (after an optional swap)
  std::vector<Dim> m, n, k, batch;
  int64_t M = 1, N = 1, K = 1, B = 1;
  bool swapped = false;   // A/B exchanged so C's fastest mode is an N mode
  bool empty = false;     // C has no elements
  bool has_b = true;      // false for unary reductions
  std::vector<ModeInfo> modes;  // classification, first-appearance order (c, a, b)
};

// Elementwise description for hadamard: dims fastest-first in C.
struct ElementwiseDesc {
  int dt[3] = {0, 0, 0};
  std::vector<Dim> dims;
  int64_t total = 1;
};

ContractionDesc lower_binary(const muscle_hip_tensor_t& c, const muscle_hip_tensor_t& a,
                             const muscle_hip_tensor_t& b);
ContractionDesc lower_unary(const muscle_hip_tensor_t& y, const muscle_hip_tensor_t& x);
ElementwiseDesc lower_hadamard(const muscle_hip_tensor_t& c, const muscle_hip_tensor_t& a,
                               const muscle_hip_tensor_t& b);

// Splits a fastest-first dim list into a fast part (lo) of product P close to
// sqrt(product) and a slow part (hi), splitting one dim at a factor boundary
// when that gets closer; offsets then factor as lo-offset + hi-offset.
// Classes of at most kSingleLevelMax points are not split (lo = d, P = total).
constexpr int64_t kSingleLevelMax = 65536;
void split_dims(const std::vector<Dim>& d, std::vector<Dim>& lo, std::vector<Dim>& hi, int64_t& P,
                int64_t single_max = kSingleLevelMax);

// Offsets of every point of a mixed-radix loop nest (fastest-first dims),
// using stride slot `which` (0 A, 1 B, 2 C).  Length = product of extents
// (1 for an empty list).
std::vector<int64_t> offset_table(const std::vector<Dim>& dims, int which);

// Byte key identifying a call's shape/stride/dtype signature for the plan cache.
std::string plan_key(int kind, int precision, const muscle_hip_tensor_t* t0,
                     const muscle_hip_tensor_t* t1, const muscle_hip_tensor_t* t2);

void check_tensor(const muscle_hip_tensor_t* t, const char* name, bool need_data = true);

}  // namespace mh
