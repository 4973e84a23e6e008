// Split-precision contraction (the precision ladder for Float64 operands).
//
// FP32 runs on the gather GEMM with FP32 MFMA arithmetic (no extra passes).
// BF16X3, FP16X2 and OZAKI_INT8 are "sliced": every operand value is
// written as a short sum of low-precision slices, the slice cross products
// run as dense low-precision GEMMs on MFMA, and an FP64 pass combines them.
#pragma once

#include "handle.hpp"

namespace mh {

// One K-segment of a slice-product group: plane-slice indices into the packed
// A and B arrays and the K range [k0, k1) in elements.
struct SliceSeg {
  int32_t a, b, k0, k1;
};
// A group of segments accumulated into one int32/FP32 output plane.
struct SliceGroup {
  int32_t first, count;
};

// Chooses slice counts / K-chunking for plan.precision and fills err_bound.
void prepare_split(Plan& p, double target_rel_error);
size_t split_workspace_bytes(const Plan& p);
void run_split(Handle& h, Plan& p, void* C, const void* A, const void* B, bool ca, bool cb, void* ws);
// Device buffers of a split plan (segments, groups, combine coefficients).
void upload_split_tables(Plan& p, hipStream_t s);

}  // namespace mh
