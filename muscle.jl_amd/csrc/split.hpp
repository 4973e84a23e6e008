// Split-precision contraction (the precision ladder for Float64 operands).
#pragma once

#include "handle.hpp"

namespace mh {

// Chooses slice counts / K-chunking for plan.precision and fills err_bound.
void prepare_split(Plan& p, double target_rel_error);
size_t split_workspace_bytes(const Plan& p);
void run_split(Handle& h, Plan& p, void* C, const void* A, const void* B, bool ca, bool cb, void* ws);

}  // namespace mh
