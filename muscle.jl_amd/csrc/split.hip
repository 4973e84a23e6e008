// Split-precision contraction path (precision ladder).
#include "split.hpp"

namespace mh {

void prepare_split(Plan& p, double target_rel_error) {
  (void)p;
  (void)target_rel_error;
  fail(MUSCLE_HIP_UNSUPPORTED, "split-precision modes are not built yet");
}

size_t split_workspace_bytes(const Plan&) { return 0; }

void run_split(Handle&, Plan&, void*, const void*, const void*, bool, bool, void*) {
  fail(MUSCLE_HIP_UNSUPPORTED, "split-precision modes are not built yet");
}

}  // namespace mh
