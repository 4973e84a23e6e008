// The library's static device scratch, used by muscle_hip_binary_einsum_noalloc
// for a contraction's offset tables and its workspace (split-K and reduction
// partial sums, operand casts).  It is a file-scope __device__ array: every
// GPU the library runs on holds its own copy, at its own address, from the
// moment the code object is loaded until the process ends.  Host code never
// passes the array's name to a kernel (on the host it is only a stand-in);
// it uses the device address that hipGetSymbolAddress returns, resolved once
// per device with that device current, and kept here.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "handle.hpp"

namespace mh {

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::vector<char*> g_scratch_ptr;  // device address of `scratch`, per device
}  // namespace

char* scratch_address(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (!g_scratch_ptr[(size_t)device]) {
    int cur = -1;
    MH_HIP_CHECK(hipGetDevice(&cur));
    if (cur != device)
      fail(MUSCLE_HIP_INVALID_ARGUMENT, "scratch set-up for device " + std::to_string(device) +
                                            " while device " + std::to_string(cur) + " is current");
    void* p = nullptr;
    MH_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(MUSCLE_HIP_HIP_ERROR, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[(size_t)device] = static_cast<char*>(p);
  }
  return g_scratch_ptr[(size_t)device];
}

}  // namespace mh
