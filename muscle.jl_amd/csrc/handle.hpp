// Library handle and cached contraction plans.
#pragma once

#include <hip/hip_runtime.h>

#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "kernels.hpp"
#include "lowering.hpp"

namespace mh {

// Offset tables of a plan, in this order; each has a lo and a hi part.
enum TabId { TAB_ROWA = 0, TAB_KA, TAB_KB, TAB_COLB, TAB_ROWC, TAB_COLC, TAB_BA, TAB_BB, TAB_BC, TAB_COUNT };

// A lowered contraction with its offset tables resident on the device.
struct Plan {
  enum Kind { EMPTY = 0, GEMM = 1, REDUCE = 2, SPLIT = 3, PERMUTE = 4 };
  int kind = EMPTY;
  int precision = MUSCLE_HIP_PRECISION_NATIVE;
  ContractionDesc d;
  // two-level offset tables (lo, hi for each TabId), concatenated; P of each
  std::vector<int64_t> htab;
  size_t tab_lo[TAB_COUNT] = {}, tab_hi[TAB_COUNT] = {};
  uint32_t tab_P[TAB_COUNT] = {};
  uint32_t tab_single[TAB_COUNT] = {};
  int64_t* dtab = nullptr;
  // device view of one table (valid once uploaded)
  OffTab tab(int id, uint32_t base = 0) const {
    OffTab t;
    t.lo = dtab + tab_lo[id];
    t.hi = dtab + tab_hi[id];
    t.div = FastDiv(tab_P[id]);
    t.base = base;
    t.single = tab_single[id];
    return t;
  }
  bool a_kmajor = false, b_kmajor = false;
  bool fp32_math = false;  // FP64 operands multiplied on FP32 MFMA (precision FP32)
  // tile GEMM split-K
  int gemm_ksplit = 1;
  int gemm_kchunk = 0;
  // Large GEMMs whose operands gather in short runs (e.g. many modes of
  // extent 2 in random orders) first copy them K-contiguous; `inner` is the
  // plan of the contraction on the packed copies (pack_a / pack_b say which).
  bool pack_a = false, pack_b = false;
  bool pack_a_kfast = false, pack_b_kfast = false;  // read order of the pack tiles
  std::unique_ptr<Plan> inner;
  // reduction kernel configuration
  int gshift = 0;
  int nsplit = 1;
  int64_t kchunk = 1;
  // split-precision configuration (sliced modes)
  int slices = 0;
  int split_mode = -1;  // 0 int8 base-255 digits (Ozaki), 1 bf16 x3, 2 fp16 x2, 3 modular (CRT)
  bool crt_split = false;  // modular form: complex products in Z_p x Z_p (two GEMMs per modulus)
  bool split_wide = false;  // sliced form on the 256 x 256 slice GEMM (large plans)
  int planes_a = 0, planes_b = 0;
  int64_t Mp = 0, Np = 0, Kp = 0;  // padded extents of the packed operands
  int esize = 1;                   // bytes per packed element
  int ngroups = 0;
  std::vector<int32_t> split_segs;    // SliceSeg records, flattened
  std::vector<int32_t> split_groups;  // SliceGroup records, flattened
  std::vector<double> split_coef;     // [2][ngroups]: real / imaginary combine weights
  void* split_dev = nullptr;          // device copy: segs | groups | coef
  size_t dev_groups_off = 0, dev_coef_off = 0;
  double err_bound = 0.0;
  std::list<std::string>::iterator lru_pos;

  Plan() = default;
  Plan(const Plan&) = delete;
  Plan& operator=(const Plan&) = delete;
  ~Plan();
};

// Size of the static device scratch of muscle_hip_binary_einsum_noalloc
// (scratch.hip), and its device address on `device` (resolved once per device,
// with that device current).
constexpr size_t kScratchBytes = size_t(32) << 20;
char* scratch_address(int device);

struct Handle {
  static constexpr size_t kMaxPlans = 8192;
  int device = 0;
  char* scratch = nullptr;  // the static scratch's address on this device
  int cus = 0;  // compute units of the device (queried once)
  hipStream_t stream = nullptr;
  // recorded on the old stream when the stream changes; the new stream waits
  // for it, so calls on different streams never share the workspace, the
  // split-K partials or a plan's tables at the same time
  hipEvent_t order_event = nullptr;
  std::mutex mu;
  std::unordered_map<std::string, std::unique_ptr<Plan>> cache;
  std::list<std::string> lru;
  void* ws = nullptr;
  size_t ws_size = 0;
  int complex_mult = 3;  // MUSCLE_HIP_OPTION_COMPLEX_MULT
  void* rb = nullptr;     // rocBLAS handle of the factorizations (created on first use)
  void* fws = nullptr;    // factorization workspace (matrix copies), grown like ws
  size_t fws_size = 0;
  static constexpr int kTuneClasses = 25;
  int tune[kTuneClasses] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1,
                            -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};  // batched tile per size class

  ~Handle();
  void ensure_workspace(size_t bytes);
  Plan* find_plan(const std::string& key);
  Plan* insert_plan(const std::string& key, std::unique_ptr<Plan> p);
};

extern thread_local std::string g_last_error;

std::unique_ptr<Plan> build_contraction_plan(const ContractionDesc& d, int precision);
// Split-K slices of a tile-GEMM plan (one slice for plans of other kinds).
void plan_gemm_ksplit(Plan& p);
void upload_tables(Plan& p, hipStream_t s);
// Enqueues one planned contraction on h.stream.  Its workspace is the
// handle's (grown as needed), or `ws` of `ws_size` bytes when given (no
// allocation then: usable under graph capture).
void run_contraction(Handle& h, Plan& p, const muscle_hip_tensor_t& c, const muscle_hip_tensor_t& a,
                     const muscle_hip_tensor_t* b, char* ws = nullptr, size_t ws_size = 0);
size_t contraction_workspace_bytes(const Plan& p, const muscle_hip_tensor_t& a, const muscle_hip_tensor_t* b);
// A plan that the caller owns (not in the handle's cache); kind 1 binary, 2 unary.
std::unique_ptr<Plan> make_plan(Handle& h, int kind, int precision, double target, const muscle_hip_tensor_t* c,
                                const muscle_hip_tensor_t* a, const muscle_hip_tensor_t* b);
Plan* get_plan(Handle& h, int kind, int precision, double target, const muscle_hip_tensor_t* c,
               const muscle_hip_tensor_t* a, const muscle_hip_tensor_t* b);
// Thin SVD (factor.cpp) and the release of its rocBLAS handle and workspace.
void svd_thin(Handle& h, const muscle_hip_tensor_t& U, const muscle_hip_tensor_t& S, const muscle_hip_tensor_t& V,
              const muscle_hip_tensor_t& A);
void qr_thin(Handle& h, const muscle_hip_tensor_t& Q, const muscle_hip_tensor_t& R, const muscle_hip_tensor_t& A);
void release_rb(Handle& h);

}  // namespace mh
