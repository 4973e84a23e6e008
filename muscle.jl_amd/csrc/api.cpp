// C ABI of libmuscle_hip: handle, plan cache, workspace and the three
// operations of Muscle.jl's contraction hot path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "handle.hpp"
#include "kernels.hpp"
#include "lowering.hpp"
#include "crt.hpp"
#include "split.hpp"

namespace mh {

thread_local std::string g_last_error;

namespace {

int64_t product_distinct(const muscle_hip_tensor_t& t) {
  int64_t p = 1;
  for (int i = 0; i < t.rank; ++i) {
    bool seen = false;
    for (int j = 0; j < i; ++j) seen = seen || t.modes[j] == t.modes[i];
    if (!seen) p *= t.extents[i];
  }
  return p;
}

// [min, max] element offsets an operand touches.
void span_of(const muscle_hip_tensor_t& t, int64_t& lo, int64_t& hi) {
  lo = hi = 0;
  for (int i = 0; i < t.rank; ++i) {
    if (t.extents[i] == 0) {
      lo = 0;
      hi = -1;
      return;
    }
    const int64_t d = (t.extents[i] - 1) * t.strides[i];
    if (d < 0) lo += d;
    else hi += d;
  }
}

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

void append_table(std::vector<int64_t>& h, std::vector<size_t>& at, const std::vector<int64_t>& t) {
  at.push_back(h.size());
  if (t.empty()) h.push_back(0);
  else h.insert(h.end(), t.begin(), t.end());
}

}  // namespace

void Handle::ensure_workspace(size_t bytes) {
  if (bytes <= ws_size) return;
  if (ws) {
    MH_HIP_CHECK(hipStreamSynchronize(stream));
    MH_HIP_CHECK(hipFree(ws));
    ws = nullptr;
    ws_size = 0;
  }
  size_t want = std::max(bytes, size_t(1) << 20);
  if (hipMalloc(&ws, want) != hipSuccess) {
    (void)hipGetLastError();
    ws = nullptr;
    fail(MUSCLE_HIP_OUT_OF_MEMORY, "workspace allocation of " + std::to_string(want) + " bytes failed");
  }
  ws_size = want;
}

Plan::~Plan() {
  if (dtab) (void)hipFree(dtab);
  if (split_dev) (void)hipFree(split_dev);
}

Handle::~Handle() {
  if (order_event) (void)hipEventDestroy(order_event);
  if (ws) (void)hipFree(ws);
  cache.clear();  // plans free their device tables
  release_rb(*this);
}

Plan* Handle::find_plan(const std::string& key) {
  auto it = cache.find(key);
  if (it == cache.end()) return nullptr;
  lru.splice(lru.begin(), lru, it->second->lru_pos);
  return it->second.get();
}

Plan* Handle::insert_plan(const std::string& key, std::unique_ptr<Plan> p) {
  if (cache.size() >= kMaxPlans) {
    // evict the least recently used plan; its tables may still be read by
    // queued kernels, so drain the device first
    MH_HIP_CHECK(hipDeviceSynchronize());
    const std::string victim = lru.back();
    lru.pop_back();
    auto it = cache.find(victim);
    if (it != cache.end()) cache.erase(it);
  }
  lru.push_front(key);
  p->lru_pos = lru.begin();
  Plan* raw = p.get();
  cache[key] = std::move(p);
  return raw;
}

void upload_tables(Plan& p, hipStream_t s) {
  const size_t bytes = p.htab.size() * sizeof(int64_t);
  if (hipMalloc(&p.dtab, bytes) != hipSuccess) {
    (void)hipGetLastError();
    p.dtab = nullptr;
    fail(MUSCLE_HIP_OUT_OF_MEMORY, "plan table allocation failed");
  }
  MH_HIP_CHECK(hipMemcpyAsync(p.dtab, p.htab.data(), bytes, hipMemcpyHostToDevice, s));
}

// Split-K for tile GEMMs with too few output tiles to fill 256 CUs: slices
// of at least 32 along K (two K-tiles; config 1 takes 15 us per call with
// them, 17.4 us with 64), partial sums bounded to 256 MiB.  Plans of other
// kinds keep one slice.
void plan_gemm_ksplit(Plan& p) {
  const ContractionDesc& d = p.d;
  p.gemm_ksplit = 1;
  p.gemm_kchunk = (int)std::min<int64_t>(d.K, INT32_MAX);
  if (p.kind != Plan::GEMM) return;
  const int64_t tiles = ((d.M + 63) / 64) * ((d.N + 63) / 64) * d.B;
  if (tiles < 512 && d.K >= 128) {
    int64_t ks = std::min<int64_t>((1024 + tiles - 1) / tiles, d.K / 32);
    const int64_t elem = (int64_t)dtype_size(d.dt[2]);
    const int64_t cap = (int64_t(256) << 20) / std::max<int64_t>(1, d.M * d.N * d.B * elem);
    ks = std::max<int64_t>(1, std::min<int64_t>({ks, 64, cap}));
    int64_t kc = (d.K + ks - 1) / ks;
    kc = (kc + 15) / 16 * 16;
    p.gemm_kchunk = (int)kc;
    p.gemm_ksplit = (int)((d.K + kc - 1) / kc);
  }
}

// Builds offset tables and picks a kernel for a lowered contraction.
std::unique_ptr<Plan> build_contraction_plan(const ContractionDesc& d, int precision) {
  std::unique_ptr<Plan> p(new Plan());
  p->d = d;
  p->precision = precision;
  if (d.empty) {
    p->kind = Plan::EMPTY;
    return p;
  }
  const int64_t lim32 = int64_t(1) << 32;
  if (d.M >= lim32 || d.N >= lim32 || d.K >= lim32 || d.B >= lim32)
    fail(MUSCLE_HIP_UNSUPPORTED, "an index class with 2^32 or more elements");
  // A unary permutation (no summed modes) whose input and output run
  // contiguous along different modes is a tiled transpose: the tile rows are
  // the output's fastest dims (at least 32 elements), the tile columns the
  // input's fastest remaining dims, whose input strides move to the B slot.
  ContractionDesc pd;
  bool permute = false;
  if (!d.has_b && d.k.empty() && d.M >= (int64_t(1) << 14) && d.m.size() >= 2) {
    std::vector<int> byx(d.m.size());
    for (size_t i = 0; i < byx.size(); ++i) byx[i] = (int)i;
    auto xs = [&](int i) { return d.m[i].s[0] < 0 ? -d.m[i].s[0] : d.m[i].s[0]; };
    std::stable_sort(byx.begin(), byx.end(), [&](int u, int v) { return xs(u) < xs(v); });
    std::vector<char> row(d.m.size(), 0), col(d.m.size(), 0);
    int64_t pr = 1;
    for (size_t i = 0; i < d.m.size() && pr < 32; ++i) {
      row[i] = 1;
      pr *= d.m[i].extent;
    }
    int64_t pc = 1;
    for (int j : byx) {
      if (pc >= 32) break;
      if (row[j]) continue;
      col[j] = 1;
      pc *= d.m[j].extent;
    }
    int first_col = -1;
    for (int j : byx)
      if (col[j]) {
        first_col = j;
        break;
      }
    if (byx[0] != 0 && first_col >= 0 && xs(first_col) < xs(0)) {
      permute = true;
      pd = d;
      pd.m.clear();
      pd.n.clear();
      for (size_t i = 0; i < d.m.size(); ++i)
        if (!col[i]) pd.m.push_back(d.m[i]);
      for (int j : byx)
        if (col[j]) {
          Dim x = d.m[j];
          x.s[1] = x.s[0];
          x.s[0] = 0;
          pd.n.push_back(x);
        }
      pd.M = pd.N = 1;
      for (auto& x : pd.m) pd.M *= x.extent;
      for (auto& x : pd.n) pd.N *= x.extent;
    }
  }
  const ContractionDesc& td = permute ? pd : d;
  // two-level tables: (class dims, operand slot) per TabId
  const std::vector<Dim>* cls[TAB_COUNT] = {&td.m, &td.k, &td.k, &td.n, &td.m, &td.n, &td.batch, &td.batch, &td.batch};
  const int slot[TAB_COUNT] = {0, 0, 1, 1, 2, 2, 0, 1, 2};
  for (int t = 0; t < TAB_COUNT; ++t) {
    std::vector<Dim> lo, hi;
    int64_t P;
    split_dims(*cls[t], lo, hi, P);
    std::vector<size_t> at;
    append_table(p->htab, at, offset_table(lo, slot[t]));
    append_table(p->htab, at, offset_table(hi, slot[t]));
    p->tab_lo[t] = at[0];
    p->tab_hi[t] = at[1];
    p->tab_P[t] = (uint32_t)P;
    p->tab_single[t] = hi.empty() ? 1u : 0u;
  }
  if (permute) {
    p->d = pd;
    p->kind = Plan::PERMUTE;
    return p;
  }

  auto min_stride = [](const std::vector<Dim>& v, int w) {
    int64_t b = INT64_MAX;
    for (auto& x : v)
      if (x.s[w] != 0) b = std::min(b, x.s[w] < 0 ? -x.s[w] : x.s[w]);
    return b;
  };
  p->a_kmajor = min_stride(d.k, 0) < min_stride(d.m, 0);
  p->b_kmajor = min_stride(d.k, 1) < min_stride(d.n, 1);

  const bool fits32 = d.M < (int64_t(1) << 31) && d.N < (int64_t(1) << 31) && d.K < (int64_t(1) << 31);
  const bool real_fp = true;
  if (precision != MUSCLE_HIP_PRECISION_NATIVE && d.has_b && is_double(d.dt[2]) && d.M >= 16 && d.N >= 16 &&
      d.K >= 16 && fits32 && d.B <= 65535) {
    p->kind = Plan::SPLIT;
  } else if (d.has_b && real_fp && fits32 &&
             ((d.M >= 32 && d.N >= 32 && d.K >= 8) || (d.K >= 16 && d.M * d.N >= 4096))) {
    // skinny products (one side under 32, e.g. 2^18 x 8 x 2^10) also go to the
    // MFMA tiles: they read each operand once, where the reduction kernel
    // would re-read the long side once per output of the short one
    p->kind = Plan::GEMM;
  } else {
    p->kind = Plan::REDUCE;
  }
  plan_gemm_ksplit(*p);
  // reduction configuration (also the fallback of the other kinds)
  const int64_t outputs = d.M * d.N * d.B;
  int gshift = 0;
  while (gshift < 6 && (int64_t(8) << gshift) < d.K) ++gshift;
  // a unary sum over modes the input holds at larger strides than the
  // outputs' fastest dim: one lane per output, so every load instruction
  // reads neighbouring elements instead of lanes striding over k
  if (!d.has_b && !d.m.empty() && !d.k.empty() && outputs >= 64) {
    int64_t kmin = INT64_MAX;
    for (auto& x : d.k) kmin = std::min<int64_t>(kmin, x.s[0] < 0 ? -x.s[0] : x.s[0]);
    const int64_t mfast = d.m[0].s[0] < 0 ? -d.m[0].s[0] : d.m[0].s[0];
    if (kmin > mfast) gshift = 0;
  }
  int64_t threads = outputs << gshift;
  int64_t nsplit = 1;
  const int64_t target = 256 * 1024;
  if (threads < target && d.K > (int64_t(32) << gshift)) {
    nsplit = std::min<int64_t>((target + threads - 1) / threads, d.K / (int64_t(32) << gshift));
    nsplit = std::max<int64_t>(1, std::min<int64_t>(nsplit, 1024));
  }
  p->gshift = gshift;
  p->nsplit = (int)nsplit;
  p->kchunk = (d.K + nsplit - 1) / nsplit;
  if (p->kchunk == 0) p->kchunk = 1;
  return p;
}

// Decides whether a GEMM plan packs its operands first, and builds the plan
// of the contraction on the packed copies.  An operand is packed when the
// contraction is large (2^27 index points or more), its gather direction
// runs contiguous for less than 64 bytes, the tiles re-read it at least four
// times (the other open side spans 256 or more), every one of its dims has a
// stride, and the copy is at most 2 GiB.
void consider_packing(Plan& p) {
  if (p.kind != Plan::GEMM || p.fp32_math || p.precision != MUSCLE_HIP_PRECISION_NATIVE) return;
  const ContractionDesc& d = p.d;
  if ((double)d.M * (double)d.N * (double)d.K * (double)d.B < (double)(int64_t(1) << 27) || d.B > 65535) return;
  auto lowest = [](const std::vector<Dim>& v, int w) {
    int64_t b = INT64_MAX;
    for (auto& x : v) b = std::min(b, x.s[w] < 0 ? -x.s[w] : x.s[w]);
    return b;
  };
  auto run = [](const std::vector<Dim>& v, int w) {
    return (!v.empty() && (v[0].s[w] == 1 || v[0].s[w] == -1)) ? v[0].extent : int64_t(1);
  };
  auto all_strided = [](const std::vector<Dim>& v, int w) {
    for (auto& x : v)
      if (x.s[w] == 0) return false;
    return true;
  };
  const size_t es[2] = {dtype_size(d.dt[0]), dtype_size(d.dt[1])};
  const std::vector<Dim>* rows[2] = {&d.m, &d.n};
  const bool kmaj[2] = {p.a_kmajor, p.b_kmajor};
  bool pack[2] = {false, false}, kfast[2] = {false, false};
  for (int w = 0; w < 2; ++w) {
    const int64_t r = kmaj[w] ? run(d.k, w) : run(*rows[w], w);
    const double bytes = (double)(w == 0 ? d.M : d.N) * (double)d.K * (double)d.B * (double)es[w];
    const int64_t other = w == 0 ? d.N : d.M;
    pack[w] = (double)r * (double)es[w] < 64.0 && other >= 256 && bytes <= (double)(int64_t(2) << 30) &&
              all_strided(d.k, w) && all_strided(*rows[w], w) && all_strided(d.batch, w);
    kfast[w] = lowest(d.k, w) < lowest(*rows[w], w);
  }
  if (!pack[0] && !pack[1]) return;
  // strides of the packed copies: [batch][rows][K], K contiguous
  ContractionDesc di = d;
  for (int w = 0; w < 2; ++w) {
    if (!pack[w]) continue;
    int64_t cur = 1;
    for (auto& x : di.k) {
      x.s[w] = cur;
      cur *= x.extent;
    }
    for (auto& x : (w == 0 ? di.m : di.n)) {
      x.s[w] = cur;
      cur *= x.extent;
    }
    for (auto& x : di.batch) {
      x.s[w] = cur;
      cur *= x.extent;
    }
  }
  std::unique_ptr<Plan> in = build_contraction_plan(di, MUSCLE_HIP_PRECISION_NATIVE);
  if (in->kind != Plan::GEMM) return;
  p.pack_a = pack[0];
  p.pack_b = pack[1];
  p.pack_a_kfast = kfast[0];
  p.pack_b_kfast = kfast[1];
  p.inner = std::move(in);
}

namespace {

struct Operand {
  const void* data;
  int dtype;
};

// Casts an operand to the compute precision inside the workspace when its
// precision differs (Float32 with Float64 -> Float64, as Julia promotes).
const void* promote_operand(Handle& h, const muscle_hip_tensor_t& t, bool want_double, char*& ws_cursor,
                            int& dtype_out) {
  dtype_out = t.dtype;
  if (is_double(t.dtype) == want_double) return t.data;
  int64_t lo, hi;
  span_of(t, lo, hi);
  const int64_t n = hi - lo + 1;
  const int dst = make_dtype(want_double, is_complex(t.dtype));
  char* buf = ws_cursor;
  ws_cursor += align_up((size_t)std::max<int64_t>(n, 1) * dtype_size(dst));
  const char* src = (const char*)t.data + lo * (int64_t)dtype_size(t.dtype);
  launch_cast(t.dtype, dst, src, buf, n, h.stream);
  dtype_out = dst;
  return buf - lo * (int64_t)dtype_size(dst);
}

size_t cast_bytes(const muscle_hip_tensor_t& t, bool want_double) {
  if (is_double(t.dtype) == want_double) return 0;
  int64_t lo, hi;
  span_of(t, lo, hi);
  return align_up((size_t)std::max<int64_t>(hi - lo + 1, 1) * dtype_size(make_dtype(want_double, is_complex(t.dtype))));
}

void check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) fail(MUSCLE_HIP_HIP_ERROR, std::string("kernel launch: ") + hipGetErrorString(e));
}

void fill_info(muscle_hip_info_t* info, const Plan& p, const muscle_hip_tensor_t* c, const muscle_hip_tensor_t* a,
               const muscle_hip_tensor_t* b) {
  if (!info) return;
  std::memset(info, 0, sizeof *info);
  const ContractionDesc& d = p.d;
  info->kernel = p.kind == Plan::GEMM      ? 1
                 : p.kind == Plan::REDUCE  ? 2
                 : p.kind == Plan::SPLIT   ? (p.split_mode == kSplitModeCrt ? (p.crt_split ? 8 : 7) : 3)
                 : p.kind == Plan::PERMUTE ? 6
                                           : 0;
  info->precision = p.fp32_math ? MUSCLE_HIP_PRECISION_FP32
                                 : (p.kind == Plan::SPLIT ? p.precision : MUSCLE_HIP_PRECISION_NATIVE);
  info->slices = p.slices;
  info->swapped = d.swapped;
  info->m = d.M;
  info->n = d.N;
  info->k = d.K;
  info->batch = d.B;
  const bool ca = is_complex(d.dt[0]), cb = d.has_b && is_complex(d.dt[1]);
  const double per = d.has_b ? ((ca && cb) ? 8.0 : (ca || cb) ? 4.0 : 2.0) : (ca ? 2.0 : 1.0);
  info->flops = per * (double)d.M * (double)d.N * (double)d.K * (double)d.B;
  double bytes = 0;
  for (const muscle_hip_tensor_t* t : {c, a, b})
    if (t) bytes += (double)product_distinct(*t) * (double)dtype_size(t->dtype);
  info->bytes = bytes;
  info->err_bound = p.err_bound;
}

}  // namespace

namespace {

// Workspace of one contraction: casts | packed A | packed B | partial sums or
// split-precision state, at these offsets.
struct WsPlan {
  size_t pack_a_off, pack_b_off, partial_off, split_off, need;
  size_t es_a, es_b;
};

WsPlan workspace_plan(const Plan& p, const muscle_hip_tensor_t* ka, const muscle_hip_tensor_t* kb) {
  const ContractionDesc& d = p.d;
  const bool dbl = is_double(d.dt[2]);
  WsPlan w;
  size_t need = cast_bytes(*ka, dbl) + (kb ? cast_bytes(*kb, dbl) : 0);
  const size_t csize = dtype_size(d.dt[2]);
  // packed operand copies (element size after the precision promotion)
  w.es_a = dtype_size(make_dtype(dbl, is_complex(ka->dtype)));
  w.es_b = kb ? dtype_size(make_dtype(dbl, is_complex(kb->dtype))) : 0;
  w.pack_a_off = need;
  if (p.pack_a) need += align_up((size_t)(d.M * d.K * d.B) * w.es_a);
  w.pack_b_off = need;
  if (p.pack_b) need += align_up((size_t)(d.N * d.K * d.B) * w.es_b);
  const Plan& q = p.inner ? *p.inner : p;
  w.partial_off = need;
  if (p.kind == Plan::REDUCE && p.nsplit > 1) need += align_up((size_t)(d.M * d.N * d.B) * csize * p.nsplit);
  if (p.kind == Plan::GEMM && q.gemm_ksplit > 1)
    need += align_up((size_t)(d.M * d.N * d.B) * csize * q.gemm_ksplit);
  w.split_off = need;
  if (p.kind == Plan::SPLIT) need += split_workspace_bytes(p);
  w.need = need;
  return w;
}

}  // namespace

size_t contraction_workspace_bytes(const Plan& p, const muscle_hip_tensor_t& a, const muscle_hip_tensor_t* b) {
  if (p.kind == Plan::EMPTY) return 0;
  const muscle_hip_tensor_t* ka = &a;
  const muscle_hip_tensor_t* kb = b;
  if (p.d.swapped) std::swap(ka, kb);
  return workspace_plan(p, ka, kb).need;
}

void run_contraction(Handle& h, Plan& p, const muscle_hip_tensor_t& c, const muscle_hip_tensor_t& a,
                     const muscle_hip_tensor_t* b, char* ws, size_t ws_size) {
  if (p.kind == Plan::EMPTY) return;
  const ContractionDesc& d = p.d;
  const bool dbl = is_double(d.dt[2]);
  // operands in kernel order (after the lowering's optional A/B exchange)
  const muscle_hip_tensor_t* ka = &a;
  const muscle_hip_tensor_t* kb = b;
  if (d.swapped) std::swap(ka, kb);
  const size_t csize = dtype_size(d.dt[2]);
  const WsPlan w = workspace_plan(p, ka, kb);
  const size_t es_a = w.es_a, es_b = w.es_b;
  Plan& q = p.inner ? *p.inner : p;  // the plan whose tables the GEMM reads
  char* base;
  if (ws) {  // a caller-provided workspace (e.g. inside a captured graph): no allocation here
    if (w.need > ws_size) fail(MUSCLE_HIP_INVALID_ARGUMENT, "contraction workspace too small");
    base = ws;
  } else {
    if (w.need) h.ensure_workspace(w.need);
    base = (char*)h.ws;
  }
  char* cursor = base;
  int dta, dtb = d.dt[1];
  const void* A = promote_operand(h, *ka, dbl, cursor, dta);
  const void* B = kb ? promote_operand(h, *kb, dbl, cursor, dtb) : nullptr;
  if (p.kind == Plan::SPLIT) {
    run_split(h, p, c.data, A, B, is_complex(dta), is_complex(dtb), base + w.split_off);
    check_launch();
    return;
  }
  if (p.kind == Plan::PERMUTE) {
    launch_permute((int)csize, PermuteArgs{A, c.data, p.tab(TAB_ROWA), p.tab(TAB_ROWC), p.tab(TAB_COLB),
                                           p.tab(TAB_COLC), d.M, d.N},
                   h.stream);
    check_launch();
    return;
  }
  if (p.kind == Plan::GEMM) {
    // the planner sends empty K (sums of nothing) to the reduction; the tile
    // kernels read their first K-tile's offsets unconditionally
    if (d.K < 1) fail(MUSCLE_HIP_INVALID_ARGUMENT, "internal: tile GEMM planned for an empty K");
    if (p.pack_a) {
      char* buf = base + w.pack_a_off;
      launch_pack_operand((int)es_a, PackOpArgs{A, buf, p.tab(TAB_ROWA), p.tab(TAB_KA), p.tab(TAB_BA), (int32_t)d.M,
                                                (int32_t)d.K, (int32_t)d.B, p.pack_a_kfast},
                          h.stream);
      check_launch();
      A = buf;
    }
    if (p.pack_b) {
      char* buf = base + w.pack_b_off;
      launch_pack_operand((int)es_b, PackOpArgs{B, buf, p.tab(TAB_COLB), p.tab(TAB_KB), p.tab(TAB_BB), (int32_t)d.N,
                                                (int32_t)d.K, (int32_t)d.B, p.pack_b_kfast},
                          h.stream);
      check_launch();
      B = buf;
    }
    GemmArgs g;
    std::memset(&g, 0, sizeof g);
    g.A = A;
    g.B = B;
    g.C = c.data;
    g.rowA = q.tab(TAB_ROWA);
    g.kA = q.tab(TAB_KA);
    g.kB = q.tab(TAB_KB);
    g.colB = q.tab(TAB_COLB);
    g.rowC = q.tab(TAB_ROWC);
    g.colC = q.tab(TAB_COLC);
    g.M = (int32_t)d.M;
    g.N = (int32_t)d.N;
    g.K = (int32_t)d.K;
    g.a_kmajor = q.a_kmajor;
    g.b_kmajor = q.b_kmajor;
    g.complex3m = h.complex_mult == 3;
    g.ksplit = q.gemm_ksplit;
    g.kchunk = q.gemm_kchunk;
    g.partial = base + w.partial_off;
    g.c_stream = csize >= 16 && (double)d.M * (double)d.N * (double)d.B * (double)csize >= kStreamStoreBytes;
    // grid.y carries the batch; split launches beyond the 65535 limit
    for (int64_t b0 = 0; b0 < d.B; b0 += 65535) {
      g.batch = (int32_t)std::min<int64_t>(65535, d.B - b0);
      g.bA = q.tab(TAB_BA, (uint32_t)b0);
      g.bB = q.tab(TAB_BB, (uint32_t)b0);
      g.bC = q.tab(TAB_BC, (uint32_t)b0);
      launch_gemm_mfma(dbl, p.fp32_math, is_complex(dta), is_complex(dtb), g, h.stream);
      check_launch();
    }
    return;
  }
  ReduceArgs r;
  std::memset(&r, 0, sizeof r);
  r.A = A;
  r.B = B;
  r.C = c.data;
  r.partial = base + w.partial_off;
  r.rowA = p.tab(TAB_ROWA);
  r.kA = p.tab(TAB_KA);
  r.kB = p.tab(TAB_KB);
  r.colB = p.tab(TAB_COLB);
  r.rowC = p.tab(TAB_ROWC);
  r.colC = p.tab(TAB_COLC);
  r.bA = p.tab(TAB_BA);
  r.bB = p.tab(TAB_BB);
  r.bC = p.tab(TAB_BC);
  r.M = d.M;
  r.N = d.N;
  r.K = d.K;
  r.batch = d.B;
  r.outputs = d.M * d.N * d.B;
  r.kchunk = p.kchunk;
  r.gshift = p.gshift;
  r.nsplit = p.nsplit;
  r.small_index = r.outputs < (int64_t(1) << 32) ? 1 : 0;
  if (d.k.size() == 1) {
    r.kstepA = d.k[0].s[0];
    r.kstepB = d.k[0].s[1];
  }
  if (r.small_index) {
    r.divN = FastDiv((uint32_t)d.N);
    r.divM = FastDiv((uint32_t)d.M);
  }
  launch_reduce(dbl, is_complex(dta), kb && is_complex(dtb), d.has_b, r, h.stream);
  check_launch();
}

std::unique_ptr<Plan> make_plan(Handle& h, int kind, int precision, double target, const muscle_hip_tensor_t* c,
                                const muscle_hip_tensor_t* a, const muscle_hip_tensor_t* b) {
  ContractionDesc d = kind == 1 ? lower_binary(*c, *a, *b) : lower_unary(*c, *a);
  std::unique_ptr<Plan> np = build_contraction_plan(d, precision);
  if (np->kind == Plan::SPLIT) prepare_split(*np, target);
  consider_packing(*np);
  if (np->kind != Plan::EMPTY) upload_tables(*np, h.stream);
  if (np->inner) upload_tables(*np->inner, h.stream);
  upload_split_tables(*np, h.stream);
  return np;
}

Plan* get_plan(Handle& h, int kind, int precision, double target, const muscle_hip_tensor_t* c,
               const muscle_hip_tensor_t* a, const muscle_hip_tensor_t* b) {
  std::string key = plan_key(kind, precision, c, a, b);
  key.append((const char*)&target, sizeof target);
  Plan* p = h.find_plan(key);
  if (p) return p;
  return h.insert_plan(key, make_plan(h, kind, precision, target, c, a, b));
}

// ---- elementwise planning

// When b (with 2^14 or more elements) runs contiguous along other modes than
// c, the product goes through 32 x 32 tiles: the tile rows are c's fastest
// dims, the columns b's fastest remaining dims (reordered into w).
void plan_hadamard_tiles(const ElementwiseDesc& e, EwArgs& w) {
  const std::vector<Dim>& d = e.dims;
  if (e.total < (int64_t(1) << 16) || d.size() < 2) return;
  auto sb = [&](size_t i) { return d[i].s[1] < 0 ? -d[i].s[1] : d[i].s[1]; };
  int64_t nb = 1;
  for (size_t i = 0; i < d.size(); ++i)
    if (d[i].s[1] != 0) nb *= d[i].extent;
  if (nb < (int64_t(1) << 14) || sb(0) == 0) return;  // small b, or b broadcast along c's fastest dim
  std::vector<size_t> byb;
  for (size_t i = 0; i < d.size(); ++i)
    if (d[i].s[1] != 0) byb.push_back(i);
  std::stable_sort(byb.begin(), byb.end(), [&](size_t u, size_t v) { return sb(u) < sb(v); });
  if (byb[0] == 0) return;  // b and c both run along c's fastest dim
  std::vector<char> row(d.size(), 0), col(d.size(), 0);
  int64_t pr = 1;
  for (size_t i = 0; i < d.size() && pr < 32; ++i) {
    row[i] = 1;
    pr *= d[i].extent;
  }
  int64_t pc = 1;
  size_t first = d.size();
  for (size_t j : byb) {
    if (pc >= 32) break;
    if (row[j]) continue;
    if (first == d.size()) first = j;
    col[j] = 1;
    pc *= d[j].extent;
  }
  if (first == d.size() || sb(first) >= sb(0)) return;
  std::vector<size_t> order;
  for (size_t i = 0; i < d.size(); ++i)
    if (row[i]) order.push_back(i);
  const int32_t Rr = (int32_t)order.size();
  for (size_t j : byb)
    if (col[j]) order.push_back(j);
  const int32_t Rc = (int32_t)order.size() - Rr;
  for (size_t i = 0; i < d.size(); ++i)
    if (!row[i] && !col[i]) order.push_back(i);
  int64_t nrows = 1, ncols = 1, nouter = 1;
  for (size_t k = 0; k < order.size(); ++k)
    (k < (size_t)Rr ? nrows : k < (size_t)(Rr + Rc) ? ncols : nouter) *= d[order[k]].extent;
  if ((double)((nrows + 31) / 32) * (double)((ncols + 31) / 32) * (double)nouter >= 2147483648.0) return;
  w.tiled = 1;
  w.Rr = Rr;
  w.Rc = Rc;
  w.nrows = nrows;
  w.ncols = ncols;
  w.nouter = nouter;
  for (size_t k = 0; k < order.size(); ++k) {
    const Dim& x = d[order[k]];
    w.ext[k] = x.extent;
    w.sA[k] = x.s[0];
    w.sB[k] = x.s[1];
    w.sC[k] = x.s[2];
    w.div[k] = FastDiv(w.small_index ? (uint32_t)x.extent : 1u);
  }
}


}  // namespace mh

using namespace mh;

template <typename F>
static int guarded(F&& f) {
  try {
    f();
    return MUSCLE_HIP_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return MUSCLE_HIP_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return MUSCLE_HIP_INVALID_ARGUMENT;
  }
}

extern "C" {

int muscle_hip_version(void) { return 100; }

const char* muscle_hip_status_string(int status) {
  switch (status) {
    case MUSCLE_HIP_OK: return "MUSCLE_HIP_OK";
    case MUSCLE_HIP_INVALID_ARGUMENT: return "MUSCLE_HIP_INVALID_ARGUMENT";
    case MUSCLE_HIP_DIM_MISMATCH: return "MUSCLE_HIP_DIM_MISMATCH";
    case MUSCLE_HIP_UNSUPPORTED: return "MUSCLE_HIP_UNSUPPORTED";
    case MUSCLE_HIP_HIP_ERROR: return "MUSCLE_HIP_HIP_ERROR";
    case MUSCLE_HIP_OUT_OF_MEMORY: return "MUSCLE_HIP_OUT_OF_MEMORY";
    case MUSCLE_HIP_NOT_READY: return "MUSCLE_HIP_NOT_READY";
    case MUSCLE_HIP_REJECTED: return "MUSCLE_HIP_REJECTED";
  }
  return "MUSCLE_HIP_UNKNOWN_STATUS";
}

const char* muscle_hip_last_error(void) { return g_last_error.c_str(); }

int muscle_hip_create(muscle_hip_handle_t* handle) {
  return guarded([&] {
    if (!handle) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null handle pointer");
    int dev = 0;
    MH_HIP_CHECK(hipGetDevice(&dev));
    std::unique_ptr<Handle> h(new Handle());
    h->device = dev;
    h->scratch = scratch_address(dev);
    *handle = reinterpret_cast<muscle_hip_handle_t>(h.release());
  });
}

int muscle_hip_destroy(muscle_hip_handle_t handle) {
  return guarded([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    if (!h) return;
    (void)hipStreamSynchronize(h->stream);
    delete h;
  });
}

int muscle_hip_set_stream(muscle_hip_handle_t handle, void* stream) {
  return guarded([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    if (!h) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    const hipStream_t ns = reinterpret_cast<hipStream_t>(stream);
    if (ns == h->stream) return;
    // Every call of a handle shares its workspace (casts, packed operands,
    // split-K partials, sliced planes) and its cached plans' tables, which
    // the first call uploads asynchronously.  Work enqueued on the new stream
    // therefore starts after all the handle has enqueued on the old one.
    if (!h->order_event) MH_HIP_CHECK(hipEventCreateWithFlags(&h->order_event, hipEventDisableTiming));
    MH_HIP_CHECK(hipEventRecord(h->order_event, h->stream));
    MH_HIP_CHECK(hipStreamWaitEvent(ns, h->order_event, 0));
    h->stream = ns;
  });
}

int muscle_hip_set_option(muscle_hip_handle_t handle, int option, int64_t value) {
  return guarded([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    if (!h) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    switch (option) {
      case MUSCLE_HIP_OPTION_COMPLEX_MULT:
        if (value != 3 && value != 4) fail(MUSCLE_HIP_INVALID_ARGUMENT, "complex multiplication must be 3 or 4");
        h->complex_mult = (int)value;
        break;
      default: fail(MUSCLE_HIP_INVALID_ARGUMENT, "unknown option");
    }
  });
}

int muscle_hip_measure_bandwidth(muscle_hip_handle_t handle, int64_t bytes, double* read_tbs, double* write_tbs) {
  return guarded([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    if (!h || !read_tbs || !write_tbs) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null handle or result pointer");
    if (bytes < (int64_t(1) << 20)) fail(MUSCLE_HIP_INVALID_ARGUMENT, "probe buffer below 1 MiB");
    bytes &= ~int64_t(15);
    std::lock_guard<std::mutex> lk(h->mu);
    void* buf = nullptr;
    if (hipMalloc(&buf, (size_t)bytes + 256) != hipSuccess) {
      (void)hipGetLastError();
      fail(MUSCLE_HIP_OUT_OF_MEMORY, "probe buffer allocation failed");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double rate[2] = {0, 0};
    hipError_t err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    for (int w = 1; w >= 0 && err == hipSuccess; --w) {  // write first: the read pass reads written memory
      uint32_t* out = reinterpret_cast<uint32_t*>((char*)buf + bytes);
      launch_probe(w == 1, buf, bytes, out, h->stream);
      err = hipEventRecord(e0, h->stream);
      for (int r = 0; r < 5 && err == hipSuccess; ++r) launch_probe(w == 1, buf, bytes, out, h->stream);
      if (err == hipSuccess) err = hipEventRecord(e1, h->stream);
      if (err == hipSuccess) err = hipEventSynchronize(e1);
      float ms = 0;
      if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
      if (err == hipSuccess) rate[w] = 5.0 * (double)bytes / (ms * 1e-3) / 1e12;
    }
    if (err == hipSuccess) err = hipGetLastError();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    if (err != hipSuccess) fail(MUSCLE_HIP_HIP_ERROR, std::string("bandwidth probe: ") + hipGetErrorString(err));
    *read_tbs = rate[0];
    *write_tbs = rate[1];
  });
}

int muscle_hip_synchronize(muscle_hip_handle_t handle) {
  return guarded([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    if (!h) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null handle");
    MH_HIP_CHECK(hipStreamSynchronize(h->stream));
  });
}

int64_t muscle_hip_plan_cache_size(muscle_hip_handle_t handle) {
  Handle* h = reinterpret_cast<Handle*>(handle);
  if (!h) return -1;
  std::lock_guard<std::mutex> lk(h->mu);
  return (int64_t)h->cache.size();
}

int muscle_hip_plan_cache_clear(muscle_hip_handle_t handle) {
  return guarded([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    if (!h) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    MH_HIP_CHECK(hipDeviceSynchronize());
    h->cache.clear();
    h->lru.clear();
  });
}

int muscle_hip_binary_einsum(muscle_hip_handle_t handle, const muscle_hip_tensor_t* c, const muscle_hip_tensor_t* a,
                             const muscle_hip_tensor_t* b, int precision, double target_rel_error,
                             muscle_hip_info_t* info) {
  return guarded([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    if (!h) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null handle");
    check_tensor(c, "C");
    check_tensor(a, "A");
    check_tensor(b, "B");
    if (precision < MUSCLE_HIP_PRECISION_NATIVE || precision > MUSCLE_HIP_PRECISION_OZAKI_INT8)
      fail(MUSCLE_HIP_INVALID_ARGUMENT, "unknown precision mode");
    if (!(target_rel_error > 0)) target_rel_error = 1e-12;
    std::lock_guard<std::mutex> lk(h->mu);
    Plan* p = get_plan(*h, 1, precision, target_rel_error, c, a, b);
    run_contraction(*h, *p, *c, *a, b);
    fill_info(info, *p, c, a, b);
  });
}

int muscle_hip_binary_einsum_noalloc(muscle_hip_handle_t handle, const muscle_hip_tensor_t* c,
                                     const muscle_hip_tensor_t* a, const muscle_hip_tensor_t* b,
                                     muscle_hip_info_t* info) {
  return guarded([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    if (!h) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null handle");
    check_tensor(c, "C");
    check_tensor(a, "A");
    check_tensor(b, "B");
    std::lock_guard<std::mutex> lk(h->mu);
    // host planning only: the plan's tables go to the scratch, not to a
    // cached device allocation, and operands are gathered where they lie
    // (no packed copies)
    std::unique_ptr<Plan> p = build_contraction_plan(lower_binary(*c, *a, *b), MUSCLE_HIP_PRECISION_NATIVE);
    if (p->kind != Plan::EMPTY && p->kind != Plan::GEMM && p->kind != Plan::REDUCE)
      fail(MUSCLE_HIP_UNSUPPORTED, "binary_einsum_noalloc: unsupported contraction plan");
    const size_t tab = p->kind == Plan::EMPTY ? 0 : align_up(p->htab.size() * sizeof(int64_t));
    const size_t need = tab + contraction_workspace_bytes(*p, *a, b);
    if (need > kScratchBytes)
      fail(MUSCLE_HIP_INVALID_ARGUMENT, "binary_einsum_noalloc: this contraction needs " + std::to_string(need) +
                                            " bytes of scratch (offset tables, partial sums, casts); the maximum is " +
                                            std::to_string(kScratchBytes));
    if (p->kind != Plan::EMPTY) {
      // stream-ordered after the previous call's kernels, which may still read the old tables
      MH_HIP_CHECK(hipMemcpyAsync(h->scratch, p->htab.data(), p->htab.size() * sizeof(int64_t),
                                  hipMemcpyHostToDevice, h->stream));
      p->dtab = reinterpret_cast<int64_t*>(h->scratch);
    }
    try {
      run_contraction(*h, *p, *c, *a, b, h->scratch + tab, kScratchBytes - tab);
    } catch (...) {
      p->dtab = nullptr;  // borrowed: the plan must not free the scratch
      throw;
    }
    p->dtab = nullptr;
    fill_info(info, *p, c, a, b);
  });
}

int muscle_hip_unary_einsum(muscle_hip_handle_t handle, const muscle_hip_tensor_t* y, const muscle_hip_tensor_t* x,
                            muscle_hip_info_t* info) {
  return guarded([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    if (!h) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null handle");
    check_tensor(y, "Y");
    check_tensor(x, "X");
    std::lock_guard<std::mutex> lk(h->mu);
    Plan* p = get_plan(*h, 2, MUSCLE_HIP_PRECISION_NATIVE, 0.0, y, x, nullptr);
    run_contraction(*h, *p, *y, *x, nullptr);
    fill_info(info, *p, y, x, nullptr);
  });
}

int muscle_hip_svd_thin(muscle_hip_handle_t handle, const muscle_hip_tensor_t* u, const muscle_hip_tensor_t* s,
                        const muscle_hip_tensor_t* v, const muscle_hip_tensor_t* a) {
  return guarded([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    if (!h) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null handle");
    if (!u || !s || !v || !a) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null tensor descriptor");
    std::lock_guard<std::mutex> lk(h->mu);
    svd_thin(*h, *u, *s, *v, *a);
  });
}

int muscle_hip_qr_thin(muscle_hip_handle_t handle, const muscle_hip_tensor_t* q, const muscle_hip_tensor_t* r,
                       const muscle_hip_tensor_t* a) {
  return guarded([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    if (!h) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null handle");
    if (!q || !r || !a) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null tensor descriptor");
    std::lock_guard<std::mutex> lk(h->mu);
    qr_thin(*h, *q, *r, *a);
  });
}

int muscle_hip_hadamard(muscle_hip_handle_t handle, const muscle_hip_tensor_t* c, const muscle_hip_tensor_t* a,
                        const muscle_hip_tensor_t* b, muscle_hip_info_t* info) {
  return guarded([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    if (!h) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null handle");
    check_tensor(c, "c");
    check_tensor(a, "a");
    check_tensor(b, "b");
    ElementwiseDesc e = lower_hadamard(*c, *a, *b);
    if ((int)e.dims.size() > kMaxEwDims)
      fail(MUSCLE_HIP_UNSUPPORTED, "hadamard: more than 24 non-fusable modes");
    if (info) {
      std::memset(info, 0, sizeof *info);
      info->kernel = 4;
      info->m = e.total;
      info->n = info->k = info->batch = 1;
      info->flops = (is_complex(a->dtype) ? (is_complex(b->dtype) ? 6.0 : 2.0) : 1.0) * (double)e.total;
      info->bytes = (double)e.total * 2 * dtype_size(a->dtype) + (double)product_distinct(*b) * dtype_size(b->dtype);
    }
    if (e.total == 0) return;
    EwArgs w;
    std::memset(&w, 0, sizeof w);
    w.C = c->data;
    w.A = a->data;
    w.B = b->data;
    w.total = e.total;
    w.R = (int32_t)e.dims.size();
    w.small_index = e.total < (int64_t(1) << 32) ? 1 : 0;
    for (int i = 0; i < w.R; ++i) {
      w.ext[i] = e.dims[i].extent;
      w.sA[i] = e.dims[i].s[0];
      w.sB[i] = e.dims[i].s[1];
      w.sC[i] = e.dims[i].s[2];
      w.div[i] = FastDiv(w.small_index ? (uint32_t)e.dims[i].extent : 1u);
    }
    if (w.R == 0) {  // every extent is 1: a single element
      w.R = 1;
      w.ext[0] = 1;
    }
    plan_hadamard_tiles(e, w);
    std::lock_guard<std::mutex> lk(h->mu);
    launch_hadamard(a->dtype, b->dtype, w, h->stream);
    check_launch();
  });
}

int muscle_hip_describe(const muscle_hip_tensor_t* c, const muscle_hip_tensor_t* a, const muscle_hip_tensor_t* b,
                        int precision, double target_rel_error, muscle_hip_info_t* info, int32_t* modes_out,
                        int32_t* class_out, int32_t* nmodes_out) {
  return guarded([&] {
    check_tensor(c, "C", false);
    check_tensor(a, "A", false);
    check_tensor(b, "B", false);
    if (precision < MUSCLE_HIP_PRECISION_NATIVE || precision > MUSCLE_HIP_PRECISION_OZAKI_INT8)
      fail(MUSCLE_HIP_INVALID_ARGUMENT, "unknown precision mode");
    ContractionDesc d = lower_binary(*c, *a, *b);
    std::unique_ptr<Plan> p = build_contraction_plan(d, precision);
    if (!(target_rel_error > 0)) target_rel_error = 1e-12;
    if (p->kind == Plan::SPLIT) prepare_split(*p, target_rel_error);
    fill_info(info, *p, c, a, b);
    if (nmodes_out) *nmodes_out = (int32_t)d.modes.size();
    for (size_t i = 0; i < d.modes.size(); ++i) {
      if (modes_out) modes_out[i] = d.modes[i].label;
      if (class_out) class_out[i] = d.modes[i].cls;
    }
  });
}

}  // extern "C"
