// Index classification and contraction lowering.
//
// Semantics followed (reference files):
//   - binary contraction: modes shared by A, B and C are batch (hyper) modes,
//     modes shared by A and B only are contracted, open modes come from their
//     own operand (ext/MuscleReactantExt.jl:94-122, ext/MuscleCUDAExt.jl:24-38);
//     with no batch modes this is exactly BackendBase's left/contract/right
//     split (src/Operations/binary_einsum.jl:77-79).
//   - unary: OMEinsum x -> y, repeated modes of x tied, modes absent from y
//     summed (ext/MuscleOMEinsumExt.jl:31-37, src/Operations/unary_einsum.jl:26-32).
//   - hadamard: C = A .* broadcast(B), modes(B) subset of modes(A)
//     (src/Operations/hadamard.jl:7-13,59-74).
#include "lowering.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace mh {

void check_tensor(const muscle_hip_tensor_t* t, const char* name, bool need_data) {
  if (!t) fail(MUSCLE_HIP_INVALID_ARGUMENT, std::string(name) + ": null tensor descriptor");
  if (!valid_dtype(t->dtype)) fail(MUSCLE_HIP_INVALID_ARGUMENT, std::string(name) + ": unknown dtype");
  if (t->rank < 0 || t->rank > 64)
    fail(MUSCLE_HIP_INVALID_ARGUMENT, std::string(name) + ": rank must be in [0, 64]");
  if (t->rank > 0 && (!t->modes || !t->extents || !t->strides))
    fail(MUSCLE_HIP_INVALID_ARGUMENT, std::string(name) + ": null modes/extents/strides");
  for (int i = 0; i < t->rank; ++i)
    if (t->extents[i] < 0) fail(MUSCLE_HIP_DIM_MISMATCH, std::string(name) + ": negative extent");
  if (need_data && !t->data) {
    int64_t n = 1;
    for (int i = 0; i < t->rank; ++i) n *= t->extents[i];
    if (n > 0) fail(MUSCLE_HIP_INVALID_ARGUMENT, std::string(name) + ": null data pointer");
  }
}

namespace {

const int64_t kInf = std::numeric_limits<int64_t>::max();

int64_t absl(int64_t v) { return v < 0 ? -v : v; }

// Registers every occurrence of a label; repeated occurrences inside one
// operand add their strides (the diagonal of that operand).
int add_occurrence(std::vector<ModeInfo>& modes, int op, int32_t label, int64_t extent, int64_t stride,
                   const char* opname) {
  for (size_t i = 0; i < modes.size(); ++i) {
    if (modes[i].label == label) {
      if (modes[i].extent != extent)
        fail(MUSCLE_HIP_DIM_MISMATCH, std::string(opname) + ": mode " + std::to_string(label) +
                                          " has extent " + std::to_string(extent) + ", elsewhere " +
                                          std::to_string(modes[i].extent));
      modes[i].s[op] += stride;
      modes[i].in[op] = true;
      return (int)i;
    }
  }
  ModeInfo m;
  m.label = label;
  m.extent = extent;
  m.s[0] = m.s[1] = m.s[2] = 0;
  m.in[0] = m.in[1] = m.in[2] = false;
  m.s[op] = stride;
  m.in[op] = true;
  m.cls = -1;
  modes.push_back(m);
  return (int)modes.size() - 1;
}

void require_unique(const muscle_hip_tensor_t& t, const char* what) {
  for (int i = 0; i < t.rank; ++i)
    for (int j = i + 1; j < t.rank; ++j)
      if (t.modes[i] == t.modes[j])
        fail(MUSCLE_HIP_INVALID_ARGUMENT,
             std::string(what) + ": mode " + std::to_string(t.modes[i]) + " repeated");
}

int64_t min_stride(const std::vector<Dim>& d, int which) {
  int64_t best = kInf;
  for (auto& x : d)
    if (x.s[which] != 0) best = std::min(best, absl(x.s[which]));
  return best;
}

// Sorts dims by |stride[key]| (tie-break |stride[key2]|) and merges neighbours
// p, q (p faster) when every operand in `ops` has s[q] == s[p] * extent(p),
// i.e. the pair walks memory like a single mode of extent e_p * e_q.
void sort_and_fuse(std::vector<Dim>& d, int key, int key2, std::initializer_list<int> ops) {
  std::stable_sort(d.begin(), d.end(), [&](const Dim& x, const Dim& y) {
    int64_t kx = absl(x.s[key]), ky = absl(y.s[key]);
    if (kx != ky) return kx < ky;
    return absl(x.s[key2]) < absl(y.s[key2]);
  });
  std::vector<Dim> out;
  for (auto& x : d) {
    if (!out.empty()) {
      Dim& p = out.back();
      bool ok = true;
      for (int o : ops)
        if (x.s[o] != p.s[o] * p.extent) ok = false;
      if (ok) {
        p.extent *= x.extent;
        continue;
      }
    }
    out.push_back(x);
  }
  d.swap(out);
}

int64_t product(const std::vector<Dim>& d) {
  int64_t p = 1;
  for (auto& x : d) p *= x.extent;
  return p;
}

}  // namespace

void split_dims(const std::vector<Dim>& d, std::vector<Dim>& lo, std::vector<Dim>& hi, int64_t& P,
                int64_t single_max) {
  lo.clear();
  hi.clear();
  int64_t total = product(d);
  P = 1;
  if (total <= 1) {
    hi = d;
    return;
  }
  if (total <= single_max) {  // small class: one level, P = total
    lo = d;
    P = total;
    return;
  }
  const double target = std::sqrt((double)total);
  size_t i = 0;
  for (; i < d.size(); ++i) {
    const Dim& x = d[i];
    if ((double)P * (double)x.extent <= target) {
      lo.push_back(x);
      P *= x.extent;
      continue;
    }
    // split x = (f, s) x (e / f, s * f) with P * f as close to target as possible
    int64_t best = 1;
    const int64_t lim = (int64_t)(target / (double)P);
    for (int64_t f = std::min<int64_t>(lim, x.extent - 1); f >= 2; --f)
      if (x.extent % f == 0) {
        best = f;
        break;
      }
    if (best > 1) {
      Dim a = x, b = x;
      a.extent = best;
      b.extent = x.extent / best;
      for (int w = 0; w < 3; ++w) b.s[w] = x.s[w] * best;
      lo.push_back(a);
      hi.push_back(b);
      P *= best;
      ++i;
    }
    break;
  }
  for (; i < d.size(); ++i) hi.push_back(d[i]);
}

std::vector<int64_t> offset_table(const std::vector<Dim>& dims, int which) {
  int64_t n = product(dims);
  std::vector<int64_t> t((size_t)std::max<int64_t>(n, 0));
  if (n <= 0) return t;
  std::vector<int64_t> idx(dims.size(), 0);
  int64_t off = 0;
  for (int64_t i = 0; i < n; ++i) {
    t[(size_t)i] = off;
    // mixed-radix increment, fastest dim first
    for (size_t d = 0; d < dims.size(); ++d) {
      off += dims[d].s[which];
      if (++idx[d] < dims[d].extent) break;
      off -= dims[d].s[which] * dims[d].extent;
      idx[d] = 0;
    }
  }
  return t;
}

ContractionDesc lower_binary(const muscle_hip_tensor_t& c, const muscle_hip_tensor_t& a,
                             const muscle_hip_tensor_t& b) {
  ContractionDesc d;
  // Output modes are unique (MuscleReactantExt.jl:92: "indices in `out` for a
  // binary einsum must be unique").
  require_unique(c, "binary_einsum output");
  const muscle_hip_tensor_t* ops[3] = {&a, &b, &c};
  const char* names[3] = {"A", "B", "C"};
  // first-appearance order (c, a, b) for the describe output
  for (int pass : {2, 0, 1}) {
    const muscle_hip_tensor_t& t = *ops[pass];
    for (int i = 0; i < t.rank; ++i)
      add_occurrence(d.modes, pass, t.modes[i], t.extents[i], t.strides[i], names[pass]);
  }
  for (auto& m : d.modes) {
    if (m.in[2]) {
      if (m.in[0] && m.in[1]) m.cls = CLS_BATCH;
      else if (m.in[0]) m.cls = CLS_M;
      else if (m.in[1]) m.cls = CLS_N;
      else
        fail(MUSCLE_HIP_INVALID_ARGUMENT, "binary_einsum: output mode " + std::to_string(m.label) +
                                              " is in neither input");
    } else {
      if (m.in[0] && m.in[1]) m.cls = CLS_K;
      else if (m.in[0]) m.cls = CLS_SUM_A;
      else m.cls = CLS_SUM_B;
    }
  }
  d.dt[0] = a.dtype;
  d.dt[1] = b.dtype;
  d.dt[2] = c.dtype;
  if (c.dtype != promote_dtype(a.dtype, b.dtype))
    fail(MUSCLE_HIP_INVALID_ARGUMENT, std::string("binary_einsum: output eltype ") + dtype_name(c.dtype) +
                                          " is not promote_type(" + dtype_name(a.dtype) + ", " +
                                          dtype_name(b.dtype) + ")");
  for (auto& m : d.modes) {
    if (m.in[2] && m.extent == 0) d.empty = true;
    if (m.extent == 1) continue;  // contributes offset 0 only
    Dim x;
    x.extent = m.extent;
    x.s[0] = m.s[0];
    x.s[1] = m.s[1];
    x.s[2] = m.s[2];
    switch (m.cls) {
      case CLS_BATCH: d.batch.push_back(x); break;
      case CLS_M: x.s[1] = 0; d.m.push_back(x); break;
      case CLS_N: x.s[0] = 0; d.n.push_back(x); break;
      default: x.s[2] = 0; d.k.push_back(x); break;
    }
  }
  // Make C's fastest-varying mode an N mode: the tile kernels give
  // consecutive lanes consecutive n, so C stores coalesce.
  int64_t sm = min_stride(d.m, 2), sn = min_stride(d.n, 2);
  if (sm < sn) {
    d.swapped = true;
    std::swap(d.m, d.n);
    std::swap(d.dt[0], d.dt[1]);
    for (auto* lst : {&d.m, &d.n, &d.k, &d.batch})
      for (auto& x : *lst) std::swap(x.s[0], x.s[1]);
  }
  sort_and_fuse(d.m, 2, 0, {0, 2});
  sort_and_fuse(d.n, 2, 1, {1, 2});
  // Order K the way the operand whose fastest mode is contracted walks memory.
  int64_t b_k = min_stride(d.k, 1), b_n = min_stride(d.n, 1);
  bool b_kmajor = b_k < b_n;
  int kkey = b_kmajor ? 1 : 0;
  sort_and_fuse(d.k, kkey, 1 - kkey, {0, 1});
  sort_and_fuse(d.batch, 2, 0, {0, 1, 2});
  d.M = product(d.m);
  d.N = product(d.n);
  d.K = product(d.k);
  d.B = product(d.batch);
  if (d.M == 0 || d.N == 0 || d.B == 0) d.empty = true;
  return d;
}

ContractionDesc lower_unary(const muscle_hip_tensor_t& y, const muscle_hip_tensor_t& x) {
  ContractionDesc d;
  d.has_b = false;
  require_unique(y, "unary_einsum output");
  for (int i = 0; i < y.rank; ++i) {
    bool found = false;
    for (int j = 0; j < x.rank; ++j) found = found || (x.modes[j] == y.modes[i]);
    if (!found)
      fail(MUSCLE_HIP_INVALID_ARGUMENT, "unary_einsum: output indices must be a subset of input indices");
  }
  if (y.dtype != x.dtype)
    fail(MUSCLE_HIP_INVALID_ARGUMENT, "unary_einsum: output eltype must equal input eltype");
  for (int i = 0; i < y.rank; ++i) add_occurrence(d.modes, 2, y.modes[i], y.extents[i], y.strides[i], "Y");
  for (int i = 0; i < x.rank; ++i) add_occurrence(d.modes, 0, x.modes[i], x.extents[i], x.strides[i], "X");
  d.dt[0] = x.dtype;
  d.dt[1] = x.dtype;
  d.dt[2] = y.dtype;
  for (auto& m : d.modes) {
    m.cls = m.in[2] ? CLS_M : CLS_SUM_A;
    if (m.in[2] && m.extent == 0) d.empty = true;
    if (m.extent == 1) continue;
    Dim dm;
    dm.extent = m.extent;
    dm.s[0] = m.s[0];
    dm.s[1] = 0;
    dm.s[2] = m.in[2] ? m.s[2] : 0;
    (m.in[2] ? d.m : d.k).push_back(dm);
  }
  sort_and_fuse(d.m, 2, 0, {0, 2});
  sort_and_fuse(d.k, 0, 0, {0});
  d.M = product(d.m);
  d.K = product(d.k);
  if (d.M == 0) d.empty = true;
  return d;
}

ElementwiseDesc lower_hadamard(const muscle_hip_tensor_t& c, const muscle_hip_tensor_t& a,
                               const muscle_hip_tensor_t& b) {
  ElementwiseDesc e;
  require_unique(a, "hadamard: a");
  require_unique(b, "hadamard: b");
  require_unique(c, "hadamard: c");
  if (c.rank != a.rank) fail(MUSCLE_HIP_INVALID_ARGUMENT, "hadamard: inds(c) must equal inds(a)");
  if (b.rank > a.rank) fail(MUSCLE_HIP_INVALID_ARGUMENT, "hadamard: inds(b) must be a subset of inds(a)");
  if (c.dtype != a.dtype) fail(MUSCLE_HIP_INVALID_ARGUMENT, "hadamard: c must have the eltype of a");
  if (!is_complex(a.dtype) && is_complex(b.dtype))
    fail(MUSCLE_HIP_UNSUPPORTED, "hadamard: complex b into real a (InexactError)");
  std::vector<ModeInfo> modes;
  for (int i = 0; i < a.rank; ++i) add_occurrence(modes, 0, a.modes[i], a.extents[i], a.strides[i], "a");
  for (int i = 0; i < c.rank; ++i) {
    size_t before = modes.size();
    add_occurrence(modes, 2, c.modes[i], c.extents[i], c.strides[i], "c");
    if (modes.size() != before) fail(MUSCLE_HIP_INVALID_ARGUMENT, "hadamard: inds(c) must equal inds(a)");
  }
  for (int i = 0; i < b.rank; ++i) {
    size_t before = modes.size();
    add_occurrence(modes, 1, b.modes[i], b.extents[i], b.strides[i], "b");
    if (modes.size() != before)
      fail(MUSCLE_HIP_INVALID_ARGUMENT, "hadamard: inds(b) must be a subset of inds(a)");
  }
  e.dt[0] = a.dtype;
  e.dt[1] = b.dtype;
  e.dt[2] = c.dtype;
  for (auto& m : modes) {
    e.total *= m.extent;
    if (m.extent == 1) continue;
    Dim x;
    x.extent = m.extent;
    x.s[0] = m.s[0];
    x.s[1] = m.s[1];
    x.s[2] = m.s[2];
    e.dims.push_back(x);
  }
  sort_and_fuse(e.dims, 2, 0, {0, 1, 2});
  return e;
}

std::string plan_key(int kind, int precision, const muscle_hip_tensor_t* t0, const muscle_hip_tensor_t* t1,
                     const muscle_hip_tensor_t* t2) {
  std::string k;
  auto put = [&](const void* p, size_t n) { k.append((const char*)p, n); };
  put(&kind, sizeof kind);
  put(&precision, sizeof precision);
  for (const muscle_hip_tensor_t* t : {t0, t1, t2}) {
    int32_t r = t ? t->rank : -1;
    put(&r, sizeof r);
    if (!t) continue;
    put(&t->dtype, sizeof t->dtype);
    for (int i = 0; i < t->rank; ++i) {
      put(&t->modes[i], sizeof(int32_t));
      put(&t->extents[i], sizeof(int64_t));
      put(&t->strides[i], sizeof(int64_t));
    }
  }
  return k;
}

}  // namespace mh
