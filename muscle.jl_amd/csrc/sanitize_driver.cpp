// Host-side checks for the address/undefined-behaviour sanitizer build of the
// library's host code (lowering, planning, batch and server).  The driver is a
// plain executable, so the sanitizer runtime comes with it and nothing has to
// be preloaded into another process.
//
//   sanitize_driver          lowering and argument checking through
//                            muscle_hip_describe (no GPU needed)
//   sanitize_driver --gpu    additionally: random contractions, a batch in all
//                            three run modes and a server with contraction-tree
//                            jobs, each against a naive host loop over every
//                            label assignment (einsum semantics: the output sums
//                            over the labels it does not carry)
//
// Exit status 0 when every check passed; the sanitizers abort on their own.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "muscle_hip.h"

namespace {

int g_failures = 0;

void expect(bool ok, const char* what, int64_t detail = 0) {
  if (ok) return;
  ++g_failures;
  std::fprintf(stderr, "FAIL: %s (%lld): %s\n", what, (long long)detail, muscle_hip_last_error());
}

// A tensor operand on the host: labels, extents and column-major strides of
// the listed order, possibly with gaps between elements.
struct Operand {
  std::vector<int32_t> modes;
  std::vector<int64_t> ext, strides;
  int64_t span = 1;  // elements from the first to one past the last
  muscle_hip_tensor_t desc(int dtype, void* data) const {
    muscle_hip_tensor_t t;
    t.dtype = dtype;
    t.rank = (int32_t)modes.size();
    t.data = data;
    t.modes = modes.data();
    t.extents = ext.data();
    t.strides = strides.data();
    return t;
  }
};

Operand make_operand(const std::vector<int32_t>& labels, const std::vector<int64_t>& extent, std::mt19937_64& rng) {
  Operand o;
  o.modes = labels;
  std::shuffle(o.modes.begin(), o.modes.end(), rng);
  int64_t s = 1 + (int64_t)(rng() % 2);  // a gap of one element now and then
  for (int32_t m : o.modes) {
    o.ext.push_back(extent[(size_t)m]);
    o.strides.push_back(s);
    s *= std::max<int64_t>(1, extent[(size_t)m]);
  }
  o.span = 1;
  for (size_t i = 0; i < o.modes.size(); ++i) o.span += (o.ext[i] - 1) * o.strides[i];
  if (o.span < 1) o.span = 1;
  return o;
}

// Roles of the labels of one contraction.
enum Role { BATCH, CONTRACT, LEFT, RIGHT, SUM_A, SUM_B, NROLES };

struct Case {
  std::vector<int64_t> extent;  // by label (label 0 unused)
  std::vector<int> role;
  Operand a, b, c;
};

Case random_case(std::mt19937_64& rng, int max_labels, int64_t max_ext, int64_t max_points) {
  Case cs;
  for (;;) {
    const int L = 1 + (int)(rng() % (uint64_t)max_labels);
    cs.extent.assign((size_t)L + 1, 1);
    cs.role.assign((size_t)L + 1, 0);
    std::vector<int32_t> la, lb, lc;
    int64_t points = 1;
    for (int l = 1; l <= L; ++l) {
      cs.extent[(size_t)l] = 1 + (int64_t)(rng() % (uint64_t)max_ext);
      const int r = (int)(rng() % 16);  // contracted and open labels are the common ones
      const int role = r < 4 ? CONTRACT : r < 8 ? LEFT : r < 12 ? RIGHT : r < 14 ? BATCH : r < 15 ? SUM_A : SUM_B;
      cs.role[(size_t)l] = role;
      points *= cs.extent[(size_t)l];
      if (role == BATCH || role == CONTRACT || role == LEFT || role == SUM_A) la.push_back(l);
      if (role == BATCH || role == CONTRACT || role == RIGHT || role == SUM_B) lb.push_back(l);
      if (role == BATCH || role == LEFT || role == RIGHT) lc.push_back(l);
    }
    if (points > max_points) continue;
    cs.a = make_operand(la, cs.extent, rng);
    cs.b = make_operand(lb, cs.extent, rng);
    cs.c = make_operand(lc, cs.extent, rng);
    // C must be dense (no gaps) so every element is written
    int64_t s = 1;
    for (size_t i = 0; i < cs.c.modes.size(); ++i) {
      cs.c.strides[i] = s;
      s *= std::max<int64_t>(1, cs.c.ext[i]);
    }
    cs.c.span = s;
    return cs;
  }
}

// ------------------------------------------------------------ no-GPU checks
void check_describe(std::mt19937_64& rng) {
  int checked = 0;
  for (int it = 0; it < 3000; ++it) {
    Case cs = random_case(rng, 9, 5, 1 << 22);
    const int precision = (int)(rng() % 5);
    const int dt = (int)(rng() % 4);  // MUSCLE_HIP_F32..C64
    muscle_hip_tensor_t c = cs.c.desc(dt, nullptr), a = cs.a.desc(dt, nullptr), b = cs.b.desc(dt, nullptr);
    muscle_hip_info_t info;
    int32_t modes[32], classes[32], nm = 0;
    const int st = muscle_hip_describe(&c, &a, &b, precision, 1e-10, &info, modes, classes, &nm);
    expect(st == MUSCLE_HIP_OK, "describe of a valid contraction", it);
    if (st != MUSCLE_HIP_OK) continue;
    // the fused extents multiply to the products of the label classes, times
    // those single-operand sums the plan folds into its K (the other operand
    // broadcast along them) instead of reducing them first
    int64_t pm = 1, pn = 1, pk = 1, pb = 1, ps = 1;
    for (size_t l = 1; l < cs.extent.size(); ++l) {
      const int r = cs.role[l];
      if (r == LEFT) pm *= cs.extent[l];
      if (r == RIGHT) pn *= cs.extent[l];
      if (r == CONTRACT) pk *= cs.extent[l];
      if (r == BATCH) pb *= cs.extent[l];
      if (r == SUM_A || r == SUM_B) ps *= cs.extent[l];
    }
    const int64_t base = pm * pn * pk * pb, fused = info.m * info.n * info.k * info.batch;
    expect(info.kernel == 0 || (fused % base == 0 && (base * ps) % fused == 0), "fused extents", it);
    // every label is classified as its role says (labels of extent 1 may merge away)
    for (int i = 0; i < nm; ++i) {
      const int l = modes[i];
      if (l <= 0 || l >= (int)cs.extent.size() || cs.extent[(size_t)l] == 1) continue;
      const int r = cs.role[(size_t)l];
      const int want_class = r == BATCH ? 0 : r == LEFT ? 1 : r == RIGHT ? 2 : r == CONTRACT ? 3 : r == SUM_A ? 4 : 5;
      expect(classes[i] == want_class, "label class", it * 100 + l);
    }
    ++checked;
  }
  // invalid inputs end in an error status, never in a crash
  Case cs = random_case(rng, 4, 3, 1 << 10);
  muscle_hip_tensor_t c = cs.c.desc(MUSCLE_HIP_C64, nullptr), a = cs.a.desc(MUSCLE_HIP_C64, nullptr),
                      b = cs.b.desc(MUSCLE_HIP_C64, nullptr);
  muscle_hip_info_t info;
  expect(muscle_hip_describe(nullptr, &a, &b, 0, 1e-10, &info, nullptr, nullptr, nullptr) != MUSCLE_HIP_OK,
         "null output descriptor");
  if (a.rank > 0) {
    std::vector<int64_t> bad(cs.a.ext);
    bad[0] += 1;
    muscle_hip_tensor_t a2 = a;
    a2.extents = bad.data();
    const bool shared = std::find(cs.b.modes.begin(), cs.b.modes.end(), cs.a.modes[0]) != cs.b.modes.end() ||
                        std::find(cs.c.modes.begin(), cs.c.modes.end(), cs.a.modes[0]) != cs.c.modes.end();
    if (shared)
      expect(muscle_hip_describe(&c, &a2, &b, 0, 1e-10, &info, nullptr, nullptr, nullptr) ==
                 MUSCLE_HIP_DIM_MISMATCH,
             "mismatched extents");
    std::vector<int64_t> neg(cs.a.ext);
    neg[0] = -3;
    a2.extents = neg.data();
    expect(muscle_hip_describe(&c, &a2, &b, 0, 1e-10, &info, nullptr, nullptr, nullptr) != MUSCLE_HIP_OK,
           "negative extent");
  }
  std::vector<int32_t> orphan(cs.c.modes);
  std::vector<int64_t> oext(cs.c.ext), ostr(cs.c.strides);
  orphan.push_back(999);
  oext.push_back(2);
  ostr.push_back(cs.c.span);
  muscle_hip_tensor_t c2 = c;
  c2.rank = (int32_t)orphan.size();
  c2.modes = orphan.data();
  c2.extents = oext.data();
  c2.strides = ostr.data();
  expect(muscle_hip_describe(&c2, &a, &b, 0, 1e-10, &info, nullptr, nullptr, nullptr) != MUSCLE_HIP_OK,
         "output label in neither operand");
  muscle_hip_tensor_t a3 = a;
  a3.dtype = 77;
  expect(muscle_hip_describe(&c, &a3, &b, 0, 1e-10, &info, nullptr, nullptr, nullptr) != MUSCLE_HIP_OK,
         "unknown dtype");
  expect(muscle_hip_describe(&c, &a, &b, 99, 1e-10, &info, nullptr, nullptr, nullptr) != MUSCLE_HIP_OK,
         "unknown precision");
  for (int s = 0; s < 9; ++s) expect(muscle_hip_status_string(s) != nullptr, "status string", s);
  for (int64_t m = 1; m < (int64_t(1) << 20); m *= 7) expect(muscle_hip_tuning_class(m, m + 3) >= 0, "size class", m);
  std::printf("describe: %d random contractions checked\n", checked);
}

// ------------------------------------------------------------ GPU checks
typedef std::complex<double> cd;

// Host values of an operand (as complex double) at its strided positions.
std::vector<cd> random_values(const Operand& o, std::mt19937_64& rng) {
  std::normal_distribution<double> nd;
  std::vector<cd> v((size_t)o.span);
  for (auto& x : v) x = cd(nd(rng), nd(rng));
  return v;
}

size_t esize(int dt) { return dt == MUSCLE_HIP_F32 ? 4 : dt == MUSCLE_HIP_F64 ? 8 : dt == MUSCLE_HIP_C32 ? 8 : 16; }
bool is_cplx(int dt) { return dt == MUSCLE_HIP_C32 || dt == MUSCLE_HIP_C64; }

std::vector<char> encode(const std::vector<cd>& v, int dt) {
  std::vector<char> out(v.size() * esize(dt));
  for (size_t i = 0; i < v.size(); ++i) {
    switch (dt) {
      case MUSCLE_HIP_F32: reinterpret_cast<float*>(out.data())[i] = (float)v[i].real(); break;
      case MUSCLE_HIP_F64: reinterpret_cast<double*>(out.data())[i] = v[i].real(); break;
      case MUSCLE_HIP_C32:
        reinterpret_cast<float*>(out.data())[2 * i] = (float)v[i].real();
        reinterpret_cast<float*>(out.data())[2 * i + 1] = (float)v[i].imag();
        break;
      default:
        reinterpret_cast<double*>(out.data())[2 * i] = v[i].real();
        reinterpret_cast<double*>(out.data())[2 * i + 1] = v[i].imag();
    }
  }
  return out;
}

cd decode(const std::vector<char>& buf, int dt, size_t i) {
  switch (dt) {
    case MUSCLE_HIP_F32: return cd(reinterpret_cast<const float*>(buf.data())[i], 0.0);
    case MUSCLE_HIP_F64: return cd(reinterpret_cast<const double*>(buf.data())[i], 0.0);
    case MUSCLE_HIP_C32:
      return cd(reinterpret_cast<const float*>(buf.data())[2 * i], reinterpret_cast<const float*>(buf.data())[2 * i + 1]);
    default:
      return cd(reinterpret_cast<const double*>(buf.data())[2 * i], reinterpret_cast<const double*>(buf.data())[2 * i + 1]);
  }
}

// C[c] = sum over the labels C lacks of A[a] B[b], every label assignment visited once.
std::vector<cd> reference(const Case& cs, const std::vector<cd>& A, const std::vector<cd>& B, bool cplx) {
  const size_t L = cs.extent.size() - 1;
  std::vector<cd> C((size_t)cs.c.span, cd(0, 0));
  std::vector<int64_t> idx(L + 1, 0);
  auto offset = [&](const Operand& o) {
    int64_t off = 0;
    for (size_t i = 0; i < o.modes.size(); ++i) off += idx[(size_t)o.modes[i]] * o.strides[i];
    return off;
  };
  for (int64_t e : cs.extent)
    if (e == 0) return C;
  for (;;) {
    cd x = A[(size_t)offset(cs.a)], y = B[(size_t)offset(cs.b)];
    if (!cplx) {
      x = cd(x.real(), 0);
      y = cd(y.real(), 0);
    }
    C[(size_t)offset(cs.c)] += x * y;
    size_t l = 1;
    for (; l <= L; ++l) {
      if (++idx[l] < cs.extent[l]) break;
      idx[l] = 0;
    }
    if (l > L) break;
  }
  return C;
}

double rel_err(const std::vector<cd>& got, const std::vector<cd>& want) {
  double num = 0, den = 0;
  for (size_t i = 0; i < got.size(); ++i) {
    num += std::norm(got[i] - want[i]);
    den += std::norm(want[i]);
  }
  return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

struct DevCase {
  Case cs;
  int dt;
  std::vector<cd> A, B, want;
  void *da = nullptr, *db = nullptr, *dc = nullptr;
  muscle_hip_tensor_t ta, tb, tc;
};

DevCase make_dev_case(std::mt19937_64& rng, int dt) {
  DevCase d;
  d.cs = random_case(rng, 7, 6, 1 << 16);
  d.dt = dt;
  d.A = random_values(d.cs.a, rng);
  d.B = random_values(d.cs.b, rng);
  d.want = reference(d.cs, d.A, d.B, is_cplx(dt));
  const auto ha = encode(d.A, dt), hb = encode(d.B, dt);
  expect(hipMalloc(&d.da, ha.size()) == hipSuccess && hipMalloc(&d.db, hb.size()) == hipSuccess &&
             hipMalloc(&d.dc, (size_t)d.cs.c.span * esize(dt)) == hipSuccess,
         "hipMalloc");
  expect(hipMemcpy(d.da, ha.data(), ha.size(), hipMemcpyHostToDevice) == hipSuccess, "copy A");
  expect(hipMemcpy(d.db, hb.data(), hb.size(), hipMemcpyHostToDevice) == hipSuccess, "copy B");
  expect(hipMemset(d.dc, 0x7f, (size_t)d.cs.c.span * esize(dt)) == hipSuccess, "poison C");
  d.ta = d.cs.a.desc(dt, d.da);
  d.tb = d.cs.b.desc(dt, d.db);
  d.tc = d.cs.c.desc(dt, d.dc);
  return d;
}

double dev_error(const DevCase& d) {
  std::vector<char> buf((size_t)d.cs.c.span * esize(d.dt));
  expect(hipMemcpy(buf.data(), d.dc, buf.size(), hipMemcpyDeviceToHost) == hipSuccess, "copy C");
  std::vector<cd> got(d.want.size());
  for (size_t i = 0; i < got.size(); ++i) got[i] = decode(buf, d.dt, i);
  return rel_err(got, d.want);
}

void free_dev_case(DevCase& d) {
  expect(hipFree(d.da) == hipSuccess && hipFree(d.db) == hipSuccess && hipFree(d.dc) == hipSuccess, "hipFree");
}

double tolerance(int dt) { return dt == MUSCLE_HIP_F32 || dt == MUSCLE_HIP_C32 ? 2e-5 : 1e-12; }

void check_gpu(std::mt19937_64& rng) {
  muscle_hip_handle_t h = nullptr;
  expect(muscle_hip_create(&h) == MUSCLE_HIP_OK, "create handle");
  if (!h) return;
  // single contractions, every dtype
  int n_single = 0;
  for (int it = 0; it < 80; ++it) {
    const int dt = (int)(rng() % 4);
    DevCase d = make_dev_case(rng, dt);
    muscle_hip_info_t info;
    const int st = muscle_hip_binary_einsum(h, &d.tc, &d.ta, &d.tb, 0, 1e-12, &info);
    expect(st == MUSCLE_HIP_OK, "binary_einsum", it);
    expect(muscle_hip_synchronize(h) == MUSCLE_HIP_OK, "synchronize", it);
    if (st == MUSCLE_HIP_OK) {
      const double e = dev_error(d);
      expect(e <= tolerance(dt), "binary_einsum result", it);
      ++n_single;
    }
    free_dev_case(d);
  }
  // one batch of ComplexF64 pairs, in every run mode
  std::vector<DevCase> pairs;
  for (int i = 0; i < 40; ++i) pairs.push_back(make_dev_case(rng, MUSCLE_HIP_C64));
  std::vector<muscle_hip_tensor_t> cs, as, bs;
  for (auto& d : pairs) {
    cs.push_back(d.tc);
    as.push_back(d.ta);
    bs.push_back(d.tb);
  }
  muscle_hip_batch_t batch = nullptr;
  expect(muscle_hip_batch_create(h, (int64_t)pairs.size(), cs.data(), as.data(), bs.data(), nullptr, &batch) ==
             MUSCLE_HIP_OK,
         "batch_create");
  if (batch) {
    for (int mode : {0, 0, 1, 2}) {
      for (auto& d : pairs)
        expect(hipMemset(d.dc, 0x7f, (size_t)d.cs.c.span * esize(d.dt)) == hipSuccess, "poison C", mode);
      expect(muscle_hip_batch_run(h, batch, mode) == MUSCLE_HIP_OK, "batch_run", mode);
      expect(muscle_hip_synchronize(h) == MUSCLE_HIP_OK, "batch synchronize", mode);
      for (size_t i = 0; i < pairs.size(); ++i) expect(dev_error(pairs[i]) <= 1e-12, "batch pair", (int64_t)i);
    }
    muscle_hip_batch_info_t bi;
    expect(muscle_hip_batch_get_info(batch, &bi) == MUSCLE_HIP_OK && bi.pairs == (int64_t)pairs.size(), "batch info");
    for (int64_t i = 0; i < (int64_t)pairs.size(); ++i) {
      muscle_hip_batch_pair_info_t pi;
      expect(muscle_hip_batch_get_pair_info(batch, i, &pi) == MUSCLE_HIP_OK, "pair info", i);
    }
    muscle_hip_batch_pair_info_t pi;
    expect(muscle_hip_batch_get_pair_info(batch, (int64_t)pairs.size(), &pi) != MUSCLE_HIP_OK, "pair index range");
    expect(muscle_hip_batch_destroy(batch) == MUSCLE_HIP_OK, "batch_destroy");
  }
  for (auto& d : pairs) free_dev_case(d);
  expect(muscle_hip_destroy(h) == MUSCLE_HIP_OK, "destroy handle");
  std::printf("gpu: %d contractions, a batch of %zu pairs in 4 runs\n", n_single, pairs.size());

  // server: chains ((x y) z) of ComplexF64 matrices, a few jobs at once
  muscle_hip_server_t srv = nullptr;
  expect(muscle_hip_server_create(int64_t(256) << 20, 4, 0, &srv) == MUSCLE_HIP_OK, "server_create");
  if (!srv) return;
  struct Job {
    int64_t d0, d1, d2, d3;
    std::vector<cd> X, Y, Z, want;
    void *dx, *dy, *dz;
    int64_t id;
  };
  std::vector<Job> jobs(12);
  const int32_t lx[2] = {1, 2}, ly[2] = {2, 3}, lz[2] = {3, 4};
  for (auto& j : jobs) {
    std::uniform_int_distribution<int> ud(1, 24);
    j.d0 = ud(rng);
    j.d1 = ud(rng);
    j.d2 = ud(rng);
    j.d3 = ud(rng);
    std::normal_distribution<double> nd;
    auto fill = [&](std::vector<cd>& v, int64_t n) {
      v.resize((size_t)n);
      for (auto& x : v) x = cd(nd(rng), nd(rng));
    };
    fill(j.X, j.d0 * j.d1);
    fill(j.Y, j.d1 * j.d2);
    fill(j.Z, j.d2 * j.d3);
    // want = (X Y) Z, column-major
    std::vector<cd> XY((size_t)(j.d0 * j.d2), cd(0, 0));
    for (int64_t a = 0; a < j.d0; ++a)
      for (int64_t c = 0; c < j.d2; ++c)
        for (int64_t b = 0; b < j.d1; ++b) XY[(size_t)(a + j.d0 * c)] += j.X[(size_t)(a + j.d0 * b)] * j.Y[(size_t)(b + j.d1 * c)];
    j.want.assign((size_t)(j.d0 * j.d3), cd(0, 0));
    for (int64_t a = 0; a < j.d0; ++a)
      for (int64_t d = 0; d < j.d3; ++d)
        for (int64_t c = 0; c < j.d2; ++c) j.want[(size_t)(a + j.d0 * d)] += XY[(size_t)(a + j.d0 * c)] * j.Z[(size_t)(c + j.d2 * d)];
    auto up = [&](const std::vector<cd>& v, void** dst) {
      const auto hb = encode(v, MUSCLE_HIP_C64);
      expect(hipMalloc(dst, hb.size()) == hipSuccess, "server leaf alloc");
      expect(hipMemcpy(*dst, hb.data(), hb.size(), hipMemcpyHostToDevice) == hipSuccess, "server leaf copy");
    };
    up(j.X, &j.dx);
    up(j.Y, &j.dy);
    up(j.Z, &j.dz);
    const int64_t ext[5] = {1, j.d0, j.d1, j.d2, j.d3};
    const int64_t sx[2] = {1, j.d0}, sy[2] = {1, j.d1}, sz[2] = {1, j.d2};
    const int64_t ex[2] = {j.d0, j.d1}, ey[2] = {j.d1, j.d2}, ez[2] = {j.d2, j.d3};
    muscle_hip_tensor_t leaves[3] = {{MUSCLE_HIP_C64, 2, j.dx, lx, ex, sx},
                                     {MUSCLE_HIP_C64, 2, j.dy, ly, ey, sy},
                                     {MUSCLE_HIP_C64, 2, j.dz, lz, ez, sz}};
    const int32_t lhs[2] = {0, 3}, rhs[2] = {1, 2}, off[3] = {0, 2, 4}, nmodes[4] = {1, 3, 1, 4};
    muscle_hip_job_t job{3, 2, leaves, lhs, rhs, off, nmodes, 5, ext};
    expect(muscle_hip_server_submit(srv, &job, &j.id) == MUSCLE_HIP_OK, "server_submit");
  }
  int done = 0;
  for (auto& j : jobs) {
    muscle_hip_job_result_t r;
    expect(muscle_hip_server_wait(srv, j.id, 120.0, &r) == MUSCLE_HIP_OK, "server_wait", j.id);
    if (r.status != MUSCLE_HIP_OK || r.result_elems != j.d0 * j.d3) {
      expect(false, "server job status", j.id);
      continue;
    }
    std::vector<char> buf((size_t)r.result_elems * 16);
    expect(muscle_hip_server_copy_result(srv, j.id, buf.data()) == MUSCLE_HIP_OK, "copy_result", j.id);
    std::vector<cd> got((size_t)r.result_elems);
    for (size_t i = 0; i < got.size(); ++i) got[i] = decode(buf, MUSCLE_HIP_C64, i);
    expect(rel_err(got, j.want) <= 1e-12, "server job result", j.id);
    expect(muscle_hip_server_release(srv, j.id) == MUSCLE_HIP_OK, "release", j.id);
    ++done;
  }
  muscle_hip_server_stats_t stats;
  expect(muscle_hip_server_get_stats(srv, &stats) == MUSCLE_HIP_OK && stats.completed == (int64_t)jobs.size(),
         "server stats");
  expect(muscle_hip_server_destroy(srv) == MUSCLE_HIP_OK, "server_destroy");
  for (auto& j : jobs) {
    expect(hipFree(j.dx) == hipSuccess && hipFree(j.dy) == hipSuccess && hipFree(j.dz) == hipSuccess, "hipFree");
  }
  std::printf("server: %d jobs checked\n", done);
}

}  // namespace

int main(int argc, char** argv) {
  std::mt19937_64 rng(20261203);
  check_describe(rng);
  if (argc > 1 && std::strcmp(argv[1], "--gpu") == 0) check_gpu(rng);
  std::printf("%s: %d failures\n", g_failures ? "FAILED" : "ok", g_failures);
  return g_failures ? 1 : 0;
}
