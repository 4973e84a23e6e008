// Thin SVD and thin QR of a tensor viewed as a matrix (the factorizations
// behind simple_update, src/Operations/simple_update.jl:52-56, and the
// canonical forms that keep tree-contraction workflows on the device).
//
// tensor_svd_thin's BackendBase (src/Operations/tensor_svd.jl:100-125)
// permutes A to (inds_u..., inds_v...), reshapes it to an m x n matrix,
// calls LAPACK's SVD and reshapes U (m x k), s (k) and conj(V) (n x k) back
// to tensors.  Here the permutation copies are this library's own unary
// (copy) plans.  The SVD of the m x n matrix M comes from the Hermitian
// eigenproblem of the augmented matrix [[0, M], [M^H, 0]] (rocSOLVER
// syevd/heevd, divide and conquer): its eigenpairs are (+-sigma_i,
// [u_i; +-v_i] / sqrt(2)), so the top k eigenvectors give U and V after each
// half is normalised.  The eigensolver is backward stable, so the singular
// values carry LAPACK's absolute accuracy (a few eps * ||M||) and the vectors
// its eps * ||M|| / gap accuracy.  Two passes of Cholesky QR then make U and V
// orthonormal to working precision (column j is only corrected with columns
// of larger singular value, so U s V^H moves by O(eps * ||M||)); singular
// values at the eigensolver's noise level (where +sigma and -sigma pairs and
// the null space mix) get a deterministic pseudo-random orthonormal
// completion instead, as LAPACK's vectors for a zero singular value are any
// orthonormal completion too.  A matrix more than twice as tall as wide is
// first reduced by QR; a wide one is factored as its transpose.  The eigen
// route replaces rocSOLVER's gesvd, whose bidiagonal QR iteration is about
// 60x slower at 512 x 512 on this GPU.
//
// tensor_qr_thin (src/Operations/tensor_qr.jl:45-68) is the same permutation
// with geqrf + orgqr/ungqr; R is the upper trapezoid of the first k rows of
// the factored matrix.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "handle.hpp"
#include "kernels.hpp"
#include "lowering.hpp"

namespace mh {
namespace {

// A column-major descriptor over a dense buffer: modes/extents given, strides
// computed fastest-first, optionally with one trailing mode of stride `last`.
struct View {
  std::vector<int32_t> modes;
  std::vector<int64_t> ext, str;
  muscle_hip_tensor_t t;
  void finish(int dtype, void* data) {
    t.dtype = dtype;
    t.rank = (int32_t)modes.size();
    t.data = data;
    t.modes = modes.data();
    t.extents = ext.data();
    t.strides = str.data();
  }
};

void copy_tensor(Handle& h, const muscle_hip_tensor_t& dst, const muscle_hip_tensor_t& src) {
  Plan* p = get_plan(h, 2, MUSCLE_HIP_PRECISION_NATIVE, 0.0, &dst, &src, nullptr);
  run_contraction(h, *p, dst, src, nullptr);
}

void ensure_fws(Handle& h, size_t bytes) {
  if (bytes <= h.fws_size) return;
  if (h.fws) {
    MH_HIP_CHECK(hipStreamSynchronize(h.stream));
    MH_HIP_CHECK(hipFree(h.fws));
    h.fws = nullptr;
    h.fws_size = 0;
  }
  if (hipMalloc(&h.fws, bytes) != hipSuccess) {
    (void)hipGetLastError();
    h.fws = nullptr;
    fail(MUSCLE_HIP_OUT_OF_MEMORY, "factorization workspace of " + std::to_string(bytes) + " bytes");
  }
  h.fws_size = bytes;
}

size_t al(size_t v) { return (v + 255) & ~size_t(255); }

rocblas_status heevd(int dt, rocblas_handle rb, int n, void* A, void* w, void* e, rocblas_int* info) {
  const rocblas_evect ev = rocblas_evect_original;
  const rocblas_fill up = rocblas_fill_upper;
  switch (dt) {
    case MUSCLE_HIP_F32:
      return rocsolver_ssyevd(rb, ev, up, n, (float*)A, n, (float*)w, (float*)e, info);
    case MUSCLE_HIP_F64:
      return rocsolver_dsyevd(rb, ev, up, n, (double*)A, n, (double*)w, (double*)e, info);
    case MUSCLE_HIP_C32:
      return rocsolver_cheevd(rb, ev, up, n, (rocblas_float_complex*)A, n, (float*)w, (float*)e, info);
    default:
      return rocsolver_zheevd(rb, ev, up, n, (rocblas_double_complex*)A, n, (double*)w, (double*)e, info);
  }
}

// C = alpha op(A) op(B) + beta C, op the identity or the conjugate transpose
rocblas_status gemm(int dt, rocblas_handle rb, bool ha, bool hb, int64_t m, int64_t n, int64_t k, double alpha,
                    const void* A, int64_t lda, const void* B, int64_t ldb, double beta, void* C, int64_t ldc) {
  const bool cplx = dt == MUSCLE_HIP_C32 || dt == MUSCLE_HIP_C64;
  const rocblas_operation h = cplx ? rocblas_operation_conjugate_transpose : rocblas_operation_transpose;
  const rocblas_operation oa = ha ? h : rocblas_operation_none, ob = hb ? h : rocblas_operation_none;
  switch (dt) {
    case MUSCLE_HIP_F32: {
      const float a = (float)alpha, b = (float)beta;
      return rocblas_sgemm(rb, oa, ob, (int)m, (int)n, (int)k, &a, (const float*)A, (int)lda, (const float*)B,
                           (int)ldb, &b, (float*)C, (int)ldc);
    }
    case MUSCLE_HIP_F64:
      return rocblas_dgemm(rb, oa, ob, (int)m, (int)n, (int)k, &alpha, (const double*)A, (int)lda, (const double*)B,
                           (int)ldb, &beta, (double*)C, (int)ldc);
    case MUSCLE_HIP_C32: {
      const rocblas_float_complex a((float)alpha, 0.f), b((float)beta, 0.f);
      return rocblas_cgemm(rb, oa, ob, (int)m, (int)n, (int)k, &a, (const rocblas_float_complex*)A, (int)lda,
                           (const rocblas_float_complex*)B, (int)ldb, &b, (rocblas_float_complex*)C, (int)ldc);
    }
    default: {
      const rocblas_double_complex a(alpha, 0.0), b(beta, 0.0);
      return rocblas_zgemm(rb, oa, ob, (int)m, (int)n, (int)k, &a, (const rocblas_double_complex*)A, (int)lda,
                           (const rocblas_double_complex*)B, (int)ldb, &b, (rocblas_double_complex*)C, (int)ldc);
    }
  }
}

// G (n x n, upper) = R with R^H R = G
rocblas_status potrf(int dt, rocblas_handle rb, int64_t n, void* G, rocblas_int* info) {
  const rocblas_fill up = rocblas_fill_upper;
  switch (dt) {
    case MUSCLE_HIP_F32:
      return rocsolver_spotrf(rb, up, (int)n, (float*)G, (int)n, info);
    case MUSCLE_HIP_F64:
      return rocsolver_dpotrf(rb, up, (int)n, (double*)G, (int)n, info);
    case MUSCLE_HIP_C32:
      return rocsolver_cpotrf(rb, up, (int)n, (rocblas_float_complex*)G, (int)n, info);
    default:
      return rocsolver_zpotrf(rb, up, (int)n, (rocblas_double_complex*)G, (int)n, info);
  }
}

// X (m x n) = X R^{-1}, R (n x n) upper triangular
rocblas_status trsm_right_upper(int dt, rocblas_handle rb, int64_t m, int64_t n, const void* R, void* X) {
  const rocblas_side sd = rocblas_side_right;
  const rocblas_fill up = rocblas_fill_upper;
  const rocblas_operation no = rocblas_operation_none;
  const rocblas_diagonal dg = rocblas_diagonal_non_unit;
  switch (dt) {
    case MUSCLE_HIP_F32: {
      const float one = 1.f;
      return rocblas_strsm(rb, sd, up, no, dg, (int)m, (int)n, &one, (const float*)R, (int)n, (float*)X, (int)m);
    }
    case MUSCLE_HIP_F64: {
      const double one = 1.0;
      return rocblas_dtrsm(rb, sd, up, no, dg, (int)m, (int)n, &one, (const double*)R, (int)n, (double*)X, (int)m);
    }
    case MUSCLE_HIP_C32: {
      const rocblas_float_complex one(1.f, 0.f);
      return rocblas_ctrsm(rb, sd, up, no, dg, (int)m, (int)n, &one, (const rocblas_float_complex*)R, (int)n,
                           (rocblas_float_complex*)X, (int)m);
    }
    default: {
      const rocblas_double_complex one(1.0, 0.0);
      return rocblas_ztrsm(rb, sd, up, no, dg, (int)m, (int)n, &one, (const rocblas_double_complex*)R, (int)n,
                           (rocblas_double_complex*)X, (int)m);
    }
  }
}

void solver_check(rocblas_status st, const char* what) {
  if (st != rocblas_status_success) fail(MUSCLE_HIP_HIP_ERROR, std::string("rocSOLVER/rocBLAS ") + what + " failed");
}

// Two passes of Cholesky QR over the columns of X (rows x cols, ld rows):
// X = X R^{-1} with R^H R = X^H X.  info[0..1] receive potrf's flags.
void orthonormalize(int dt, rocblas_handle rb, void* X, int64_t rows, int64_t cols, void* G, rocblas_int* info) {
  for (int pass = 0; pass < 2; ++pass) {
    solver_check(gemm(dt, rb, true, false, cols, cols, rows, 1.0, X, rows, X, rows, 0.0, G, cols), "gemm");
    solver_check(potrf(dt, rb, cols, G, info + pass), "potrf");
    solver_check(trsm_right_upper(dt, rb, rows, cols, G, X), "trsm");
  }
}

rocblas_status geqrf(int dt, rocblas_handle rb, int m, int n, void* A, void* tau) {
  switch (dt) {
    case MUSCLE_HIP_F32:
      return rocsolver_sgeqrf(rb, m, n, (float*)A, m, (float*)tau);
    case MUSCLE_HIP_F64:
      return rocsolver_dgeqrf(rb, m, n, (double*)A, m, (double*)tau);
    case MUSCLE_HIP_C32:
      return rocsolver_cgeqrf(rb, m, n, (rocblas_float_complex*)A, m, (rocblas_float_complex*)tau);
    default:
      return rocsolver_zgeqrf(rb, m, n, (rocblas_double_complex*)A, m, (rocblas_double_complex*)tau);
  }
}

// the first k columns of the m x m unitary factor, in place over A (ld m)
rocblas_status orgqr(int dt, rocblas_handle rb, int m, int k, void* A, void* tau) {
  switch (dt) {
    case MUSCLE_HIP_F32:
      return rocsolver_sorgqr(rb, m, k, k, (float*)A, m, (float*)tau);
    case MUSCLE_HIP_F64:
      return rocsolver_dorgqr(rb, m, k, k, (double*)A, m, (double*)tau);
    case MUSCLE_HIP_C32:
      return rocsolver_cungqr(rb, m, k, k, (rocblas_float_complex*)A, m, (rocblas_float_complex*)tau);
    default:
      return rocsolver_zungqr(rb, m, k, k, (rocblas_double_complex*)A, m, (rocblas_double_complex*)tau);
  }
}

rocblas_handle solver_on_stream(Handle& h) {
  if (!h.rb) {
    rocblas_handle rb = nullptr;
    if (rocblas_create_handle(&rb) != rocblas_status_success) fail(MUSCLE_HIP_HIP_ERROR, "rocblas_create_handle");
    h.rb = rb;
  }
  if (rocblas_set_stream((rocblas_handle)h.rb, h.stream) != rocblas_status_success)
    fail(MUSCLE_HIP_HIP_ERROR, "rocblas_set_stream");
  return (rocblas_handle)h.rb;
}

// A's extent of `mode`; fails unless A holds the mode exactly once
int64_t extent_in(const muscle_hip_tensor_t& A, int32_t mode, const char* what) {
  int64_t e = -1;
  for (int i = 0; i < A.rank; ++i)
    if (A.modes[i] == mode) {
      if (e >= 0) fail(MUSCLE_HIP_INVALID_ARGUMENT, std::string(what) + ": repeated mode in A");
      e = A.extents[i];
    }
  if (e < 0) fail(MUSCLE_HIP_INVALID_ARGUMENT, std::string(what) + ": a row mode of a factor is not a mode of A");
  return e;
}

// Appends T's modes [begin, end) -- a factor's modes other than the bond -- to
// the matrix view `mat` and returns the product of their extents.
int64_t add_rows(View& mat, const muscle_hip_tensor_t& A, const muscle_hip_tensor_t& T, int begin, int end,
                 int32_t bond, const char* what) {
  int64_t prod = 1;
  for (int i = begin; i < end; ++i) {
    if (T.modes[i] == bond) fail(MUSCLE_HIP_INVALID_ARGUMENT, std::string(what) + ": the bond mode is a mode of A");
    const int64_t e = extent_in(A, T.modes[i], what);
    if (e != T.extents[i]) fail(MUSCLE_HIP_DIM_MISMATCH, std::string(what) + ": a factor's extents differ from A's");
    mat.modes.push_back(T.modes[i]);
    mat.ext.push_back(e);
    prod *= e;
  }
  return prod;
}

// The matrix view covers A's modes once each; strides column-major.
void finish_matrix(View& mat, const muscle_hip_tensor_t& A, int64_t m, int64_t n, void* data, const char* what) {
  if ((int)mat.modes.size() != A.rank)
    fail(MUSCLE_HIP_INVALID_ARGUMENT, std::string(what) + ": the factors' modes must cover A's");
  for (size_t i = 0; i < mat.modes.size(); ++i)
    for (size_t j = i + 1; j < mat.modes.size(); ++j)
      if (mat.modes[i] == mat.modes[j])
        fail(MUSCLE_HIP_INVALID_ARGUMENT, std::string(what) + ": the factors share a mode");
  if (m >= (int64_t(1) << 31) || n >= (int64_t(1) << 31) || (double)m * (double)n >= 9.0e18)
    fail(MUSCLE_HIP_UNSUPPORTED, std::string(what) + ": matrix too large");
  int64_t st = 1;
  for (int64_t e : mat.ext) {
    mat.str.push_back(st);
    st *= e;
  }
  mat.finish(A.dtype, data);
}

}  // namespace

void release_rb(Handle& h) {
  if (h.fws) (void)hipFree(h.fws);
  if (h.rb) (void)rocblas_destroy_handle((rocblas_handle)h.rb);
  h.fws = nullptr;
  h.fws_size = 0;
  h.rb = nullptr;
}

// U[u..., s], S[s], V[v..., s] from A[u ∪ v]: see muscle_hip_svd_thin.
void svd_thin(Handle& h, const muscle_hip_tensor_t& U, const muscle_hip_tensor_t& S, const muscle_hip_tensor_t& V,
              const muscle_hip_tensor_t& A) {
  check_tensor(&A, "A");
  check_tensor(&U, "U");
  check_tensor(&S, "S");
  check_tensor(&V, "V");
  if (U.rank < 2 || V.rank < 2 || S.rank != 1)
    fail(MUSCLE_HIP_INVALID_ARGUMENT, "svd: U and V need their row modes and a trailing bond mode, S the bond mode");
  const int32_t sm = S.modes[0];
  if (U.modes[U.rank - 1] != sm || V.modes[V.rank - 1] != sm)
    fail(MUSCLE_HIP_INVALID_ARGUMENT, "svd: the bond mode must be the last mode of U and of V and the mode of S");
  const int dt = A.dtype;
  if (dt != U.dtype || dt != V.dtype || S.dtype != (is_double(dt) ? MUSCLE_HIP_F64 : MUSCLE_HIP_F32))
    fail(MUSCLE_HIP_INVALID_ARGUMENT, "svd: U and V have A's element type, S its real type");
  // A's modes split into U's (u) and V's (v) row modes, each once
  View mat;
  const int64_t m = add_rows(mat, A, U, 0, U.rank - 1, sm, "svd");
  const int64_t n = add_rows(mat, A, V, 0, V.rank - 1, sm, "svd");
  const int64_t k = std::min(m, n);
  if (U.extents[U.rank - 1] != k || V.extents[V.rank - 1] != k || S.extents[0] != k)
    fail(MUSCLE_HIP_DIM_MISMATCH, "svd: the bond extent must be min(rows of U, rows of V)");
  finish_matrix(mat, A, m, n, nullptr, "svd");
  if (k == 0) return;

  // Factor the tall orientation: M (m x n) when m >= n, else M^T, whose
  // factors are M's with U and V exchanged (M^T = U' s V'^H gives
  // M = conj(V') s conj(U')^H).  Either way the row-group tensor RT receives
  // the staged left vectors and the column-group tensor CT the conjugated
  // right ones.
  const bool tall = m >= n;
  const muscle_hip_tensor_t& RT = tall ? U : V;
  const muscle_hip_tensor_t& CT = tall ? V : U;
  const int64_t rows = tall ? m : n, cols = k;
  const bool reduce = rows > 2 * cols;  // QR first, so that the eigenproblem is 2 cols wide
  const int64_t top = reduce ? cols : rows, nev = top + cols;
  if (nev >= (int64_t(1) << 31)) fail(MUSCLE_HIP_UNSUPPORTED, "svd: matrix too large");
  const size_t es = dtype_size(dt), rs = dtype_size(S.dtype);
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t o = off;
    off += al(bytes);
    return o;
  };
  const size_t o_h = carve((size_t)(nev * nev) * es), o_w = carve((size_t)nev * rs), o_e = carve((size_t)nev * rs),
               o_u = carve((size_t)(rows * k) * es), o_ut = reduce ? carve((size_t)(top * k) * es) : o_u,
               o_v = carve((size_t)(cols * k) * es), o_g = carve((size_t)(k * k) * es),
               o_q = reduce ? carve((size_t)(rows * cols) * es) : 0, o_tau = reduce ? carve((size_t)k * es) : 0,
               o_s = carve((size_t)k * rs), o_info = carve(8 * sizeof(rocblas_int));
  ensure_fws(h, off);
  char* w = (char*)h.fws;
  rocblas_handle rb = solver_on_stream(h);
  rocblas_int* info = (rocblas_int*)(w + o_info);

  // the matrix, written by a copy plan straight into the augmented matrix's
  // upper-right block (or into the QR buffer first); the rest of the
  // augmented matrix is zero
  View mv;
  auto add_group = [&](const muscle_hip_tensor_t& T, int64_t stride) {
    for (int i = 0; i < T.rank - 1; ++i) {
      mv.modes.push_back(T.modes[i]);
      mv.ext.push_back(T.extents[i]);
      mv.str.push_back(stride);
      stride *= T.extents[i];
    }
  };
  add_group(RT, 1);
  add_group(CT, reduce ? rows : nev);
  launch_zero(w + o_h, (int64_t)((size_t)(nev * nev) * es), h.stream);
  MH_HIP_CHECK(hipGetLastError());
  if (!reduce) {
    mv.finish(dt, w + o_h + (size_t)(rows * nev) * es);
    copy_tensor(h, mv.t, A);
  } else {
    mv.finish(dt, w + o_q);
    copy_tensor(h, mv.t, A);
    solver_check(geqrf(dt, rb, (int)rows, (int)cols, w + o_q, w + o_tau), "geqrf");
    launch_upper_trapezoid((int)es, w + o_q, w + o_h + (size_t)(top * nev) * es, rows, cols, cols, nev, h.stream);
    MH_HIP_CHECK(hipGetLastError());
    solver_check(orgqr(dt, rb, (int)rows, (int)cols, w + o_q, w + o_tau), "orgqr");
  }
  solver_check(heevd(dt, rb, (int)nev, w + o_h, w + o_w, w + o_e, info), "syevd/heevd");
  launch_split_pairs(dt, w + o_h, w + o_w, nev, top, cols, k, w + o_ut, w + o_v, w + o_s, h.stream);
  MH_HIP_CHECK(hipGetLastError());

  // the eigenvalues decide which columns sit at the eigensolver's noise level
  std::vector<char> wh((size_t)nev * rs);
  rocblas_int info0 = 0;
  MH_HIP_CHECK(hipMemcpyAsync(wh.data(), w + o_w, wh.size(), hipMemcpyDeviceToHost, h.stream));
  MH_HIP_CHECK(hipMemcpyAsync(&info0, info, sizeof info0, hipMemcpyDeviceToHost, h.stream));
  MH_HIP_CHECK(hipStreamSynchronize(h.stream));
  if (info0 != 0) fail(MUSCLE_HIP_HIP_ERROR, "svd: the eigensolver did not converge");
  auto lam = [&](int64_t i) -> double {  // i-th largest eigenvalue
    const int64_t j = nev - 1 - i;
    return rs == 8 ? ((const double*)wh.data())[j] : (double)((const float*)wh.data())[j];
  };
  const double eps = rs == 8 ? 0x1p-52 : 0x1p-23;
  const double tol = 8.0 * std::sqrt((double)nev) * eps * std::max(lam(0), 0.0);
  int64_t r = 0;
  while (r < k && lam(r) > tol) ++r;
  if (r < k) {
    launch_random_unit(dt, w + o_ut + (size_t)(r * top) * es, (k - r) * top, 0x5EED0001ull, h.stream);
    launch_random_unit(dt, w + o_v + (size_t)(r * cols) * es, (k - r) * cols, 0x5EED0002ull, h.stream);
    MH_HIP_CHECK(hipGetLastError());
  }
  orthonormalize(dt, rb, w + o_ut, top, k, w + o_g, info + 1);
  orthonormalize(dt, rb, w + o_v, cols, k, w + o_g, info + 3);
  if (reduce)  // U = Q U_R
    solver_check(gemm(dt, rb, false, false, rows, k, cols, 1.0, w + o_q, rows, w + o_ut, cols, 0.0, w + o_u, rows),
                 "gemm");

  // staged factors (column-major, bond last) -> the caller's layouts
  View uv, vv, sv;
  auto staged = [&](View& t, const muscle_hip_tensor_t& T, int64_t ld, size_t o) {
    int64_t st = 1;
    for (int i = 0; i < T.rank - 1; ++i) {
      t.modes.push_back(T.modes[i]);
      t.ext.push_back(T.extents[i]);
      t.str.push_back(st);
      st *= T.extents[i];
    }
    t.modes.push_back(sm);
    t.ext.push_back(k);
    t.str.push_back(ld);
    t.finish(dt, w + o);
  };
  staged(uv, RT, rows, o_u);
  staged(vv, CT, cols, o_v);
  copy_tensor(h, RT, uv.t);
  copy_tensor(h, CT, vv.t);
  sv.modes.push_back(sm);
  sv.ext.push_back(k);
  sv.str.push_back(1);
  sv.finish(S.dtype, w + o_s);
  copy_tensor(h, S, sv.t);
  rocblas_int infos[4] = {0, 0, 0, 0};
  MH_HIP_CHECK(hipMemcpyAsync(infos, info + 1, sizeof infos, hipMemcpyDeviceToHost, h.stream));
  MH_HIP_CHECK(hipStreamSynchronize(h.stream));
  for (rocblas_int f : infos)
    if (f != 0) fail(MUSCLE_HIP_HIP_ERROR, "svd: the singular vectors lost rank in the Cholesky QR");
}

// Q[q..., v], R[v, r...] from A[q ∪ r]: see muscle_hip_qr_thin.
void qr_thin(Handle& h, const muscle_hip_tensor_t& Q, const muscle_hip_tensor_t& R, const muscle_hip_tensor_t& A) {
  check_tensor(&A, "A");
  check_tensor(&Q, "Q");
  check_tensor(&R, "R");
  if (Q.rank < 2 || R.rank < 2)
    fail(MUSCLE_HIP_INVALID_ARGUMENT, "qr: Q needs its row modes and a trailing bond mode, R a leading bond mode");
  const int32_t bond = Q.modes[Q.rank - 1];
  if (R.modes[0] != bond) fail(MUSCLE_HIP_INVALID_ARGUMENT, "qr: the bond mode must be Q's last mode and R's first");
  if (A.dtype != Q.dtype || A.dtype != R.dtype) fail(MUSCLE_HIP_INVALID_ARGUMENT, "qr: Q and R have A's element type");
  View mat, qview, rview;
  const int64_t m = add_rows(mat, A, Q, 0, Q.rank - 1, bond, "qr");
  const int64_t n = add_rows(mat, A, R, 1, R.rank, bond, "qr");
  const int64_t k = std::min(m, n);
  if (Q.extents[Q.rank - 1] != k || R.extents[0] != k)
    fail(MUSCLE_HIP_DIM_MISMATCH, "qr: the bond extent must be min(rows of Q, columns of R)");
  finish_matrix(mat, A, m, n, nullptr, "qr");
  if (k == 0) return;

  const size_t es = dtype_size(A.dtype);
  const size_t o_mat = 0, o_r = o_mat + al((size_t)(m * n) * es), o_tau = o_r + al((size_t)(k * n) * es),
               total = o_tau + al((size_t)k * es);
  ensure_fws(h, total);
  char* w = (char*)h.fws;
  rocblas_handle rb = solver_on_stream(h);
  mat.t.data = w + o_mat;
  copy_tensor(h, mat.t, A);
  solver_check(geqrf(A.dtype, rb, (int)m, (int)n, w + o_mat, w + o_tau), "geqrf");
  // R (k x n): the upper trapezoid, before orgqr overwrites the matrix
  launch_upper_trapezoid((int)es, w + o_mat, w + o_r, m, k, n, k, h.stream);
  MH_HIP_CHECK(hipGetLastError());
  rview.modes.push_back(bond);
  rview.ext.push_back(k);
  rview.str.push_back(1);
  int64_t sr = k;
  for (int i = 1; i < R.rank; ++i) {
    rview.modes.push_back(R.modes[i]);
    rview.ext.push_back(R.extents[i]);
    rview.str.push_back(sr);
    sr *= R.extents[i];
  }
  rview.finish(A.dtype, w + o_r);
  copy_tensor(h, R, rview.t);
  solver_check(orgqr(A.dtype, rb, (int)m, (int)k, w + o_mat, w + o_tau), "orgqr");
  // Q (m x k, ld m): modes (q..., v)
  qview.modes.assign(mat.modes.begin(), mat.modes.begin() + (Q.rank - 1));
  qview.ext.assign(mat.ext.begin(), mat.ext.begin() + (Q.rank - 1));
  qview.str.assign(mat.str.begin(), mat.str.begin() + (Q.rank - 1));
  qview.modes.push_back(bond);
  qview.ext.push_back(k);
  qview.str.push_back(m);
  qview.finish(A.dtype, w + o_mat);
  copy_tensor(h, Q, qview.t);
}

}  // namespace mh
