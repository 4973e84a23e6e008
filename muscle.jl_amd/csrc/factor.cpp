// Thin SVD and thin QR of a tensor viewed as a matrix (the factorizations
// behind simple_update, src/Operations/simple_update.jl:52-56, and the
// canonical forms that keep tree-contraction workflows on the device).
//
// tensor_svd_thin's BackendBase (src/Operations/tensor_svd.jl:100-125)
// permutes A to (inds_u..., inds_v...), reshapes it to an m x n matrix,
// calls LAPACK's SVD and reshapes U (m x k), s (k) and conj(V) (n x k) back
// to tensors.  Here the permutation copies are this library's own unary
// (copy) plans, and the SVD of the m x n matrix is rocSOLVER's gesvd on the
// handle's stream; U and V^H come out as the column-major m x k and k x n
// matrices that the output tensors then view.  tensor_qr_thin
// (src/Operations/tensor_qr.jl:45-68) is the same with geqrf + orgqr/ungqr;
// R is the upper trapezoid of the first k rows of the factored matrix.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
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

rocblas_status gesvd(int dt, rocblas_handle rb, int m, int n, void* A, void* S, void* U, void* V, void* E,
                     rocblas_int* info) {
  const int k = std::min(m, n);
  switch (dt) {
    case MUSCLE_HIP_F32:
      return rocsolver_sgesvd(rb, rocblas_svect_singular, rocblas_svect_singular, m, n, (float*)A, m, (float*)S,
                              (float*)U, m, (float*)V, k, (float*)E, rocblas_outofplace, info);
    case MUSCLE_HIP_F64:
      return rocsolver_dgesvd(rb, rocblas_svect_singular, rocblas_svect_singular, m, n, (double*)A, m, (double*)S,
                              (double*)U, m, (double*)V, k, (double*)E, rocblas_outofplace, info);
    case MUSCLE_HIP_C32:
      return rocsolver_cgesvd(rb, rocblas_svect_singular, rocblas_svect_singular, m, n, (rocblas_float_complex*)A,
                              m, (float*)S, (rocblas_float_complex*)U, m, (rocblas_float_complex*)V, k, (float*)E,
                              rocblas_outofplace, info);
    default:
      return rocsolver_zgesvd(rb, rocblas_svect_singular, rocblas_svect_singular, m, n, (rocblas_double_complex*)A,
                              m, (double*)S, (rocblas_double_complex*)U, m, (rocblas_double_complex*)V, k,
                              (double*)E, rocblas_outofplace, info);
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
  if (A.dtype != U.dtype || A.dtype != V.dtype || S.dtype != (is_double(A.dtype) ? MUSCLE_HIP_F64 : MUSCLE_HIP_F32))
    fail(MUSCLE_HIP_INVALID_ARGUMENT, "svd: U and V have A's element type, S its real type");
  // A's modes split into U's (u) and V's (v) row modes, each once
  View mat, uview, vview, sview;
  const int64_t m = add_rows(mat, A, U, 0, U.rank - 1, sm, "svd");
  const int64_t n = add_rows(mat, A, V, 0, V.rank - 1, sm, "svd");
  const int64_t k = std::min(m, n);
  if (U.extents[U.rank - 1] != k || V.extents[V.rank - 1] != k || S.extents[0] != k)
    fail(MUSCLE_HIP_DIM_MISMATCH, "svd: the bond extent must be min(rows of U, rows of V)");

  const size_t es = dtype_size(A.dtype), rs = dtype_size(S.dtype);
  const size_t o_mat = 0, o_u = o_mat + al((size_t)(m * n) * es), o_v = o_u + al((size_t)(m * k) * es),
               o_s = o_v + al((size_t)(k * n) * es), o_e = o_s + al((size_t)k * rs), o_info = o_e + al((size_t)k * rs),
               total = o_info + 256;
  finish_matrix(mat, A, m, n, nullptr, "svd");
  if (k == 0) return;
  ensure_fws(h, total);
  char* w = (char*)h.fws;
  rocblas_handle rb = solver_on_stream(h);

  // A -> column-major m x n matrix, rows = U's row modes (first fastest), columns = V's
  mat.t.data = w + o_mat;
  copy_tensor(h, mat.t, A);

  const rocblas_status rs_ = gesvd(A.dtype, rb, (int)m, (int)n, w + o_mat, w + o_s, w + o_u,
                                   w + o_v, w + o_e, (rocblas_int*)(w + o_info));
  if (rs_ != rocblas_status_success) fail(MUSCLE_HIP_HIP_ERROR, "rocsolver gesvd failed to launch");
  rocblas_int info = 0;
  MH_HIP_CHECK(hipMemcpyAsync(&info, w + o_info, sizeof info, hipMemcpyDeviceToHost, h.stream));
  MH_HIP_CHECK(hipStreamSynchronize(h.stream));
  if (info != 0) fail(MUSCLE_HIP_HIP_ERROR, "svd: the bidiagonal QR iteration did not converge");

  // U (m x k): modes (u..., s), strides column-major then m
  uview.modes.assign(mat.modes.begin(), mat.modes.begin() + (U.rank - 1));
  uview.ext.assign(mat.ext.begin(), mat.ext.begin() + (U.rank - 1));
  uview.str.assign(mat.str.begin(), mat.str.begin() + (U.rank - 1));
  uview.modes.push_back(sm);
  uview.ext.push_back(k);
  uview.str.push_back(m);
  uview.finish(A.dtype, w + o_u);
  copy_tensor(h, U, uview.t);
  // V^H (k x n, leading dimension k): element (s, v) = conj(V[v, s]), which is
  // exactly the reference's Vt[v..., s]
  vview.modes.push_back(sm);
  vview.ext.push_back(k);
  vview.str.push_back(1);
  int64_t sv = k;
  for (int i = 0; i < V.rank - 1; ++i) {
    vview.modes.push_back(V.modes[i]);
    vview.ext.push_back(V.extents[i]);
    vview.str.push_back(sv);
    sv *= V.extents[i];
  }
  vview.finish(A.dtype, w + o_v);
  copy_tensor(h, V, vview.t);
  sview.modes.push_back(sm);
  sview.ext.push_back(k);
  sview.str.push_back(1);
  sview.finish(S.dtype, w + o_s);
  copy_tensor(h, S, sview.t);
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
  if (geqrf(A.dtype, rb, (int)m, (int)n, w + o_mat, w + o_tau) != rocblas_status_success)
    fail(MUSCLE_HIP_HIP_ERROR, "rocsolver geqrf failed to launch");
  // R (k x n): the upper trapezoid, before orgqr overwrites the matrix
  launch_upper_trapezoid((int)es, w + o_mat, w + o_r, m, k, n, h.stream);
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
  if (orgqr(A.dtype, rb, (int)m, (int)k, w + o_mat, w + o_tau) != rocblas_status_success)
    fail(MUSCLE_HIP_HIP_ERROR, "rocsolver orgqr failed to launch");
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
