/*
 * Brute-force einsum in long double — TEST INFRASTRUCTURE ONLY (the oracle's
 * independent check; never linked into libmuscle_hip).
 *
 *   C[c-modes] = sum over every other label of A[a-modes] * B[b-modes]
 *
 * by walking the full index space of all distinct labels, exactly the nested
 * loop the reference's own tests use as ground truth
 * (test/integration/cuda.jl:169-180, test/integration/omeinsum.jl:225-236:
 *  C[i,k,l] += A[i,j,k] * B[k,l,j]).  A label repeated inside one operand
 * simply reads that operand on the diagonal.  With has_b == 0 the product
 * factor is 1 (unary einsum).  Inputs and output are complex double,
 * interleaved (re, im), addressed by element strides; accumulation is in
 * 80-bit long double, which is why it can pin the FP64 oracle on small cases.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int brute_einsum(int nlab, const int64_t* ext,
                 int na, const int32_t* la, const int64_t* sa, const double* a,
                 int has_b, int nb, const int32_t* lb, const int64_t* sb, const double* b,
                 int nc, const int32_t* lc, const int64_t* sc, double* c) {
  /* per-label stride contribution of each operand (sums for repeated labels) */
  int64_t* wa = (int64_t*)calloc((size_t)nlab + 1, sizeof(int64_t));
  int64_t* wb = (int64_t*)calloc((size_t)nlab + 1, sizeof(int64_t));
  int64_t* wc = (int64_t*)calloc((size_t)nlab + 1, sizeof(int64_t));
  int64_t* idx = (int64_t*)calloc((size_t)nlab + 1, sizeof(int64_t));
  int64_t csize = 1, total = 1, i;
  long double *acc_re, *acc_im;
  int d;
  if (!wa || !wb || !wc || !idx) return 1;
  for (d = 0; d < na; ++d) wa[la[d]] += sa[d];
  if (has_b)
    for (d = 0; d < nb; ++d) wb[lb[d]] += sb[d];
  for (d = 0; d < nc; ++d) wc[lc[d]] += sc[d];
  /* the output span: offsets are assumed non-negative and dense enough */
  {
    int64_t hi = 0;
    for (d = 0; d < nc; ++d) hi += (ext[lc[d]] - 1) * sc[d];
    csize = nc ? hi + 1 : 1;
    for (d = 0; d < nc; ++d)
      if (ext[lc[d]] == 0) csize = 0;
  }
  for (d = 0; d < nlab; ++d) total *= ext[d];
  acc_re = (long double*)calloc((size_t)(csize > 0 ? csize : 1), sizeof(long double));
  acc_im = (long double*)calloc((size_t)(csize > 0 ? csize : 1), sizeof(long double));
  if (!acc_re || !acc_im) return 1;
  for (i = 0; i < total; ++i) {
    int64_t oa = 0, ob = 0, oc = 0;
    long double xr, xi, yr = 1.0L, yi = 0.0L;
    for (d = 0; d < nlab; ++d) {
      oa += idx[d] * wa[d];
      ob += idx[d] * wb[d];
      oc += idx[d] * wc[d];
    }
    xr = a[2 * oa];
    xi = a[2 * oa + 1];
    if (has_b) {
      yr = b[2 * ob];
      yi = b[2 * ob + 1];
    }
    acc_re[oc] += xr * yr - xi * yi;
    acc_im[oc] += xr * yi + xi * yr;
    for (d = 0; d < nlab; ++d) {
      if (++idx[d] < ext[d]) break;
      idx[d] = 0;
    }
  }
  for (i = 0; i < csize; ++i) {
    c[2 * i] = (double)acc_re[i];
    c[2 * i + 1] = (double)acc_im[i];
  }
  free(wa);
  free(wb);
  free(wc);
  free(idx);
  free(acc_re);
  free(acc_im);
  return 0;
}
