"""CPU oracle for the factorization path (TEST INFRASTRUCTURE ONLY: used by
tests/ and bench.py's cpu_baseline leg, never by the product).

Restates BackendBase's tensor_svd_thin (src/Operations/tensor_svd.jl:100-125:
permutedims to (inds_u..., inds_v...), column-major reshape, LAPACK SVD,
U and conj(V) reshaped back), tensor_svd_trunc (:160-215), tensor_qr_thin
(src/Operations/tensor_qr.jl:45-68) and simple_update
(src/Operations/simple_update.jl:37-87) in numpy.  Labels are plain hashable
values.  numpy.linalg.svd and Julia's LinearAlgebra.svd both call LAPACK's
divide-and-conquer gesdd.

How it is pinned (tests/test_factor_oracle.py): the known answers of the
reference's test/unit/operations/simple_update.jl (identity and CNOT gates on
the |00>+|01>+|10>+|11> MPS, normalized, truncated to one, absorbed) and
tensor_svd_thin.jl (index and shape rules, U s Vt = A, isometric U and Vt),
tensor_svd_trunc.jl (the cut rules, discarded weight = reconstruction error)
and tensor_qr_thin.jl (index and shape rules, Q R = A, isometric Q), plus
singular values against the eigenvalues of A^H A, exact rank-1 tensors, and
R against the Cholesky factor of A^H A (R^H R = A^H A, |diag R| fixed).
"""
from __future__ import annotations

import numpy as np

from .backend_base import ArgumentError, binary_einsum_general, permutedims, reshape_f


def factorinds(all_inds, left, right):
    """src/Index.jl:41-63."""
    left, right = list(left), list(right)
    if set(left) & set(right):
        raise ArgumentError("left and right indices must be disjoint")
    if not left:
        left = [i for i in all_inds if i not in right]
    elif not right:
        right = [i for i in all_inds if i not in left]
    if not left or not right:
        raise ArgumentError("no right-indices left in factorization")
    if not all(i in all_inds for i in left + right):
        raise ArgumentError("indices must be in the tensor's")
    return left, right


def tensor_svd_thin_base(A: np.ndarray, inds_a, inds_u=(), inds_v=(), ind_s="svd"):
    """(U, inds_u + [s]), (s, [s]), (Vt, inds_v + [s]) of src/Operations/tensor_svd.jl:100-125."""
    inds_a = list(inds_a)
    inds_u, inds_v = factorinds(inds_a, inds_u, inds_v)
    if set(inds_u) & set(inds_v) or set(inds_u) | set(inds_v) != set(inds_a):
        raise ArgumentError("inds_u and inds_v must split the indices")
    if ind_s in inds_a:
        raise ArgumentError("the bond index is already an index of A")
    left = [A.shape[inds_a.index(i)] for i in inds_u]
    right = [A.shape[inds_a.index(i)] for i in inds_v]
    Amat = reshape_f(permutedims(A, inds_a, inds_u + inds_v), (int(np.prod(left)), int(np.prod(right))))
    U, s, Vh = np.linalg.svd(Amat, full_matrices=False)  # Julia: U, s, V with A = U S V'
    k = s.shape[0]
    Ut = reshape_f(U, tuple(left) + (k,))
    Vt = reshape_f(Vh.T, tuple(right) + (k,))  # conj(V) with V = Vh^H is Vh^T
    return (Ut, inds_u + [ind_s]), (s, [ind_s]), (Vt, inds_v + [ind_s])


def tensor_svd_trunc_base(A: np.ndarray, inds_a, inds_u=(), inds_v=(), ind_s="svd", threshold=None, maxdim=None):
    """src/Operations/tensor_svd.jl:160-215: cut at maxdim, then at the first value below
    threshold * norm(s) (that value is kept: findfirst gives the last index of `keep`)."""
    (U, iu), (s, is_), (Vt, iv) = tensor_svd_thin_base(A, inds_a, inds_u, inds_v, ind_s)
    k = s.shape[0]
    if maxdim is not None:
        k = min(k, maxdim)
    if threshold is not None:
        cut = np.linalg.norm(s) * threshold
        below = np.nonzero(s[:k] < cut)[0]
        if below.size:
            k = int(below[0]) + 1
    return (U[..., :k], iu), (s[:k], is_), (Vt[..., :k], iv)


def tensor_qr_thin_base(A: np.ndarray, inds_a, inds_q=(), inds_r=(), ind_virtual="qr"):
    """(Q, inds_q + [v]), (R, [v] + inds_r) of src/Operations/tensor_qr.jl:45-68.  numpy's qr is
    LAPACK geqrf + orgqr/ungqr; Julia's qr is geqrt, the same Householder reflectors and R."""
    inds_a = list(inds_a)
    if ind_virtual in inds_a:
        raise ArgumentError("the virtual bond is already an index of A")
    inds_q, inds_r = factorinds(inds_a, inds_q, inds_r)
    if set(inds_q) | set(inds_r) != set(inds_a):
        raise ArgumentError("inds_q and inds_r must cover the indices")
    left = [A.shape[inds_a.index(i)] for i in inds_q]
    right = [A.shape[inds_a.index(i)] for i in inds_r]
    Amat = reshape_f(permutedims(A, inds_a, inds_q + inds_r), (int(np.prod(left)), int(np.prod(right))))
    Q, R = np.linalg.qr(Amat, mode="reduced")
    k = Q.shape[1]
    return (reshape_f(Q, tuple(left) + (k,)), inds_q + [ind_virtual]), \
        (reshape_f(R, (k,) + tuple(right)), [ind_virtual] + inds_r)


def simple_update_base(A, ia, ind_pa, B, ib, ind_pb, ind_bond, G, ig, ind_pga, ind_pgb, normalize=False,
                       absorb="none", maxdim=None):
    """src/Operations/simple_update.jl:37-87 (absorb: none | u | v | equally)."""
    ab_inds = [i for i in ia if i != ind_bond] + [i for i in ib if i != ind_bond]
    AB = binary_einsum_general(ab_inds, A, ia, B, ib)
    th_inds = [i for i in ab_inds if i not in (ind_pa, ind_pb)] + [i for i in ig if i not in (ind_pa, ind_pb)]
    theta = binary_einsum_general(th_inds, AB, ab_inds, G, ig)
    th_inds = [ind_pa if i == ind_pga else ind_pb if i == ind_pgb else i for i in th_inds]
    inds_u = [i for i in ia if i != ind_bond]
    inds_v = [i for i in ib if i != ind_bond]
    (U, iu), (s, is_), (V, iv) = tensor_svd_thin_base(theta, th_inds, inds_u, inds_v, ind_bond)
    if maxdim is not None:
        k = min(int(maxdim), s.shape[0])
        U, s, V = U[..., :k], s[:k], V[..., :k]
    if normalize:
        s = s / np.linalg.norm(s)
    if absorb == "none":
        return (U, iu), (s, is_), (V, iv)
    if absorb == "u":
        U = U * s
    elif absorb == "v":
        V = V * s
    elif absorb == "equally":
        U, V = U * np.sqrt(s), V * np.sqrt(s)
    return (U, iu), (V, iv)
