"""CPU oracle for Muscle.jl's contraction hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the ``muscle`` package or
libmuscle_hip.so) imports, links or calls this module; only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg use it, as
the checker and as the timed CPU baseline.

It restates, in numpy, the reference's in-core backend ``BackendBase``
(permutedims + BLAS gemm + permutedims) and the index algebra of the front
ends.  Arrays are indexed in index-label order; Julia's column-major
``reshape`` is reproduced with ``order="F"``.  Every function cites the
reference lines it follows (paths relative to the Muscle.jl source tree).

How each function is pinned (tests/test_oracle.py):
  * front_binary / front_unary: the expected ``inds``/``size`` tables of
    test/unit/operations/binary_einsum.jl:4-155 and
    test/integration/omeinsum.jl:6-100, ported as known-answer tests.
  * binary_einsum_base: exact closed forms on ``ones`` (K, 12, 1, 2*A,
    binary_einsum.jl:4-119), the textbook identity of :135-149, the
    brute-force nested loop of test/integration/cuda.jl:169-180, and the C
    long-double brute-force einsum (oracle/brute.c) on seeded random inputs.
  * binary_einsum_general (hyperindices): ``3*ones`` batch KATs
    (test/integration/cuda.jl:126-145, omeinsum.jl:188-207), the brute-force
    loop of omeinsum.jl:225-236 and oracle/brute.c.
  * unary_einsum_base: closed forms of test/integration/omeinsum.jl:6-100
    (axis sum, full sum, diagonal, trace) and oracle/brute.c.
  * hadamard_base: closed forms of test/unit/operations/hadamard.jl:4-119 and
    oracle/brute.c, including 3-cycle orders of ``b`` where the reference's
    own ``findperm`` use is wrong (hadamard.jl:68-69).
No function is left "parity unpinned".
"""
from __future__ import annotations

import math
from typing import Sequence

import numpy as np


class ArgumentError(ValueError):
    """Julia's ArgumentError (raised by @argcheck in the reference)."""


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch."""


# ---------------------------------------------------------------- index algebra
def unique(xs: Sequence) -> list:
    """Julia ``unique``: first occurrences, order preserved."""
    out = []
    for x in xs:
        if x not in out:
            out.append(x)
    return out


def intersect(x: Sequence, *ys: Sequence) -> list:
    """Julia ``∩(x, ys...)``: unique elements of ``x`` present in every ``y``, in x-order."""
    return [e for e in unique(x) if all(e in y for y in ys)]


def union(x: Sequence, y: Sequence) -> list:
    """Julia ``x ∪ y``: unique elements, x-order then new elements of y."""
    return unique(list(x) + list(y))


def setdiff(x: Sequence, y: Sequence) -> list:
    """Julia ``setdiff(x, y)``: unique elements of ``x`` not in ``y``, x-order."""
    return [e for e in unique(x) if e not in y]


def symdiff(x: Sequence, y: Sequence) -> list:
    return [e for e in unique(x) if e not in y] + [e for e in unique(y) if e not in x]


def nonunique(xs: Sequence) -> list:
    """Repeated elements in order of their second occurrence (src/Utils/Utils.jl:5-25)."""
    seen, repeated = [], []
    for x in xs:
        if x in seen:
            if x not in repeated:
                repeated.append(x)
        else:
            seen.append(x)
    return repeated


def front_binary(inds_a, inds_b, dims=None, out=None) -> list:
    """Output indices chosen by binary_einsum(a, b; dims, out) (binary_einsum.jl:33-40)."""
    if dims is None:
        dims = intersect(inds_a, inds_b)
    inds_sum = intersect(dims, inds_a, inds_b)
    if out is None:
        return setdiff(union(inds_a, inds_b), inds_sum)
    return list(out)


def front_unary(inds_x, dims=None, out=None) -> list:
    """Output indices chosen by unary_einsum(x; dims, out) (unary_einsum.jl:26-32)."""
    if dims is None:
        dims = nonunique(inds_x)
    inds_sum = intersect(dims, inds_x)
    if out is None:
        return setdiff(inds_x, inds_sum)
    return list(out)


# ---------------------------------------------------------------- array helpers
def permutedims(x: np.ndarray, inds_x: Sequence, inds_to: Sequence) -> np.ndarray:
    """permutedims by Index: position of each target index by ``findfirst`` (Tensor.jl:316-319)."""
    if x.ndim == 0:
        return x
    perm = [list(inds_x).index(i) for i in inds_to]
    return np.transpose(x, perm)


def reshape_f(x: np.ndarray, shape) -> np.ndarray:
    """Julia ``reshape`` (column-major)."""
    return np.reshape(x, tuple(int(s) for s in shape), order="F")


def prod(xs) -> int:
    return int(math.prod(int(v) for v in xs))


def _sizes(x: np.ndarray, inds_x, which):
    return [x.shape[list(inds_x).index(i)] for i in which]


# ---------------------------------------------------------------- BackendBase
def binary_einsum_base(inds_c, a: np.ndarray, inds_a, b: np.ndarray, inds_b) -> np.ndarray:
    """binary_einsum(::BackendBase, inds_c, a, b) (src/Operations/binary_einsum.jl:76-96).

    permute A to [left; contract], B to [contract; right], one gemm, reshape,
    permute to inds_c.  Rejects hyperindices exactly as :82-83 does.
    """
    inds_a, inds_b, inds_c = list(inds_a), list(inds_b), list(inds_c)
    contract = intersect(inds_a, inds_b)
    left = setdiff(inds_a, contract)
    right = setdiff(inds_b, contract)
    if any(i in contract for i in inds_c):
        raise ArgumentError("`BackendBase` can't deal with hyperindices")
    if set(inds_c) != set(symdiff(inds_a, inds_b)):
        raise ArgumentError("`BackendBase` can't deal with hyperindices")
    sl = _sizes(a, inds_a, left)
    sr = _sizes(b, inds_b, right)
    sc = _sizes(a, inds_a, contract)
    a_mat = reshape_f(permutedims(a, inds_a, left + contract), (prod(sl), prod(sc)))
    b_mat = reshape_f(permutedims(b, inds_b, contract + right), (prod(sc), prod(sr)))
    c_mat = a_mat @ b_mat
    c = reshape_f(c_mat, sl + sr)
    return permutedims(c, left + right, inds_c)


def binary_einsum_base_inplace(inds_c, a, inds_a, b, inds_b) -> np.ndarray:
    """binary_einsum!(::BackendBase, c, a, b) (binary_einsum.jl:98-121): inds(c) must be [left; right]."""
    contract = intersect(inds_a, inds_b)
    left = setdiff(inds_a, contract)
    right = setdiff(inds_b, contract)
    if list(inds_c) != left + right:
        raise ArgumentError("inds(c) == [inds_left; inds_right] must hold")
    return binary_einsum_base(inds_c, a, inds_a, b, inds_b)


# ---------------------------------------------------------------- hyperindex extension
def binary_einsum_general(inds_c, a: np.ndarray, inds_a, b: np.ndarray, inds_b) -> np.ndarray:
    """General binary contraction with batch (hyper) modes.

    Semantics of ext/MuscleCUDAExt.jl:22-41 / ext/MuscleOMEinsumExt.jl:40-59,
    lowered as ext/MuscleReactantExt.jl:94-122 does (batching dims =
    a∩b∩c, contracting dims = (a∩b)∖c) and computed the BackendBase way
    (permute + gemm), looped over the batch.
    Repeated modes inside an operand and modes of one operand absent from
    everything else are first reduced with unary_einsum_base (OMEinsum's
    unary pre-pass).
    """
    inds_a, inds_b, inds_c = list(inds_a), list(inds_b), list(inds_c)
    if len(unique(inds_c)) != len(inds_c):
        raise ArgumentError("indices in `out` for a binary einsum must be unique")
    if any(i not in inds_a and i not in inds_b for i in inds_c):
        raise ArgumentError("`out` must be a subset of the union of the indices")
    keep_a = [i for i in unique(inds_a) if i in inds_b or i in inds_c]
    if keep_a != inds_a:
        a = unary_einsum_base(keep_a, a, inds_a)
        inds_a = keep_a
    keep_b = [i for i in unique(inds_b) if i in inds_a or i in inds_c]
    if keep_b != inds_b:
        b = unary_einsum_base(keep_b, b, inds_b)
        inds_b = keep_b
    batch = [i for i in inds_a if i in inds_b and i in inds_c]
    contract = [i for i in inds_a if i in inds_b and i not in inds_c]
    left = [i for i in inds_a if i not in inds_b]
    right = [i for i in inds_b if i not in inds_a]
    sl, sk, sr = _sizes(a, inds_a, left), _sizes(a, inds_a, contract), _sizes(b, inds_b, right)
    sb = _sizes(a, inds_a, batch)
    if _sizes(b, inds_b, contract) != sk or _sizes(b, inds_b, batch) != sb:
        raise DimensionMismatch("shared index sizes differ")
    M, K, N, NB = prod(sl), prod(sk), prod(sr), prod(sb)
    a3 = reshape_f(permutedims(a, inds_a, left + contract + batch), (M, K, NB))
    b3 = reshape_f(permutedims(b, inds_b, contract + right + batch), (K, N, NB))
    dt = np.result_type(a.dtype, b.dtype)
    c3 = np.empty((M, N, NB), dtype=dt, order="F")
    for t in range(NB):
        c3[:, :, t] = a3[:, :, t] @ b3[:, :, t]
    c = reshape_f(c3, sl + sr + sb)
    return permutedims(c, left + right + batch, inds_c)


# ---------------------------------------------------------------- unary
def unary_einsum_base(inds_y, x: np.ndarray, inds_x) -> np.ndarray:
    """OMEinsum's unary rule x -> y (ext/MuscleOMEinsumExt.jl:31-37).

    Repeated modes of x are tied (the strided diagonal: one axis whose stride
    is the sum of the repeated axes' strides), modes absent from y are summed,
    and the rest is permuted to y's order.
    """
    inds_x, inds_y = list(inds_x), list(inds_y)
    if not all(i in inds_x for i in inds_y):
        raise ArgumentError("Output indices must be a subset of input indices")
    if len(unique(inds_y)) != len(inds_y):
        raise ArgumentError("repeated output indices are not supported")
    xc = np.ascontiguousarray(x)
    uniq = unique(inds_x)
    shape, strides = [], []
    for lab in uniq:
        axes = [p for p, l in enumerate(inds_x) if l == lab]
        ext = {xc.shape[p] for p in axes}
        if len(ext) != 1:
            raise DimensionMismatch(f"repeated index {lab!r} has different sizes")
        shape.append(ext.pop())
        strides.append(sum(xc.strides[p] for p in axes))
    diag = np.lib.stride_tricks.as_strided(xc, shape=shape, strides=strides, writeable=False)
    sum_axes = tuple(p for p, lab in enumerate(uniq) if lab not in inds_y)
    s = diag.sum(axis=sum_axes) if sum_axes else np.array(diag)
    remaining = [lab for lab in uniq if lab in inds_y]
    return np.asarray(permutedims(np.asarray(s), remaining, inds_y)).astype(x.dtype, copy=False)


# ---------------------------------------------------------------- hadamard
def hadamard_base(a: np.ndarray, inds_a, b: np.ndarray, inds_b):
    """hadamard(::BackendBase, a, b) (src/Operations/hadamard.jl:34-77).

    Returns (c, inds_c).  Swaps so ndims(a) >= ndims(b) (:36), requires
    inds(b) ⊆ inds(a) (:10), output has a's inds and a's eltype (copy(a),
    :38), b is permuted to a's relative order and broadcast (:59-74).  The
    permutation used is the one that takes inds(b) to that order; the
    reference passes findperm's inverse (:68-69), which agrees only for
    involutive orders.
    """
    inds_a, inds_b = list(inds_a), list(inds_b)
    if a.ndim < b.ndim:
        a, b, inds_a, inds_b = b, a, inds_b, inds_a
    if not all(i in inds_a for i in inds_b):
        raise ArgumentError("inds(b) ⊆ inds(a) must hold")
    if np.iscomplexobj(b) and not np.iscomplexobj(a):
        raise ArgumentError("InexactError: complex b into real a")
    if b.ndim == 0:
        return (a * b).astype(a.dtype), inds_a
    shape_b = [b.shape[inds_b.index(i)] if i in inds_b else 1 for i in inds_a]
    order = [i for i in inds_a if i in inds_b]
    bb = permutedims(b, inds_b, order) if b.ndim > 1 else b
    bb = reshape_f(bb, shape_b)
    return (a * bb).astype(a.dtype), inds_a
