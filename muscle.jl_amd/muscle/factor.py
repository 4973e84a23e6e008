"""tensor_svd_thin, tensor_svd_trunc, tensor_qr_thin and simple_update on the
HIP backend.

Front ends mirror src/Operations/tensor_svd.jl:59-68 and :160-215,
src/Operations/tensor_qr.jl:27-68 (with factorinds, src/Index.jl:41-63) and
src/Operations/simple_update.jl:16-87; the device work is muscle_hip_svd_thin
(permutation copies + the eigensolver SVD), muscle_hip_qr_thin (permutation
copies + geqrf/orgqr) and this library's binary_einsum / hadamard!.

The in-place forms tensor_svd_thin_(U, s, V, A) and tensor_qr_thin_(Q, R, A)
take the factors' row modes from the caller's tensors, bond excluded (the
reference's BackendBase forms, tensor_svd.jl:127-151 and tensor_qr.jl:70-90,
pass the bond to factorinds as well).
"""
from __future__ import annotations

import itertools

import torch

from . import _lib
from .backend import BackendHIP, DomainROCm, choose_backend, register_rule
from .errors import ArgumentError
from .ops import _descs, _device_of, _handle_for, binary_einsum, hadamard_
from .tensor import Index, Tensor, _as_index, empty_col_major

_fresh = itertools.count()


def _gensym(name: str) -> Index:
    return Index((name, next(_fresh)))


def _as_index_list(x) -> list:
    if x is None:
        return []
    if isinstance(x, (Index, str)):
        return [_as_index(x)]
    return [_as_index(i) for i in x]


def factorinds(all_inds, left, right):
    """factorinds (src/Index.jl:41-63): fill the empty side with the rest."""
    left, right = _as_index_list(left), _as_index_list(right)
    if set(left) & set(right):
        raise ArgumentError(f"left ({left}) and right {right} indices must be disjoint")
    if not left:
        left = [i for i in all_inds if i not in right]
    elif not right:
        right = [i for i in all_inds if i not in left]
    if not left or not right:
        raise ArgumentError("no right-indices left in factorization")
    if not all(i in all_inds for i in left + right):
        raise ArgumentError(f"indices must be in {all_inds}")
    return left, right


class AbsorbBehavior:
    """``abstract type AbsorbBehavior`` (src/Operations/simple_update.jl:9)."""


class DontAbsorb(AbsorbBehavior):
    pass


class AbsorbU(AbsorbBehavior):
    pass


class AbsorbV(AbsorbBehavior):
    pass


class AbsorbEqually(AbsorbBehavior):
    pass


register_rule("tensor_svd_thin", (DomainROCm,), BackendHIP())
register_rule("tensor_svd_thin_", (DomainROCm, DomainROCm, DomainROCm, DomainROCm), BackendHIP())
register_rule("tensor_svd_trunc", (DomainROCm,), BackendHIP())
register_rule("tensor_qr_thin", (DomainROCm,), BackendHIP())
register_rule("tensor_qr_thin_", (DomainROCm, DomainROCm, DomainROCm), BackendHIP())
register_rule("simple_update", (DomainROCm, DomainROCm, DomainROCm), BackendHIP())

_REAL = {torch.complex128: torch.float64, torch.complex64: torch.float32, torch.float64: torch.float64,
         torch.float32: torch.float32}


def hip_tensor_svd_thin(backend: BackendHIP, A: Tensor, inds_u=(), inds_v=(), ind_s=None):
    ind_s = _gensym("svd") if ind_s is None else _as_index(ind_s)
    if not inds_u and not inds_v:
        raise ArgumentError("either `inds_u` or `inds_v` must be specified")
    inds_u, inds_v = factorinds(A.inds(), inds_u, inds_v)
    if set(inds_u) & set(inds_v) or set(inds_u) | set(inds_v) != set(A.inds()):
        raise ArgumentError("`inds_u` and `inds_v` must split the indices of A")
    if ind_s in A.inds():
        raise ArgumentError(f"the bond index {ind_s} is already an index of A")
    su = [A.size(i) for i in inds_u]
    sv = [A.size(i) for i in inds_v]
    k = min(_prod(su), _prod(sv))
    dev = A.device
    U = Tensor(empty_col_major(su + [k], A.dtype, dev), inds_u + [ind_s])
    S = Tensor(empty_col_major([k], _REAL[A.dtype], dev), [ind_s])
    V = Tensor(empty_col_major(sv + [k], A.dtype, dev), inds_v + [ind_s])
    return hip_tensor_svd_thin_(backend, U, S, V, A)


def _prod(xs) -> int:
    out = 1
    for e in xs:
        out *= int(e)
    return out


def _hip_only(op: str, *arrays) -> BackendHIP:
    backend = choose_backend(op, *arrays)
    if not isinstance(backend, BackendHIP):
        raise ArgumentError(f"`{op}` not implemented or not loaded for backend {backend!r}")
    return backend


def hip_tensor_svd_thin_(backend: BackendHIP, U: Tensor, S: Tensor, V: Tensor, A: Tensor):
    """U (u..., s), S (s), V (v..., s) from A, into caller-owned tensors."""
    h = _handle_for(_device_of(U, S, V, A))
    du, ds, dv, da = _descs(U, S, V, A)
    _lib.check(_lib.lib.muscle_hip_svd_thin(h, du.ref, ds.ref, dv.ref, da.ref))
    return U, S, V


def tensor_svd_thin_(U: Tensor, S: Tensor, V: Tensor, A: Tensor):
    """``tensor_svd_thin!(U, s, V, A)`` (src/Operations/tensor_svd.jl:127-151)."""
    backend = _hip_only("tensor_svd_thin_", U.parent(), S.parent(), V.parent(), A.parent())
    return hip_tensor_svd_thin_(backend, U, S, V, A)


def tensor_svd_thin(A: Tensor, inds_u=(), inds_v=(), ind_s=None):
    """tensor_svd_thin(A; inds_u, inds_v, ind_s) (src/Operations/tensor_svd.jl:59-68, :100-125):
    U (inds_u..., ind_s), s (ind_s), Vt (inds_v..., ind_s) with A = U * s * Vt."""
    backend = _hip_only("tensor_svd_thin", A.parent())
    return hip_tensor_svd_thin(backend, A, inds_u, inds_v, ind_s)


def hip_tensor_svd_trunc(backend: BackendHIP, A: Tensor, inds_u=(), inds_v=(), ind_s=None, threshold=None,
                         maxdim=None):
    U, S, V = hip_tensor_svd_thin(backend, A, inds_u, inds_v, ind_s)
    ind_s = S.inds()[0]
    s = S.parent().detach().cpu()  # the cut is chosen on the host, as the reference does
    k = s.shape[0]
    if maxdim is not None:
        k = min(k, int(maxdim))
    if threshold is not None:
        cut = float(torch.linalg.vector_norm(s)) * float(threshold)
        below = torch.nonzero(s[:k] < cut)
        if below.numel():
            k = int(below[0, 0]) + 1  # findfirst: the first value below the cut is kept
    return _view_bond(U, ind_s, k), _view_bond(S, ind_s, k), _view_bond(V, ind_s, k)


def tensor_svd_trunc(A: Tensor, inds_u=(), inds_v=(), ind_s=None, threshold=None, maxdim=None):
    """tensor_svd_trunc(A; inds_u, inds_v, ind_s, threshold, maxdim) (src/Operations/tensor_svd.jl:160-215):
    the thin SVD cut to at most ``maxdim`` values and at the first value below ``threshold * norm(s)``;
    U, s and Vt are views of the thin factors."""
    backend = _hip_only("tensor_svd_trunc", A.parent())
    return hip_tensor_svd_trunc(backend, A, inds_u, inds_v, ind_s, threshold, maxdim)


def hip_tensor_qr_thin_(backend: BackendHIP, Q: Tensor, R: Tensor, A: Tensor):
    """Q (q..., v), R (v, r...) from A, into caller-owned tensors."""
    h = _handle_for(_device_of(Q, R, A))
    dq, dr, da = _descs(Q, R, A)
    _lib.check(_lib.lib.muscle_hip_qr_thin(h, dq.ref, dr.ref, da.ref))
    return Q, R


def tensor_qr_thin_(Q: Tensor, R: Tensor, A: Tensor):
    """``tensor_qr_thin!(Q, R, A)`` (src/Operations/tensor_qr.jl:70-90)."""
    backend = _hip_only("tensor_qr_thin_", Q.parent(), R.parent(), A.parent())
    return hip_tensor_qr_thin_(backend, Q, R, A)


def hip_tensor_qr_thin(backend: BackendHIP, A: Tensor, inds_q=(), inds_r=(), ind_virtual=None):
    ind_virtual = _gensym("qr") if ind_virtual is None else _as_index(ind_virtual)
    if ind_virtual in A.inds():
        raise ArgumentError(f"new virtual bond name ({ind_virtual}) cannot be already be present")
    inds_q, inds_r = factorinds(A.inds(), inds_q, inds_r)
    if set(inds_q) | set(inds_r) != set(A.inds()):
        raise ArgumentError("`inds_q` and `inds_r` must cover the indices of A")
    sq = [A.size(i) for i in inds_q]
    sr = [A.size(i) for i in inds_r]
    k = min(_prod(sq), _prod(sr))
    Q = Tensor(empty_col_major(sq + [k], A.dtype, A.device), inds_q + [ind_virtual])
    R = Tensor(empty_col_major([k] + sr, A.dtype, A.device), [ind_virtual] + inds_r)
    return hip_tensor_qr_thin_(backend, Q, R, A)


def tensor_qr_thin(A: Tensor, inds_q=(), inds_r=(), ind_virtual=None):
    """tensor_qr_thin(A; inds_q, inds_r, ind_virtual) (src/Operations/tensor_qr.jl:27-68):
    Q (inds_q..., ind_virtual) with orthonormal columns and R (ind_virtual, inds_r...) upper
    trapezoidal, A = Q * R."""
    backend = _hip_only("tensor_qr_thin", A.parent())
    return hip_tensor_qr_thin(backend, A, inds_q, inds_r, ind_virtual)


def _view_bond(t: Tensor, ind: Index, n: int) -> Tensor:
    """``view(t, ind => 1:n)`` (src/Tensor.jl:334-356)."""
    return Tensor(t.parent().narrow(t.dim(ind), 0, n), t.inds())


def _replace(t: Tensor, old_new: dict) -> Tensor:
    """``replace(t, old => new...)`` (src/Tensor.jl:137)."""
    return Tensor(t.parent(), [old_new.get(i, i) for i in t.inds()])


def hip_simple_update(backend: BackendHIP, A, ind_pa, B, ind_pb, ind_bond, G, ind_pga, ind_pgb, normalize=False,
                      absorb=None, maxdim=None):
    absorb = absorb or DontAbsorb()
    ind_pa, ind_pb, ind_bond = _as_index(ind_pa), _as_index(ind_pb), _as_index(ind_bond)
    ind_pga, ind_pgb = _as_index(ind_pga), _as_index(ind_pgb)
    theta = binary_einsum(binary_einsum(A, B, dims=[ind_bond]), G, dims=[ind_pa, ind_pb])
    theta = _replace(theta, {ind_pga: ind_pa, ind_pgb: ind_pb})
    inds_u = [i for i in A.inds() if i != ind_bond]
    inds_v = [i for i in B.inds() if i != ind_bond]
    U, S, V = hip_tensor_svd_thin(backend, theta, inds_u, inds_v, ind_bond)
    if maxdim is not None:
        k = min(int(maxdim), S.size(ind_bond))
        U, S, V = _view_bond(U, ind_bond, k), _view_bond(S, ind_bond, k), _view_bond(V, ind_bond, k)
    if normalize:
        s = S.parent()
        s.div_(torch.linalg.vector_norm(s))  # normalize!(S): k singular values, device-side
    if isinstance(absorb, DontAbsorb):
        return U, S, V
    if isinstance(absorb, AbsorbU):
        U = hadamard_(U, U, S)
    elif isinstance(absorb, AbsorbV):
        V = hadamard_(V, V, S)
    elif isinstance(absorb, AbsorbEqually):
        s_sqrt = Tensor(torch.sqrt(S.parent()), S.inds())
        U = hadamard_(U, U, s_sqrt)
        V = hadamard_(V, V, s_sqrt)
    else:
        raise ArgumentError(f"unknown absorb behavior {absorb!r}")
    return U, V


def simple_update(A: Tensor, ind_physical_a, B: Tensor, ind_physical_b, ind_bond_ab, G: Tensor, ind_physical_g_a,
                  ind_physical_g_b, normalize=False, absorb=None, maxdim=None):
    """simple_update (src/Operations/simple_update.jl:16-87): apply the two-site gate G to the bond
    A - B; returns (U, s, V), or (U, V) with s absorbed (AbsorbU / AbsorbV / AbsorbEqually)."""
    backend = _hip_only("simple_update", A.parent(), B.parent(), G.parent())
    return hip_simple_update(backend, A, ind_physical_a, B, ind_physical_b, ind_bond_ab, G, ind_physical_g_a,
                             ind_physical_g_b, normalize=normalize, absorb=absorb, maxdim=maxdim)
