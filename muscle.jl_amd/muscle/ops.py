"""The hot-path operations: front ends (index algebra + dispatch) and the
BackendHIP methods, which call libmuscle_hip through its C ABI.

Front ends follow src/Operations/binary_einsum.jl:33-70, hadamard.jl:7-32 and
unary_einsum.jl:26-50.  Julia's ``f!`` functions are spelled ``f_`` here.
"""
from __future__ import annotations

import ctypes
import threading

import torch

from . import _lib
from .backend import Backend, BackendHIP, choose_backend
from .errors import ArgumentError
from .tensor import Index, Tensor, empty_col_major

# ---------------------------------------------------------------- index algebra
# Julia's set operations on index vectors are order-preserving and unique.


def _unique(xs):
    out = []
    for x in xs:
        if x not in out:
            out.append(x)
    return out


def _intersect(x, *ys):
    return [e for e in _unique(x) if all(e in y for y in ys)]


def _setdiff(x, y):
    return [e for e in _unique(x) if e not in y]


def nonunique(xs) -> list:
    """Repeated elements, in order of their second occurrence (src/Utils/Utils.jl:5-25)."""
    seen, repeated = [], []
    for x in xs:
        if x in seen:
            if x not in repeated:
                repeated.append(x)
        else:
            seen.append(x)
    return repeated


def _idx_list(xs):
    if xs is None:
        return None
    if isinstance(xs, (Index, str)):
        xs = [xs]
    return [x if isinstance(x, Index) else Index(x) for x in xs]


# ---------------------------------------------------------------- per-thread call info
class _Last(threading.local):
    info: dict | None = None


_last = _Last()


def last_info() -> dict | None:
    """What the library reported for this thread's last call (kernel, M/N/K/batch, flops, bytes, ...)."""
    return _last.info


# ---------------------------------------------------------------- C ABI plumbing
def _device_of(*ts: Tensor) -> int:
    devs = {t.device for t in ts}
    if len(devs) != 1:
        raise ArgumentError(f"all tensors must be on one device, got {devs}")
    dev = devs.pop()
    if dev.type != "cuda":
        raise ArgumentError("BackendHIP operates on device (ROCm) arrays")
    return dev.index if dev.index is not None else torch.cuda.current_device()


def _descs(*ts: Tensor):
    labels = {}
    for t in ts:
        for i in t._inds:
            labels.setdefault(i, len(labels) + 1)
    out = []
    for t in ts:
        d = t._data
        code = _lib.DTYPE_CODE.get(d.dtype)
        if code is None:
            raise ArgumentError(f"unsupported eltype {d.dtype}")
        out.append(_lib.Desc(code, d.data_ptr(), [labels[i] for i in t._inds], d.shape, d.stride()))
    return out


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


_slots = threading.local()


def _call_descs(*ts: Tensor):
    """Descriptors for one synchronous library call, filled into this thread's
    reusable slots (valid until the thread's next call; use _descs for
    descriptors that must outlive it)."""
    pool = getattr(_slots, "pool", None)
    if pool is None:
        pool = _slots.pool = [_lib.DescSlot() for _ in range(4)]
    if len(ts) > len(pool) or any(len(t._inds) > 64 for t in ts):
        return _descs(*ts)
    labels = {}
    for t in ts:
        for i in t._inds:
            labels.setdefault(i, len(labels) + 1)
    out = []
    for t, slot in zip(ts, pool):
        d = t._data
        code = _lib.DTYPE_CODE.get(d.dtype)
        if code is None:
            raise ArgumentError(f"unsupported eltype {d.dtype}")
        out.append(slot.fill(code, d.data_ptr(), [labels[i] for i in t._inds], d.shape, d.stride()))
    return out


def _handle_for(dev: int):
    """The thread's handle for `dev` (an index or a torch.device), pointed at
    torch's current stream there."""
    if isinstance(dev, torch.device):
        dev = dev.index if dev.index is not None else torch.cuda.current_device()
    h = _lib.handle(dev)
    stream = _raw_stream(dev) if _raw_stream is not None else torch.cuda.current_stream(dev).cuda_stream
    _lib.set_stream(dev, h, stream)
    return h


def measure_bandwidth(gib: float = 4.0, device=None) -> dict:
    """HBM rates (TB/s) this library's kernels reach on `device`: 16-byte
    stores, then loads, streamed over a fresh buffer of `gib` GiB."""
    dev = torch.cuda.current_device() if device is None else device
    h = _handle_for(dev)
    r, w = ctypes.c_double(), ctypes.c_double()
    _lib.check(_lib.lib.muscle_hip_measure_bandwidth(h, int(gib * 2 ** 30), ctypes.byref(r), ctypes.byref(w)))
    return {"read_tbs": r.value, "write_tbs": w.value, "buffer_gib": gib}


def _record(info: _lib.Info) -> None:
    _last.info = info.as_dict()


# ---------------------------------------------------------------- BackendHIP methods
def hip_binary_einsum_(backend: BackendHIP, c: Tensor, a: Tensor, b: Tensor) -> Tensor:
    dev = _device_of(c, a, b)
    h = _handle_for(dev)
    dc, da, db = _call_descs(c, a, b)
    info = _lib.Info()
    prec = _lib.PRECISION.get(backend.precision)
    if prec is None:
        raise ArgumentError(f"unknown precision {backend.precision!r}")
    _lib.set_option(dev, h, 0, int(backend.complex_mult))
    _lib.check(_lib.lib.muscle_hip_binary_einsum(h, dc.ref, da.ref, db.ref, prec, float(backend.target_rel_error),
                                                 ctypes.byref(info)))
    _record(info)
    return c


def hip_binary_einsum(backend: BackendHIP, inds_c, a: Tensor, b: Tensor) -> Tensor:
    """binary_einsum(::BackendHIP, inds_c, a, b): allocate C with the promoted eltype
    (as ext/MuscleCUDAExt.jl:10-20 does) and contract into it."""
    size = {}
    for t in (a, b):
        for i, n in zip(t.inds(), t.shape):
            size.setdefault(i, n)
    inds_c = _idx_list(inds_c)
    for i in inds_c:
        if i not in size:
            raise ArgumentError(f"output index {i} is in neither input")
    T = torch.promote_types(a.dtype, b.dtype)
    c = Tensor(empty_col_major([size[i] for i in inds_c], T, a.device), inds_c)
    return hip_binary_einsum_(backend, c, a, b)


def hip_hadamard_(backend: BackendHIP, c: Tensor, a: Tensor, b: Tensor) -> Tensor:
    dev = _device_of(c, a, b)
    h = _handle_for(dev)
    dc, da, db = _call_descs(c, a, b)
    info = _lib.Info()
    _lib.check(_lib.lib.muscle_hip_hadamard(h, dc.ref, da.ref, db.ref, ctypes.byref(info)))
    _record(info)
    return c


def hip_hadamard(backend: BackendHIP, a: Tensor, b: Tensor) -> Tensor:
    """hadamard(::BackendHIP, a, b): C is a fresh array with a's inds and eltype (copy(a), hadamard.jl:38)."""
    if a.ndim < b.ndim:
        a, b = b, a
    c = Tensor(empty_col_major(a.shape, a.dtype, a.device), a.inds())
    return hip_hadamard_(backend, c, a, b)


def hip_unary_einsum_(backend: BackendHIP, y: Tensor, x: Tensor) -> Tensor:
    dev = _device_of(y, x)
    h = _handle_for(dev)
    dy, dx = _call_descs(y, x)
    info = _lib.Info()
    _lib.check(_lib.lib.muscle_hip_unary_einsum(h, dy.ref, dx.ref, ctypes.byref(info)))
    _record(info)
    return y


def hip_unary_einsum(backend: BackendHIP, inds_y, x: Tensor) -> Tensor:
    """unary_einsum(::BackendHIP, inds_y, x): y = similar(x) with y's sizes (MuscleOMEinsumExt.jl:25-29)."""
    inds_y = _idx_list(inds_y)
    if not all(i in x.inds() for i in inds_y):
        raise ArgumentError("Output indices must be a subset of input indices")
    y = Tensor(empty_col_major([x.size(i) for i in inds_y], x.dtype, x.device), inds_y)
    return hip_unary_einsum_(backend, y, x)


_METHODS = {
    ("binary_einsum", BackendHIP): hip_binary_einsum,
    ("binary_einsum!", BackendHIP): hip_binary_einsum_,
    ("hadamard", BackendHIP): hip_hadamard,
    ("hadamard!", BackendHIP): hip_hadamard_,
    ("unary_einsum", BackendHIP): hip_unary_einsum,
    ("unary_einsum!", BackendHIP): hip_unary_einsum_,
}


def _method(op: str, backend: Backend):
    fn = _METHODS.get((op, type(backend)))
    if fn is None:
        # the reference's fallback methods (binary_einsum.jl:53-55,72-74, hadamard.jl:15-17,31-33,
        # unary_einsum.jl:39-41,48-50)
        raise ArgumentError(f"`{op}` not implemented or not loaded for backend {backend!r}")
    return fn


# ---------------------------------------------------------------- front ends
def binary_einsum(a: Tensor, b: Tensor, dims=None, out=None) -> Tensor:
    """binary_einsum(a, b; dims=∩(inds(a), inds(b)), out=nothing) (binary_einsum.jl:33-51)."""
    dims = _idx_list(dims)
    if dims is None:
        dims = _intersect(a.inds(), b.inds())
    inds_sum = _intersect(dims, a.inds(), b.inds())
    out = _idx_list(out)
    inds_c = _setdiff(_unique(a.inds() + b.inds()), inds_sum) if out is None else out
    backend = choose_backend("binary_einsum", a.parent(), b.parent())
    return _method("binary_einsum", backend)(backend, inds_c, a, b)


def binary_einsum_(c: Tensor, a: Tensor, b: Tensor) -> Tensor:
    """binary_einsum!(c, a, b) (binary_einsum.jl:57-70); any mode order of c is allowed."""
    backend = choose_backend("binary_einsum!", c.parent(), a.parent(), b.parent())
    _method("binary_einsum!", backend)(backend, c, a, b)
    return c


def hadamard(a: Tensor, b: Tensor) -> Tensor:
    """hadamard(a, b) (hadamard.jl:7-13): b is broadcast into a; operands swap if ndims(a) < ndims(b)."""
    if a.ndim < b.ndim:
        return hadamard(b, a)
    if not all(i in a.inds() for i in b.inds()):
        raise ArgumentError("inds(b) ⊆ inds(a) must hold")
    backend = choose_backend("hadamard", a.parent(), b.parent())
    return _method("hadamard", backend)(backend, a, b)


def hadamard_(c: Tensor, a: Tensor, b: Tensor) -> Tensor:
    """hadamard!(c, a, b) (hadamard.jl:19-28).  When ndims(a) < ndims(b) the operands are
    swapped; the reference's hadamard(c, b, a) call at :21 is a typo for that."""
    if a.ndim < b.ndim:
        return hadamard_(c, b, a)
    if c.inds() != a.inds():
        raise ArgumentError("inds(c) == inds(a) must hold")
    if not all(i in a.inds() for i in b.inds()):
        raise ArgumentError("inds(b) ⊆ inds(a) must hold")
    backend = choose_backend("hadamard!", c.parent(), a.parent(), b.parent())
    return _method("hadamard!", backend)(backend, c, a, b)


def unary_einsum(x: Tensor, dims=None, out=None) -> Tensor:
    """unary_einsum(x; dims=nonunique(inds(x)), out=nothing) (unary_einsum.jl:26-36)."""
    dims = _idx_list(dims)
    if dims is None:
        dims = nonunique(x.inds())
    inds_sum = _intersect(dims, x.inds())
    out = _idx_list(out)
    inds_y = _setdiff(x.inds(), inds_sum) if out is None else out
    backend = choose_backend("unary_einsum", x.parent())
    return _method("unary_einsum", backend)(backend, inds_y, x)


def unary_einsum_(y: Tensor, x: Tensor) -> Tensor:
    """unary_einsum!(y, x) (unary_einsum.jl:43-46)."""
    backend = choose_backend("unary_einsum!", y.parent(), x.parent())
    _method("unary_einsum!", backend)(backend, y, x)
    return y
