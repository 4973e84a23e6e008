"""muscle — Python mirror of Muscle.jl's contraction API over libmuscle_hip.

    from muscle import Tensor, Index, binary_einsum, hadamard, unary_einsum
    a = Tensor(x_gpu, ["i", "j", "k"]); b = Tensor(y_gpu, ["k", "l", "j"])
    c = binary_einsum(a, b, dims=["j", "k"])          # -> inds (i, l)

Names follow src/Muscle.jl:16-36; Julia's in-place ``f!`` is ``f_``.  Device
arrays (torch tensors on a ROCm device) dispatch to BackendHIP, which runs the
hand-written gfx950 kernels of libmuscle_hip.so.
"""
from .backend import (Backend, BackendBase, BackendCuTENSOR, BackendDagger, BackendDefault, BackendHIP,
                      BackendOMEinsum, BackendReactant, BackendStrided, Domain, DomainHost, DomainROCm,
                      choose_backend, choose_backend_rule, domain, with_backend)
from .errors import ArgumentError, DimensionMismatch, MuscleHipError, UnsupportedError
from .ops import (binary_einsum, binary_einsum_, hadamard, hadamard_, last_info, measure_bandwidth, nonunique,
                  unary_einsum, unary_einsum_)
from .tensor import Index, Tensor, col_major_strides, empty_col_major, inds, parent
from .describe import describe
from .factor import (AbsorbBehavior, AbsorbEqually, AbsorbU, AbsorbV, DontAbsorb, factorinds, simple_update,
                     tensor_qr_thin, tensor_qr_thin_, tensor_svd_thin, tensor_svd_thin_, tensor_svd_trunc)

__all__ = [
    "tensor_svd_thin", "tensor_svd_thin_", "tensor_svd_trunc", "tensor_qr_thin", "tensor_qr_thin_", "simple_update", "factorinds", "AbsorbBehavior", "DontAbsorb", "AbsorbU", "AbsorbV",
    "AbsorbEqually", "Tensor", "Index", "inds", "parent", "binary_einsum", "binary_einsum_", "hadamard", "hadamard_",
    "unary_einsum", "unary_einsum_", "nonunique", "Backend", "BackendBase", "BackendHIP", "BackendOMEinsum",
    "BackendDefault", "BackendStrided", "BackendCuTENSOR", "BackendReactant", "BackendDagger", "Domain",
    "DomainHost", "DomainROCm", "domain", "with_backend", "choose_backend", "choose_backend_rule",
    "ArgumentError", "DimensionMismatch", "UnsupportedError", "MuscleHipError", "last_info", "measure_bandwidth", "describe",
    "col_major_strides", "empty_col_major",
]
