"""Exceptions mirroring the Julia ones the reference raises."""


class ArgumentError(ValueError):
    """Julia ArgumentError: @argcheck failures (binary_einsum.jl:82-83, hadamard.jl:10,23-24)
    and the "not implemented or not loaded" backend fallbacks (binary_einsum.jl:53-55)."""


class DimensionMismatch(ValueError):
    """Julia DimensionMismatch (Tensor constructor, src/Tensor.jl:16-24)."""


class UnsupportedError(ArgumentError):
    """A combination the HIP backend does not implement (status MUSCLE_HIP_UNSUPPORTED)."""


class MuscleHipError(RuntimeError):
    """HIP runtime failure, out of memory, or a rejected request (Julia ErrorException)."""

    def __init__(self, msg: str, status: int = 0):
        super().__init__(msg)
        self.status = status
