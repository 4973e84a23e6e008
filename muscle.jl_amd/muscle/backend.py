"""Backend dispatch: Backend singletons, task-scoped override, memory Domains
and per-operation rules (src/Backend.jl:4-36, src/Domain.jl:4-14, the
choose_backend_rule lines of each src/Operations/*.jl file).

The HIP backend plugs in the way a Muscle.jl package extension does: a new
Domain for device arrays (``DomainROCm``, the analogue of
``Domain(::Type{<:CuArray}) = DomainCUDA()`` in ext/MuscleCUDAExt.jl:7), a new
Backend (``BackendHIP``) and rules mapping ROCm domains to it.  Host arrays
keep the reference's rules (BackendBase / BackendOMEinsum); those backends
are not part of this library, so calling them raises the reference's own
"not implemented or not loaded" ArgumentError.
"""
from __future__ import annotations

import contextlib
import contextvars
from dataclasses import dataclass

import torch


class Backend:
    """``abstract type Backend`` (src/Backend.jl:4)."""

    def __repr__(self):
        return f"{type(self).__name__}()"

    def __eq__(self, other):
        return type(self) is type(other) and self.__dict__ == other.__dict__

    def __hash__(self):
        return hash(type(self))


class BackendDefault(Backend):
    pass


class BackendBase(Backend):
    pass


class BackendStrided(Backend):
    pass


class BackendOMEinsum(Backend):
    pass


class BackendCuTENSOR(Backend):
    pass


class BackendReactant(Backend):
    pass


class BackendDagger(Backend):
    pass


@dataclass(eq=True, frozen=True)
class BackendHIP(Backend):
    """The MI355X backend.  ``precision`` selects the arithmetic of binary_einsum
    ("native", "fp32", "bf16x3", "fp16x2", "ozaki_int8"); ``target_rel_error``
    sets the Ozaki slice count; ``complex_mult`` picks 3M or 4M complex
    products.  Like the keyword arguments BackendCuTENSOR
    forwards to cuTENSOR (ext/MuscleCUDAExt.jl:22,37), they only tune the
    library call."""

    precision: str = "native"
    target_rel_error: float = 1e-12
    complex_mult: int = 3  # complex x complex: 3 real products (Gauss) or 4

    def __repr__(self):
        if self.precision == "native" and self.complex_mult == 3:
            return "BackendHIP()"
        return (f"BackendHIP(precision={self.precision!r}, target_rel_error={self.target_rel_error:g}, "
                f"complex_mult={self.complex_mult})")


# ---------------------------------------------------------------- Domains
class Domain:
    def __repr__(self):
        return f"{type(self).__name__}()"

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self))


class DomainHost(Domain):
    pass


class DomainROCm(Domain):
    pass


def domain(x) -> Domain:
    """Memory domain of an array (src/Domain.jl:11-14): device tensors are ROCm."""
    if isinstance(x, torch.Tensor) and x.device.type == "cuda":
        return DomainROCm()
    return DomainHost()


# ---------------------------------------------------------------- override + rules
CURRENT_BACKEND: contextvars.ContextVar = contextvars.ContextVar("CURRENT_BACKEND")


@contextlib.contextmanager
def with_backend(backend: Backend):
    """Task-scoped backend override (``with_backend``, src/Backend.jl:16-18)."""
    token = CURRENT_BACKEND.set(backend)
    try:
        yield backend
    finally:
        CURRENT_BACKEND.reset(token)


_RULES: dict = {}


def choose_backend_rule(op: str, *domains: Domain):
    rule = _RULES.get((op,) + tuple(type(d) for d in domains))
    if rule is None:
        from .errors import ArgumentError
        raise ArgumentError(f"no backend rule for {op} on {domains}")
    return rule


def register_rule(op: str, domains: tuple, backend: Backend) -> None:
    _RULES[(op,) + tuple(domains)] = backend


def choose_backend(op: str, *arrays) -> Backend:
    """Override first, then the per-op rule on the arrays' Domains (src/Backend.jl:28-36)."""
    cur = CURRENT_BACKEND.get(None)
    if cur is not None:
        return cur
    return choose_backend_rule(op, *(domain(a) for a in arrays))


# reference rules for host arrays (binary_einsum.jl:20, :28; hadamard.jl:4-5; unary_einsum.jl:20,23)
register_rule("binary_einsum", (DomainHost, DomainHost), BackendBase())
register_rule("binary_einsum!", (DomainHost, DomainHost, DomainHost), BackendBase())
register_rule("hadamard", (DomainHost, DomainHost), BackendBase())
register_rule("hadamard!", (DomainHost, DomainHost, DomainHost), BackendBase())
register_rule("unary_einsum", (DomainHost,), BackendOMEinsum())
register_rule("unary_einsum!", (DomainHost, DomainHost), BackendOMEinsum())
# the HIP extension's rules; hadamard needs its own, more specific than the
# generic (::Domain, ::Domain) -> BackendBase rule of hadamard.jl:4-5
register_rule("binary_einsum", (DomainROCm, DomainROCm), BackendHIP())
register_rule("binary_einsum!", (DomainROCm, DomainROCm, DomainROCm), BackendHIP())
register_rule("hadamard", (DomainROCm, DomainROCm), BackendHIP())
register_rule("hadamard!", (DomainROCm, DomainROCm, DomainROCm), BackendHIP())
register_rule("unary_einsum", (DomainROCm,), BackendHIP())
register_rule("unary_einsum!", (DomainROCm, DomainROCm), BackendHIP())
