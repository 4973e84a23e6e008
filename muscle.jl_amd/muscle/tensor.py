"""Index and Tensor: the data model of Muscle.jl (src/Index.jl, src/Tensor.jl).

A Tensor is an array plus one Index label per dimension.  The array is a
torch tensor (device memory on the GPU: the role ROCArray plays for the Julia
extension) with arbitrary element strides, like a Julia strided array or view.
Freshly allocated outputs are column-major, as Julia allocates them.
"""
from __future__ import annotations

from typing import Iterable, Sequence

import torch

from .errors import DimensionMismatch


class Index:
    """``Index{T}`` holding an arbitrary hashable tag (src/Index.jl:3-5)."""

    __slots__ = ("tag", "_hash")

    def __init__(self, tag):
        if isinstance(tag, Index):
            tag = tag.tag
        self.tag = tag
        # the index algebra hashes every label many times per call; an
        # unhashable tag makes an Index that compares but cannot be hashed
        try:
            self._hash = hash(("Index", tag))
        except TypeError:
            self._hash = None

    def __eq__(self, other):
        if self is other:
            return True
        if not isinstance(other, Index):
            return False
        if self._hash is not None and other._hash is not None and self._hash != other._hash:
            return False
        return self.tag == other.tag

    def __hash__(self):
        if self._hash is None:
            raise TypeError(f"unhashable Index tag: {self.tag!r}")
        return self._hash

    def __repr__(self):
        return f"index<{self.tag}>"


def _as_index(i) -> Index:
    return i if isinstance(i, Index) else Index(i)


def col_major_strides(shape: Sequence[int]) -> tuple:
    s, out = 1, []
    for n in shape:
        out.append(s)
        s *= max(int(n), 1)
    return tuple(out)


def empty_col_major(shape: Sequence[int], dtype: torch.dtype, device) -> torch.Tensor:
    """Uninitialised column-major array (Julia's ``similar``)."""
    shape = tuple(int(n) for n in shape)
    return torch.empty_strided(shape, col_major_strides(shape), dtype=dtype, device=device)


class Tensor:
    """``Tensor{T,N,A}``: data + inds (src/Tensor.jl:11-32)."""

    __slots__ = ("_data", "_inds")

    def __init__(self, data, inds: Iterable = ()):
        if not isinstance(data, torch.Tensor):
            data = torch.as_tensor(data)
        inds = tuple(_as_index(i) for i in inds)
        if data.dim() != len(inds):
            raise DimensionMismatch(f"ndims(data) = {data.dim()} must be equal to length(inds) = {len(inds)}")
        for p, i in enumerate(inds):
            for q in range(p + 1, len(inds)):
                if inds[q] == i and data.shape[p] != data.shape[q]:
                    raise DimensionMismatch(f"size of repeated index {i} differs: {data.shape[p]} vs {data.shape[q]}")
        self._data = data
        self._inds = inds

    # Julia accessors
    def parent(self) -> torch.Tensor:
        return self._data

    @property
    def data(self) -> torch.Tensor:
        return self._data

    def inds(self) -> list:
        return list(self._inds)

    @property
    def ndim(self) -> int:
        return len(self._inds)

    @property
    def shape(self) -> tuple:
        return tuple(self._data.shape)

    @property
    def dtype(self) -> torch.dtype:
        return self._data.dtype

    @property
    def device(self) -> torch.device:
        return self._data.device

    def size(self, ind=None):
        if ind is None:
            return self.shape
        return self._data.shape[self.dim(ind)]

    def dim(self, ind) -> int:
        """Position of an index (src/Tensor.jl ``dim``)."""
        return self._inds.index(_as_index(ind))

    def strides(self) -> tuple:
        return tuple(self._data.stride())

    def permutedims(self, perm) -> "Tensor":
        """permutedims by Index list or by integer positions (src/Tensor.jl:302-319)."""
        if self.ndim == 0:
            return self
        perm = list(perm)
        if perm and not isinstance(perm[0], int):
            perm = [self.dim(i) for i in perm]
        return Tensor(self._data.permute(*perm), [self._inds[p] for p in perm])

    def __repr__(self):
        return f"Tensor({tuple(self.shape)}, inds={list(self._inds)}, dtype={self.dtype}, device={self.device})"

    def isapprox(self, other: "Tensor", rtol: float | None = None, atol: float = 0.0) -> bool:
        """Permutation-aware, norm-wise ``isapprox`` (src/Tensor.jl:117-121): aligns ``other``'s
        dimensions to this tensor's inds with ``findperm``, then tests
        norm(x - y) <= max(atol, rtol * max(norm(x), norm(y))); default rtol = sqrt(eps)."""
        if set(self._inds) != set(other._inds) or len(self._inds) != len(other._inds):
            return False
        y = other.permutedims(self.inds()).parent() if self.ndim else other.parent()
        x = self._data
        if rtol is None:
            eps = torch.finfo(torch.promote_types(x.dtype, y.dtype)).eps
            rtol = 0.0 if atol > 0 else eps ** 0.5
        xd, yd = x.to(torch.complex128), y.to(x.device).to(torch.complex128)
        nd = torch.linalg.vector_norm(xd - yd).item()
        nx, ny = torch.linalg.vector_norm(xd).item(), torch.linalg.vector_norm(yd).item()
        return nd <= max(atol, rtol * max(nx, ny))


def inds(t: Tensor) -> list:
    return t.inds()


def parent(t: Tensor) -> torch.Tensor:
    return t.parent()
