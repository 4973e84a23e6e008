"""Helpers for GPU parity tests: numpy (oracle side) <-> device Tensors."""
from __future__ import annotations

import numpy as np
import torch

import muscle

NP2TORCH = {np.float32: torch.float32, np.float64: torch.float64, np.complex64: torch.complex64,
            np.complex128: torch.complex128}


def dev_tensor(x: np.ndarray, inds, device="cuda") -> muscle.Tensor:
    """Copy a numpy array to the GPU keeping its (column-major) strides."""
    x = np.asarray(x)
    t = torch.from_numpy(np.asfortranarray(x) if x.ndim else x.copy())
    t = t.to(device)  # preserve_format keeps the column-major strides
    return muscle.Tensor(t, inds)


def host(t: muscle.Tensor) -> np.ndarray:
    torch.cuda.synchronize()
    return t.parent().cpu().numpy()


def rel_err(got, ref) -> float:
    """Norm-wise relative error ||got - ref|| / ||ref|| (Julia isapprox's measure)."""
    got = np.asarray(got, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    den = np.linalg.norm(ref.ravel())
    num = np.linalg.norm((got - ref).ravel())
    return float(num / den) if den > 0 else float(num)


def randn(rng, shape, dtype=np.complex128, scale=1.0) -> np.ndarray:
    n = int(np.prod(shape)) if len(shape) else 1
    if np.issubdtype(dtype, np.complexfloating):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    else:
        x = rng.standard_normal(n)
    return np.reshape((scale * x).astype(dtype), shape, order="F")


# tolerance per arithmetic type (norm-wise relative): FP64 paths 1e-12, FP32 paths 1e-5
TOL = {np.float64: 1e-12, np.complex128: 1e-12, np.float32: 2e-5, np.complex64: 2e-5}
