"""ctypes wrapper of oracle/brute.c — TEST INFRASTRUCTURE ONLY.

``brute_einsum(inds_c, a, inds_a, b, inds_b)`` returns the exact-ish (80-bit
accumulated) einsum of numpy inputs, indexed in ``inds_c`` order.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libbrute.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        lib.brute_einsum.argtypes = [ctypes.c_int, P, ctypes.c_int, P, P, P, ctypes.c_int, ctypes.c_int, P, P, P,
                                     ctypes.c_int, P, P, P]
        lib.brute_einsum.restype = ctypes.c_int
        _lib = lib
    return _lib


def _ptr(x: np.ndarray):
    return x.ctypes.data_as(ctypes.c_void_p)


def brute_einsum(inds_c, a: np.ndarray, inds_a, b: np.ndarray | None = None, inds_b=None) -> np.ndarray:
    lib = _load()
    inds_a = list(inds_a)
    inds_c = list(inds_c)
    inds_b = list(inds_b) if b is not None else []
    labels = []
    for i in inds_c + inds_a + inds_b:
        if i not in labels:
            labels.append(i)
    ext = {}
    for arr, inds in ((a, inds_a), (b, inds_b)):
        if arr is None:
            continue
        for p, i in enumerate(inds):
            if ext.setdefault(i, arr.shape[p]) != arr.shape[p]:
                raise ValueError(f"extent mismatch for {i!r}")
    for i in inds_c:
        if i not in ext:
            raise ValueError(f"output index {i!r} not in inputs")
    ext_arr = np.array([ext[i] for i in labels], dtype=np.int64)

    def prep(arr, inds):
        z = np.ascontiguousarray(np.asarray(arr, dtype=np.complex128))
        lab = np.array([labels.index(i) for i in inds], dtype=np.int32)
        st = np.array([s // 16 for s in z.strides], dtype=np.int64)
        return z, lab, st

    za, laa, sa = prep(a, inds_a)
    if b is not None:
        zb, lbb, sb = prep(b, inds_b)
    else:
        zb, lbb, sb = np.zeros(1, np.complex128), np.zeros(1, np.int32), np.zeros(1, np.int64)
    shape_c = [ext[i] for i in inds_c]
    zc = np.zeros(shape_c, dtype=np.complex128, order="F")
    lcc = np.array([labels.index(i) for i in inds_c], dtype=np.int32)
    sc = np.array([s // 16 for s in zc.strides], dtype=np.int64)
    rc = lib.brute_einsum(len(labels), _ptr(ext_arr), len(inds_a), _ptr(laa), _ptr(sa), _ptr(za),
                          1 if b is not None else 0, len(inds_b), _ptr(lbb), _ptr(sb), _ptr(zb),
                          len(inds_c), _ptr(lcc), _ptr(sc), _ptr(zc))
    if rc != 0:
        raise MemoryError("brute_einsum failed")
    dt = np.result_type(a.dtype, b.dtype) if b is not None else a.dtype
    if not np.issubdtype(dt, np.complexfloating):
        return zc.real.astype(dt)
    return zc.astype(dt)
