"""Known-answer tests ported from the reference's own test files.

Each case is data: inputs (numpy, column-major), index labels, the
front-end keyword arguments and the expected ``inds``/``size``/values the
reference asserts.  tests/test_oracle.py runs them against the CPU oracle,
tests/test_gpu_*.py against the HIP backend.
"""
from __future__ import annotations

import numpy as np


def ones(*shape, dtype=np.float64):
    return np.ones(shape, dtype=dtype, order="F")


def fill(v, dtype=np.float64):
    return np.array(v, dtype=dtype)


# --------------------------------------------------------------- binary_einsum
# (name, A, inds_a, B, inds_b, kwargs, expected inds, expected value or None, base_ok)
# base_ok=False: BackendBase must raise ArgumentError (hyperindex); the GPU
# semantics (test/integration/cuda.jl) give the listed value instead.
def binary_kats():
    cases = []
    # test/unit/operations/binary_einsum.jl:4-30 "matmul"
    A, B = ones(2, 3), ones(3, 4)
    for kw, out_inds, val in [({}, ["i", "k"], 3 * ones(2, 4)),
                              ({"out": ["i", "k"]}, ["i", "k"], 3 * ones(2, 4)),
                              ({"out": ["k", "i"]}, ["k", "i"], 3 * ones(4, 2)),
                              ({"dims": ["j"]}, ["i", "k"], 3 * ones(2, 4))]:
        cases.append(("matmul", A, ["i", "j"], B, ["j", "k"], kw, out_inds, val, True))
    # :32-58 "inner product"
    A, B = ones(3, 4), ones(4, 3)
    for kw in [{}, {"out": []}, {"dims": ["i", "j"]}, {"dims": ["j", "i"]}]:
        cases.append(("inner product", A, ["i", "j"], B, ["j", "i"], kw, [], fill(12.0), True))
    # :60-94 "outer product"
    A, B = ones(2, 3), ones(4, 5)
    size = {"i": 2, "j": 3, "k": 4, "l": 5}
    for out in [None, ["i", "j", "k", "l"], ["k", "l", "i", "j"], ["l", "k", "j", "i"], ["l", "i", "k", "j"],
                ["j", "i", "k", "l"]]:
        kw = {} if out is None else {"out": out}
        o = out or ["i", "j", "k", "l"]
        cases.append(("outer product", A, ["i", "j"], B, ["k", "l"], kw, o, ones(*[size[x] for x in o]), True))
    # :96-119 "scale"
    A, al = ones(2, 3), fill(2.0)
    cases.append(("scale", A, ["i", "j"], al, [], {}, ["i", "j"], 2.0 * A, True))
    cases.append(("scale", al, [], A, ["i", "j"], {}, ["i", "j"], 2.0 * A, True))
    cases.append(("scale", A, ["i", "j"], al, [], {"out": ["j", "i"]}, ["j", "i"], 2.0 * A.T, True))
    cases.append(("scale", al, [], A, ["i", "j"], {"out": ["j", "i"]}, ["j", "i"], 2.0 * A.T, True))
    # :121-133 "batch matmul" (BackendBase throws; GPU expectations test/integration/cuda.jl:126-150)
    A, B = ones(2, 3, 6), ones(3, 4, 6)
    for out, shp in [(["i", "k", "batch"], (2, 4, 6)), (["k", "i", "batch"], (4, 2, 6)),
                     (["batch", "i", "k"], (6, 2, 4))]:
        cases.append(("batch matmul", A, ["i", "j", "batch"], B, ["j", "k", "batch"], {"out": out}, out,
                      3 * ones(*shp), False))
    # dims=[j] keeps :batch as a hyperindex; the front end orders it [i, batch, k]
    # (cuda.jl:147-150 marks its [i, k, batch] expectation broken=true)
    cases.append(("batch matmul dims", A, ["i", "j", "batch"], B, ["j", "k", "batch"], {"dims": ["j"]},
                  ["i", "batch", "k"], 3 * ones(2, 6, 4), False))
    return cases


def manual_case(T):
    """test/unit/operations/binary_einsum.jl:135-155 (and cuda.jl:152-181)."""
    A = ones(2, 3, 4, dtype=T)
    B = ones(4, 5, 3, dtype=T)
    return A, ["i", "j", "k"], B, ["k", "l", "j"]


def manual_expected_full(A, B):
    """reshape(A, 2, 12) * reshape(permutedims(B, [3, 1, 2]), 12, 5) (binary_einsum.jl:145-149)."""
    A_mat = np.reshape(A, (2, 12), order="F")
    B_mat = np.reshape(np.transpose(B, (2, 0, 1)), (12, 5), order="F")
    return A_mat @ B_mat


def manual_expected_hyper(A, B):
    """The 4-nested loop C[i,k,l] += A[i,j,k] * B[k,l,j] (cuda.jl:170-180, omeinsum.jl:225-236)."""
    C = np.zeros((2, 4, 5), dtype=np.result_type(A, B))
    for i in range(2):
        for j in range(3):
            for k in range(4):
                for l in range(5):
                    C[i, k, l] += A[i, j, k] * B[k, l, j]
    return C


# --------------------------------------------------------------- hadamard
def hadamard_kats():
    """test/unit/operations/hadamard.jl:4-119: (name, a, inds_a, b, inds_b, expected)."""
    a = ones(2, 3, 4)
    ia = ["i", "j", "k"]
    out = [("scalar", a, ia, fill(2.0), [], 2.0 * a)]
    for lab, vec in [("i", [1.0, 2.0]), ("j", [1.0, 2.0, 3.0]), ("k", [1.0, 2.0, 3.0, 4.0])]:
        v = np.array(vec)
        shape = [1, 1, 1]
        shape[ia.index(lab)] = len(vec)
        out.append((f"vector {lab}", a, ia, v, [lab], a * v.reshape(shape, order="F")))
    b = np.array([[1.0, 2, 3], [4, 5, 6]], order="F")
    out.append(("matrix ij", a, ia, b, ["i", "j"], a * b[:, :, None]))
    out.append(("tensor", a, ia, 2 * ones(2, 3, 4), ia, 2 * a))
    return out


# --------------------------------------------------------------- unary_einsum
def unary_kats():
    """test/integration/omeinsum.jl:6-100: (name, x, inds_x, kwargs, expected inds, expected)."""
    A = ones(2, 3, 4)
    cases = [
        ("axis sum dims", A, ["i", "j", "k"], {"dims": ["i"]}, ["j", "k"], 2 * ones(3, 4)),
        ("axis sum out", A, ["i", "j", "k"], {"out": ["j", "k"]}, ["j", "k"], 2 * ones(3, 4)),
        ("full sum dims", A, ["i", "j", "k"], {"dims": ["i", "j", "k"]}, [], fill(24.0)),
        ("full sum out", A, ["i", "j", "k"], {"out": []}, [], fill(24.0)),
        ("diagonal ii->i", ones(2, 2), ["i", "i"], {"out": ["i"]}, ["i"], ones(2)),
        ("diagonal iji->ij", ones(2, 3, 2), ["i", "j", "i"], {"out": ["i", "j"]}, ["i", "j"], ones(2, 3)),
        ("trace default", ones(2, 2), ["i", "i"], {}, [], fill(2.0)),
        ("trace dims", ones(2, 2), ["i", "i"], {"dims": ["i"]}, [], fill(2.0)),
        ("trace out", ones(2, 2), ["i", "i"], {"out": []}, [], fill(2.0)),
        ("iji->j default", ones(2, 3, 2), ["i", "j", "i"], {}, ["j"], 2 * ones(3)),
        ("iji->j dims", ones(2, 3, 2), ["i", "j", "i"], {"dims": ["i"]}, ["j"], 2 * ones(3)),
        ("iji->j out", ones(2, 3, 2), ["i", "j", "i"], {"out": ["j"]}, ["j"], 2 * ones(3)),
    ]
    return cases


# in-place forms of the unary KATs: (name, x, inds_x, inds_y, expected)
def unary_inplace_kats():
    return [
        ("axis sum !", ones(2, 3, 4), ["i", "j", "k"], ["j", "k"], 2 * ones(3, 4)),
        ("full sum !", ones(2, 3, 4), ["i", "j", "k"], [], fill(24.0)),
        ("diagonal !", ones(2, 2), ["i", "i"], ["i"], ones(2)),
        ("iji->ij !", ones(2, 3, 2), ["i", "j", "i"], ["i", "j"], ones(2, 3)),
        ("trace !", ones(2, 2), ["i", "i"], [], fill(2.0)),
        ("iji->j !", ones(2, 3, 2), ["i", "j", "i"], ["j"], 2 * ones(3)),
    ]
