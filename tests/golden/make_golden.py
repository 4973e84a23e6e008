"""Writes the golden fixtures under tests/golden/ (run from the repo root:
python tests/golden/make_golden.py).

The reference cannot be run here (no Julia toolchain), so every expected
output is the oracle's (oracle/backend_base.py), cross-checked at generation
time against the long-double brute-force einsum (oracle/brute.c) to 1e-13.
Inputs that are too large to store (config 1) are regenerated from their seed
by muscle.workloads; the fixture keeps their first elements to detect any
generator drift.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "muscle.jl_amd")]

from oracle import backend_base as bb  # noqa: E402
from oracle.brute import brute_einsum  # noqa: E402
from muscle.workloads import config1_inputs  # noqa: E402


def checked(ref, brute):
    err = np.linalg.norm((ref - brute).ravel()) / np.linalg.norm(brute.ravel())
    assert err < 1e-13, err
    return ref


def main():
    A, ia, B, ib, dims, ic = config1_inputs(0)
    C = bb.binary_einsum_base(bb.front_binary(ia, ib, dims=dims), A, ia, B, ib)
    C = checked(C, brute_einsum(ic, A, ia, B, ib))
    np.savez_compressed(os.path.join(HERE, "config1_c64.npz"), c=C, a_head=A.ravel(order="F")[:8],
                        b_head=B.ravel(order="F")[:8])

    rng = np.random.default_rng(2024)
    def cn(shape):
        n = int(np.prod(shape))
        return np.reshape(rng.standard_normal(n) + 1j * rng.standard_normal(n), shape, order="F")
    # hyperindex (batch) contraction C[i,k,l] = sum_j A[i,j,k] B[k,l,j]
    A, B = cn((5, 6, 7)), cn((7, 8, 6))
    C = checked(bb.binary_einsum_general(["i", "k", "l"], A, ["i", "j", "k"], B, ["k", "l", "j"]),
                brute_einsum(["i", "k", "l"], A, ["i", "j", "k"], B, ["k", "l", "j"]))
    np.savez_compressed(os.path.join(HERE, "hyper_c64.npz"), a=A, b=B, c=C)
    # unary: trace + diagonal "i j i k -> k j"
    X = cn((4, 5, 4, 3))
    Y = checked(bb.unary_einsum_base(["k", "j"], X, ["i", "j", "i", "k"]),
                brute_einsum(["k", "j"], X, ["i", "j", "i", "k"]))
    np.savez_compressed(os.path.join(HERE, "unary_c64.npz"), x=X, y=Y)
    # hadamard with a 3-cycle order of b
    a, b = cn((2, 3, 4)), cn((4, 2, 3))
    c, _ = bb.hadamard_base(a, ["i", "j", "k"], b, ["k", "i", "j"])
    c = checked(c, brute_einsum(["i", "j", "k"], a, ["i", "j", "k"], b, ["k", "i", "j"]))
    np.savez_compressed(os.path.join(HERE, "hadamard_c64.npz"), a=a, b=b, c=c)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
