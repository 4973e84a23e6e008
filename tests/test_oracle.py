"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

* the reference's known-answer tests (tests/kats.py) on every oracle function;
* the C long-double brute-force einsum (oracle/brute.c) on seeded random
  inputs, an independent implementation (nested loops, no permutes, no BLAS);
* invariants: output permutation, linearity;
* planted mistakes in the oracle (a transposed operand, a dropped term, the
  reference's own inverse-permutation quirk) are caught by these checks.
"""
import numpy as np
import pytest

from oracle import backend_base as bb
from oracle.brute import brute_einsum
from tests import kats


def rand(rng, shape, dtype=np.complex128):
    n = int(np.prod(shape)) if len(shape) else 1
    if np.issubdtype(dtype, np.complexfloating):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    else:
        x = rng.standard_normal(n)
    return np.reshape(x.astype(dtype), shape, order="F")


# ------------------------------------------------------------ KATs: binary
@pytest.mark.parametrize("case", kats.binary_kats(), ids=lambda c: c[0])
def test_binary_kats_oracle(case):
    name, A, ia, B, ib, kw, out_inds, val, base_ok = case
    inds_c = bb.front_binary(ia, ib, **kw)
    assert inds_c == out_inds
    if base_ok:
        C = bb.binary_einsum_base(inds_c, A, ia, B, ib)
    else:
        with pytest.raises(bb.ArgumentError):
            bb.binary_einsum_base(inds_c, A, ia, B, ib)
        C = bb.binary_einsum_general(inds_c, A, ia, B, ib)
    assert C.shape == val.shape
    assert np.array_equal(C, val)  # exact, as the reference asserts `==`
    # the general (hyperindex-aware) restatement agrees everywhere
    assert np.array_equal(bb.binary_einsum_general(inds_c, A, ia, B, ib), val)


@pytest.mark.parametrize("T", [np.float64, np.complex128])
def test_manual_kat_oracle(T):
    A, ia, B, ib = kats.manual_case(T)
    C = bb.binary_einsum_base(bb.front_binary(ia, ib, dims=["j", "k"]), A, ia, B, ib)
    assert np.allclose(C, kats.manual_expected_full(A, B))
    with pytest.raises(bb.ArgumentError):
        bb.binary_einsum_base(bb.front_binary(ia, ib, dims=["j"]), A, ia, B, ib)
    inds_c = bb.front_binary(ia, ib, dims=["j"])
    assert inds_c == ["i", "k", "l"]
    assert np.allclose(bb.binary_einsum_general(inds_c, A, ia, B, ib), kats.manual_expected_hyper(A, B))
    # same identities on seeded random values (ones cannot see transposes)
    rng = np.random.default_rng(7)
    A, B = rand(rng, (2, 3, 4), T), rand(rng, (4, 5, 3), T)
    assert np.allclose(bb.binary_einsum_base(["i", "l"], A, ia, B, ib), kats.manual_expected_full(A, B),
                       rtol=1e-13, atol=1e-13)
    assert np.allclose(bb.binary_einsum_general(["i", "k", "l"], A, ia, B, ib), kats.manual_expected_hyper(A, B),
                       rtol=1e-13, atol=1e-13)


def test_inplace_requires_left_right_order():
    A, B = kats.ones(2, 3), kats.ones(3, 4)
    assert np.array_equal(bb.binary_einsum_base_inplace(["i", "k"], A, ["i", "j"], B, ["j", "k"]), 3 * kats.ones(2, 4))
    with pytest.raises(bb.ArgumentError):
        bb.binary_einsum_base_inplace(["k", "i"], A, ["i", "j"], B, ["j", "k"])


# ------------------------------------------------------------ KATs: hadamard / unary
@pytest.mark.parametrize("case", kats.hadamard_kats(), ids=lambda c: c[0])
def test_hadamard_kats_oracle(case):
    name, a, ia, b, ib, expected = case
    c, ic = bb.hadamard_base(a, ia, b, ib)
    assert ic == ia and c.dtype == a.dtype
    assert np.allclose(c, expected)
    # swapped argument order gives the same (hadamard.jl:9)
    c2, ic2 = bb.hadamard_base(b, ib, a, ia)
    assert ic2 == ia and np.allclose(c2, expected)


@pytest.mark.parametrize("case", kats.unary_kats(), ids=lambda c: c[0])
def test_unary_kats_oracle(case):
    name, x, ix, kw, out_inds, expected = case
    iy = bb.front_unary(ix, **kw)
    assert iy == out_inds
    y = bb.unary_einsum_base(iy, x, ix)
    assert y.shape == expected.shape and np.allclose(y, expected)


def test_nonunique():
    assert bb.nonunique(["i", "j", "i"]) == ["i"]
    assert bb.nonunique(["a", "b", "b", "a", "a", "c"]) == ["b", "a"]
    assert bb.nonunique(["i", "j"]) == []


# ------------------------------------------------------------ brute force pins
BINARY_CASES = [
    # (inds_a, shape_a, inds_b, shape_b, inds_c)
    (["i", "j", "k"], (3, 4, 5), ["k", "l", "j"], (5, 2, 4), ["i", "l"]),
    (["i", "j", "k"], (3, 4, 5), ["k", "l", "j"], (5, 2, 4), ["l", "i"]),
    (["a", "b"], (3, 4), ["c", "d"], (2, 5), ["d", "a", "c", "b"]),
    (["i", "j"], (3, 4), ["j", "i"], (4, 3), []),
    (["i", "j", "h"], (3, 4, 2), ["j", "k", "h"], (4, 5, 2), ["h", "k", "i"]),
    (["i", "j", "k"], (2, 3, 4), ["k", "l", "j"], (4, 5, 3), ["i", "k", "l"]),
    (["p", "q", "r", "s"], (2, 3, 2, 3), ["s", "q", "t"], (3, 3, 4), ["t", "p", "r"]),
]


@pytest.mark.parametrize("case", BINARY_CASES, ids=lambda c: "".join(c[0]) + "," + "".join(c[2]) + "->" + "".join(c[4]))
@pytest.mark.parametrize("T", [np.float64, np.complex128])
def test_binary_oracle_vs_brute(case, T):
    ia, sa, ib, sb, ic = case
    rng = np.random.default_rng(11)
    A, B = rand(rng, sa, T), rand(rng, sb, T)
    ref = brute_einsum(ic, A, ia, B, ib)
    got = bb.binary_einsum_general(ic, A, ia, B, ib)
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-13)
    if not any(i in ia and i in ib for i in ic):
        np.testing.assert_allclose(bb.binary_einsum_base(ic, A, ia, B, ib), ref, rtol=1e-12, atol=1e-13)


UNARY_CASES = [
    (["i", "j", "k"], (3, 4, 5), ["k", "i"]),
    (["i", "j", "i"], (3, 4, 3), ["j"]),
    (["i", "j", "i"], (3, 4, 3), ["i", "j"]),
    (["i", "i"], (5, 5), []),
    (["a", "b", "a", "c", "b"], (2, 3, 2, 4, 3), ["c", "a"]),
    (["i", "j"], (6, 7), ["j", "i"]),
]


@pytest.mark.parametrize("case", UNARY_CASES, ids=lambda c: "".join(c[0]) + "->" + "".join(c[2]))
def test_unary_oracle_vs_brute(case):
    ix, sx, iy = case
    rng = np.random.default_rng(3)
    X = rand(rng, sx)
    np.testing.assert_allclose(bb.unary_einsum_base(iy, X, ix), brute_einsum(iy, X, ix), rtol=1e-12, atol=1e-13)


HADAMARD_CASES = [
    (["i", "j", "k"], (2, 3, 4), ["k", "i", "j"]),  # 3-cycle: the reference's findperm use is wrong here
    (["i", "j", "k"], (2, 3, 4), ["j", "k", "i"]),
    (["i", "j", "k"], (2, 3, 4), ["k", "i"]),
    (["i", "j", "k", "l"], (2, 3, 4, 5), ["l", "j"]),
    (["i", "j"], (3, 4), []),
]


@pytest.mark.parametrize("case", HADAMARD_CASES, ids=lambda c: "".join(c[0]) + "*" + "".join(c[2]))
def test_hadamard_oracle_vs_brute(case):
    ia, sa, ib = case
    rng = np.random.default_rng(5)
    size = dict(zip(ia, sa))
    A = rand(rng, sa)
    B = rand(rng, tuple(size[i] for i in ib))
    c, ic = bb.hadamard_base(A, ia, B, ib)
    np.testing.assert_allclose(c, brute_einsum(ia, A, ia, B, ib), rtol=1e-14, atol=0)


# ------------------------------------------------------------ invariants
def test_output_permutation_and_linearity():
    rng = np.random.default_rng(2)
    ia, ib = ["i", "j", "k"], ["k", "l", "j"]
    A, A2, B = rand(rng, (3, 4, 5)), rand(rng, (3, 4, 5)), rand(rng, (5, 6, 4))
    C = bb.binary_einsum_base(["i", "l"], A, ia, B, ib)
    CT = bb.binary_einsum_base(["l", "i"], A, ia, B, ib)
    np.testing.assert_array_equal(C.T, CT)
    s = 0.3 - 1.1j
    lin = bb.binary_einsum_base(["i", "l"], A + s * A2, ia, B, ib)
    np.testing.assert_allclose(lin, C + s * bb.binary_einsum_base(["i", "l"], A2, ia, B, ib), rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------ planted mistakes
def test_planted_mistakes_are_caught(monkeypatch):
    """A plausible oracle bug must fail one of the pins above."""
    rng = np.random.default_rng(9)
    ia, sa, ib, sb, ic = BINARY_CASES[0]
    A, B = rand(rng, sa), rand(rng, sb)
    ref = brute_einsum(ic, A, ia, B, ib)

    # 1. A's contracted modes linearised in row-major order while B's are
    #    column-major (a transposed K ordering of one operand)
    real_reshape = bb.reshape_f
    calls = []

    def first_call_row_major(x, shape):
        calls.append(1)
        return np.reshape(x, tuple(shape), order="C" if len(calls) == 1 else "F")

    monkeypatch.setattr(bb, "reshape_f", first_call_row_major)
    assert not np.allclose(bb.binary_einsum_base(ic, A, ia, B, ib), ref)
    monkeypatch.setattr(bb, "reshape_f", real_reshape)
    # 1b. an output permutation applied backwards
    real_perm = bb.permutedims
    monkeypatch.setattr(bb, "permutedims", lambda x, fr, to: real_perm(x, fr, to[::-1] if len(to) == 2 and to != fr else to))
    got = bb.binary_einsum_base(["l", "i"], A, ia, B, ib)
    assert got.shape != ref.T.shape or not np.allclose(got, ref.T)
    monkeypatch.setattr(bb, "permutedims", real_perm)

    # 2. dropped term in the complex product: real parts only
    assert not np.allclose(bb.binary_einsum_base(ic, A.real, ia, B, ib), ref)

    # 3. the reference's inverse-permutation quirk for hadamard (findperm passed to permutedims)
    a = rand(rng, (2, 3, 4))
    b = rand(rng, (4, 2, 3))  # inds (k, i, j): a 3-cycle
    good, _ = bb.hadamard_base(a, ["i", "j", "k"], b, ["k", "i", "j"])
    assert np.allclose(good, brute_einsum(["i", "j", "k"], a, ["i", "j", "k"], b, ["k", "i", "j"]))
    # findperm(from=inds(b), to=a-order) = position of each b index in a's order (src/Index.jl:20-34)
    findperm = [["i", "j", "k"].index(x) for x in ["k", "i", "j"]]
    bq = np.transpose(b, findperm)  # what hadamard.jl:68-69 does with it
    assert bq.shape == (3, 4, 2)  # instead of (2, 3, 4): the quirk is real for a 3-cycle
    correct = [["k", "i", "j"].index(x) for x in ["i", "j", "k"]]
    assert np.transpose(b, correct).shape == (2, 3, 4)
