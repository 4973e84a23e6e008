"""The factorization oracle pinned on the CPU.

Known answers of the reference's test/unit/operations/simple_update.jl (an MPS
of |00> + |11> with a 2x2 identity tensor per site, identity and CNOT gates,
normalized, truncated to one, absorbed) and tensor_svd_thin.jl (argument
errors, index order and shapes, U s Vt = A, isometric U and Vt); singular
values against the eigenvalues of A^H A; an exact rank-1 tensor; and a planted
transposition of Vt that the reconstruction check must catch.
"""
import numpy as np
import pytest

from oracle import backend_base as bb
from oracle import factor as of

R2 = np.sqrt(2.0)
GA = (np.eye(2), ["s1", "bond"])
GB = (np.eye(2), ["s2", "bond"])
G_INDS = ["s1", "s2", "g1", "g2"]
OP_ID = np.eye(4).reshape(2, 2, 2, 2, order="F")
OP_CX = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=float).reshape(2, 2, 2, 2,
                                                                                                  order="F")


def su(op, **kw):
    return of.simple_update_base(GA[0], GA[1], "s1", GB[0], GB[1], "s2", "bond", op, G_INDS, "g1", "g2", **kw)


def norm_of_state(U, s, V):
    psi = bb.binary_einsum_general(["s1", "s2"], U * s, ["s1", "bond"], V, ["s2", "bond"])
    return float(np.real(np.vdot(psi, psi)))


def test_apply_identity():
    (U, iu), (s, is_), (V, iv) = su(OP_ID)
    assert iu == ["s1", "bond"] and iv == ["s2", "bond"] and is_ == ["bond"]
    assert np.allclose(U, GA[0]) and np.allclose(V, GB[0])
    assert np.array_equal(s, [1.0, 1.0])
    assert np.isclose(norm_of_state(U, s, V), 2.0)


def test_apply_cx():
    (U, _), (s, _), (V, _) = su(OP_CX)
    assert np.allclose(U, GA[0])
    assert np.allclose(V, np.array([[1, -1], [1, 1]]) / R2)
    assert np.allclose(s, [R2, 0.0], atol=1e-15)
    assert np.isclose(norm_of_state(U, s, V), 2.0)


def test_apply_identity_normalize():
    (U, _), (s, _), (V, _) = su(OP_ID, normalize=True)
    assert np.allclose(U, GA[0]) and np.allclose(V, GB[0])
    assert np.allclose(s, [1 / R2, 1 / R2])
    assert np.isclose(norm_of_state(U, s, V), 1.0)


def test_apply_cx_normalize():
    (U, _), (s, _), (V, _) = su(OP_CX, normalize=True)
    assert np.allclose(U, GA[0])
    assert np.allclose(V, np.array([[1, -1], [1, 1]]) / R2)
    assert np.allclose(s, [1.0, 0.0], atol=1e-15)
    assert np.isclose(norm_of_state(U, s, V), 1.0)


def test_truncate_to_one():
    (U, _), (s, _), (V, _) = su(OP_ID, maxdim=1)
    assert np.allclose(U, GA[0][:, :1]) and np.allclose(V, GB[0][:, :1]) and np.allclose(s, [1.0])
    (U, _), (s, _), (V, _) = su(OP_CX, maxdim=1)
    assert np.allclose(U, GA[0][:, :1])
    assert V.shape == (2, 1) and np.allclose(V, np.array([[1], [1]]) / R2)
    assert np.allclose(s, [R2])


def test_absorb():
    ab_norm = np.linalg.norm(bb.binary_einsum_general(["s1", "s2"], GA[0], GA[1], GB[0], GB[1]))
    (U, _), (V, _) = su(OP_ID, absorb="u")
    assert np.isclose(np.linalg.norm(U), ab_norm)
    (U, _), (V, _) = su(OP_ID, absorb="v")
    assert np.isclose(np.linalg.norm(V), ab_norm)
    (U, _), (V, _) = su(OP_ID, absorb="equally")
    assert np.allclose(U, GA[0]) and np.allclose(V, GB[0])


A_INDS = ["i", "j", "k", "l"]


def rand_c(rng, shape):
    return rng.random(shape) + 1j * rng.random(shape)


def isisometry(t, inds, ind):
    """src/Tensor.jl:520-538: t contracted with conj(t) over all but `ind` is the identity."""
    a = list(inds)
    b = [("b", i) if i == ind else i for i in inds]
    got = bb.binary_einsum_general([ind, ("b", ind)], t, a, np.conj(t), b)
    return np.allclose(got, np.eye(t.shape[inds.index(ind)]), atol=1e-12)


def test_svd_thin_argument_errors():
    A = rand_c(np.random.default_rng(0), (2, 4, 6, 8))
    for kw in ({}, {"inds_u": ["z"]}, {"inds_v": ["z"]}, {"inds_u": A_INDS}, {"inds_v": A_INDS},
               {"inds_u": ["i"], "ind_s": "j"}):
        with pytest.raises(bb.ArgumentError):
            of.tensor_svd_thin_base(A, A_INDS, **kw)


def test_svd_thin_shapes_reconstruction_isometry():
    A = rand_c(np.random.default_rng(1), (2, 4, 6, 8))
    (U, iu), (s, is_), (Vt, iv) = of.tensor_svd_thin_base(A, A_INDS, inds_u=["i", "j"], ind_s="x")
    assert iu == ["i", "j", "x"] and is_ == ["x"] and iv == ["k", "l", "x"]
    assert U.shape == (2, 4, 8) and s.shape == (8,) and Vt.shape == (6, 8, 8)
    back = bb.binary_einsum_general(A_INDS, U * s, iu, Vt, iv)
    assert np.allclose(back, A, rtol=0, atol=1e-12 * np.abs(A).max())
    assert isisometry(U, iu, "x") and isisometry(Vt, iv, "x")
    # a transposed Vt (the conjugation dropped) no longer reconstructs A
    bad = bb.binary_einsum_general(A_INDS, U * s, iu, np.conj(Vt), iv)
    assert not np.allclose(bad, A, atol=1e-6)


@pytest.mark.parametrize("seed,shape,split", [(2, (3, 5, 7), 1), (3, (16, 2, 2, 3), 2), (4, (40, 6), 1)])
def test_singular_values_are_eigenvalue_roots(seed, shape, split):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    inds = [f"m{d}" for d in range(len(shape))]
    (U, iu), (s, _), (Vt, iv) = of.tensor_svd_thin_base(A, inds, inds_v=inds[split:], ind_s="x")
    m = int(np.prod(shape[:split]))
    mat = A.reshape(m, -1, order="F")
    ev = np.sort(np.linalg.eigvalsh(mat.conj().T @ mat))[::-1][:min(mat.shape)]
    assert np.allclose(s ** 2, np.clip(ev, 0, None), rtol=1e-10, atol=1e-10 * ev.max())
    assert np.all(np.diff(s) <= 0)


def test_rank_one_tensor():
    rng = np.random.default_rng(5)
    u = rng.standard_normal((3, 4)) + 1j * rng.standard_normal((3, 4))
    v = rng.standard_normal(5) + 1j * rng.standard_normal(5)
    A = u[:, :, None] * v[None, None, :]
    (U, iu), (s, _), (Vt, iv) = of.tensor_svd_thin_base(A, ["a", "b", "c"], inds_u=["a", "b"], ind_s="x")
    assert s.shape == (5,)
    assert np.isclose(s[0], np.linalg.norm(u) * np.linalg.norm(v)) and np.allclose(s[1:], 0, atol=1e-12)
    # the first left vector is u / |u| up to a phase
    ov = np.vdot(U[..., 0].ravel(order="F"), u.ravel(order="F")) / np.linalg.norm(u)
    assert np.isclose(abs(ov), 1.0)


def test_factorinds():
    assert of.factorinds(["a", "b", "c"], ["a"], []) == (["a"], ["b", "c"])
    assert of.factorinds(["a", "b", "c"], [], ["c"]) == (["a", "b"], ["c"])
    with pytest.raises(bb.ArgumentError):
        of.factorinds(["a", "b"], ["a"], ["a"])


def test_svd_trunc_known_answers():
    """test/unit/operations/tensor_svd_trunc.jl."""
    rng = np.random.default_rng(6)
    A = rand_c(rng, (2, 4, 6, 8))
    for kw in ({}, {"inds_u": ["z"]}, {"inds_v": ["z"]}, {"inds_u": A_INDS}, {"inds_v": A_INDS},
               {"inds_u": ["i"], "ind_s": "j"}):
        with pytest.raises(bb.ArgumentError):
            of.tensor_svd_trunc_base(A, A_INDS, **kw)
    (U, iu), (s, _), (Vt, iv) = of.tensor_svd_trunc_base(A, A_INDS, inds_u=["i", "j"], ind_s="x")
    assert U.shape == (2, 4, 8) and s.shape == (8,) and Vt.shape == (6, 8, 8)
    assert np.allclose(bb.binary_einsum_general(A_INDS, U * s, iu, Vt, iv), A)
    A = rand_c(rng, (200, 100))
    (_, _), (sfull, _), (_, _) = of.tensor_svd_trunc_base(A, ["i", "j"], inds_u=["i"], ind_s="x")
    assert of.tensor_svd_trunc_base(A, ["i", "j"], inds_u=["i"], ind_s="x", maxdim=20)[1][0].shape == (20,)
    (U, iu), (s, _), (Vt, iv) = of.tensor_svd_trunc_base(A, ["i", "j"], inds_u=["i"], ind_s="x", threshold=1e-3,
                                                          maxdim=80)
    err = np.linalg.norm(bb.binary_einsum_general(["i", "j"], U * s, iu, Vt, iv) - A)
    assert np.isclose(err, np.linalg.norm(sfull[s.shape[0]:]))


def test_svd_trunc_threshold_keeps_first_value_below():
    """findfirst(<(cut), s) is the last kept index: the first value under the cut stays."""
    s_true = np.array([4.0, 2.0, 1.0, 1e-3, 1e-4, 1e-5])
    rng = np.random.default_rng(7)
    qa, _ = np.linalg.qr(rng.standard_normal((9, 6)))
    qb, _ = np.linalg.qr(rng.standard_normal((6, 6)))
    A = (qa * s_true) @ qb.T
    (_, _), (s, _), (_, _) = of.tensor_svd_trunc_base(A, ["i", "j"], inds_u=["i"], ind_s="x", threshold=1e-2)
    # norm(s) ~ 4.58: the cut is ~0.0458, the first value below it (1e-3) is kept
    assert s.shape == (4,) and np.isclose(s[-1], 1e-3)


def test_qr_thin_known_answers():
    """test/unit/operations/tensor_qr_thin.jl."""
    A = rand_c(np.random.default_rng(8), (2, 4, 6, 8))
    for kw in ({}, {"inds_q": ["z"]}, {"inds_r": ["z"]}, {"inds_q": A_INDS}, {"inds_r": A_INDS},
               {"inds_q": ["i"], "ind_virtual": "j"}):
        with pytest.raises(bb.ArgumentError):
            of.tensor_qr_thin_base(A, A_INDS, **kw)
    (Q, iq), (R, ir) = of.tensor_qr_thin_base(A, A_INDS, inds_q=["i", "j"], ind_virtual="x")
    assert iq == ["i", "j", "x"] and ir == ["x", "k", "l"]
    assert Q.shape == (2, 4, 8) and R.shape == (8, 6, 8)
    assert np.allclose(bb.binary_einsum_general(A_INDS, Q, iq, R, ir), A)
    assert isisometry(Q, iq, "x")


@pytest.mark.parametrize("shape,split", [((12, 5), 1), ((5, 12), 1), ((3, 4, 2, 3), 2)])
def test_qr_r_is_cholesky_factor(shape, split):
    rng = np.random.default_rng(9)
    A = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    inds = [f"m{d}" for d in range(len(shape))]
    (Q, iq), (R, ir) = of.tensor_qr_thin_base(A, inds, inds_q=inds[:split], ind_virtual="x")
    m = int(np.prod(shape[:split]))
    mat = A.reshape(m, -1, order="F")
    k = min(mat.shape)
    r = R.reshape(k, -1, order="F")
    assert np.allclose(np.tril(r, -1), 0)
    assert np.allclose(r.conj().T @ r, mat.conj().T @ mat)  # Q^H Q = I
    # the leading k x k block of R: R^H R = A1^H A1 for the first k columns A1 of the matrix
    a1 = mat[:, :k]
    chol = np.linalg.cholesky(a1.conj().T @ a1)
    assert np.allclose(np.abs(np.diag(r)[:k]), np.abs(np.diag(chol)))
    # LAPACK's convention: diag(R) is real
    assert np.allclose(np.diag(r).imag, 0)
