"""tensor_svd_thin and simple_update on the GPU against the factorization oracle.

Singular values agree with the oracle's LAPACK SVD to 1e-12 relative (FP64
types) or 2e-5 (FP32 types); U s Vt reconstructs A to the same tolerance; U
and Vt are isometries over the bond; each singular vector of a simple
singular value matches the oracle's up to a phase.  simple_update replays
the known answers of the reference's test/unit/operations/simple_update.jl
and compares random MPS bond updates (truncated and absorbed) with the
oracle through gauge-invariant products.
"""
import numpy as np
import pytest
import torch

import muscle
from oracle import backend_base as bb
from oracle import factor as of
from tests.gpu_util import dev_tensor, host, randn, rel_err

pytestmark = pytest.mark.gpu

TOLS = {np.float64: 1e-12, np.complex128: 1e-12, np.float32: 2e-5, np.complex64: 2e-5}
I = muscle.Index


def check_svd(A_np, inds, inds_u=(), inds_v=(), dtype=np.complex128, permute=None):
    tol = TOLS[dtype]
    A = dev_tensor(A_np.astype(dtype), [I(i) for i in inds])
    if permute is not None:  # a strided (non column-major) operand
        A = A.permutedims([I(inds[p]) for p in permute])
    U, S, Vt = muscle.tensor_svd_thin(A, inds_u=[I(i) for i in inds_u], inds_v=[I(i) for i in inds_v],
                                      ind_s=I("x"))
    A_np = A_np.astype(dtype).astype(np.complex128)  # the oracle factors the same rounded values
    (Uo, iu), (so, _), (Vo, iv) = of.tensor_svd_thin_base(A_np, inds, inds_u, inds_v, "x")
    assert U.inds() == [I(i) for i in iu] and Vt.inds() == [I(i) for i in iv] and S.inds() == [I("x")]
    u, s, v = host(U), host(S), host(Vt)
    assert u.shape == Uo.shape and v.shape == Vo.shape and s.shape == so.shape
    assert rel_err(s, so) <= tol
    back = bb.binary_einsum_general(list(inds), u * s, iu, v, iv)
    assert rel_err(back, A_np) <= 4 * tol
    k = s.shape[0]
    for t, n in ((u, u.size // k), (v, v.size // k)):
        m = t.reshape(n, k, order="F")
        assert np.abs(m.conj().T @ m - np.eye(k)).max() <= 10 * tol * max(1, np.sqrt(n))
    # singular vectors of well separated singular values: the oracle's up to a phase
    gap = np.minimum(np.r_[np.inf, so[:-1] - so[1:]], np.r_[so[:-1] - so[1:], np.inf])
    for x in np.nonzero(gap > 1e-3 * so[0])[0][:8]:
        for got, ref in ((u[..., x], Uo[..., x]), (v[..., x], Vo[..., x])):
            assert abs(abs(np.vdot(ref.ravel(), got.ravel())) - 1) <= 100 * tol
    return U, S, Vt


@pytest.mark.parametrize("dtype", [np.complex128, np.float64, np.complex64, np.float32])
def test_svd_thin_matches_oracle(dtype):
    rng = np.random.default_rng(11)
    A = randn(rng, (2, 4, 6, 8), dtype=np.complex128 if np.iscomplexobj(dtype(0)) else np.float64)
    check_svd(A, ["i", "j", "k", "l"], inds_u=["i", "j"], dtype=dtype)
    check_svd(A, ["i", "j", "k", "l"], inds_v=["j", "l"], dtype=dtype)  # interleaved split


@pytest.mark.parametrize("shape,inds_u", [((300, 40), ["a"]), ((40, 300), ["a"]), ((8, 16, 2, 16), ["c", "a"]),
                                          ((64, 2, 64), ["a", "b"]), ((1, 7, 5), ["a"])])
def test_svd_thin_shapes(shape, inds_u):
    rng = np.random.default_rng(len(shape) * 100 + shape[0])
    inds = ["a", "b", "c", "d"][:len(shape)]
    check_svd(randn(rng, shape), inds, inds_u=inds_u)


def test_svd_thin_strided_operand():
    rng = np.random.default_rng(3)
    A = randn(rng, (6, 5, 4))
    check_svd(A, ["p", "q", "r"], inds_u=["r", "p"], permute=[2, 0, 1])


def test_svd_thin_bond_256():
    """An MPS bond of dimension 256 and physical dimension 2: theta is 512 x 512."""
    rng = np.random.default_rng(4)
    check_svd(randn(rng, (256, 2, 2, 256)), ["l", "p", "q", "r"], inds_u=["l", "p"])


def isometry_err(t, k):
    m = t.reshape(-1, k, order="F")
    return float(np.abs(m.conj().T @ m - np.eye(k)).max())


@pytest.mark.parametrize("dtype", [np.complex128, np.float64])
def test_svd_thin_rank_deficient(dtype):
    """Rank 8 in 64 x 48 (a null space the vectors must still span orthonormally)
    and a spectrum falling from 1 to 1e-10 (vectors of tiny singular values)."""
    rng = np.random.default_rng(31)
    x, y = randn(rng, (64, 8), dtype), randn(rng, (8, 48), dtype)
    for A_np in (x @ y, (x @ y).T.copy()):
        A = dev_tensor(A_np, [I("a"), I("b")])
        U, S, Vt = muscle.tensor_svd_thin(A, inds_u=[I("a")], ind_s=I("x"))
        u, s, v = host(U), host(S), host(Vt)
        so = np.linalg.svd(A_np, compute_uv=False)
        assert np.abs(s - so).max() <= 1e-13 * so[0]
        assert isometry_err(u, 48) <= 1e-12 and isometry_err(v, 48) <= 1e-12
        assert rel_err((u * s) @ v.T, A_np) <= 1e-13
    q1 = np.linalg.qr(randn(rng, (96, 80), dtype))[0]
    q2 = np.linalg.qr(randn(rng, (80, 80), dtype))[0]
    sig = 10.0 ** (-np.arange(80) / 8)
    A_np = (q1 * sig) @ q2.conj().T
    A = dev_tensor(A_np, [I("a"), I("b")])
    U, S, Vt = muscle.tensor_svd_thin(A, inds_u=[I("a")], ind_s=I("x"))
    u, s, v = host(U), host(S), host(Vt)
    assert np.abs(s - sig).max() <= 1e-14
    assert isometry_err(u, 80) <= 1e-12 and isometry_err(v, 80) <= 1e-12
    assert rel_err((u * s) @ v.T, A_np) <= 1e-13
    # the leading vectors (relative gaps ~ 0.25) match the exact ones up to a phase
    for x_ in range(10):
        assert abs(abs(np.vdot(q1[:, x_], u[:, x_])) - 1) <= 1e-12
        assert abs(abs(np.vdot(q2[:, x_].conj(), v[:, x_])) - 1) <= 1e-12


def test_svd_thin_zero_matrix():
    A = dev_tensor(np.zeros((5, 3)), [I("a"), I("b")])
    U, S, Vt = muscle.tensor_svd_thin(A, inds_u=[I("a")], ind_s=I("x"))
    assert np.array_equal(host(S), np.zeros(3))
    assert isometry_err(host(U), 3) <= 1e-14 and isometry_err(host(Vt), 3) <= 1e-14


def test_svd_thin_argument_errors():
    A = dev_tensor(randn(np.random.default_rng(0), (2, 4, 6, 8)), [I(i) for i in "ijkl"])
    for kw in ({}, {"inds_u": [I("z")]}, {"inds_v": [I("z")]}, {"inds_u": A.inds()}, {"inds_v": A.inds()},
               {"inds_u": [I("i")], "ind_s": I("j")}):
        with pytest.raises(muscle.ArgumentError):
            muscle.tensor_svd_thin(A, **kw)


R2 = np.sqrt(2.0)
OP_ID = np.eye(4).reshape(2, 2, 2, 2, order="F")
OP_CX = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], float).reshape(2, 2, 2, 2, order="F")
S1, S2, BOND, G1, G2 = I(("site", 1, "cut", 1)), I(("site", 2, "cut", 1)), I(("bond", 1, 2)), \
    I(("site", 1, "cut", 2)), I(("site", 2, "cut", 2))


def gpu_su(op, **kw):
    ga = dev_tensor(np.eye(2), [S1, BOND])
    gb = dev_tensor(np.eye(2), [S2, BOND])
    g = dev_tensor(op, [S1, S2, G1, G2])
    return muscle.simple_update(ga, S1, gb, S2, BOND, g, G1, G2, **kw)


def same_up_to_sign(got, ref, tol=1e-12):
    """Columns (last axis) equal up to a sign each."""
    for x in range(ref.shape[-1]):
        g, r = got[..., x].ravel(), ref[..., x].ravel()
        if np.abs(g - r).max() > tol and np.abs(g + r).max() > tol:
            return False
    return True


def test_simple_update_known_answers():
    U, s, V = gpu_su(OP_ID)
    assert U.inds() == [S1, BOND] and V.inds() == [S2, BOND] and s.inds() == [BOND]
    assert same_up_to_sign(host(U), np.eye(2)) and same_up_to_sign(host(V), np.eye(2))
    assert np.allclose(host(s), [1, 1], atol=1e-15)
    U, s, V = gpu_su(OP_CX)
    assert same_up_to_sign(host(U), np.eye(2))
    assert same_up_to_sign(host(V), np.array([[1, -1], [1, 1]]) / R2)
    assert np.allclose(host(s), [R2, 0], atol=1e-14)
    U, s, V = gpu_su(OP_ID, normalize=True)
    assert np.allclose(host(s), [1 / R2, 1 / R2], atol=1e-15)
    U, s, V = gpu_su(OP_CX, normalize=True)
    assert np.allclose(host(s), [1, 0], atol=1e-14)
    U, s, V = gpu_su(OP_CX, maxdim=1)
    assert U.size(BOND) == V.size(BOND) == s.size(BOND) == 1
    assert same_up_to_sign(host(V), np.array([[1], [1]]) / R2) and np.allclose(host(s), [R2])
    U, V = gpu_su(OP_ID, absorb=muscle.AbsorbU())
    assert np.isclose(np.linalg.norm(host(U)), R2)
    U, V = gpu_su(OP_ID, absorb=muscle.AbsorbV())
    assert np.isclose(np.linalg.norm(host(V)), R2)
    U, V = gpu_su(OP_ID, absorb=muscle.AbsorbEqually())
    assert same_up_to_sign(host(U), np.eye(2)) and same_up_to_sign(host(V), np.eye(2))


@pytest.mark.parametrize("chi,d,maxdim,absorb", [(16, 2, None, None), (32, 2, 32, "equally"), (24, 3, 20, "u"),
                                                 (64, 2, 48, "v")])
def test_simple_update_random_bond(chi, d, maxdim, absorb):
    rng = np.random.default_rng(chi + d)
    a = randn(rng, (chi, d, chi))  # (left, physical, bond)
    b = randn(rng, (chi, d, chi))  # (bond, physical, right)
    g = randn(rng, (d, d, d, d))
    ia, ib, ig = ["l", "pa", "bond"], ["bond", "pb", "r"], ["pa", "pb", "ga", "gb"]
    ora = of.simple_update_base(a, ia, "pa", b, ib, "pb", "bond", g, ig, "ga", "gb", normalize=True,
                                absorb=absorb or "none", maxdim=maxdim)
    A, B, G = (dev_tensor(x, [I(i) for i in li]) for x, li in ((a, ia), (b, ib), (g, ig)))
    kind = {None: None, "u": muscle.AbsorbU(), "v": muscle.AbsorbV(), "equally": muscle.AbsorbEqually()}[absorb]
    got = muscle.simple_update(A, I("pa"), B, I("pb"), I("bond"), G, I("ga"), I("gb"), normalize=True,
                               absorb=kind, maxdim=maxdim)
    out = ["l", "pa", "pb", "r"]
    if absorb is None:
        (Uo, iu), (so, _), (Vo, iv) = ora
        U, S, V = got
        assert rel_err(host(S), so) <= 1e-12
        ref = bb.binary_einsum_general(out, Uo * so, iu, Vo, iv)
        prod = bb.binary_einsum_general(out, host(U) * host(S), iu, host(V), iv)
    else:
        (Uo, iu), (Vo, iv) = ora
        U, V = got
        ref = bb.binary_einsum_general(out, Uo, iu, Vo, iv)
        prod = bb.binary_einsum_general(out, host(U), iu, host(V), iv)
    assert U.inds() == [I(i) for i in iu] and V.inds() == [I(i) for i in iv]
    assert U.size(I("bond")) == min(maxdim or 10 ** 9, chi * d)
    assert rel_err(prod, ref) <= 1e-11


def check_qr(A_np, inds, inds_q=(), inds_r=(), dtype=np.complex128):
    tol = TOLS[dtype]
    A = dev_tensor(A_np.astype(dtype), [I(i) for i in inds])
    Q, R = muscle.tensor_qr_thin(A, inds_q=[I(i) for i in inds_q], inds_r=[I(i) for i in inds_r],
                                 ind_virtual=I("x"))
    A_np = A_np.astype(dtype).astype(np.complex128)
    (Qo, iq), (Ro, ir) = of.tensor_qr_thin_base(A_np, inds, inds_q, inds_r, "x")
    assert Q.inds() == [I(i) for i in iq] and R.inds() == [I(i) for i in ir]
    q, r = host(Q), host(R)
    assert q.shape == Qo.shape and r.shape == Ro.shape
    assert rel_err(bb.binary_einsum_general(list(inds), q, iq, r, ir), A_np) <= 4 * tol
    k = r.shape[0]
    qm, rm = q.reshape(-1, k, order="F"), r.reshape(k, -1, order="F")
    assert np.abs(qm.conj().T @ qm - np.eye(k)).max() <= 10 * tol * max(1, np.sqrt(qm.shape[0]))
    assert np.count_nonzero(np.tril(rm, -1)) == 0  # exact zeros below the diagonal
    # LAPACK's Householder convention on both sides: the same Q and R
    assert rel_err(r, Ro) <= 100 * tol and rel_err(q, Qo) <= 100 * tol


@pytest.mark.parametrize("dtype", [np.complex128, np.float64, np.complex64, np.float32])
def test_qr_thin_matches_oracle(dtype):
    rng = np.random.default_rng(21)
    A = randn(rng, (2, 4, 6, 8), dtype=np.complex128 if np.iscomplexobj(dtype(0)) else np.float64)
    check_qr(A, ["i", "j", "k", "l"], inds_q=["i", "j"], dtype=dtype)
    check_qr(A, ["i", "j", "k", "l"], inds_r=["i", "k"], dtype=dtype)


@pytest.mark.parametrize("shape,inds_q", [((300, 40), ["a"]), ((40, 300), ["a"]), ((256, 2, 256), ["a", "b"])])
def test_qr_thin_shapes(shape, inds_q):
    rng = np.random.default_rng(shape[0] + shape[-1])
    check_qr(randn(rng, shape), ["a", "b", "c"][:len(shape)], inds_q=inds_q)


def test_qr_thin_argument_errors():
    A = dev_tensor(randn(np.random.default_rng(0), (2, 4, 6, 8)), [I(i) for i in "ijkl"])
    for kw in ({}, {"inds_q": [I("z")]}, {"inds_r": [I("z")]}, {"inds_q": A.inds()}, {"inds_r": A.inds()},
               {"inds_q": [I("i")], "ind_virtual": I("j")}):
        with pytest.raises(muscle.ArgumentError):
            muscle.tensor_qr_thin(A, **kw)


def test_svd_trunc_matches_oracle():
    rng = np.random.default_rng(22)
    A_np = randn(rng, (200, 100))
    A = dev_tensor(A_np, [I("i"), I("j")])
    _, sfull, _ = muscle.tensor_svd_trunc(A, inds_u=[I("i")], ind_s=I("x"))
    assert sfull.size(I("x")) == 100
    for kw in ({"maxdim": 20}, {"threshold": 1e-3, "maxdim": 80}, {"threshold": 0.05}):
        U, s, Vt = muscle.tensor_svd_trunc(A, inds_u=[I("i")], ind_s=I("x"), **kw)
        (Uo, iu), (so, _), (Vo, iv) = of.tensor_svd_trunc_base(A_np, ["i", "j"], ["i"], (), "x", **kw)
        assert s.size(I("x")) == so.shape[0] and U.size(I("x")) == Vt.size(I("x")) == so.shape[0]
        assert rel_err(host(s), so) <= 1e-12
        back = bb.binary_einsum_general(["i", "j"], host(U) * host(s), iu, host(Vt), iv)
        assert np.isclose(np.linalg.norm(back - A_np), np.linalg.norm(host(sfull)[so.shape[0]:]), rtol=1e-9)


def test_in_place_forms():
    rng = np.random.default_rng(23)
    A_np = randn(rng, (6, 5, 4))
    A = dev_tensor(A_np, [I("p"), I("q"), I("r")])
    dev = A.device
    U = muscle.Tensor(torch.empty((4, 6, 5), dtype=torch.complex128, device=dev), [I("r"), I("p"), I("x")])
    S = muscle.Tensor(torch.empty((5,), dtype=torch.float64, device=dev), [I("x")])
    V = muscle.Tensor(torch.empty((5, 5), dtype=torch.complex128, device=dev), [I("q"), I("x")])
    muscle.tensor_svd_thin_(U, S, V, A)  # row-major (torch default) outputs
    (Uo, iu), (so, _), (Vo, iv) = of.tensor_svd_thin_base(A_np, ["p", "q", "r"], ["r", "p"], ["q"], "x")
    assert rel_err(host(S), so) <= 1e-12
    back = bb.binary_einsum_general(["p", "q", "r"], host(U) * host(S), iu, host(V), iv)
    assert rel_err(back, A_np) <= 1e-12
    Q = muscle.Tensor(torch.empty((5, 4, 20), dtype=torch.complex128, device=dev), [I("q"), I("r"), I("y")])
    R = muscle.Tensor(torch.empty((6, 6), dtype=torch.complex128, device=dev).t()[:, :], [I("y"), I("p")])
    with pytest.raises(muscle.DimensionMismatch):  # the bond must be min(20, 6) = 6
        muscle.tensor_qr_thin_(Q, R, A)
    Q = muscle.Tensor(torch.empty((5, 4, 6), dtype=torch.complex128, device=dev), [I("q"), I("r"), I("y")])
    muscle.tensor_qr_thin_(Q, R, A)
    (Qo, iq), (Ro, ir) = of.tensor_qr_thin_base(A_np, ["p", "q", "r"], ["q", "r"], ["p"], "y")
    assert rel_err(host(R), Ro) <= 1e-10 and rel_err(host(Q), Qo) <= 1e-10


@pytest.mark.parametrize("shape", [(1, 1), (5, 1), (1, 7), (300, 1)])
def test_factor_degenerate_shapes(shape):
    """k = 1 factorizations: a column, a row, a single element, a tall vector
    (QR-reduced before the eigensolver)."""
    rng = np.random.default_rng(41)
    A_np = randn(rng, shape)
    A = dev_tensor(A_np, [I("a"), I("b")])
    U, S, Vt = muscle.tensor_svd_thin(A, inds_u=[I("a")], ind_s=I("x"))
    u, s, v = host(U), host(S), host(Vt)
    assert s.shape == (1,) and np.isclose(s[0], np.linalg.norm(A_np), rtol=1e-13)
    assert rel_err((u * s) @ v.T, A_np) <= 1e-13
    Q, R = muscle.tensor_qr_thin(A, inds_q=[I("a")], ind_virtual=I("y"))
    q, r = host(Q), host(R)
    assert rel_err(q @ r, A_np) <= 1e-13 and np.isclose(np.linalg.norm(q), 1.0)


def test_simple_update_maxdim_above_rank():
    rng = np.random.default_rng(42)
    a, b = randn(rng, (4, 2, 3)), randn(rng, (3, 2, 5))
    g = randn(rng, (2, 2, 2, 2))
    A = dev_tensor(a, [I("l"), I("pa"), I("bond")])
    B = dev_tensor(b, [I("bond"), I("pb"), I("r")])
    G = dev_tensor(g, [I("pa"), I("pb"), I("ga"), I("gb")])
    U, S, V = muscle.simple_update(A, I("pa"), B, I("pb"), I("bond"), G, I("ga"), I("gb"), maxdim=100)
    assert S.size(I("bond")) == 8  # min(4 * 2, 2 * 5)
    (Uo, iu), (so, _), (Vo, iv) = of.simple_update_base(a, ["l", "pa", "bond"], "pa", b, ["bond", "pb", "r"], "pb",
                                                        "bond", g, ["pa", "pb", "ga", "gb"], "ga", "gb", maxdim=100)
    assert rel_err(host(S), so) <= 1e-12
