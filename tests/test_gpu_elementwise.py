"""hadamard / hadamard! and unary_einsum / unary_einsum! on the GPU against
the CPU oracle and the reference's known answers."""
import os

import numpy as np
import pytest
import torch

import muscle
from oracle import backend_base as bb
from tests import kats
from tests.gpu_util import dev_tensor, host, randn, rel_err

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------ hadamard
@pytest.mark.parametrize("case", kats.hadamard_kats(), ids=lambda c: c[0])
def test_hadamard_kats(case):
    name, a_np, ia, b_np, ib, expected = case
    a, b = dev_tensor(a_np, ia), dev_tensor(b_np, ib)
    c = muscle.hadamard(a, b)
    assert c.inds() == a.inds() and c.shape == a.shape
    assert np.array_equal(host(c), expected)
    # hadamard!(a, a, b) returns the same object (test/unit/operations/hadamard.jl:14-17)
    a2 = dev_tensor(a_np, ia)
    out = muscle.hadamard_(a2, a2, b)
    assert out is a2
    assert np.array_equal(host(a2), expected)
    # argument order does not matter (hadamard.jl:9)
    assert np.array_equal(host(muscle.hadamard(b, a)), expected)


HCASES = [
    (["i", "j", "k"], (2, 3, 4), ["k", "i", "j"]),  # 3-cycle
    (["i", "j", "k"], (2, 3, 4), ["j", "k", "i"]),
    (["i", "j", "k", "l"], (9, 10, 11, 12), ["l", "j"]),
    (["i", "j", "k"], (64, 65, 33), ["j"]),
    (["i", "j"], (1000, 999), ["j", "i"]),
    (["i", "j", "k"], (128, 64, 32), ["i", "j", "k"]),
]


@pytest.mark.parametrize("ta, tb", [(np.complex128, np.complex128), (np.complex128, np.float64),
                                    (np.float64, np.float64), (np.float32, np.float64),
                                    (np.complex64, np.complex128), (np.complex64, np.float32)])
@pytest.mark.parametrize("case", HCASES, ids=lambda c: "".join(c[0]) + "*" + "".join(c[2]))
def test_hadamard_parity(case, ta, tb):
    ia, sa, ib = case
    rng = np.random.default_rng(31)
    size = dict(zip(ia, sa))
    A = randn(rng, sa, ta)
    B = randn(rng, tuple(size[i] for i in ib), tb)
    ref, _ = bb.hadamard_base(A, ia, B, ib)
    c = muscle.hadamard(dev_tensor(A, ia), dev_tensor(B, ib))
    C = host(c)
    assert C.dtype == ta  # a's eltype (copy(a), hadamard.jl:38)
    # one product per element: FP64 results are identical up to the final rounding
    tol = 1e-15 if ta in (np.float64, np.complex128) else 1e-6
    assert rel_err(C, ref) <= tol


def test_hadamard_golden_3cycle():
    g = np.load(os.path.join(GOLDEN, "hadamard_c64.npz"))
    c = muscle.hadamard(dev_tensor(g["a"], ["i", "j", "k"]), dev_tensor(g["b"], ["k", "i", "j"]))
    assert rel_err(host(c), g["c"]) <= 1e-15


def test_hadamard_errors():
    a = dev_tensor(np.ones((2, 3)), ["i", "j"])
    with pytest.raises(muscle.ArgumentError):
        muscle.hadamard(a, dev_tensor(np.ones(4), ["x"]))  # inds(b) not a subset
    c = muscle.Tensor(torch.zeros((3, 2), dtype=torch.float64, device="cuda"), ["j", "i"])
    with pytest.raises(muscle.ArgumentError):
        muscle.hadamard_(c, a, dev_tensor(np.ones(2), ["i"]))  # inds(c) != inds(a)
    with pytest.raises(muscle.ArgumentError):
        muscle.hadamard(a, dev_tensor(np.ones(2, dtype=np.complex128), ["i"]))  # complex into real a


def test_hadamard_strided_output_view():
    rng = np.random.default_rng(3)
    A, B = randn(rng, (50, 60)), randn(rng, (60,))
    base = torch.zeros((100, 60), dtype=torch.complex128, device="cuda")
    c = muscle.Tensor(base[::2], ["i", "j"])
    muscle.hadamard_(c, dev_tensor(A, ["i", "j"]), dev_tensor(B, ["j"]))
    assert rel_err(host(c), A * B[None, :]) <= 1e-15
    assert torch.count_nonzero(base[1::2]).item() == 0


# ------------------------------------------------------------ unary_einsum
@pytest.mark.parametrize("case", kats.unary_kats(), ids=lambda c: c[0])
def test_unary_kats(case):
    name, x_np, ix, kw, out_inds, expected = case
    y = muscle.unary_einsum(dev_tensor(x_np, ix), **kw)
    assert [i.tag for i in y.inds()] == out_inds
    assert np.array_equal(host(y), expected)


@pytest.mark.parametrize("case", kats.unary_inplace_kats(), ids=lambda c: c[0])
def test_unary_inplace_kats(case):
    name, x_np, ix, iy, expected = case
    y = dev_tensor(np.zeros(expected.shape), iy)
    out = muscle.unary_einsum_(y, dev_tensor(x_np, ix))
    assert out is y
    assert np.array_equal(host(y), expected)


UCASES = [
    (["i", "j", "k"], (33, 44, 55), ["k", "i"]),
    (["i", "j", "i"], (30, 40, 30), ["j"]),
    (["i", "j", "i"], (30, 40, 30), ["i", "j"]),
    (["i", "i"], (500, 500), []),
    (["a", "b", "a", "c", "b"], (6, 7, 6, 8, 7), ["c", "a"]),
    (["i", "j"], (300, 400), ["j", "i"]),
    (["i", "j"], (2048, 2048), []),  # full sum, split over workgroups
    (["i", "j", "k"], (256, 8, 256), ["j"]),
]


@pytest.mark.parametrize("T", [np.complex128, np.float64, np.complex64, np.float32])
@pytest.mark.parametrize("case", UCASES, ids=lambda c: "".join(c[0]) + "->" + "".join(c[2]))
def test_unary_parity(case, T):
    ix, sx, iy = case
    rng = np.random.default_rng(17)
    X = randn(rng, sx, T)
    ref = bb.unary_einsum_base(iy, X.astype(np.complex128), ix)
    y = muscle.unary_einsum(dev_tensor(X, ix), out=iy)
    Y = host(y)
    assert Y.dtype == T
    tol = 1e-13 if T in (np.float64, np.complex128) else 1e-5
    assert rel_err(Y, ref) <= tol


HTCASES = [
    (["i", "j", "k"], (64, 40, 48), ["k", "i", "j"]),
    (["i", "j", "k"], (128, 8, 160), ["k", "i"]),  # b broadcast along j, runs along k
    (["i", "j", "k", "l"], (33, 17, 29, 5), ["l", "k", "j", "i"]),  # ragged
    ([f"x{i}" for i in range(16)], (2,) * 16, list(np.random.default_rng(4).permutation([f"x{i}" for i in range(16)]))),
]


@pytest.mark.parametrize("ta, tb", [(np.complex128, np.complex128), (np.complex128, np.float64),
                                    (np.float32, np.float32), (np.complex64, np.complex128)])
@pytest.mark.parametrize("case", HTCASES, ids=lambda c: "".join(c[0])[:12] + "*" + "".join(c[2])[:12])
def test_hadamard_tiled(case, ta, tb):
    """b (>= 2^14 elements) running contiguous along other modes than a and c: the 32 x 32
    tile path.  Elementwise products are exact against the oracle's in the same precision."""
    ia, sa, ib = case
    rng = np.random.default_rng(23)
    A = randn(rng, sa, ta)
    sb = tuple(sa[ia.index(i)] for i in ib)
    B = randn(rng, sb, tb)
    ref, _ = bb.hadamard_base(A.astype(np.complex128), ia, B.astype(np.complex128), ib)
    c = muscle.hadamard(dev_tensor(A, ia), dev_tensor(B, ib))
    got = host(c)
    assert got.dtype == ta
    tol = 1e-15 if ta == np.complex128 and tb in (np.complex128, np.float64) else 1e-6
    assert rel_err(got, ref) <= tol


DIMS2 = [f"x{i}" for i in range(16)]
PCASES = [
    (["i", "j", "k"], (64, 40, 48), ["k", "j", "i"]),
    (["i", "j", "k"], (64, 40, 48), ["j", "k", "i"]),
    (["i", "j", "k"], (33, 1, 997), ["k", "j", "i"]),  # ragged, an extent-1 mode
    (["i", "j", "i", "k"], (40, 50, 40, 30), ["k", "i", "j"]),  # diagonal and permutation
    (DIMS2, (2,) * 16, list(np.random.default_rng(3).permutation(DIMS2))),  # dims 2, random order
]


@pytest.mark.parametrize("T", [np.complex128, np.float64, np.complex64, np.float32])
@pytest.mark.parametrize("case", PCASES, ids=lambda c: "".join(c[0])[:12] + "->" + "".join(c[2])[:12])
def test_unary_permutation_tiled(case, T):
    """Pure permutations (no summed modes) of >= 2^14 elements whose input and output
    run contiguous along different modes go to the tiled transpose (kernel 6): a copy,
    so the result is bit-exact."""
    ix, sx, iy = case
    X = randn(np.random.default_rng(19), sx, T)
    ref = bb.unary_einsum_base(iy, X, ix)
    y = muscle.unary_einsum(dev_tensor(X, ix), out=iy)
    assert muscle.last_info()["kernel"] == 6
    assert np.array_equal(host(y), ref)


def test_unary_permutation_into_strided_view():
    X = randn(np.random.default_rng(20), (48, 36, 20), np.complex128)
    big = torch.zeros((36, 3, 48, 20), dtype=torch.complex128, device="cuda")
    yv = muscle.Tensor(big[:, 1, :, :], ["j", "i", "k"])  # a strided (non-contiguous) output view
    muscle.unary_einsum_(yv, dev_tensor(X, ["i", "j", "k"]))
    assert muscle.last_info()["kernel"] == 6  # y runs contiguous along k, x along i
    got = big.cpu().numpy()
    assert np.array_equal(got[:, 1], np.transpose(X, (1, 0, 2)))
    assert not got[:, 0].any() and not got[:, 2].any()


@pytest.mark.parametrize("T", [np.complex128, np.float32])
@pytest.mark.parametrize("case", [(["i", "j", "k"], (256, 64, 128), ["i", "j"]),
                                  (["i", "j", "k"], (128, 64, 33), ["j", "i"]),
                                  (["i", "j", "k"], (4, 8, 4096), ["j", "i"])],
                         ids=lambda c: "".join(c[0]) + "->" + "".join(c[2]))
def test_unary_sum_over_slow_modes(case, T):
    """Sums over modes the input holds at large strides (one lane per output)."""
    ix, sx, iy = case
    X = randn(np.random.default_rng(21), sx, T)
    ref = bb.unary_einsum_base(iy, X.astype(np.complex128), ix)
    y = muscle.unary_einsum(dev_tensor(X, ix), out=iy)
    tol = 1e-13 if T in (np.float64, np.complex128) else 1e-5
    assert rel_err(host(y), ref) <= tol


def test_unary_golden():
    g = np.load(os.path.join(GOLDEN, "unary_c64.npz"))
    y = muscle.unary_einsum(dev_tensor(g["x"], ["i", "j", "i", "k"]), out=["k", "j"])
    assert rel_err(host(y), g["y"]) <= 1e-14


def test_hyper_golden():
    g = np.load(os.path.join(GOLDEN, "hyper_c64.npz"))
    c = muscle.binary_einsum(dev_tensor(g["a"], ["i", "j", "k"]), dev_tensor(g["b"], ["k", "l", "j"]), dims=["j"])
    assert [i.tag for i in c.inds()] == ["i", "k", "l"]
    assert rel_err(host(c), g["c"]) <= 1e-13


def test_unary_errors():
    x = dev_tensor(np.ones((2, 3)), ["i", "j"])
    with pytest.raises(muscle.ArgumentError):
        muscle.unary_einsum(x, out=["x"])


def test_hadamard_tiled_strided_views():
    """a and c non-contiguous views (i fast, strided columns), b with k fast."""
    rng = np.random.default_rng(25)
    X = randn(rng, (130, 3, 120), np.complex128)
    B = randn(rng, (120, 130), np.complex128)
    a = muscle.Tensor(dev_tensor(X, ["i", "s", "k"]).parent()[:, 1, :], ["i", "k"])
    big = torch.zeros((120, 2, 130), dtype=torch.complex128, device="cuda")
    c = muscle.Tensor(big[:, 1, :].t(), ["i", "k"])
    muscle.hadamard_(c, a, dev_tensor(B, ["k", "i"]))
    got = big.cpu().numpy()
    assert rel_err(got[:, 1, :].T, X[:, 1, :] * B.T) <= 1e-15
    assert not got[:, 0, :].any()
