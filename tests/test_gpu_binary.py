"""binary_einsum / binary_einsum! on the GPU against the CPU oracle.

Bar: exact for the reference's `ones` known answers; norm-wise relative error
<= 1e-12 for Float64/ComplexF64 (the FP64 MFMA path), <= 2e-5 for the FP32
path, on seeded inputs.  Full-size config-2 results are checked on sampled
entries against the oracle's dot products and through a size-independent
identity (sum over n of C equals A times the row sums of B).
"""
import zlib

import numpy as np
import pytest
import torch

import muscle
from oracle import backend_base as bb
from tests import kats
from tests.gpu_util import TOL, dev_tensor, host, randn, rel_err

pytestmark = pytest.mark.gpu


def _run(A, ia, B, ib, **kw):
    a, b = dev_tensor(A, ia), dev_tensor(B, ib)
    c = muscle.binary_einsum(a, b, **kw)
    return c, host(c)


# ------------------------------------------------------------ known answers
@pytest.mark.parametrize("case", kats.binary_kats(), ids=lambda c: c[0])
def test_binary_kats(case):
    name, A, ia, B, ib, kw, out_inds, val, base_ok = case
    c, C = _run(A, ia, B, ib, **kw)
    assert [i.tag for i in c.inds()] == out_inds
    assert C.shape == val.shape
    assert np.array_equal(C, val), (name, C, val)


@pytest.mark.parametrize("T", [np.float64, np.complex128])
def test_manual_kat(T):
    A, ia, B, ib = kats.manual_case(T)
    c, C = _run(A, ia, B, ib, dims=["j", "k"])
    assert [i.tag for i in c.inds()] == ["i", "l"]
    assert np.allclose(C, kats.manual_expected_full(A, B))
    c, C = _run(A, ia, B, ib, dims=["j"])
    assert [i.tag for i in c.inds()] == ["i", "k", "l"]
    assert np.allclose(C, kats.manual_expected_hyper(A, B))


# ------------------------------------------------------------ seeded parity
CASES = [
    # name, inds_a, shape_a, inds_b, shape_b, inds_c
    ("config1", ["i", "j", "k"], (64, 32, 48), ["k", "l", "j"], (48, 96, 32), ["i", "l"]),
    ("config1 out^T", ["i", "j", "k"], (64, 32, 48), ["k", "l", "j"], (48, 96, 32), ["l", "i"]),
    ("ragged", ["i", "j"], (67, 45), ["j", "k"], (45, 93), ["i", "k"]),
    ("ragged2", ["i", "j"], (33, 17), ["k", "j"], (129, 17), ["k", "i"]),
    ("split-K ragged", ["i", "j"], (70, 1001), ["j", "k"], (1001, 50), ["i", "k"]),
    ("split-K batched", ["i", "j", "h"], (40, 777, 3), ["j", "k", "h"], (777, 33, 3), ["k", "h", "i"]),
    ("k-major A", ["j", "i"], (40, 70), ["j", "k"], (40, 50), ["i", "k"]),
    ("interleaved d=4", ["a1", "c1", "a3", "c2", "a5", "c3"], (4,) * 6, ["c1", "b1", "c2", "b2", "c3", "b3"], (4,) * 6,
     ["a1", "b1", "a3", "b2", "a5", "b3"]),
    ("interleaved d=8", ["a1", "c1", "a3", "c2", "a5", "c3"], (8,) * 6, ["c1", "b1", "c2", "b2", "c3", "b3"], (8,) * 6,
     ["a1", "b1", "a3", "b2", "a5", "b3"]),
    ("hyper", ["a1", "c1", "h", "a2", "c2"], (4, 4, 8, 8, 8), ["c1", "b1", "h", "c2", "b2"], (4, 4, 8, 8, 8),
     ["a1", "b1", "h", "a2", "b2"]),
    ("hyper batch-first out", ["i", "j", "h"], (40, 36, 3), ["j", "k", "h"], (36, 50, 3), ["h", "k", "i"]),
    ("matvec", ["i", "j"], (300, 200), ["j"], (200,), ["i"]),
    ("vecmat", ["j"], (200,), ["j", "k"], (200, 300), ["k"]),
    ("inner", ["i", "j"], (120, 130), ["j", "i"], (130, 120), []),
    ("outer", ["i", "j"], (20, 30), ["k"], (40,), ["k", "i", "j"]),
    ("dims 2 tensors", ["a", "b", "c", "d", "e"], (2,) * 5, ["c", "d", "e", "f", "g"], (2,) * 5, ["g", "a", "f", "b"]),
    ("single-operand sum", ["i", "j", "s"], (30, 40, 5), ["j", "k"], (40, 35), ["i", "k"]),
    ("diagonal operand", ["i", "j", "i"], (20, 30, 20), ["j", "k"], (30, 40), ["i", "k"]),
]


@pytest.mark.parametrize("T", [np.complex128, np.float64, np.complex64, np.float32])
@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_binary_parity(case, T):
    name, ia, sa, ib, sb, ic = case
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    A, B = randn(rng, sa, T), randn(rng, sb, T)
    ref = bb.binary_einsum_general(ic, A.astype(np.complex128), ia, B.astype(np.complex128), ib)
    c, C = _run(A, ia, B, ib, out=ic)
    assert C.dtype == T
    assert rel_err(C, ref) <= TOL[T], (name, rel_err(C, ref))


@pytest.mark.parametrize("mult", [3, 4])
@pytest.mark.parametrize("T", [np.complex128, np.complex64])
def test_complex_3m_and_4m(mult, T):
    rng = np.random.default_rng(21)
    A, B = randn(rng, (130, 70, 3), T), randn(rng, (70, 3, 150), T)
    ref = bb.binary_einsum_base(["l", "i"], A.astype(np.complex128), ["i", "j", "k"], B.astype(np.complex128),
                                ["j", "k", "l"])
    with muscle.with_backend(muscle.BackendHIP(complex_mult=mult)):
        c = muscle.binary_einsum(dev_tensor(A, ["i", "j", "k"]), dev_tensor(B, ["j", "k", "l"]), out=["l", "i"])
    assert muscle.last_info()["kernel"] == 1
    assert rel_err(host(c), ref) <= TOL[T]


def test_config1_golden():
    """Config 1 against the committed golden output (tests/golden/config1_c64.npz)."""
    import os
    from muscle.workloads import config1_inputs
    A, ia, B, ib, dims, ic = config1_inputs(0)
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "config1_c64.npz"))
    assert np.allclose(A.ravel(order="F")[:8], g["a_head"], rtol=0, atol=0)
    c, C = _run(A, ia, B, ib, dims=dims)
    assert [i.tag for i in c.inds()] == ic
    err = rel_err(C, g["c"])
    assert err <= 1e-12, err
    info = muscle.last_info()
    assert info["kernel"] == 1 and info["flops"] == 75497472.0


@pytest.mark.parametrize("ta, tb", [(np.float64, np.complex128), (np.complex128, np.float64),
                                    (np.float32, np.float64), (np.float32, np.complex64),
                                    (np.complex64, np.float64)])
def test_mixed_eltypes_promote(ta, tb):
    rng = np.random.default_rng(4)
    A, B = randn(rng, (50, 40), ta), randn(rng, (40, 60), tb)
    c, C = _run(A, ["i", "j"], B, ["j", "k"])
    T = np.result_type(ta, tb)
    assert C.dtype == T
    ref = bb.binary_einsum_base(["i", "k"], A.astype(T), ["i", "j"], B.astype(T), ["j", "k"])
    assert rel_err(C, ref) <= TOL[T.type]


def test_strided_views():
    rng = np.random.default_rng(8)
    big = randn(rng, (80, 90), np.complex128)
    Bn = randn(rng, (45, 70), np.complex128)
    at = torch.from_numpy(np.asfortranarray(big)).cuda()[::2, 1::2]  # (40, 45) view, strides (2, 160)
    a = muscle.Tensor(at, ["i", "j"])
    b = dev_tensor(Bn, ["j", "k"])
    c = muscle.binary_einsum(a, b)
    ref = big[::2, 1::2] @ Bn
    assert rel_err(host(c), ref) <= 1e-12


def test_inplace_any_output_order():
    rng = np.random.default_rng(12)
    A, B = randn(rng, (30, 20, 10)), randn(rng, (10, 25, 20))
    a, b = dev_tensor(A, ["i", "j", "k"]), dev_tensor(B, ["k", "l", "j"])
    c = muscle.Tensor(torch.zeros((25, 30), dtype=torch.complex128, device="cuda"), ["l", "i"])
    out = muscle.binary_einsum_(c, a, b)
    assert out is c
    ref = bb.binary_einsum_base(["l", "i"], A, ["i", "j", "k"], B, ["k", "l", "j"])
    assert rel_err(host(c), ref) <= 1e-12


def test_empty_and_zero_extents():
    a = dev_tensor(np.ones((0, 3)), ["i", "j"])
    b = dev_tensor(np.ones((3, 4)), ["j", "k"])
    c = muscle.binary_einsum(a, b)
    assert c.shape == (0, 4)
    a = dev_tensor(np.ones((2, 0)), ["i", "j"])
    b = dev_tensor(np.ones((0, 4)), ["j", "k"])
    c = muscle.binary_einsum(a, b)
    assert np.array_equal(host(c), np.zeros((2, 4)))  # sum over an empty mode


def test_full_contraction_large_k():
    """Inner product to a 0-dim tensor over 2^22 elements: the split-K path."""
    rng = np.random.default_rng(13)
    A = randn(rng, (2048, 2048), np.complex128)
    B = randn(rng, (2048, 2048), np.complex128)
    c, C = _run(A, ["i", "j"], B, ["i", "j"])
    ref = np.sum(A * B)
    assert abs(C - ref) / abs(ref) < 1e-12


def test_errors_map_to_julia_exceptions():
    a = dev_tensor(np.ones((2, 3)), ["i", "j"])
    b = dev_tensor(np.ones((4, 5)), ["j", "k"])
    with pytest.raises(muscle.DimensionMismatch):
        muscle.binary_einsum(a, b)
    b = dev_tensor(np.ones((3, 5)), ["j", "k"])
    c = muscle.Tensor(torch.zeros((2, 5), dtype=torch.float32, device="cuda"), ["i", "k"])
    with pytest.raises(muscle.ArgumentError):
        muscle.binary_einsum_(c, a, b)  # eltype is not the promotion


# ------------------------------------------------------------ full-size config 2
@pytest.mark.slow
@pytest.mark.parametrize("T", [np.complex128, np.complex64])
def test_config2_full_size(T):
    from muscle.workloads import config2_inputs
    A, ia, B, ib, ic = config2_inputs(16, seed=1, dtype=T)
    a, b = dev_tensor(A, ia), dev_tensor(B, ib)
    c = muscle.binary_einsum(a, b, out=ic)
    C = host(c)
    # sampled entries against the oracle's dot products
    Am = np.reshape(np.transpose(A, (0, 2, 4, 1, 3, 5)), (4096, 4096), order="F").astype(np.complex128)
    Bm = np.reshape(np.transpose(B, (0, 2, 4, 1, 3, 5)), (4096, 4096), order="F").astype(np.complex128)
    Cm = np.reshape(np.transpose(C, (0, 2, 4, 1, 3, 5)), (4096, 4096), order="F")
    rng = np.random.default_rng(0)
    ms, ns = rng.integers(0, 4096, 64), rng.integers(0, 4096, 64)
    ref = np.array([Am[m, :] @ Bm[:, n] for m, n in zip(ms, ns)])
    got = Cm[ms, ns]
    assert rel_err(got, ref) <= TOL[T]
    # size-independent identity: sum_n C[m, n] = A[m, :] . (sum_n B[:, n])
    lhs = Cm.astype(np.complex128).sum(axis=1)
    rhs = Am @ Bm.sum(axis=1)
    assert rel_err(lhs, rhs) <= (1e-10 if T == np.complex128 else 1e-4)


# ------------------------------------------------------------ packed operands
def _dims2_case(seed, n_left, n_right, n_con, n_hyp=0):
    """Circuit-like contraction: every mode of extent 2, random mode orders."""
    rng = np.random.default_rng(seed)
    L = [f"l{i}" for i in range(n_left)]
    R = [f"r{i}" for i in range(n_right)]
    C = [f"c{i}" for i in range(n_con)]
    H = [f"h{i}" for i in range(n_hyp)]
    ia = list(rng.permutation(L + C + H))
    ib = list(rng.permutation(C + H + R))
    ic = list(rng.permutation(L + R + H))
    return ia, ib, ic


# 2^27+ index points and 2-element contiguous runs: the operands are packed
# K-contiguous before the GEMM
@pytest.mark.parametrize("T", [np.complex128, np.complex64, np.float64, np.float32])
@pytest.mark.parametrize("shape", [(9, 9, 9, 0), (10, 8, 9, 1), (11, 9, 8, 0)], ids=str)
def test_dims2_permuted_large(shape, T):
    ia, ib, ic = _dims2_case(sum(shape) * 7 + 1, *shape)
    rng = np.random.default_rng(21)
    A, B = randn(rng, (2,) * len(ia), T), randn(rng, (2,) * len(ib), T)
    ref = bb.binary_einsum_general(ic, A.astype(np.complex128), ia, B.astype(np.complex128), ib)
    c, got = _run(A, ia, B, ib, out=ic)
    assert rel_err(got, ref) <= TOL[T], rel_err(got, ref)
    c2, got2 = _run(A, ia, B, ib, out=ic)  # the cached plan
    assert np.array_equal(got, got2)


@pytest.mark.parametrize("ta, tb", [(np.float64, np.complex128), (np.complex64, np.float64)])
def test_dims2_permuted_mixed_eltypes(ta, tb):
    ia, ib, ic = _dims2_case(5, 9, 9, 9)
    rng = np.random.default_rng(22)
    A, B = randn(rng, (2,) * len(ia), ta), randn(rng, (2,) * len(ib), tb)
    ref = bb.binary_einsum_general(ic, A.astype(np.complex128), ia, B.astype(np.complex128), ib)
    c, got = _run(A, ia, B, ib, out=ic)
    assert rel_err(got, ref) <= 1e-12


# skinny products: one open side under 32 (the MFMA tiles run partly empty)
SKINNY = [
    (["i", "j"], (8, 1024), ["j", "k"], (1024, 4096), ["i", "k"]),
    (["i", "j"], (4096, 512), ["j", "k"], (512, 4), ["k", "i"]),
    (["j"], (2048,), ["j", "k"], (2048, 8192), ["k"]),
    (["i", "j", "h"], (16, 300, 3), ["j", "k", "h"], (300, 777, 3), ["h", "k", "i"]),
]


@pytest.mark.parametrize("T", [np.complex128, np.float32])
@pytest.mark.parametrize("case", SKINNY, ids=lambda c: "x".join(map(str, c[1] + c[3])))
def test_skinny_products(case, T):
    ia, sa, ib, sb, ic = case
    rng = np.random.default_rng(23)
    A, B = randn(rng, sa, T), randn(rng, sb, T)
    ref = bb.binary_einsum_general(ic, A.astype(np.complex128), ia, B.astype(np.complex128), ib)
    c, got = _run(A, ia, B, ib, out=ic)
    assert muscle.last_info()["kernel"] == 1
    assert rel_err(got, ref) <= TOL[T], rel_err(got, ref)


def test_handle_follows_device_forms_and_streams():
    """One handle per thread and device whichever way the device is named, and
    calls follow torch's current stream (the handle's stream is set again
    whenever it changes)."""
    from muscle.ops import _handle_for
    h = [_handle_for(0), _handle_for(torch.device("cuda", 0)), _handle_for(torch.device("cuda"))]
    assert len({x.value for x in h}) == 1
    rng = np.random.default_rng(31)
    A, B = randn(rng, (96, 80)), randn(rng, (80, 72))
    a, b = dev_tensor(A, ["i", "j"]), dev_tensor(B, ["j", "k"])
    ref = A @ B
    side = torch.cuda.Stream()
    torch.cuda.synchronize()  # the inputs' copies are done before the side stream reads them
    for stream in (side, torch.cuda.default_stream(), side):
        with torch.cuda.stream(stream):
            c = muscle.binary_einsum(a, b)
            got = c.parent().cpu().numpy()  # ordered after the call on this stream
        assert rel_err(got, ref) <= 1e-13


def test_stream_switch_orders_shared_workspace():
    """Calls of one handle on two streams, with no synchronization between them:
    both run split-K (their partial sums share the handle's workspace), so the
    call on the second stream must start only after the first stream's call.
    Repeated with alternating streams; every result against the oracle."""
    rng = np.random.default_rng(33)
    K = 1 << 17
    ops = []
    for _ in range(2):
        A, B = randn(rng, (128, K)), randn(rng, (K, 96))
        ops.append((dev_tensor(A, ["i", "j"]), dev_tensor(B, ["j", "k"]), A @ B))
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    outs = []
    for rep in range(6):
        stream = side if rep % 2 == 0 else torch.cuda.default_stream()
        a, b, ref = ops[rep % 2]
        with torch.cuda.stream(stream):
            c = muscle.binary_einsum(a, b)
            c.parent().record_stream(stream)
        outs.append((c, ref))
    torch.cuda.synchronize()
    for c, ref in outs:
        assert rel_err(c.parent().cpu().numpy(), ref) <= 1e-12


@pytest.mark.parametrize("T", [np.complex128, np.float32])
def test_contracted_extent_zero(T):
    """A contracted mode of extent 0 sums nothing: C is all zeros (C's memory
    starts as NaN, so every element is written)."""
    a = dev_tensor(np.zeros((3, 0), dtype=T), ["i", "j"])
    b = dev_tensor(np.zeros((0, 4), dtype=T), ["j", "k"])
    c = muscle.Tensor(torch.full((3, 4), float("nan"), dtype=a.dtype, device="cuda"), ["i", "k"])
    muscle.binary_einsum_(c, a, b)
    assert np.array_equal(host(c), np.zeros((3, 4), dtype=T))


def test_bandwidth_probe():
    """The HBM probe reports positive rates below the nominal peak and rejects
    a buffer under 1 MiB."""
    r = muscle.measure_bandwidth(1.0)
    assert 0.3 < r["read_tbs"] < 8.0 and 0.3 < r["write_tbs"] < 8.0, r
    with pytest.raises(muscle.ArgumentError):
        muscle.measure_bandwidth(0.5 / 1024)
