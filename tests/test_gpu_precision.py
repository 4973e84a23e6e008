"""The precision ladder of binary_einsum for Float64 / ComplexF64 operands,
against the FP64 CPU oracle (BackendBase restatement).

Tolerances (norm-wise relative, ||C - C_oracle|| / ||C_oracle||):
  fp32        2e-5   (FP32 operands and accumulation; K up to 1536)
  bf16x3      2e-5   (three bf16 terms, FP32 accumulation per 1024-long K chunk)
  fp16x2      2e-4   (two fp16 terms of row/column-scaled values)
  ozaki_int8  the requested target: 1e-10 (5 base-255 digits) and 1e-12 (6)
"""
import numpy as np
import pytest

import muscle
from muscle.workloads import ladder_inputs
from oracle import backend_base as bb
from tests.gpu_util import dev_tensor, host, randn, rel_err

pytestmark = pytest.mark.gpu

TOL = {"fp32": 2e-5, "bf16x3": 2e-5, "fp16x2": 2e-4}

SHAPES = [
    ("config1", ["i", "j", "k"], (64, 32, 48), ["k", "l", "j"], (48, 96, 32), ["i", "l"]),
    ("ragged", ["i", "j"], (70, 45), ["j", "k"], (45, 93), ["k", "i"]),
    ("hyper", ["i", "j", "h"], (40, 36, 3), ["j", "k", "h"], (36, 50, 3), ["h", "k", "i"]),
    ("interleaved d=8", ["a1", "c1", "a3", "c2", "a5", "c3"], (8,) * 6, ["c1", "b1", "c2", "b2", "c3", "b3"], (8,) * 6,
     ["a1", "b1", "a3", "b2", "a5", "b3"]),
    # M and N above 512 and not multiples of the 256-wide sliced tile
    ("ragged wide", ["i", "j"], (600, 130), ["j", "k"], (130, 530), ["k", "i"]),
]


def run(A, ia, B, ib, ic, **backend):
    with muscle.with_backend(muscle.BackendHIP(**backend)):
        c = muscle.binary_einsum(dev_tensor(A, ia), dev_tensor(B, ib), out=ic)
    return host(c), muscle.last_info()


@pytest.mark.parametrize("mode", ["fp32", "bf16x3", "fp16x2"])
@pytest.mark.parametrize("T", [np.complex128, np.float64])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: s[0])
def test_ladder_rungs(shape, T, mode):
    name, ia, sa, ib, sb, ic = shape
    rng = np.random.default_rng(5)
    A, B = randn(rng, sa, T), randn(rng, sb, T)
    ref = bb.binary_einsum_general(ic, A, ia, B, ib)
    C, info = run(A, ia, B, ib, ic, precision=mode)
    assert C.dtype == T
    assert info["precision"] == muscle._lib.PRECISION[mode]
    assert rel_err(C, ref) <= TOL[mode], (mode, rel_err(C, ref))


@pytest.mark.parametrize("target, S", [(1e-10, 5), (1e-12, 6)])
@pytest.mark.parametrize("T", [np.complex128, np.float64])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: s[0])
def test_ozaki_int8(shape, T, target, S):
    name, ia, sa, ib, sb, ic = shape
    rng = np.random.default_rng(6)
    A, B = randn(rng, sa, T), randn(rng, sb, T)
    ref = bb.binary_einsum_general(ic, A, ia, B, ib)
    C, info = run(A, ia, B, ib, ic, precision="ozaki_int8", target_rel_error=target)
    assert info["kernel"] == 3 and info["slices"] == S
    assert rel_err(C, ref) <= target, rel_err(C, ref)


# Large shapes (M*N >= 512^2, M*N*K*batch >= 1e9, K < 2^17) take the modular
# (CRT) form of the Ozaki rung: int8 products modulo 9-13 small moduli on this
# library's int8 MFMA GEMM, recombined by the Chinese remainder theorem.
# Complex x complex products split in Z_p x Z_p (kernel 8, two GEMMs per
# modulus) down to the split table's reach (about 7e-14 at K = 1024); 3e-14
# needs the 3M form (kernel 7, three GEMMs per modulus).
# Ragged extents (padding to the 256-wide tiles), a batch (hyper) mode, an
# M-contiguous and a K-contiguous operand, and a second call on the cached plan.
CRT_TARGETS = [(1e-10, 8), (1e-12, 8), (3e-14, 7)]
CRT_SHAPES = [
    ("square 1024", ["i", "j"], (1024, 1024), ["j", "k"], (1024, 1024), ["i", "k"]),
    ("ragged", ["i", "j"], (1000, 1000), ["j", "k"], (1000, 1100), ["k", "i"]),
    ("hyper", ["i", "j", "h"], (1024, 512, 2), ["j", "k", "h"], (512, 1024, 2), ["h", "k", "i"]),
    ("k-major A", ["j", "i"], (1100, 1030), ["j", "k"], (1100, 990), ["i", "k"]),
]


# The same large shapes on the sliced rungs: plans with M, N >= 256, M*N >= 512^2
# and M*N*K*batch >= 1e9 run their slice products on the 256 x 256 slice GEMM
# (32x32 MFMAs).  Ozaki int8 at 3e-15 is beyond the modular form's reach
# (2^50 integers): 7 base-255 slices on the same GEMM in int8.
@pytest.mark.parametrize("mode", ["bf16x3", "fp16x2", "ozaki_3e-15"])
@pytest.mark.parametrize("T", [np.complex128, np.float64])
@pytest.mark.parametrize("shape", CRT_SHAPES, ids=lambda s: s[0])
def test_wide_slice_rungs(shape, T, mode):
    name, ia, sa, ib, sb, ic = shape
    rng = np.random.default_rng(12)
    A, B = randn(rng, sa, T), randn(rng, sb, T)
    ref = bb.binary_einsum_general(ic, A, ia, B, ib)
    if mode == "ozaki_3e-15":
        C, info = run(A, ia, B, ib, ic, precision="ozaki_int8", target_rel_error=3e-15)
        assert info["kernel"] == 3 and info["slices"] == 7, info
        tol = 3e-15
    else:
        C, info = run(A, ia, B, ib, ic, precision=mode)
        assert info["kernel"] == 3 and info["precision"] == muscle._lib.PRECISION[mode], info
        tol = max(TOL[mode], 2 * info["err_bound"])
    assert C.dtype == T
    assert rel_err(C, ref) <= tol, (mode, rel_err(C, ref), tol)


@pytest.mark.parametrize("target, kernel", CRT_TARGETS)
@pytest.mark.parametrize("shape", CRT_SHAPES, ids=lambda s: s[0])
def test_ozaki_int8_crt(shape, target, kernel):
    name, ia, sa, ib, sb, ic = shape
    rng = np.random.default_rng(8)
    A, B = randn(rng, sa, np.complex128), randn(rng, sb, np.complex128)
    ref = bb.binary_einsum_general(ic, A, ia, B, ib)
    first = None
    for call in range(2):
        C, info = run(A, ia, B, ib, ic, precision="ozaki_int8", target_rel_error=target)
        assert info["kernel"] == kernel and 9 <= info["slices"] <= 13, info
        assert info["err_bound"] <= target
        assert rel_err(C, ref) <= target, (call, rel_err(C, ref))
        if first is None:
            first = C
        else:
            assert np.array_equal(first, C)  # deterministic


@pytest.mark.parametrize("ta, tb", [(np.float64, np.float64), (np.float64, np.complex128),
                                    (np.complex128, np.float64)])
def test_ozaki_int8_crt_real_and_mixed(ta, tb):
    rng = np.random.default_rng(9)
    A, B = randn(rng, (1030, 1000), ta), randn(rng, (1000, 1040), tb)
    C, info = run(A, ["i", "j"], B, ["j", "k"], ["i", "k"], precision="ozaki_int8", target_rel_error=1e-10)
    assert info["kernel"] == 7
    assert rel_err(C, A @ B) <= 1e-10


@pytest.mark.parametrize("dist", ["loguniform", "unitrows"])
def test_ozaki_int8_crt_distributions(dist):
    """The log-uniform (1e-4..1e4) and unit-row distributions of the ladder at a
    CRT-sized shape, with an all-zero row of A and column of B (their C row and
    column must be exactly zero)."""
    rng = np.random.default_rng(10)
    n = 1024
    if dist == "loguniform":
        A = (10.0 ** rng.uniform(-4, 4, (n, n))) * np.exp(2j * np.pi * rng.uniform(size=(n, n)))
        B = (10.0 ** rng.uniform(-4, 4, (n, n))) * np.exp(2j * np.pi * rng.uniform(size=(n, n)))
    else:
        A, B = randn(rng, (n, n)), randn(rng, (n, n))
        A /= np.linalg.norm(A, axis=1, keepdims=True)
        B /= np.linalg.norm(B, axis=0, keepdims=True)
    A[5, :] = 0
    B[:, 7] = 0
    ref = A @ B
    for target, kernel in CRT_TARGETS:
        C, info = run(A, ["i", "j"], B, ["j", "k"], ["i", "k"], precision="ozaki_int8", target_rel_error=target)
        assert info["kernel"] == kernel
        assert rel_err(C, ref) <= target, (dist, target, rel_err(C, ref))
        assert not C[5, :].any() and not C[:, 7].any()


@pytest.mark.parametrize("dt, kernel", [(np.complex128, 8), (np.float64, 7)])
def test_ozaki_int8_crt_nonfinite(dt, kernel):
    """An Inf in a row of A and a NaN in a column of B (after finite values in
    that row / column) give NaN across that row / column of C; every other
    entry keeps the target accuracy."""
    rng = np.random.default_rng(11)
    n = 1024
    A, B = randn(rng, (n, n), dt), randn(rng, (n, n), dt)
    A[3, 700] = np.inf
    B[900, 9] = np.nan
    C, info = run(A, ["i", "j"], B, ["j", "k"], ["i", "k"], precision="ozaki_int8", target_rel_error=1e-10)
    assert info["kernel"] == kernel
    assert np.isnan(C[3, :]).all() and np.isnan(C[:, 9]).all()
    keep_r = np.arange(n) != 3
    keep_c = np.arange(n) != 9
    A2, B2 = A.copy(), B.copy()
    A2[3, 700] = 0
    B2[900, 9] = 0
    ref = (A2 @ B2)[np.ix_(keep_r, keep_c)]
    got = C[np.ix_(keep_r, keep_c)]
    assert np.isfinite(got).all()
    assert rel_err(got, ref) <= 1e-10


@pytest.mark.parametrize("ta, tb", [(np.float64, np.complex128), (np.complex128, np.float64)])
def test_ozaki_mixed_real_complex(ta, tb):
    rng = np.random.default_rng(7)
    A, B = randn(rng, (70, 90), ta), randn(rng, (90, 60), tb)
    ref = A @ B
    C, info = run(A, ["i", "j"], B, ["j", "k"], ["i", "k"], precision="ozaki_int8", target_rel_error=1e-10)
    assert info["kernel"] == 3
    assert rel_err(C, ref) <= 1e-10


RUNGS = [("native", 1e-12), ("fp32", 0), ("bf16x3", 0), ("fp16x2", 0), ("ozaki_int8", 1e-10),
         ("ozaki_int8", 1e-12)]


@pytest.mark.parametrize("rung", RUNGS, ids=lambda r: f"{r[0]}_{r[1]:g}" if r[1] else r[0])
@pytest.mark.parametrize("dist", ["normal", "loguniform", "unitrows"])
def test_ladder_distributions_config1(dist, rung):
    """Every rung on the config-1 shapes under the three value distributions of
    the precision ladder (config 3).  The norm-wise error against the FP64
    oracle stays within: the error model the library states for the call
    (FP32, BF16x3, FP16x2); the requested target (Ozaki int8, whose stated
    0.4 * 255^-(S-1) only picks the slice count S: log-uniform data lands at
    1.0-1.2 times it); 1e-12 (native FP64)."""
    mode, target = rung
    A, ia, B, ib, dims, ic = ladder_inputs(dist, 0)
    ref = bb.binary_einsum_base(ic, A, ia, B, ib)
    C, info = run(A, ia, B, ib, ic, precision=mode, target_rel_error=target or 1e-12)
    err = rel_err(C, ref)
    if mode == "native":
        assert info["precision"] == 0 and err <= target, err
        return
    assert info["precision"] == muscle._lib.PRECISION[mode]
    bound = info["err_bound"]
    if mode == "ozaki_int8":
        assert bound <= target and err <= target, (mode, err, target)
        return
    assert 0 < bound <= 1e-5 and err <= bound, (mode, err, bound)


def test_float32_operands_ignore_ladder():
    rng = np.random.default_rng(8)
    A, B = randn(rng, (64, 64), np.complex64), randn(rng, (64, 64), np.complex64)
    C, info = run(A, ["i", "j"], B, ["j", "k"], ["i", "k"], precision="ozaki_int8")
    assert info["precision"] == 0  # FP32 data: native arithmetic
    assert rel_err(C, A.astype(np.complex128) @ B.astype(np.complex128)) <= 2e-5
