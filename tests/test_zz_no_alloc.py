"""muscle_hip_binary_einsum_noalloc (the entry point that allocates no device
memory) against muscle_hip_binary_einsum, both called through the C ABI.

Two calls in a row share the static scratch, the second with a smaller input;
each result must equal the allocating entry point's result bit for bit (same
plan, same kernels) and match the FP64 oracle to 1e-12.  Device memory in use
does not change over repeated calls, and a contraction that needs more than
the scratch's 32 MiB is refused before anything runs.
"""
import ctypes

import numpy as np
import pytest
import torch

import muscle
from muscle import _lib
from muscle.ops import _descs, _handle_for
from oracle import backend_base as bb
from tests.gpu_util import dev_tensor, randn, rel_err

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 32 << 20


def _out(a, b, ic):
    size = dict(zip(a.inds(), a.shape)) | dict(zip(b.inds(), b.shape))
    T = torch.promote_types(a.dtype, b.dtype)
    x = torch.full([size[muscle.Index(i)] for i in ic][::-1], float("nan"), dtype=T, device="cuda")
    return muscle.Tensor(x.permute(*reversed(range(len(ic)))) if ic else x, ic)  # column-major


def _call(fn_name, c, a, b):
    h = _handle_for(torch.device("cuda", 0))
    dc, da, db = _descs(c, a, b)
    info = _lib.Info()
    fn = getattr(_lib.lib, fn_name)
    if fn_name == "muscle_hip_binary_einsum":
        st = fn(h, dc.ref, da.ref, db.ref, 0, 1e-12, ctypes.byref(info))
    else:
        st = fn(h, dc.ref, da.ref, db.ref, ctypes.byref(info))
    return st, info.as_dict()


CASES = [
    # config 1 (split-K tile GEMM: tables and partial sums in the scratch)
    ("config1", ["i", "j", "k"], (64, 32, 48), ["k", "l", "j"], (48, 96, 32), ["i", "l"], np.complex128, 1),
    # smaller: a split-K GEMM with a batch mode
    ("split-K batched", ["i", "j", "h"], (40, 777, 3), ["j", "k", "h"], (777, 33, 3), ["k", "h", "i"],
     np.complex128, 1),
    # smaller still: a reduction with K-split partial sums (inner product)
    ("inner", ["i", "j"], (120, 130), ["j", "i"], (130, 120), [], np.float64, 2),
    # mixed precision: the Float32 operand is cast into the scratch
    ("mixed", ["i", "j"], (70, 90), ["j", "k"], (90, 60), ["i", "k"], (np.float32, np.complex128), 1),
]


def _inputs(case, seed):
    name, ia, sa, ib, sb, ic, T, kernel = case
    ta, tb = T if isinstance(T, tuple) else (T, T)
    rng = np.random.default_rng(seed)
    A, B = randn(rng, sa, ta), randn(rng, sb, tb)
    return A, B


def test_noalloc_matches_allocating_entry_point_twice_in_a_row():
    torch.cuda.synchronize()
    results = []
    for rep in range(2):  # the second pass finds the scratch holding the last call's contents
        for k, case in enumerate(CASES):
            name, ia, sa, ib, sb, ic, T, kernel = case
            A, B = _inputs(case, 100 * rep + k)
            a, b = dev_tensor(A, ia), dev_tensor(B, ib)
            c0, c1 = _out(a, b, ic), _out(a, b, ic)
            st0, info0 = _call("muscle_hip_binary_einsum", c0, a, b)
            st1, info1 = _call("muscle_hip_binary_einsum_noalloc", c1, a, b)
            assert st0 == 0 and st1 == 0, (name, _lib.lib.muscle_hip_last_error())
            assert info1["kernel"] == info0["kernel"] == kernel, (name, info0, info1)
            torch.cuda.synchronize()
            g0, g1 = c0.parent().cpu().numpy(), c1.parent().cpu().numpy()
            ref = bb.binary_einsum_general(ic, A.astype(np.complex128), ia, B.astype(np.complex128), ib)
            assert np.array_equal(g0, g1), name
            assert rel_err(g1, ref) <= 1e-12, (name, rel_err(g1, ref))
            results.append(name)
    assert len(results) == 2 * len(CASES)


def test_noalloc_second_call_smaller_input():
    """Config 1, then immediately (no synchronization) a smaller split-K
    contraction through the same scratch: both results correct."""
    big, small = CASES[0], CASES[1]
    outs = []
    for case, seed in ((big, 7), (small, 8)):
        name, ia, sa, ib, sb, ic, T, kernel = case
        A, B = _inputs(case, seed)
        a, b = dev_tensor(A, ia), dev_tensor(B, ib)
        ref = bb.binary_einsum_general(ic, A, ia, B, ib)
        outs.append((a, b, ic, ref))
    torch.cuda.synchronize()
    got = []
    for a, b, ic, ref in outs:
        c = _out(a, b, ic)
        st, _ = _call("muscle_hip_binary_einsum_noalloc", c, a, b)
        assert st == 0, _lib.lib.muscle_hip_last_error()
        got.append((c, ref))
    torch.cuda.synchronize()
    for c, ref in got:
        assert rel_err(c.parent().cpu().numpy(), ref) <= 1e-12


def test_noalloc_allocates_nothing_and_refuses_oversized():
    A, B = _inputs(CASES[0], 3)
    name, ia, sa, ib, sb, ic, T, kernel = CASES[0]
    a, b = dev_tensor(A, ia), dev_tensor(B, ib)
    c = _out(a, b, ic)
    st, _ = _call("muscle_hip_binary_einsum_noalloc", c, a, b)  # warm: kernels loaded
    assert st == 0
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(20):
        st, _ = _call("muscle_hip_binary_einsum_noalloc", c, a, b)
        assert st == 0
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0, (free0, free1)
    # 512 x 512 ComplexF64 output over 16 K slices: 64 MiB of partial sums
    rng = np.random.default_rng(4)
    Ab, Bb = randn(rng, (512, 4096)), randn(rng, (4096, 512))
    ab, bb_ = dev_tensor(Ab, ["i", "j"]), dev_tensor(Bb, ["j", "k"])
    cb = _out(ab, bb_, ["i", "k"])
    st, _ = _call("muscle_hip_binary_einsum_noalloc", cb, ab, bb_)
    assert st == _lib.INVALID_ARGUMENT
    msg = _lib.lib.muscle_hip_last_error().decode()
    assert "scratch" in msg and str(SCRATCH_BYTES) in msg, msg
    torch.cuda.synchronize()
    assert torch.isnan(cb.parent()).all()  # nothing ran
