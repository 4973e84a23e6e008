"""Host-side logic that needs no GPU: the C ABI loads and exports what
include/*.h declares, index classification / lowering (muscle_hip_describe),
the Python mirror's data model and dispatch."""
import glob
import os
import re

import numpy as np
import pytest
import torch

import muscle
from muscle import _lib
from muscle.describe import describe

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        names |= set(re.findall(r"MUSCLE_HIP_API\s+[\w\s\*]+?\b(muscle_hip_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(_lib.lib, n)]
    assert not missing, missing


def test_status_strings():
    assert _lib.lib.muscle_hip_status_string(0) == b"MUSCLE_HIP_OK"
    assert _lib.lib.muscle_hip_status_string(2) == b"MUSCLE_HIP_DIM_MISMATCH"
    assert _lib.lib.muscle_hip_version() >= 100


# ---------------------------------------------------------------- lowering
def test_config1_lowering():
    d = describe(["i", "l"], [64, 96], ["i", "j", "k"], [64, 32, 48], ["k", "l", "j"], [48, 96, 32])
    assert d["classes"] == {"i": "left", "l": "right", "j": "contract", "k": "contract"}
    # C is column-major (i fastest): the lowering exchanges A and B so C's fastest mode is an N mode
    assert d["swapped"] == 1 and (d["m"], d["n"], d["k"], d["batch"]) == (96, 64, 1536, 1)
    assert d["kernel_name"] == "gemm_mfma"
    assert d["flops"] == 8 * 64 * 96 * 1536
    assert d["bytes"] == 16 * (64 * 32 * 48 + 48 * 96 * 32 + 64 * 96)


def test_config2_lowering_interleaved():
    a = ["a1", "c1", "a3", "c2", "a5", "c3"]
    b = ["c1", "b1", "c2", "b2", "c3", "b3"]
    c = ["a1", "b1", "a3", "b2", "a5", "b3"]
    d = describe(c, [16] * 6, a, [16] * 6, b, [16] * 6)
    assert (d["m"], d["n"], d["k"]) == (4096, 4096, 4096)
    assert all(d["classes"][x] == "contract" for x in ["c1", "c2", "c3"])


def test_fusion_of_adjacent_modes():
    # A(i,j,k,l) col-major, B(k,l,m): (k,l) fuse to one K mode of 20, (i,j) to one M mode of 6
    d = describe(["i", "j", "m"], [2, 3, 7], ["i", "j", "k", "l"], [2, 3, 4, 5], ["k", "l", "m"], [4, 5, 7])
    assert (d["m"], d["n"], d["k"]) in [(6, 7, 20), (7, 6, 20)]


def test_hyperindex_classification():
    d = describe(["i", "k", "l"], [2, 4, 5], ["i", "j", "k"], [2, 3, 4], ["k", "l", "j"], [4, 5, 3])
    assert d["classes"] == {"i": "left", "k": "batch", "l": "right", "j": "contract"}
    assert d["batch"] == 4 and d["k"] == 3


def test_single_operand_sum_and_diagonal():
    # j appears only in A and not in C: summed over A alone
    d = describe(["i", "k"], [2, 4], ["i", "j"], [2, 3], ["k"], [4])
    assert d["classes"]["j"] == "sum_a"
    # i repeated in A: its diagonal
    d = describe(["i", "k"], [3, 4], ["i", "i"], [3, 3], ["k"], [4])
    assert d["classes"]["i"] == "left" and d["m"] * d["n"] == 12


def test_scalar_and_inner_product_lowering():
    d = describe([], [], ["i", "j"], [3, 4], ["j", "i"], [4, 3])
    assert (d["m"], d["n"], d["k"], d["batch"]) == (1, 1, 12, 1)
    assert d["kernel_name"] == "reduce"
    d = describe(["i", "j"], [2, 3], ["i", "j"], [2, 3], [], [])
    assert d["k"] == 1 and d["m"] * d["n"] == 6


@pytest.mark.parametrize("bad, exc", [
    ((["i", "i"], [2, 2], ["i", "j"], [2, 3], ["j", "i"], [3, 2]), muscle.ArgumentError),  # repeated out
    ((["x"], [2], ["i", "j"], [2, 3], ["j", "i"], [3, 2]), muscle.ArgumentError),  # out not in inputs
    ((["i"], [2], ["i", "j"], [2, 3], ["j", "i"], [4, 2]), muscle.DimensionMismatch),  # extents differ
])
def test_lowering_errors(bad, exc):
    with pytest.raises(exc):
        describe(*bad)


def test_promotion_rule_enforced():
    with pytest.raises(muscle.ArgumentError):
        describe(["i", "k"], [2, 4], ["i", "j"], [2, 3], ["j", "k"], [3, 4], dtypes=("f64", "c64", "f64"))
    d = describe(["i", "k"], [2, 4], ["i", "j"], [2, 3], ["j", "k"], [3, 4], dtypes=("c64", "f64", "c64"))
    assert d["flops"] == 4 * 2 * 4 * 3


@pytest.mark.parametrize("target, S", [(1e-5, 3), (1e-7, 4), (1e-10, 5), (1e-12, 6), (1e-14, 7)])
def test_ozaki_slice_count_from_target(target, S):
    """Smallest S with 0.4 * 255^-(S-1) <= target (the calibrated base-255 error model)."""
    shapes = (["i", "k"], [64, 64], ["i", "j"], [64, 128], ["j", "k"], [128, 64])
    d = describe(*shapes, precision="ozaki_int8", target_rel_error=target)
    assert d["kernel_name"] == "split" and d["slices"] == S
    assert d["err_bound"] <= target


# (target, dtype, kernel, moduli n, bits T) at K = 4096.  3M table 256, 255, 253, ...:
# T = 35, 39, 43, 46, 50 for n = 9..13.  Split table 241, 233, 229, 221, 205, ...
# (square roots of -1): T = 34, 37, 41, 44, 48 for n = 9..13.
CRT_COUNTS = [(1e-8, "f64", "split_crt", 9, 35), (1e-10, "f64", "split_crt", 10, 39),
              (1e-12, "f64", "split_crt", 12, 46), (1e-13, "f64", "split_crt", 13, 50),
              (1e-8, "c64", "split_crt_gaussian", 9, 34), (1e-10, "c64", "split_crt_gaussian", 11, 41),
              (1e-12, "c64", "split_crt_gaussian", 13, 48),
              # beyond the split table's 13 moduli: 3M
              (1e-13, "c64", "split_crt", 13, 50)]


@pytest.mark.parametrize("target, dt, kernel, n, bits", CRT_COUNTS)
def test_ozaki_crt_modulus_count_from_target(target, dt, kernel, n, bits):
    """Large shapes take the modular (CRT) form: the smallest modulus count n whose
    error model sqrt(K / 3) 2^-T meets the target, where 2^(2T) stays below half
    the moduli's product; complex x complex products split in Z_p x Z_p where
    those moduli reach the target."""
    s = [16] * 6
    d = describe(["a1", "b1", "a3", "b2", "a5", "b3"], s, ["a1", "c1", "a3", "c2", "a5", "c3"], s,
                 ["c1", "b1", "c2", "b2", "c3", "b3"], s, dtypes=(dt, dt, dt), precision="ozaki_int8",
                 target_rel_error=target)
    assert d["kernel_name"] == kernel and d["slices"] == n, d
    assert d["err_bound"] == pytest.approx((4096 / 3) ** 0.5 * 2.0 ** -bits, rel=1e-12)
    assert d["err_bound"] <= target


def test_crt_moduli_tables():
    """The two modulus tables from first principles: pairwise coprime, greedy from
    256 down, and (split table) every prime factor 1 mod 4, i.e. -1 a square."""
    import math

    def greedy(ok):
        ps = []
        for c in range(256, 64, -1):
            if ok(c) and all(math.gcd(c, q) == 1 for q in ps):
                ps.append(c)
        return ps

    def primes_of(c):
        f, d = [], 2
        while d * d <= c:
            while c % d == 0:
                f.append(d)
                c //= d
            d += 1
        return f + ([c] if c > 1 else [])

    t3 = greedy(lambda c: True)
    ts = greedy(lambda c: all(q % 4 == 1 for q in primes_of(c)))
    assert t3[:6] == [256, 255, 253, 251, 247, 241]
    assert ts[:6] == [241, 233, 229, 221, 205, 197]
    for p in ts[:13]:
        assert any((x * x + 1) % p == 0 for x in range(p)), p
    # the bit counts the library reports follow from these tables
    K = 4096

    def bits(ps, n):
        L = sum(math.log2(p) for p in ps[:n])
        T = int((L - 1) // 2)
        while T > 0 and not 2 * math.log2(2.0 ** T + 0.5 * math.sqrt(2 * K)) + 1 < L - 1e-9:
            T -= 1
        return T

    assert [bits(t3, n) for n in range(9, 14)] == [35, 39, 43, 46, 50]
    assert [bits(ts, n) for n in range(9, 14)] == [34, 37, 41, 44, 48]


def test_ozaki_crt_falls_back_to_slicing():
    # too small for the CRT GEMM tiles, and a target beyond 13 moduli (2^50 integers)
    small = (["i", "k"], [64, 96], ["i", "j"], [64, 1536], ["j", "k"], [1536, 96])
    assert describe(*small, precision="ozaki_int8", target_rel_error=1e-10)["kernel_name"] == "split"
    big = (["i", "k"], [1024, 1024], ["i", "j"], [1024, 1024], ["j", "k"], [1024, 1024])
    assert describe(*big, precision="ozaki_int8", target_rel_error=1e-10)["kernel_name"] == "split_crt_gaussian"
    assert describe(*big, dtypes=("f64",) * 3, precision="ozaki_int8",
                    target_rel_error=1e-10)["kernel_name"] == "split_crt"
    d = describe(*big, precision="ozaki_int8", target_rel_error=1e-16)
    assert d["kernel_name"] == "split" and d["slices"] == 7


def test_ladder_modes_lowering():
    for mode in ("fp32", "bf16x3", "fp16x2"):
        d = describe(["i", "k"], [64, 64], ["i", "j"], [64, 128], ["j", "k"], [128, 64], precision=mode)
        assert d["kernel_name"] == ("gemm_mfma" if mode == "fp32" else "split")
        assert d["precision"] == muscle._lib.PRECISION[mode] or mode == "fp32"


def test_strided_view_lowering():
    # A is a view with non-unit strides: (i stride 2, j stride 20)
    d = describe(["i", "k"], [5, 6], ["i", "j"], [5, 7], ["j", "k"], [7, 6],
                 strides=(None, [2, 20], None))
    assert (d["m"] * d["n"], d["k"]) == (30, 7)


# ---------------------------------------------------------------- data model + dispatch
def test_tensor_constructor_checks():
    with pytest.raises(muscle.DimensionMismatch):
        muscle.Tensor(torch.ones(2, 3), ["i"])
    with pytest.raises(muscle.DimensionMismatch):
        muscle.Tensor(torch.ones(2, 3), ["i", "i"])
    t = muscle.Tensor(torch.ones(2, 2), ["i", "i"])
    assert t.inds() == [muscle.Index("i"), muscle.Index("i")]
    assert muscle.Index("i") == muscle.Index(muscle.Index("i"))
    assert repr(muscle.Index("i")) == "index<i>"


def test_host_arrays_dispatch_to_reference_backends():
    a = muscle.Tensor(torch.ones(2, 3, dtype=torch.float64), ["i", "j"])
    b = muscle.Tensor(torch.ones(3, 4, dtype=torch.float64), ["j", "k"])
    assert muscle.choose_backend("binary_einsum", a.parent(), b.parent()) == muscle.BackendBase()
    assert muscle.choose_backend("unary_einsum", a.parent()) == muscle.BackendOMEinsum()
    # not part of this library: the reference's own "not implemented or not loaded" error
    with pytest.raises(muscle.ArgumentError, match="not implemented or not loaded"):
        muscle.binary_einsum(a, b)
    with pytest.raises(muscle.ArgumentError, match="not implemented or not loaded"):
        muscle.unary_einsum(a, dims=["i"])


def test_with_backend_override_is_scoped():
    a = muscle.Tensor(torch.ones(2), ["i"])
    with muscle.with_backend(muscle.BackendHIP(precision="ozaki_int8")) as be:
        assert muscle.choose_backend("binary_einsum", a.parent(), a.parent()) is be
        with muscle.with_backend(muscle.BackendBase()):
            assert muscle.choose_backend("hadamard", a.parent(), a.parent()) == muscle.BackendBase()
        assert muscle.choose_backend("hadamard", a.parent(), a.parent()) is be
    assert muscle.choose_backend("binary_einsum", a.parent(), a.parent()) == muscle.BackendBase()


def test_device_rules_select_hip():
    assert muscle.choose_backend_rule("binary_einsum", muscle.DomainROCm(), muscle.DomainROCm()) == muscle.BackendHIP()
    assert muscle.choose_backend_rule("hadamard!", muscle.DomainROCm(), muscle.DomainROCm(),
                                      muscle.DomainROCm()) == muscle.BackendHIP()
    assert muscle.choose_backend_rule("unary_einsum", muscle.DomainROCm()) == muscle.BackendHIP()


def test_hip_backend_rejects_host_arrays():
    a = muscle.Tensor(torch.ones(2, 3, dtype=torch.float64), ["i", "j"])
    with muscle.with_backend(muscle.BackendHIP()):
        with pytest.raises(muscle.ArgumentError, match="device"):
            muscle.binary_einsum(a, a)


def test_isapprox_is_permutation_aware():
    x = torch.arange(6, dtype=torch.float64).reshape(2, 3)
    a = muscle.Tensor(x, ["i", "j"])
    b = muscle.Tensor(x.T.contiguous(), ["j", "i"])
    assert a.isapprox(b)
    assert not a.isapprox(muscle.Tensor(x.T.contiguous() + 1, ["j", "i"]))


def test_call_descriptor_slots_match_fresh_descriptors():
    """The reusable per-thread descriptor slots of synchronous calls carry the
    same dtype, rank, pointer, mode labels, extents and strides as freshly
    built descriptors, also after a refill with other tensors."""
    from muscle.ops import _call_descs, _descs
    a = muscle.Tensor(torch.randn(4, 3, 5, dtype=torch.complex128).permute(2, 0, 1), ["k", "i", "j"])
    b = muscle.Tensor(torch.randn(5, 6, dtype=torch.complex128), ["k", "l"])
    c = muscle.Tensor(torch.empty(4, 3, 6, dtype=torch.complex128), ["i", "j", "l"])
    x = muscle.Tensor(torch.randn(7, dtype=torch.float32), ["q"])
    for ts in ((c, a, b), (x,), (c, a, b)):
        for slot, fresh in zip(_call_descs(*ts), _descs(*ts)):
            s, f = slot.desc, fresh.desc
            assert (s.dtype, s.rank, s.data) == (f.dtype, f.rank, f.data)
            for name in ("modes", "extents", "strides"):
                assert [getattr(s, name)[i] for i in range(s.rank)] == [getattr(f, name)[i] for i in range(f.rank)]


def test_index_hash_and_equality():
    I = muscle.Index
    assert I("a") == I("a") and I("a") != I("b") and hash(I(("site", 1))) == hash(I(("site", 1)))
    assert {I("a"): 1}[I("a")] == 1 and I(I("a")) == I("a")
    odd = I([1, 2])  # an unhashable tag: compares, cannot be hashed (as before)
    assert odd == I([1, 2]) and odd != I("a")
    with pytest.raises(TypeError):
        hash(odd)
