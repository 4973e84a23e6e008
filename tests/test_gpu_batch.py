"""Batched contractions (one persistent, graph-captured launch) against the
CPU oracle, pair by pair: norm-wise relative error <= 1e-12 for ComplexF64 /
Float64 and <= 2e-5 for ComplexF32.  Graph replays are compared with the
eager launch after host work in between (every replay is checked)."""
import numpy as np
import pytest
import torch

import muscle
from muscle.batch import ContractionBatch, tuning_class, tuning_load, tuning_save
from muscle.workloads import circuit_pairs
from oracle import backend_base as bb
from tests.gpu_util import dev_tensor, host, randn, rel_err

pytestmark = pytest.mark.gpu

GENERAL = [
    (["i", "j", "k"], (64, 32, 48), ["k", "l", "j"], (48, 96, 32), ["i", "l"]),
    (["i", "j"], (67, 45), ["j", "k"], (45, 93), ["k", "i"]),
    (["i", "j", "h"], (40, 36, 3), ["j", "k", "h"], (36, 50, 3), ["h", "k", "i"]),
    (["i", "j"], (300, 20), ["j"], (20,), ["i"]),
    (["j"], (200,), ["j", "k"], (200, 300), ["k"]),
    (["i", "j"], (120, 130), ["j", "i"], (130, 120), []),
    (["i", "j", "s"], (30, 40, 5), ["j", "k"], (40, 35), ["i", "k"]),
    (["i", "j", "i"], (20, 30, 20), ["j", "k"], (30, 40), ["i", "k"]),
    (["a", "b"], (5, 7), ["c"], (9,), ["c", "a", "b"]),
    (["i", "j"], (3, 40), ["j", "k"], (40, 2500), ["i", "k"]),  # skinny: an output side of 3
    (["k", "j", "h"], (2200, 300, 2), ["j", "i", "h"], (300, 5, 2), ["h", "k", "i"]),  # skinny, hyper mode
]


def make_pairs(rng, T, n_circuit=120, max_rank=11, seed=3):
    cases = []
    for ia, ib, ic in circuit_pairs(n_circuit, seed=seed, ranks=(4, max_rank), out_ranks=(0, 14)):
        cases.append((ia, (2,) * len(ia), ib, (2,) * len(ib), ic))
    cases += GENERAL
    pairs, refs = [], []
    for ia, sa, ib, sb, ic in cases:
        A, B = randn(rng, sa, T), randn(rng, sb, T)
        a, b = dev_tensor(A, ia), dev_tensor(B, ib)
        size = dict(zip(ia, sa)) | dict(zip(ib, sb))
        c = muscle.Tensor(muscle.empty_col_major([size[x] for x in ic], a.dtype, a.device), ic)
        pairs.append((c, a, b))
        refs.append(bb.binary_einsum_general(ic, A.astype(np.complex128), ia, B.astype(np.complex128), ib))
    return pairs, refs


def check(pairs, refs, tol):
    torch.cuda.synchronize()
    worst = max(rel_err(host(c), r) for (c, a, b), r in zip(pairs, refs))
    assert worst <= tol, worst


@pytest.mark.parametrize("T, tol", [(np.complex128, 1e-12), (np.float64, 1e-12), (np.complex64, 2e-5)])
def test_batch_matches_oracle(T, tol):
    rng = np.random.default_rng(11)
    pairs, refs = make_pairs(rng, T)
    batch = ContractionBatch(pairs)
    info = batch.info()
    assert info["pairs"] == len(pairs) and info["items"] > info["shallow_items"] + info["skinny_items"]
    assert info["shallow_items"] > 0 and info["skinny_items"] > 0  # all three kernels run
    batch.run("persistent")
    check(pairs, refs, tol)
    assert batch.info()["graph"] == 1
    for c, a, b in pairs:
        c.parent().zero_()
    batch.run("loop")
    check(pairs, refs, tol)
    batch.close()


def test_graph_replays_equal_eager():
    rng = np.random.default_rng(12)
    pairs, refs = make_pairs(rng, np.complex128, n_circuit=60)
    batch = ContractionBatch(pairs)
    batch.run("persistent_nograph")
    torch.cuda.synchronize()
    eager = [c.parent().clone() for c, a, b in pairs]
    for rep in range(3):
        for c, a, b in pairs:
            c.parent().fill_(float("nan"))
        _ = torch.randn(1000, device="cuda").sum().item()  # unrelated eager work between replays
        batch.run("persistent")
        torch.cuda.synchronize()
        for (c, a, b), e in zip(pairs, eager):
            assert torch.equal(c.parent(), e), rep
    check(pairs, refs, 1e-12)


@pytest.mark.parametrize("tile", ["64x64", "32x128", "128x32"])
def test_every_tile_shape(tile):
    rng = np.random.default_rng(13)
    pairs, refs = make_pairs(rng, np.complex128, n_circuit=60)
    batch = ContractionBatch(pairs, tiles=[tile] * len(pairs))
    batch.run()
    check(pairs, refs, 1e-12)


@pytest.mark.parametrize("mult", [3, 4])
def test_complex_products(mult):
    rng = np.random.default_rng(14)
    pairs, refs = make_pairs(rng, np.complex128, n_circuit=40)
    ContractionBatch(pairs, complex_mult=mult).run()
    check(pairs, refs, 1e-12)


def test_batch_keeps_its_complex_product_form():
    """A batch runs in the product form the handle had when it was built, also
    after the handle's option changes (graph replays and the plain loop agree
    bit for bit with a batch built and run under that form alone)."""
    from muscle import _lib
    from muscle.ops import _handle_for
    rng = np.random.default_rng(17)
    pairs, refs = make_pairs(rng, np.complex128, n_circuit=40)
    dev = pairs[0][1].device.index or 0
    out = {}
    for mult in (4, 3):
        batch = ContractionBatch(pairs, complex_mult=mult)
        _lib.set_option(dev, _handle_for(dev), 0, 7 - mult)  # the other form, set after the build
        for mode in ("persistent", "persistent", "loop"):
            for c, a, b in pairs:
                c.parent().fill_(float("nan"))
            batch.run(mode)
            torch.cuda.synchronize()
            got = [c.parent().clone() for c, a, b in pairs]
            key = (mult, mode)
            if key in out:
                assert all(torch.equal(x, y) for x, y in zip(out[key], got)), key
            out[key] = got
        check(pairs, refs, 1e-12)
        batch.close()
    # 3M and 4M give different roundings: the form was applied, not ignored
    assert any(not torch.equal(x, y) for x, y in zip(out[(3, "persistent")], out[(4, "persistent")]))


def test_pair_routes_cover_the_launches():
    rng = np.random.default_rng(18)
    pairs, refs = make_pairs(rng, np.complex128, n_circuit=60)
    batch = ContractionBatch(pairs)
    info = batch.info()
    routes = [batch.pair_info(i) for i in range(len(pairs))]
    assert sum(r["route"] == "shallow" for r in routes) > 0
    assert sum(r["route"] == "skinny" for r in routes) > 0
    for r in routes:
        if r["route"] == "skinny":
            assert min(r["m"], r["n"]) <= 16 and r["k"] <= 1024
    for r in routes:
        assert r["route"] in ("empty", "tiles", "split_k", "shallow", "solo", "skinny")
        assert (r["tile"] >= 0) == (r["route"] in ("tiles", "split_k"))
        assert (r["ksplit"] > 1) == (r["route"] == "split_k")
        assert r["size_class"] == tuning_class(r["m"], r["n"])
    with pytest.raises(muscle.ArgumentError):
        batch.pair_info(len(pairs))
    assert info["pairs"] == len(pairs)
    batch.close()


def test_autotune_and_tuning_file(tmp_path):
    rng = np.random.default_rng(15)
    pairs, refs = make_pairs(rng, np.complex128, n_circuit=80)
    batch = ContractionBatch(pairs)
    batch.autotune(samples=8, reps=2)
    path = str(tmp_path / "tune.txt")
    tuning_save(path)
    lines = [l for l in open(path) if not l.startswith("#")]
    assert lines, "autotuner recorded no size class"
    tuning_load(path)
    tuned = ContractionBatch(pairs)  # AUTO tiles now come from the tuning table
    tuned.run()
    check(pairs, refs, 1e-12)
    assert 0 <= tuning_class(1, 1) < 25 and tuning_class(1000, 1000) == 24


def test_batch_rejects_mixed_eltypes():
    a = dev_tensor(np.ones((4, 4)), ["i", "j"])
    b = dev_tensor(np.ones((4, 4), dtype=np.complex128), ["j", "k"])
    c = muscle.Tensor(torch.zeros((4, 4), dtype=torch.complex128, device="cuda"), ["i", "k"])
    with pytest.raises(muscle.UnsupportedError):
        ContractionBatch([(c, a, b)])


def test_largest_pairs_run_as_single_launches():
    """Pairs of 2^29 or more index points leave the persistent launch for a
    single-contraction launch of their own (packed operands, plain GEMM tiles)
    inside the same graph; they match the oracle and, bit for bit, the plain loop."""
    rng = np.random.default_rng(16)
    pairs, refs = make_pairs(rng, np.complex128, n_circuit=30)
    small_items = ContractionBatch(pairs).info()["items"]
    labels = [f"y{i}" for i in range(30)]
    for s in range(2):
        perm = list(rng.permutation(labels))
        shared, left, right = perm[:10], perm[10:20], perm[20:]
        ia = list(rng.permutation(left + shared))
        ib = list(rng.permutation(shared + right))
        ic = list(rng.permutation(left + right))
        A, B = randn(rng, (2,) * 20, np.complex128), randn(rng, (2,) * 20, np.complex128)
        a, b = dev_tensor(A, ia), dev_tensor(B, ib)
        c = muscle.Tensor(muscle.empty_col_major([2] * 20, a.dtype, a.device), ic)
        pairs.append((c, a, b))
        refs.append(bb.binary_einsum_general(ic, A, ia, B, ib))
    batch = ContractionBatch(pairs)
    info = batch.info()
    assert info["items"] == small_items  # the two 2^30-point pairs are not in the work list
    assert info["flops"] >= 2 * 8.0 * 2.0 ** 30
    batch.run("persistent_nograph")
    torch.cuda.synchronize()
    eager = [c.parent().clone() for c, a, b in pairs]
    check(pairs, refs, 1e-12)
    for rep in range(2):
        for c, a, b in pairs:
            c.parent().fill_(float("nan"))
        batch.run("persistent")
        torch.cuda.synchronize()
        for (c, a, b), e in zip(pairs, eager):
            assert torch.equal(c.parent(), e), rep
    batch.run("loop")
    torch.cuda.synchronize()
    for (c, a, b), e in zip(pairs[-2:], eager[-2:]):
        assert torch.equal(c.parent(), e)
    batch.close()


SHALLOW = [
    (["i", "j"], (67, 16), ["j", "k"], (16, 93), ["k", "i"]),  # K = 16, ragged tiles, C along N
    (["i", "j"], (67, 17), ["j", "k"], (17, 93), ["i", "k"]),  # K = 17: the MFMA kernel
    (["a", "b"], (130, 9), ["b", "c"], (9, 70), ["a", "c"]),  # two K passes, C along M
    (["j", "i"], (12, 100), ["k", "j"], (40, 12), ["i", "k"]),  # A contiguous along K
    (["i", "j", "h"], (40, 3, 5), ["j", "k", "h"], (3, 50, 5), ["h", "k", "i"]),  # hyper modes
    (["i"], (300,), ["k"], (200,), ["k", "i"]),  # outer product (K = 1)
    (["i", "j"], (5, 2), ["j", "k"], (2, 3), ["i", "k"]),  # one partial tile
]


@pytest.mark.parametrize("T, tol", [(np.complex128, 1e-12), (np.float64, 1e-12), (np.complex64, 2e-5),
                                    (np.float32, 2e-5)])
@pytest.mark.parametrize("tile", ["64x64", "32x128", "128x32"])
def test_shallow_pairs(T, tol, tile):
    """Pairs with K <= 16 run in the vector-FMA kernel (one 512-thread workgroup
    per output tile); ragged tiles, both C orientations, K-major operands,
    hyper modes and two K passes match the oracle."""
    rng = np.random.default_rng(17)
    pairs, refs = [], []
    for ia, sa, ib, sb, ic in SHALLOW:
        A, B = randn(rng, sa, T), randn(rng, sb, T)
        a, b = dev_tensor(A, ia), dev_tensor(B, ib)
        size = dict(zip(ia, sa)) | dict(zip(ib, sb))
        c = muscle.Tensor(muscle.empty_col_major([size[x] for x in ic], a.dtype, a.device), ic)
        pairs.append((c, a, b))
        refs.append(bb.binary_einsum_general(ic, A.astype(np.complex128), ia, B.astype(np.complex128), ib))
    batch = ContractionBatch(pairs, tiles=[tile] * len(pairs))
    info = batch.info()
    assert 0 < info["shallow_items"] < info["items"]
    batch.run()
    check(pairs, refs, tol)
    for c, a, b in pairs:
        c.parent().fill_(float("nan"))
    batch.run("persistent_nograph")
    check(pairs, refs, tol)
    batch.close()


def test_shallow_launch_beyond_one_tile_per_workgroup():
    """A shallow launch of more than 2^20 tiles runs the looping kernel (a
    workgroup takes a second tile, whose offsets and first K pass it loads
    before storing the first): 65 Float32 pairs of 8192 x 8192 outputs with
    K = 4 (1,064,960 tiles of one C orientation) match a plain torch fp32 sum
    of the K outer products (elementwise, no library GEMM and its precision
    modes)."""
    g = torch.Generator(device="cuda").manual_seed(23)
    n, k, npairs = 8192, 4, 65
    As = [torch.randn(n, k, device="cuda", generator=g) for _ in range(5)]
    Bs = [torch.randn(k, n, device="cuda", generator=g) for _ in range(7)]
    pairs = []
    for p in range(npairs):
        a = muscle.Tensor(As[p % 5].t().contiguous().t(), ["i", "j"])  # column-major, as Julia holds it
        b = muscle.Tensor(Bs[p % 7].t().contiguous().t(), ["j", "k"])
        c = muscle.Tensor(muscle.empty_col_major([n, n], torch.float32, torch.device("cuda")), ["i", "k"])
        pairs.append((c, a, b))
    batch = ContractionBatch(pairs, tiles=["64x64"] * npairs)
    info = batch.info()
    assert info["shallow_items"] == npairs * (n // 64) ** 2 > 1 << 20
    for rep in range(2):
        for c, a, b in pairs:
            c.parent().fill_(float("nan"))
        batch.run()
        torch.cuda.synchronize()
        for p, (c, a, b) in enumerate(pairs):
            A, B = As[p % 5], Bs[p % 7]
            ref = A[:, :1] * B[:1, :]
            for j in range(1, k):
                ref += A[:, j:j + 1] * B[j:j + 1, :]
            err = (torch.linalg.norm(c.parent() - ref) / torch.linalg.norm(ref)).item()
            assert err < 1e-6, (rep, p, err)
    batch.close()


SPLIT = [
    (["i", "j"], (8, 4096), ["j", "k"], (4096, 4), ["i", "k"]),  # tiny output, long K
    (["i", "j", "h"], (70, 1500, 1), ["j", "k", "h"], (1500, 40, 1), ["h", "k", "i"]),  # ragged tiles
    (["i", "j", "h"], (30, 700, 2), ["j", "k", "h"], (700, 20, 2), ["h", "k", "i"]),  # hyper mode
    (["a", "b"], (130, 2000), ["b", "c"], (2000, 66), ["c", "a"]),  # more tiles: not split
    (["i", "j"], (6, 5), ["j", "k"], (5, 7), ["i", "k"]),  # shallow
    (["i", "j"], (40, 20), ["j", "k"], (20, 30), ["k", "i"]),  # short K, not split
]


@pytest.mark.parametrize("T, tol", [(np.complex128, 1e-12), (np.float64, 1e-12), (np.complex64, 2e-5),
                                    (np.float32, 2e-5)])
@pytest.mark.parametrize("tile", ["64x64", "32x128", "128x32"])
def test_split_k_pairs(T, tol, tile):
    """Pairs of at most 2 output tiles and at least 32 K-tiles are cut into K
    slices (their own launch) whose partial tiles the last-arriving slice sums
    in slice order: results match the oracle, graph replays equal the eager
    launch bit for bit, and a pair's result is the same alone."""
    rng = np.random.default_rng(18)
    pairs, refs = [], []
    for ia, sa, ib, sb, ic in SPLIT:
        A, B = randn(rng, sa, T), randn(rng, sb, T)
        a, b = dev_tensor(A, ia), dev_tensor(B, ib)
        size = dict(zip(ia, sa)) | dict(zip(ib, sb))
        c = muscle.Tensor(muscle.empty_col_major([size[x] for x in ic], a.dtype, a.device), ic)
        pairs.append((c, a, b))
        refs.append(bb.binary_einsum_general(ic, A.astype(np.complex128), ia, B.astype(np.complex128), ib))
    batch = ContractionBatch(pairs, tiles=[tile] * len(pairs))
    info = batch.info()
    assert info["split_items"] > 0 and info["shallow_items"] > 0
    batch.run("persistent_nograph")
    check(pairs, refs, tol)
    eager = [c.parent().clone() for c, a, b in pairs]
    for rep in range(2):
        for c, a, b in pairs:
            c.parent().fill_(float("nan"))
        batch.run("persistent")
        torch.cuda.synchronize()
        for (c, a, b), e in zip(pairs, eager):
            assert torch.equal(c.parent(), e), rep
    batch.close()
    c, a, b = pairs[0]  # the split pair alone gives the same bits as in the batch
    alone = ContractionBatch([(c, a, b)], tiles=[tile])
    assert alone.info()["split_items"] > 0
    c.parent().fill_(float("nan"))
    alone.run()
    torch.cuda.synchronize()
    assert torch.equal(c.parent(), eager[0])
    alone.close()


def test_batch_contracted_extent_zero():
    """In a batch, a pair whose contracted extent is 0 writes zeros (shallow
    kernel with no K pass) beside a normal pair."""
    rng = np.random.default_rng(19)
    A, B = randn(rng, (20, 30)), randn(rng, (30, 40))
    pairs = [(muscle.Tensor(torch.full((5, 6), float("nan"), dtype=torch.complex128, device="cuda"), ["i", "k"]),
              dev_tensor(np.zeros((5, 0), dtype=np.complex128), ["i", "j"]),
              dev_tensor(np.zeros((0, 6), dtype=np.complex128), ["j", "k"])),
             (muscle.Tensor(muscle.empty_col_major([20, 40], torch.complex128, "cuda"), ["i", "k"]),
              dev_tensor(A, ["i", "j"]), dev_tensor(B, ["j", "k"]))]
    batch = ContractionBatch(pairs)
    batch.run()
    torch.cuda.synchronize()
    assert np.array_equal(host(pairs[0][0]), np.zeros((5, 6)))
    assert rel_err(host(pairs[1][0]), A @ B) <= 1e-12
    batch.close()


SKINNY = [
    (["i", "j"], (1, 37), ["j", "k"], (37, 2100), ["i", "k"]),  # short side 1, one row of C
    (["i", "j"], (2100, 5), ["j", "k"], (5, 3), ["i", "k"]),  # short side 3 along C's contiguous direction
    (["j", "i"], (300, 16), ["k", "j"], (2500, 300), ["i", "k"]),  # short side 16, K = 300: three staged chunks
    (["i", "j", "h"], (6, 40, 3), ["j", "k", "h"], (40, 1100, 3), ["h", "k", "i"]),  # hyper modes
    (["i", "j"], (2, 1024), ["j", "k"], (1024, 2000), ["k", "i"]),  # the largest skinny K
    (["i", "j"], (9, 1025), ["j", "k"], (1025, 40), ["i", "k"]),  # K = 1025: the MFMA route (K slices)
    (["i"], (4000,), ["k"], (4,), ["k", "i"]),  # outer product (K = 1)
    (["a", "b"], (5, 7), ["b", "c"], (7, 5), ["a", "c"]),  # one tile: stays on the shallow route
    (["i", "j", "s"], (12, 30, 3), ["j", "k"], (30, 2100), ["i", "k"]),  # index summed in A alone
    (["i", "j", "i"], (10, 30, 10), ["j", "k"], (30, 2100), ["i", "k"]),  # diagonal of A
    (["i", "j"], (15, 9), ["j", "k"], (9, 5000), ["i", "k"]),  # ragged last run of the long side
]


@pytest.mark.parametrize("T, tol", [(np.complex128, 1e-12), (np.float64, 1e-12), (np.complex64, 2e-5),
                                    (np.float32, 2e-5)])
def test_skinny_pairs(T, tol):
    """Pairs with an output side of at most 16, K <= 1024 and at least 16 MFMA
    tiles run in the skinny vector-FMA kernel (whole short side per workgroup,
    long-side operand streamed into registers): short sides 1-16 on either side
    of C, K from 1 to 1024 over several staged chunks, hyper, summed and
    diagonal modes match the oracle; graph replays equal the eager launch bit
    for bit, and a pair's result is the same alone as in the batch."""
    rng = np.random.default_rng(21)
    pairs, refs = [], []
    for ia, sa, ib, sb, ic in SKINNY:
        A, B = randn(rng, sa, T), randn(rng, sb, T)
        a, b = dev_tensor(A, ia), dev_tensor(B, ib)
        size = dict(zip(ia, sa)) | dict(zip(ib, sb))
        c = muscle.Tensor(muscle.empty_col_major([size[x] for x in ic], a.dtype, a.device), ic)
        pairs.append((c, a, b))
        refs.append(bb.binary_einsum_general(ic, A.astype(np.complex128), ia, B.astype(np.complex128), ib))
    batch = ContractionBatch(pairs)
    routes = [batch.pair_info(i)["route"] for i in range(len(pairs))]
    assert routes == ["skinny"] * 5 + ["split_k", "skinny", "shallow"] + ["skinny"] * 3, routes
    assert batch.info()["skinny_items"] > 0
    batch.run("persistent_nograph")
    check(pairs, refs, tol)
    eager = [c.parent().clone() for c, a, b in pairs]
    for rep in range(2):
        for c, a, b in pairs:
            c.parent().fill_(float("nan"))
        _ = torch.randn(1000, device="cuda").sum().item()  # unrelated eager work between replays
        batch.run("persistent")
        torch.cuda.synchronize()
        for (c, a, b), e in zip(pairs, eager):
            assert torch.equal(c.parent(), e), rep
    batch.run("loop")
    check(pairs, refs, tol)
    batch.close()
    for k in (0, 2, 4):  # alone, the same bits as in the batch
        c, a, b = pairs[k]
        alone = ContractionBatch([(c, a, b)])
        c.parent().fill_(float("nan"))
        alone.run()
        torch.cuda.synchronize()
        assert torch.equal(c.parent(), eager[k]), k
        alone.close()
