"""Served contraction-tree jobs (continuous batching) against the CPU oracle.

Every job's root is compared with the FP64 oracle evaluating the same tree
node by node (norm-wise relative error <= 1e-12 for ComplexF64 jobs, <= 2e-5
for ComplexF32); a job's result is the same alone and in a mix of other jobs
(<= 1e-13 relative, and the batched path is checked bit for bit); admission
refuses a job larger than the pool and queues jobs that do not fit yet."""
import numpy as np
import pytest
import torch

import muscle
from muscle.serve import LeafBank, Server, tree_nodes
from muscle.workloads import contraction_tree
from oracle import backend_base as bb
from tests.gpu_util import dev_tensor, host, randn, rel_err

pytestmark = pytest.mark.gpu

TOL = {"c64": 1e-12, "c32": 2e-5}


def small_trees(seeds, dtype=None, n_nodes=(20, 60), size_log2=(4, 10)):
    out = []
    for s in seeds:
        n = int(np.random.default_rng(s).integers(n_nodes[0], n_nodes[1] + 1))
        out.append(contraction_tree(s, n_nodes=n, size_log2=size_log2, dtype=dtype))
    return out


def oracle_root(spec, leaves):
    vals = [host(t).astype(np.complex128) for t in leaves]
    labs = [list(l) for l in spec.leaves]
    for a, b, o in spec.nodes:
        vals.append(bb.binary_einsum_general(list(o), vals[a], labs[a], vals[b], labs[b]))
        labs.append(list(o))
    return vals[-1]


def job_bytes(spec):
    es = 8 if spec.dtype == "c32" else 16
    return sum(((es << len(o)) + 255) // 256 * 256 for _, _, o in spec.nodes)


def run_jobs(srv, bank, specs, seed=0):
    rng = np.random.default_rng(seed)
    prepared = [(spec, srv.prepare(bank.leaves(spec, rng), tree_nodes(spec))) for spec in specs]
    jobs = [(spec, pj.leaves, srv.submit_prepared(pj)) for spec, pj in prepared]
    out = []
    for spec, leaves, jid in jobs:
        res = srv.wait(jid, timeout=120.0)
        assert res["status"] == 0, res
        out.append((spec, leaves, srv.result(jid), res))
        srv.release(jid)
    return out


def test_served_jobs_match_oracle():
    bank = LeafBank(seed=1)
    specs = small_trees(range(100, 108))  # mixed ComplexF32 / ComplexF64 jobs in one server
    assert {s.dtype for s in specs} == {"c32", "c64"}
    srv = Server(pool_gib=4)
    for spec, leaves, got, res in run_jobs(srv, bank, specs):
        err = rel_err(host(got), oracle_root(spec, leaves))
        assert err <= TOL[spec.dtype], (spec.dtype, err)
        assert res["ticks"] >= 1 and res["done_s"] >= res["admit_s"] >= res["submit_s"]
    st = srv.stats()
    assert st["completed"] == len(specs) and st["rejected"] == 0
    assert st["launched_pairs"] == sum(len(s.nodes) for s in specs)
    srv.close()


@pytest.mark.parametrize("dtype", ["c64", "c32"])
def test_alone_equals_in_mix(dtype):
    bank = LeafBank(seed=2)
    target = small_trees([200], dtype=dtype, n_nodes=(40, 40), size_log2=(4, 12))[0]
    others = small_trees(range(201, 213), n_nodes=(100, 200), size_log2=(8, 16))  # a few ms each
    srv = Server(pool_gib=4)
    (_, leaves, alone, _), = run_jobs(srv, bank, [target], seed=5)
    # the same job (same leaf tensors) submitted in the middle of 12 others
    mix = others[:6] + [target] + others[6:]
    rng = np.random.default_rng(9)
    prepared = [srv.prepare(leaves if spec is target else bank.leaves(spec, rng), tree_nodes(spec))
                for spec in mix]
    jobs = [srv.submit_prepared(pj) for pj in prepared]  # back to back, so the jobs share ticks
    results, spans = [], []
    for jid in jobs:
        r = srv.wait(jid, timeout=120.0)
        assert r["status"] == 0
        spans.append((r["admit_s"], r["done_s"]))
        results.append(srv.result(jid))
        srv.release(jid)
    in_mix = results[6]
    t0, t1 = spans[6]
    overlapping = sum(1 for k, (a0, a1) in enumerate(spans) if k != 6 and a0 < t1 and t0 < a1)
    assert overlapping >= 3, spans  # the target ran while other jobs were on the GPU
    a, m = host(alone), host(in_mix)
    assert rel_err(m, a) <= (1e-13 if dtype == "c64" else 1e-6)
    assert np.array_equal(m, a), "batched results depend on what shares the launch"
    # the plain per-call loop gives the same tensor within the stated tolerance
    ser = Server(pool_gib=4, mode="loop")
    jid = ser.submit(leaves, tree_nodes(target))
    assert ser.wait(jid, timeout=120.0)["status"] == 0
    s = host(ser.result(jid))
    ser.release(jid)
    assert rel_err(s, a) <= TOL[dtype]
    assert rel_err(a, oracle_root(target, leaves)) <= TOL[dtype]
    for x in (srv, ser):
        x.close()


def test_rejects_job_larger_than_pool():
    bank = LeafBank(seed=3)
    spec = small_trees([300], dtype="c64", size_log2=(12, 14))[0]
    assert job_bytes(spec) > 2 ** 20
    srv = Server(pool_gib=1.0 / 1024)  # 1 MiB
    leaves = bank.leaves(spec, np.random.default_rng(0))
    with pytest.raises(muscle.MuscleHipError) as e:
        srv.submit(leaves, tree_nodes(spec))
    assert e.value.status == 7  # MUSCLE_HIP_REJECTED
    st = srv.stats()
    assert st["rejected"] == 1 and st["completed"] == 0
    srv.close()


def test_admission_queues_jobs_until_memory_frees():
    bank = LeafBank(seed=4)
    specs = small_trees(range(400, 406), dtype="c64", size_log2=(8, 13))
    need = [job_bytes(s) for s in specs]
    pool = int(max(need) * 1.3)  # room for about one job at a time
    srv = Server(pool_gib=pool / 2 ** 30)
    out = run_jobs(srv, bank, specs)
    for spec, leaves, got, res in out:
        assert rel_err(host(got), oracle_root(spec, leaves)) <= 1e-12
    st = srv.stats()
    assert st["completed"] == len(specs) and st["pool_peak_bytes"] <= pool
    # the jobs were submitted back to back but never all admitted together
    spans = [(r["admit_s"], r["done_s"]) for _, _, _, r in out]
    most = max(sum(1 for a, d in spans if a <= t < d) for t, _ in spans)
    assert most < len(specs), spans
    assert max(r["admit_s"] - r["submit_s"] for _, _, _, r in out) > 0
    srv.close()


def operand_inds(spec, leaves, k):
    nl = len(leaves)
    return leaves[k].inds() if k < nl else [muscle.Index(i) for i in spec.nodes[k - nl][2]]


def test_invalid_jobs():
    bank = LeafBank(seed=5)
    spec = small_trees([500], dtype="c64")[0]
    leaves = bank.leaves(spec, np.random.default_rng(0))
    nodes = tree_nodes(spec)
    srv = Server(pool_gib=1)
    bad = list(nodes)
    bad[0] = (len(leaves) + 3, bad[0][1], bad[0][2])  # operand after the node
    with pytest.raises(muscle.ArgumentError):
        srv.submit(leaves, bad)
    bad = list(nodes)
    bad[0] = (bad[0][0], bad[0][1], [muscle.Index("not-a-leaf-index")])
    with pytest.raises(muscle.ArgumentError):
        srv.submit(leaves, bad)
    mixed = [muscle.Tensor(t.parent().to(torch.complex64), t.inds()) if k == 0 else t for k, t in enumerate(leaves)]
    with pytest.raises(muscle.ArgumentError):
        srv.submit(mixed, nodes)
    # an output index that neither operand has: found when the job is planned,
    # reported by wait
    bad = list(nodes)
    foreign = next(i for t in leaves for i in t.inds()
                   if i not in operand_inds(spec, leaves, bad[0][0]) + operand_inds(spec, leaves, bad[0][1]))
    bad[0] = (bad[0][0], bad[0][1], list(bad[0][2]) + [foreign])
    jid = srv.submit(leaves, bad)
    with pytest.raises(muscle.ArgumentError):
        srv.wait(jid, timeout=120.0)
    srv.release(jid)
    jid = srv.submit(leaves, nodes)  # the server still works
    assert srv.wait(jid, timeout=120.0)["status"] == 0
    srv.release(jid)
    srv.close()


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_job_with_single_launch_nodes(dtype):
    """A node of 2^30 index points runs as a single-contraction launch inside the
    server's tick (the others in the persistent launch); the root matches the oracle."""
    rng = np.random.default_rng(7)
    a = [f"a{i}" for i in range(10)]
    s = [f"s{i}" for i in range(10)]
    b = [f"b{i}" for i in range(10)]
    c = [f"c{i}" for i in range(5)]
    specs = [(list(rng.permutation(a + s)), 20), (list(rng.permutation(s + b)), 20), (b[:5] + c, 10)]
    vals, leaves = [], []
    for labs, r in specs:
        x = randn(rng, (2,) * r, dtype)
        vals.append(x)
        leaves.append(dev_tensor(x, [muscle.Index(l) for l in labs]))
    mid = list(rng.permutation(a + b))
    out = a + b[5:] + c
    nodes = [(0, 1, [muscle.Index(x) for x in mid]), (3, 2, [muscle.Index(x) for x in out])]
    srv = Server(pool_gib=2)
    jid = srv.submit(leaves, nodes)
    r = srv.wait(jid, timeout=120.0)
    assert r["status"] == 0, r
    got = host(srv.result(jid))
    srv.release(jid)
    srv.close()
    m = bb.binary_einsum_general(mid, vals[0].astype(np.complex128), specs[0][0], vals[1].astype(np.complex128),
                                 specs[1][0])
    ref = bb.binary_einsum_general(out, m, mid, vals[2].astype(np.complex128), specs[2][0])
    assert rel_err(got, ref) <= (1e-12 if dtype == np.complex128 else 2e-5)
