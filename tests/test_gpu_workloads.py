"""Parity at the full size of the benchmark workloads, against the CPU oracle.

  config 2, hyperindex variant: batch 64 x (M = N = K = 512), ComplexF64 and
      ComplexF32.  256 sampled output entries against the oracle's dot products
      of the same inputs, and the size-independent identity
      sum_n C[m, n, h] = A[m, :, h] . (sum_n B[:, n, h]) over every (m, h).
  config 4: the 10,000-pair dims-2 workload exactly as the benchmark builds it
      (muscle.workloads.circuit_batch_workload, autotuned tile shapes), both
      element types.  A seeded sample of pairs that covers every route
      (persistent tiles, K slices, shallow, skinny, solo) and every size class; each
      GPU output is checked against the oracle run on the GPU's own inputs.
  config 5: served jobs at the specified sizes (trees of 20-300 contractions,
      every tensor log-uniform in 2^8..2^28 elements), two per element type,
      run together in one server; each root against the oracle evaluating the
      same tree node by node on the same leaves.  The oracle's cost decides
      which jobs: the two of each element type, among the benchmark's first
      128 job seeds, whose node-by-node evaluation moves the fewest elements.

Bars: 1e-12 norm-wise relative error for ComplexF64, 2e-5 for ComplexF32.
"""
import numpy as np
import pytest
import torch

import muscle
from muscle.batch import ContractionBatch
from muscle.workloads import circuit_batch_workload, config2h_inputs, contraction_tree
from oracle import backend_base as bb
from tests.gpu_util import dev_tensor, host, rel_err

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

TOL = {"c64": 1e-12, "c32": 2e-5}


# ------------------------------------------------------------ config 2, hyperindex variant
@pytest.mark.parametrize("T", [np.complex128, np.complex64])
def test_config2_hyper_full_size(T):
    A, ia, B, ib, ic = config2h_inputs(1, seed=2, dtype=T)
    c = muscle.binary_einsum(dev_tensor(A, ia), dev_tensor(B, ib), out=ic)
    info = muscle.last_info()
    assert (info["batch"], info["m"], info["n"], info["k"]) == (64, 512, 512, 512) and info["kernel"] == 1
    C = host(c)
    assert [i.tag for i in c.inds()] == ic and C.shape == (16, 16, 64, 32, 32)
    # matrix views per batch: A[(a1, a2), (c1, c2), h], B[(c1, c2), (b1, b2), h], C[(a1, a2), (b1, b2), h]
    Am = np.reshape(np.transpose(A, (0, 3, 1, 4, 2)), (512, 512, 64), order="F").astype(np.complex128)
    Bm = np.reshape(np.transpose(B, (0, 3, 1, 4, 2)), (512, 512, 64), order="F").astype(np.complex128)
    Cm = np.reshape(np.transpose(C, (0, 3, 1, 4, 2)), (512, 512, 64), order="F").astype(np.complex128)
    rng = np.random.default_rng(4)
    ms, ns, hs = rng.integers(0, 512, 256), rng.integers(0, 512, 256), rng.integers(0, 64, 256)
    ref = np.array([Am[m, :, h] @ Bm[:, n, h] for m, n, h in zip(ms, ns, hs)])
    tol = TOL["c64" if T == np.complex128 else "c32"]
    assert rel_err(Cm[ms, ns, hs], ref) <= tol
    lhs = Cm.sum(axis=1)
    rhs = np.einsum("mkh,kh->mh", Am, Bm.sum(axis=1))
    assert rel_err(lhs, rhs) <= (1e-10 if T == np.complex128 else 1e-4)


# ------------------------------------------------------------ config 4
def _sample(batches, wl, seed, per_group=3, total=240):
    """Seeded pair sample: up to per_group pairs of every (route, size class)
    present, topped up at random to `total`."""
    groups = {}
    for r, bt in enumerate(batches):
        for j, i in enumerate(wl.round_pairs[r]):
            q = bt.pair_info(j)
            groups.setdefault((q["route"], q["size_class"]), []).append(i)
    rng = np.random.default_rng(seed)
    pick = set()
    for key in sorted(groups):
        g = groups[key]
        pick.update(int(x) for x in rng.choice(g, size=min(per_group, len(g)), replace=False))
    rest = [i for i in range(len(wl.pairs)) if i not in pick]
    pick.update(int(x) for x in rng.choice(rest, size=max(0, total - len(pick)), replace=False))
    return pick, groups


@pytest.mark.parametrize("dtype", ["c64", "c32"])
def test_config4_full_workload(dtype):
    wl = circuit_batch_workload(dtype, 10000, seed=0, budget_gib=96.0)
    probe = ContractionBatch(wl.rounds[0])
    probe.autotune(samples=16, reps=2)
    probe.close()
    batches = [ContractionBatch(ts) for ts in wl.rounds]
    pick, groups = _sample(batches, wl, seed=7)
    routes = {k[0] for k in groups}
    assert {"tiles", "split_k", "shallow", "skinny", "solo"} <= routes, routes
    print(f"config4 {dtype}: {len(wl.rounds)} rounds, {len(pick)} sampled pairs, "
          f"{len(groups)} (route, class) groups", flush=True)
    got = {}
    for r, bt in enumerate(batches):
        wl.arena.fill_(float("nan"))  # every sampled output must be written by this round
        bt.run("persistent")
        torch.cuda.synchronize()
        for j, i in enumerate(wl.round_pairs[r]):
            if i in pick:
                got[i] = wl.rounds[r][j][0].parent().cpu().numpy()
    assert all(bt.info()["graph"] == 1 for bt in batches)
    where = {i: (r, j) for r, idx in enumerate(wl.round_pairs) for j, i in enumerate(idx)}
    worst = {}
    for n_done, i in enumerate(sorted(pick)):
        if n_done % 40 == 0:
            print(f"config4 {dtype}: oracle pair {n_done}/{len(pick)}", flush=True)
        r, j = where[i]
        c, a, b = wl.rounds[r][j]
        ia, ib, ic = wl.pairs[i]
        ref = bb.binary_einsum_general(ic, a.parent().cpu().numpy().astype(np.complex128), ia,
                                       b.parent().cpu().numpy().astype(np.complex128), ib)
        route = batches[r].pair_info(j)["route"]
        err = rel_err(got[i], ref)
        worst[route] = max(worst.get(route, 0.0), err)
        assert err <= TOL[dtype], (i, route, batches[r].pair_info(j), err)
    print(f"config4 {dtype}: worst error per route {worst}", flush=True)
    for bt in batches:
        bt.close()
    del wl
    torch.cuda.empty_cache()


# ------------------------------------------------------------ config 5
SERVED_JOBS = {"c64": (1051, 1067), "c32": (1115, 1007)}


def _oracle_tree(spec, leaves):
    """The tree evaluated node by node on the host (each value freed after its last use)."""
    nl = len(spec.leaves)
    last = {}
    for n, (a, b, _) in enumerate(spec.nodes):
        last[a] = last[b] = n
    vals = {k: host(t).astype(np.complex128) for k, t in enumerate(leaves)}
    labs = {k: list(l) for k, l in enumerate(spec.leaves)}
    for n, (a, b, o) in enumerate(spec.nodes):
        vals[nl + n] = bb.binary_einsum_general(list(o), vals[a], labs[a], vals[b], labs[b])
        labs[nl + n] = list(o)
        for x in (a, b):
            if last[x] == n:
                del vals[x]
    return vals[nl + len(spec.nodes) - 1]


def test_config5_full_size_jobs():
    from muscle.serve import LeafBank, Server, tree_nodes
    specs = [contraction_tree(s, size_log2=(8, 28)) for dt in ("c64", "c32") for s in SERVED_JOBS[dt]]
    assert [s.dtype for s in specs] == ["c64", "c64", "c32", "c32"]
    assert max(len(o) for s in specs for _, _, o in s.nodes) >= 26  # intermediates of 2^26+ elements
    bank = LeafBank(seed=9)
    bank.reserve(specs)
    srv = Server(pool_gib=96)
    rng = np.random.default_rng(9)
    prepared = [srv.prepare(bank.leaves(s, rng), tree_nodes(s)) for s in specs]
    ids = [srv.submit_prepared(pj) for pj in prepared]  # one mixed ComplexF32 / ComplexF64 batch of jobs
    roots = []
    for jid in ids:
        res = srv.wait(jid, timeout=300.0)
        assert res["status"] == 0, res
        roots.append(host(srv.result(jid)))
        srv.release(jid)
    srv.close()
    print("config5: GPU jobs done, oracle trees next", flush=True)
    for spec, pj, got in zip(specs, prepared, roots):
        ref = _oracle_tree(spec, pj.leaves)
        err = rel_err(got, ref)
        print(f"config5 job {spec.dtype} nodes={len(spec.nodes)}: rel err {err:.2e}", flush=True)
        assert err <= TOL[spec.dtype], (spec.dtype, len(spec.nodes), err)
    del bank
    torch.cuda.empty_cache()
