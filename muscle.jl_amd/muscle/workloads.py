"""Synthetic workloads of the contraction hot path (shapes and seeded inputs).

  config1  plain ComplexF64 binary_einsum: A(i,j,k) = (64,32,48),
           B(k,l,j) = (48,96,32), dims = [j,k] -> C(i,l) 64x96; real and
           imaginary parts iid N(0,1)/sqrt(1536).
  config2  rank-6 tensors with every extent 16, contracted modes at positions
           (2,4,6) of A and (1,3,5) of B, output (a1,b1,a3,b2,a5,b3):
           M = N = K = 4096 on interleaved layouts; entries CN(0,1)/sqrt(K).
  config2h hyperindex variant: batch 64 with M = N = K = 512 per batch.

Inputs are column-major like Julia arrays.  The numpy generators draw, for
each operand in turn, all real parts then all imaginary parts from
numpy.random.default_rng(seed).standard_normal, in column-major order.
"""
from __future__ import annotations

import math

import numpy as np


def _cn(rng, shape, scale, dtype=np.complex128):
    n = int(np.prod(shape))
    re = rng.standard_normal(n)
    im = rng.standard_normal(n)
    z = (re + 1j * im) * scale
    return np.reshape(z.astype(dtype), shape, order="F")


def config1_inputs(seed: int = 0):
    """(A, inds_a, B, inds_b, dims, inds_c) for the plain ComplexF64 case."""
    rng = np.random.default_rng(seed)
    s = 1.0 / math.sqrt(1536)
    A = _cn(rng, (64, 32, 48), s)
    B = _cn(rng, (48, 96, 32), s)
    return A, ["i", "j", "k"], B, ["k", "l", "j"], ["j", "k"], ["i", "l"]


CONFIG2_A = ["a1", "c1", "a3", "c2", "a5", "c3"]
CONFIG2_B = ["c1", "b1", "c2", "b2", "c3", "b3"]
CONFIG2_C = ["a1", "b1", "a3", "b2", "a5", "b3"]


def config2_shapes(d: int = 16):
    return [d] * 6, [d] * 6, [d] * 6


def config2_inputs(d: int = 16, seed: int = 0, dtype=np.complex128):
    """Rank-6 interleaved case at extent d per mode (d = 16: M = N = K = 4096)."""
    rng = np.random.default_rng(seed)
    K = d ** 3
    s = 1.0 / math.sqrt(2 * K)  # CN(0,1): E|z|^2 = 1
    A = _cn(rng, (d,) * 6, s, dtype)
    B = _cn(rng, (d,) * 6, s, dtype)
    return A, list(CONFIG2_A), B, list(CONFIG2_B), list(CONFIG2_C)


CONFIG2H_A = ["a1", "c1", "h", "a2", "c2"]
CONFIG2H_B = ["c1", "b1", "h", "c2", "b2"]
CONFIG2H_C = ["a1", "b1", "h", "a2", "b2"]


def config2h_sizes(scale: int = 1):
    """Extents of the hyperindex case; scale = 1 gives batch 64, M = N = K = 512."""
    return {"a1": 16 // scale, "c1": 16 // scale, "h": 64 // scale, "a2": 32 // scale, "c2": 32 // scale,
            "b1": 16 // scale, "b2": 32 // scale}


def config2h_inputs(scale: int = 1, seed: int = 0, dtype=np.complex128):
    sz = config2h_sizes(scale)
    rng = np.random.default_rng(seed)
    K = sz["c1"] * sz["c2"]
    s = 1.0 / math.sqrt(2 * K)
    A = _cn(rng, tuple(sz[i] for i in CONFIG2H_A), s, dtype)
    B = _cn(rng, tuple(sz[i] for i in CONFIG2H_B), s, dtype)
    return A, list(CONFIG2H_A), B, list(CONFIG2H_B), list(CONFIG2H_C)


# ------------------------------------------------------------ precision-ladder distributions
def ladder_inputs(dist: str, seed: int = 0):
    """config-1 shapes under three value distributions:
    "normal"     iid N(0,1) real and imaginary parts;
    "loguniform" |z| log-uniform in [1e-4, 1e4], phase uniform;
    "unitrows"   iid normal, then every row of A's matrix view (A[i, (j,k)]) and
                 every column of B's ((k,j), l) scaled to unit 2-norm."""
    rng = np.random.default_rng(seed)
    sa, sb = (64, 32, 48), (48, 96, 32)
    if dist == "normal":
        A, B = _cn(rng, sa, 1.0), _cn(rng, sb, 1.0)
    elif dist == "loguniform":
        def lu(shape):
            n = int(np.prod(shape))
            mag = 10.0 ** rng.uniform(-4, 4, n)
            ph = rng.uniform(0, 2 * np.pi, n)
            return np.reshape(mag * np.exp(1j * ph), shape, order="F")
        A, B = lu(sa), lu(sb)
    elif dist == "unitrows":
        A, B = _cn(rng, sa, 1.0), _cn(rng, sb, 1.0)
        A = A / np.sqrt(np.sum(np.abs(A) ** 2, axis=(1, 2), keepdims=True))
        B = B / np.sqrt(np.sum(np.abs(B) ** 2, axis=(0, 2), keepdims=True))
    else:
        raise ValueError(dist)
    return A, ["i", "j", "k"], B, ["k", "l", "j"], ["j", "k"], ["i", "l"]


# ------------------------------------------------------------ batched uneven pairs
def circuit_pairs(n: int = 10000, seed: int = 0, ranks=(4, 22), contracted=(1, 10), hyper=(0, 2),
                  out_ranks=(4, 24)):
    """Index structure of n pairwise contractions of a circuit-like network:
    every index has dimension 2; tensor ranks uniform in `ranks`; 1-10
    contracted and 0-2 hyper (batch) indices per pair; output ranks kept in
    `out_ranks` (element counts 2^4..2^24).  Index orders are random
    permutations, as contraction trees produce them.  Returns a list of
    (inds_a, inds_b, inds_c) with labels unique within a pair."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        ra, rb = (int(x) for x in rng.integers(ranks[0], ranks[1] + 1, 2))
        c = int(rng.integers(contracted[0], contracted[1] + 1))
        h = int(rng.integers(hyper[0], hyper[1] + 1))
        L, R = ra - c - h, rb - c - h
        if L < 0 or R < 0 or not (out_ranks[0] <= L + R + h <= out_ranks[1]):
            continue
        lab = iter(range(10 ** 6))
        left = [f"l{next(lab)}" for _ in range(L)]
        right = [f"r{next(lab)}" for _ in range(R)]
        con = [f"c{next(lab)}" for _ in range(c)]
        hyp = [f"h{next(lab)}" for _ in range(h)]
        ia = [x for x in rng.permutation(left + con + hyp)]
        ib = [x for x in rng.permutation(con + hyp + right)]
        ic = [x for x in rng.permutation(left + right + hyp)]
        out.append((ia, ib, ic))
    return out


class CircuitBatchWorkload:
    """The batched uneven-contraction workload on the device: `pairs` (index
    structure from circuit_pairs) split into `rounds`, each a list of
    (c, a, b) muscle.Tensors whose outputs together stay within the output
    budget.  Inputs are column-major dims-2 views at seeded offsets into one
    random pool (read-only, as a tree reuses its tensors); every round's
    outputs are views into one arena, so a round overwrites the previous
    round's outputs."""

    def __init__(self, pairs, rounds, round_pairs, pool, arena):
        self.pairs, self.rounds, self.round_pairs, self.pool, self.arena = pairs, rounds, round_pairs, pool, arena


def circuit_batch_workload(dtype: str = "c64", n: int = 10000, seed: int = 0, budget_gib: float = 96.0,
                           device=None) -> CircuitBatchWorkload:
    """n pairwise contractions of a circuit-like network (every index of dimension 2,
    ranks 4-22, 1-10 contracted and 0-2 hyper indices, outputs of 2^4..2^24
    elements), ComplexF64 ("c64") or ComplexF32 ("c32"), entries iid complex normal
    from a seeded device generator.  round_pairs[r] lists the pair indices of round r."""
    import torch

    from .tensor import Tensor
    dev = device if device is not None else torch.device("cuda", 0)
    tdt = torch.complex128 if dtype == "c64" else torch.complex64
    esz = 16 if dtype == "c64" else 8
    pairs = circuit_pairs(n, seed)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    pool_n = 1 << 23
    pool = torch.randn(pool_n, dtype=tdt, device=dev, generator=g)
    rng = np.random.default_rng(seed + 1)
    budget = int(budget_gib * 2 ** 30) // esz
    round_idx, cur, used = [], [], 0
    for i, p in enumerate(pairs):
        sz = 1 << len(p[2])
        if cur and used + sz > budget:
            round_idx.append(cur)
            cur, used = [], 0
        cur.append(i)
        used += sz
    round_idx.append(cur)
    arena_n = max(sum(1 << len(pairs[i][2]) for i in r) for r in round_idx)
    arena = torch.empty(arena_n, dtype=tdt, device=dev)

    def dims2(base, rank, off):
        return base.as_strided((2,) * rank, tuple(1 << i for i in range(rank)), off)

    rounds = []
    for r in round_idx:
        ts, off = [], 0
        for i in r:
            ia, ib, ic = pairs[i]
            a = Tensor(dims2(pool, len(ia), int(rng.integers(0, pool_n - (1 << len(ia)) + 1))), ia)
            b = Tensor(dims2(pool, len(ib), int(rng.integers(0, pool_n - (1 << len(ib)) + 1))), ib)
            c = Tensor(dims2(arena, len(ic), off), ic)
            off += 1 << len(ic)
            ts.append((c, a, b))
        rounds.append(ts)
    return CircuitBatchWorkload(pairs, rounds, round_idx, pool, arena)


def pair_flops(pair, complex_=True) -> float:
    ia, ib, ic = pair
    labels = set(ia) | set(ib)
    return (8.0 if complex_ else 2.0) * 2.0 ** len(labels)


# ------------------------------------------------------------ served contraction trees
class TreeSpec:
    """Index structure of one contraction-tree job.

    leaves: list of index-label lists; nodes: list of (lhs, rhs, out_labels)
    in topological order, operands numbered leaves first then nodes; the last
    node is the root.  Every label has dimension 2; dtype is "c32" or "c64"."""

    def __init__(self, leaves, nodes, dtype):
        self.leaves, self.nodes, self.dtype = leaves, nodes, dtype

    def sizes_log2(self):
        return [len(n[2]) for n in self.nodes]


def contraction_tree(seed: int, n_nodes=None, size_log2=(8, 28), max_contracted=10, hyper=(0, 1),
                     dtype=None) -> TreeSpec:
    """A seeded random contraction tree of circuit-like tensors (every index of
    dimension 2) with 20-300 pairwise contractions and a ComplexF32 or
    ComplexF64 dtype.  Built top-down from the root: every tensor of the tree
    (intermediates and leaves) gets a rank drawn uniformly from size_log2, so
    sizes are log-uniform in 2^8..2^28; an expansion splits a node's open
    indices between its two children, keeps 0-1 of them as a hyper (batch)
    index on both, and adds as many new contracted indices (1..max_contracted)
    as the children's ranks need."""
    rng = np.random.default_rng(seed)
    lo, hi = size_log2
    if n_nodes is None:
        n_nodes = int(rng.integers(20, 301))
    if dtype is None:
        dtype = "c32" if rng.integers(0, 2) == 0 else "c64"
    counter = [0]

    def fresh(k):
        out = list(range(counter[0], counter[0] + k))
        counter[0] += k
        return out

    root = {"out": fresh(int(rng.integers(lo, hi + 1))), "kids": None}
    frontier, internal = [root], []
    while len(internal) < n_nodes:
        node = frontier.pop(int(rng.integers(len(frontier))))
        O = node["out"]
        r = len(O)
        h = min(int(rng.integers(hyper[0], hyper[1] + 1)), r)
        H = [O[i] for i in rng.choice(r, h, replace=False)] if h else []
        rest = [x for x in O if x not in H]
        tl, tr = (int(x) for x in rng.integers(lo, hi + 1, 2))
        c = int(np.clip(round((tl + tr - r - h) / 2), 1, max_contracted))
        nl = int(np.clip(tl - c - h, 0, len(rest)))
        C = fresh(c)
        perm = [rest[i] for i in rng.permutation(len(rest))]
        L, R = perm[:nl], perm[nl:]
        left = {"out": [int(x) for x in rng.permutation(L + C + H)], "kids": None}
        right = {"out": [int(x) for x in rng.permutation(R + C + H)], "kids": None}
        node["kids"] = (left, right)
        internal.append(node)
        frontier += [left, right]
    leaves = frontier
    ids = {id(n): i for i, n in enumerate(leaves)}
    nodes = []

    def emit(n):
        if n["kids"] is None:
            return ids[id(n)]
        a = emit(n["kids"][0])
        b = emit(n["kids"][1])
        nodes.append((a, b, list(n["out"])))
        return len(leaves) + len(nodes) - 1

    import sys
    old = sys.getrecursionlimit()
    sys.setrecursionlimit(max(old, 10000))
    try:
        emit(root)
    finally:
        sys.setrecursionlimit(old)
    return TreeSpec([list(n["out"]) for n in leaves], nodes, dtype)


def leaf_scale_exponents(spec: TreeSpec) -> list:
    """Integer e per leaf such that leaves with iid entries of E|z|^2 = 4^-e keep
    every node's expected E|z|^2 near 1.

    A node contracting c indices of dimension 2 multiplies its operands'
    second moments and 2^c; the factor 2^-c is shared equally by the leaves of
    each child subtree (2^(-c/2) per subtree), and the per-leaf log2 sums are
    rounded to the nearest even number (so the scale 2^-e is a power of two)."""
    nl, nn = len(spec.leaves), len(spec.nodes)
    labels = [set(l) for l in spec.leaves] + [set(n[2]) for n in spec.nodes]
    count = [1] * nl + [0] * nn
    for n, (a, b, _) in enumerate(spec.nodes):
        count[nl + n] = count[a] + count[b]
    share = [0.0] * (nl + nn)  # log2 second moment to add to every leaf below
    for n in range(nn - 1, -1, -1):
        a, b, _ = spec.nodes[n]
        k = nl + n
        c = len((labels[a] & labels[b]) - labels[k])
        for ch in (a, b):
            share[ch] += share[k] - c / 2.0 / count[ch]
    return [int(round(-share[i] / 2.0)) for i in range(nl)]
