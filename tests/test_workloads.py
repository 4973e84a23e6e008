"""Host-side checks of the synthetic served workload: contraction trees are
well formed, their sizes spread over the stated range, and the per-leaf scale
exponents keep every intermediate's second moment near 1 (checked both from
the moment recursion and numerically with the oracle on a small tree)."""
import numpy as np
import pytest

from muscle.workloads import contraction_tree, leaf_scale_exponents
from oracle import backend_base as bb


def operand_labels(spec, k):
    nl = len(spec.leaves)
    return spec.leaves[k] if k < nl else spec.nodes[k - nl][2]


@pytest.mark.parametrize("seed", range(6))
def test_trees_are_well_formed(seed):
    spec = contraction_tree(seed)
    nl, nn = len(spec.leaves), len(spec.nodes)
    assert 20 <= nn <= 300 and nl == nn + 1 and spec.dtype in ("c32", "c64")
    used = []
    for n, (a, b, out) in enumerate(spec.nodes):
        assert a < nl + n and b < nl + n and a != b  # operands precede the node
        used += [a, b]
        la, lb = operand_labels(spec, a), operand_labels(spec, b)
        for labs in (la, lb, out):
            assert len(set(labs)) == len(labs)
        assert set(out) <= set(la) | set(lb)
        assert len((set(la) & set(lb)) - set(out)) >= 1  # something is contracted
        # no index summed inside one operand: a label of one operand only is kept
        assert (set(la) ^ set(lb)) <= set(out)
    assert sorted(used) == list(range(nl + nn - 1))  # a tree: every operand but the root used once


def test_sizes_spread_log_uniformly():
    ranks = []
    for seed in range(12):
        spec = contraction_tree(seed)
        ranks += [len(o) for _, _, o in spec.nodes] + [len(l) for l in spec.leaves]
    ranks = np.array(ranks)
    assert ranks.min() >= 1 and ranks.max() <= 30
    q = np.percentile(ranks, [10, 50, 90])
    assert 8 <= q[0] <= 13 and 15 <= q[1] <= 21 and 23 <= q[2] <= 29, q


def moments(spec, exps):
    """log2 E|z|^2 of every node when leaf k has E|z|^2 = 4^-exps[k]."""
    nl = len(spec.leaves)
    m = [-2.0 * e for e in exps]
    for n, (a, b, out) in enumerate(spec.nodes):
        c = len((set(operand_labels(spec, a)) & set(operand_labels(spec, b))) - set(out))
        m.append(m[a] + m[b] + c)
    return np.array(m[nl:])


@pytest.mark.parametrize("seed", range(8))
def test_leaf_scales_keep_intermediates_near_unit(seed):
    spec = contraction_tree(1000 + seed)
    scaled = moments(spec, leaf_scale_exponents(spec))
    unscaled = moments(spec, [0] * len(spec.leaves))
    assert np.abs(scaled).max() <= 48, np.abs(scaled).max()  # far inside ComplexF32's range
    assert unscaled.max() > 200  # without the scales the root overflows


def test_leaf_scales_numerically():
    spec = contraction_tree(7, n_nodes=40, size_log2=(2, 7), dtype="c64")
    rng = np.random.default_rng(0)
    vals, labs = [], []
    for inds, e in zip(spec.leaves, leaf_scale_exponents(spec)):
        n = 1 << len(inds)
        z = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2) * 2.0 ** -e
        vals.append(z.reshape((2,) * len(inds), order="F"))
        labs.append(list(inds))
    for a, b, o in spec.nodes:
        vals.append(bb.binary_einsum_general(list(o), vals[a], labs[a], vals[b], labs[b]))
        labs.append(list(o))
    root_ms = np.mean(np.abs(vals[-1]) ** 2)
    assert 2.0 ** -30 < root_ms < 2.0 ** 30, root_ms
