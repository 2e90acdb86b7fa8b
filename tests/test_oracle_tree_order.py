"""The oracle on trees whose node order is not depth-first.

The reference only requires parents before children (parentIndex[n] < n,
docs/architecture.md:39-48), and its own tests and benchmarks build balanced
trees breadth-first: parentIndex[i] = (i - 1) // 2 (tests/test_oracle.py:59,
benchmarks/run_benchmarks.py:119-128).  Two independent pins:

* exhaustive enumeration on a 9-node breadth-first tree (the smallest such
  tree in which an internal node's first child is not the next internal node
  in index order);
* relabelling invariance: every output on a breadth-first tree equals the
  output on the same tree renumbered depth-first (alignment rows and per-node
  outputs permuted alike).
"""
import numpy as np
import pytest

from oracle import subby_oracle as O
from tests.test_oracle_invariants import _enumerate, _I_integral, _models


def heap_tree(R, rng, lo=0.02, hi=0.5):
    """Balanced tree in breadth-first order, as the reference builds it."""
    parent = np.array([-1] + [(i - 1) // 2 for i in range(1, R)], dtype=np.int32)
    t = rng.uniform(lo, hi, size=R)
    t[0] = 0.0
    return O.Tree(parent, t)


def dfs_relabel(tree):
    """(tree renumbered in depth-first preorder, order): new node k is old node order[k]."""
    parent = np.asarray(tree.parentIndex)
    R = len(parent)
    kids = [[] for _ in range(R)]
    for n in range(1, R):
        kids[parent[n]].append(n)
    order, st = [], [0]
    while st:
        v = st.pop()
        order.append(v)
        st.extend(reversed(kids[v]))
    order = np.array(order)
    new = np.empty(R, dtype=np.int64)
    new[order] = np.arange(R)
    p2 = np.array([-1] + [new[parent[o]] for o in order[1:]], dtype=np.int32)
    return O.Tree(p2, np.asarray(tree.distanceToParent)[order]), order


def test_heap_tree_is_not_depth_first():
    t = heap_tree(9, np.random.default_rng(0))
    t2, order = dfs_relabel(t)
    assert not np.array_equal(order, np.arange(9))
    assert np.all(t2.parentIndex[1:] < np.arange(1, 9))


@pytest.mark.parametrize('A', [3])
def test_heap_ordered_tree_by_enumeration(A):
    rng = np.random.default_rng(41)
    tree = heap_tree(9, rng)
    R, C = 9, 5
    aln = rng.integers(-1, A + 2, size=(R, C))
    aln[:, 0] = rng.integers(0, A, size=R)
    for name, m in _models(A, rng):
        if name != 'gtr':
            continue
        P, root, pair, M, Q = _enumerate(aln, tree, m)
        np.testing.assert_allclose(O.LogLike(aln, tree, m), np.log(P), rtol=1e-11)
        np.testing.assert_allclose(O.RootProb(aln, tree, m), root, atol=1e-11)
        exp_cnt = np.zeros((A, A, C))
        for n in range(1, R):
            I = _I_integral(Q, tree.distanceToParent[n])
            W = pair[n] / M[n][:, :, None]
            E = np.einsum('abc,abij->ijc', W, I)
            off = ~np.eye(A, dtype=bool)
            exp_cnt += np.where(off[..., None], Q[..., None] * E, E)
        np.testing.assert_allclose(O.Counts(aln, tree, m, branch_mask=None), exp_cnt, atol=1e-9)


@pytest.mark.parametrize('R', [15, 31, 63])
def test_relabelling_invariance(R):
    rng = np.random.default_rng(R)
    tree = heap_tree(R, rng)
    t2, order = dfs_relabel(tree)
    pi = rng.dirichlet(np.full(4, 5.0))
    m = O.hky85_diag(2.0, pi)
    C = 60
    aln = rng.integers(-1, 6, size=(R, C))
    aln[np.asarray(tree.parentIndex)[1:]] = 4            # internal rows unobserved
    aln2 = aln[order]
    for mask in ('auto', None):
        c1, l1, r1 = O.counts_loglike_root(aln, tree, m, branch_mask=mask)
        c2, l2, r2 = O.counts_loglike_root(aln2, t2, m, branch_mask=mask)
        np.testing.assert_allclose(l1, l2, rtol=1e-13, atol=1e-12)
        np.testing.assert_allclose(c1, c2, rtol=1e-11, atol=1e-13)
        np.testing.assert_allclose(r1, r2, atol=1e-13)
    mk = rng.random((R, C)) < 0.6
    np.testing.assert_allclose(O.Counts(aln, tree, m, branch_mask=mk), O.Counts(aln2, t2, m, branch_mask=mk[order]),
                               rtol=1e-11, atol=1e-13)
    b1 = O.BranchCounts(aln, tree, m, branch_mask='auto')
    b2 = O.BranchCounts(aln2, t2, m, branch_mask='auto')
    np.testing.assert_allclose(b1[order], b2, rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(O.node_posterior(aln, tree, m)[order], O.node_posterior(aln2, t2, m), atol=1e-12)
    np.testing.assert_allclose(O.branch_posterior(aln, tree, m)[order], O.branch_posterior(aln2, t2, m), atol=1e-12)
    g1, _ = O.loglike_branch_grad(aln, tree, m)
    g2, _ = O.loglike_branch_grad(aln2, t2, m)
    np.testing.assert_allclose(np.asarray(g1)[order], g2, rtol=1e-10, atol=1e-12)
