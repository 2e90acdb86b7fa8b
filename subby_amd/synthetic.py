"""Seeded synthetic trees and alignments (numpy.random.Generator / PCG64).

Trees: Yule process (split a uniformly chosen leaf), relabelled to DFS
preorder so parentIndex[n] < n, root branch length 0, other branches
Exp(mean).  Alignments: root state ~ pi, child state ~ M(t)[parent state]
(the simulation of the reference's tests/test_api.py:34-71), internal rows
set to the ungapped-unobserved token A, then leaf cells replaced by the gap
token -1 or the ambiguous token A at the requested rates.
"""
from __future__ import annotations

import numpy as np

from .models import Tree, DiagModel, IrrevDiagModel


def _sub_matrices(model, t) -> np.ndarray:
    """(R, A, A) M(t) of a DiagModel (symmetrised basis) or an IrrevDiagModel
    (Re V e^{mu t} V^{-1}), clipped at 0 for sampling."""
    t = np.asarray(t, dtype=np.float64)
    if isinstance(model, IrrevDiagModel):
        mu = np.asarray(model.eigenvalues, dtype=np.complex128)
        V = np.asarray(model.eigenvectors, dtype=np.complex128)
        Vi = np.asarray(model.eigenvectors_inv, dtype=np.complex128)
        M = np.einsum('ak,rk,kb->rab', V, np.exp(mu[None, :] * t[:, None]), Vi).real
    else:
        mu, V, pi = (np.asarray(x, dtype=np.float64) for x in model)
        s = np.sqrt(pi)
        M = np.einsum('ak,rk,bk->rab', V, np.exp(mu[None, :] * t[:, None]), V) / s[None, :, None] \
            * s[None, None, :]
    return np.clip(M, 0.0, None)


def random_tree(n_leaves: int, rng, mean_branch: float = 0.1) -> Tree:
    """Random binary tree with n_leaves leaves (2*n_leaves-1 nodes), preorder."""
    if n_leaves < 1:
        raise ValueError('n_leaves must be >= 1')
    children = {0: []}
    leaves = [0]
    nxt = 1
    while len(leaves) < n_leaves:
        k = int(rng.integers(len(leaves)))
        v = leaves[k]
        a, b = nxt, nxt + 1
        nxt += 2
        children[v] = [a, b]
        children[a] = []
        children[b] = []
        leaves[k] = a
        leaves.append(b)
    # DFS preorder relabelling
    order, parent_of = [], {0: -1}
    stack = [0]
    while stack:
        v = stack.pop()
        order.append(v)
        for c in reversed(children[v]):
            parent_of[c] = v
            stack.append(c)
    new = {v: i for i, v in enumerate(order)}
    R = len(order)
    parent = np.full(R, -1, dtype=np.int32)
    for v in order[1:]:
        parent[new[v]] = new[parent_of[v]]
    dist = rng.exponential(mean_branch, size=R)
    dist[0] = 0.0
    return Tree(parent, dist.astype(np.float64))


def sub_matrix(model, t: float) -> np.ndarray:
    return _sub_matrices(model, [t])[0]


def simulate(tree: Tree, model, C: int, rng, gap_frac: float = 0.0,
             ambig_frac: float = 0.0, dtype=np.int8) -> np.ndarray:
    """(R, C) tokens simulated down the tree; internal nodes get token A."""
    parent = np.asarray(tree.parentIndex)
    R = len(parent)
    pi = np.asarray(model.pi, dtype=np.float64)
    A = len(pi)
    M = _sub_matrices(model, tree.distanceToParent)
    cum = np.cumsum(M / M.sum(2, keepdims=True), axis=2)
    cum[..., -1] = 1.0
    states = np.empty((R, C), dtype=np.int16)
    u = rng.random((R, C))
    cpi = np.cumsum(pi)
    cpi[-1] = 1.0
    states[0] = np.searchsorted(cpi, u[0], side='right')
    for n in range(1, R):
        ps = states[parent[n]]
        if A <= 8:
            states[n] = (u[n][:, None] >= cum[n][ps]).sum(1)
        else:
            out = np.empty(C, dtype=np.int16)
            for a in np.unique(ps):
                idx = np.nonzero(ps == a)[0]
                out[idx] = np.searchsorted(cum[n, a], u[n, idx], side='right')
            states[n] = out
    np.minimum(states, A - 1, out=states)
    is_internal = np.zeros(R, dtype=bool)
    is_internal[parent[1:]] = True
    aln = states
    aln[is_internal] = A
    if gap_frac > 0 or ambig_frac > 0:
        g = rng.random((R, C))
        leaf = ~is_internal[:, None]
        aln[leaf & (g < gap_frac)] = -1
        aln[leaf & (g >= gap_frac) & (g < gap_frac + ambig_frac)] = A
    return aln.astype(dtype)


def dirichlet(alpha, rng):
    return rng.dirichlet(np.asarray(alpha, dtype=np.float64))
