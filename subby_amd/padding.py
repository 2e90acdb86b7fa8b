"""Column and node padding helpers (subby/jax/_utils.py:10-27, 104-229).

The reference pads alignments (gap columns) and trees (dummy leaves under the
root) to a few bin sizes so that JAX does not recompile for every shape.  The
kernels here take any R and C, so padding is never needed; these helpers keep
the reference's API for callers that already pad.  Gap columns are neutral
(logL = 0, zero masked counts, root posterior = pi).  Trailing dummy leaves
(parent 0, length 0, all-gap rows) are recognised by the API and dropped
before a call, since they make the root non-binary; they change nothing in
LogLike, RootProb or Counts, as in the reference (upward-only semantics).
"""
from __future__ import annotations

import numpy as np
import torch

from .models import Tree

# Geometric bin sizes ~1.19^k (O(log n) distinct sizes).
GEOM_BINS = sorted(set(max(1, round(1.19 ** k)) for k in range(120)))


def pad_to_geom_bin(n: int) -> int:
    """Smallest geometric bin >= n; beyond the table the next power of two."""
    for b in GEOM_BINS:
        if b >= n:
            return b
    return int(2 ** np.ceil(np.log2(max(n, 1))))


def _cat(a, pad, axis):
    if isinstance(a, torch.Tensor):
        return torch.cat([a, torch.as_tensor(pad, dtype=a.dtype, device=a.device)], dim=axis)
    return np.concatenate([a, pad.astype(np.asarray(a).dtype)], axis=axis)


def pad_alignment(alignment, bin_size: int = 128, bin_fn=None):
    """Pad columns with gap tokens (-1) up to a multiple of bin_size, or to
    bin_fn(C).  Returns (padded alignment, original C)."""
    C = alignment.shape[1]
    Cp = bin_fn(C) if bin_fn is not None else -(-C // bin_size) * bin_size
    if Cp == C:
        return alignment, C
    return _cat(alignment, np.full((alignment.shape[0], Cp - C), -1), 1), C


def pad_tree(tree, n_nodes_padded: int):
    """Append dummy leaves under the root with zero branch length."""
    parent = np.asarray(tree.parentIndex)
    dist = np.asarray(tree.distanceToParent, dtype=np.float64)
    n = n_nodes_padded - len(parent)
    if n <= 0:
        return tree
    return Tree(np.concatenate([parent, np.zeros(n, dtype=parent.dtype)]),
                np.concatenate([dist, np.zeros(n)]))


def pad_tree_and_alignment(tree, alignment, node_bin_fn=None, col_bin_fn=None):
    """Pad nodes (dummy leaves, gap rows) and columns (gap columns) to bins.
    Returns (tree, alignment, original R, original C)."""
    node_bin_fn = node_bin_fn or pad_to_geom_bin
    col_bin_fn = col_bin_fn or pad_to_geom_bin
    R, C = alignment.shape
    Rp, Cp = node_bin_fn(R), col_bin_fn(C)
    tree = pad_tree(tree, Rp)
    if Rp > R:
        alignment = _cat(alignment, np.full((Rp - R, C), -1), 0)
    if Cp > C:
        alignment = _cat(alignment, np.full((Rp, Cp - C), -1), 1)
    return tree, alignment, R, C


def unpad_columns(result, C_original: int):
    """Drop padding columns (last axis)."""
    return result[..., :C_original]


def strip_dummy_leaves(parent, dist, aln) -> int:
    """Number of nodes to keep: trailing leaves under the root with zero length
    and an all-gap (-1) row, the dummies of pad_tree, are dropped when the root
    has more than two children.  ``aln`` is the (R, C) alignment tensor."""
    R = len(parent)
    if R < 4 or int(np.sum(parent[1:] == 0)) <= 2:
        return R
    is_parent = np.zeros(R, dtype=bool)
    is_parent[parent[1:]] = True
    keep = R
    while keep > 3 and parent[keep - 1] == 0 and dist[keep - 1] == 0.0 and not is_parent[keep - 1]:
        keep -= 1
    if keep < R and aln.shape[1] and not bool((aln[keep:] == -1).all()):
        return R
    return keep
