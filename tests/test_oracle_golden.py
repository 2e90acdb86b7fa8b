"""The FP64 oracle against every array of the reference's golden fixtures.

The fixtures were written by the reference's own oracle
(scripts/generate_golden_tests.py:54-112) and hold inputs, all intermediates
(M, U, logNormU, D, logNormD, J, U~, D~, C_eigen) and outputs (logLike,
unmasked counts, root posterior, Steiner branch mask).
"""
import numpy as np
import pytest

from oracle import subby_oracle as O
from tests.golden_cases import CASES, load

TOL = 1e-12


def _inputs(g):
    m = O.DiagModel(g['inputs.model.eigenvalues'], g['inputs.model.eigenvectors'], g['inputs.model.pi'])
    t = O.Tree(g['inputs.parentIndex'], g['inputs.distanceToParent'])
    return g['inputs.alignment'], t, m


def _close(a, b, tol=TOL):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(1.0, np.abs(b).max())
    assert a.shape == b.shape
    err = np.abs(a - b).max() / scale
    assert err <= tol, err


@pytest.mark.parametrize('case', CASES)
def test_intermediates(case):
    g = load(case)
    aln, tree, m = _inputs(g)
    M, U, lnU, ll, D, lnD = O.inside_outside(aln, tree, m)
    _close(M, g['intermediates.sub_matrices'])
    _close(U, g['intermediates.U'])
    _close(lnU, g['intermediates.logNormU'])
    _close(D, g['intermediates.D'])
    _close(lnD, g['intermediates.logNormD'])
    J = O.compute_J(m.eigenvalues, tree.distanceToParent)
    _close(J, g['intermediates.J'])
    Ut, Dt = O.eigenbasis_project(U, D, m)
    _close(Ut, g['intermediates.U_tilde'])
    _close(Dt, g['intermediates.D_tilde'])
    Ce = O.accumulate_C(Dt, Ut, J, lnU, lnD, ll, tree.parentIndex)
    _close(Ce, g['intermediates.C_eigen'])


@pytest.mark.parametrize('case', CASES)
def test_outputs(case):
    g = load(case)
    aln, tree, m = _inputs(g)
    _close(O.LogLike(aln, tree, m), g['outputs.logLike'])
    _close(O.Counts(aln, tree, m, branch_mask=None), g['outputs.counts'])
    _close(O.RootProb(aln, tree, m), g['outputs.root_prob'])
    np.testing.assert_array_equal(O.compute_branch_mask(aln, tree.parentIndex, len(m.pi)),
                                  g['outputs.branch_mask'])


@pytest.mark.parametrize('case', ['5node_jc4', '7node_jc64', '7node_jc4_mixed_gaps', '7node_jc4_all_gaps'])
def test_f81_fast_path_equals_golden_counts(case):
    """F81/JC closed form == eigensub counts (exact identity, f81_fast.py:7-141)."""
    g = load(case)
    aln, tree, m = _inputs(g)
    _close(O.Counts(aln, tree, m, f81_fast=True, branch_mask=None), g['outputs.counts'], 1e-11)


def test_golden_counts_are_unmasked_and_dwell_sums():
    """Sum_i counts[i,i,c] = sum_{n>=1} t_n in every golden (unmasked) column."""
    for case in CASES:
        g = load(case)
        cnt = g['outputs.counts']
        T = g['inputs.distanceToParent'][1:].sum()
        np.testing.assert_allclose(np.einsum('iic->c', cnt), T, rtol=1e-12)


def test_mixture_loglikes():
    """Rate-scaled JC log-likelihoods of the mixture golden (oracle.py:1454-1467)."""
    g = load('7node_mixture_jc4') if False else None
    import os
    from tests.golden_cases import GOLDEN
    d = np.load(os.path.join(GOLDEN, '7node_mixture_jc4.npz'))
    base = O.DiagModel(d['inputs.base_model.eigenvalues'], d['inputs.base_model.eigenvectors'],
                       d['inputs.base_model.pi'])
    tree = O.Tree(d['inputs.parentIndex'], d['inputs.distanceToParent'])
    for k, r in enumerate(d['inputs.rates']):
        m = O.DiagModel(base.eigenvalues * r, base.eigenvectors, base.pi)
        _close(O.LogLike(d['inputs.alignment'], tree, m), d['outputs.log_likes'][k])
