"""Host-side model builders and synthetic data (product code) vs the oracle."""
import numpy as np
import pytest

from oracle import subby_oracle as O
from subby_amd import models as M
from subby_amd import synthetic as S


def test_models_match_oracle():
    pi = np.array([0.3, 0.2, 0.25, 0.25])
    for a, b in [(M.hky85_diag(2.0, pi), O.hky85_diag(2.0, pi)),
                 (M.jukes_cantor_model(20), O.jukes_cantor_model(20)),
                 (M.f81_model(np.arange(1, 7) / 21.0), O.f81_model(np.arange(1, 7) / 21.0))]:
        for x, y in zip(a, b):
            np.testing.assert_allclose(x, y, atol=1e-14)
    np.testing.assert_allclose(M.hky85_rate_matrix(2.0, pi), O.hky85_rate_matrix(2.0, pi), atol=1e-15)
    Q = M.reversible_rate_matrix(np.random.default_rng(0).lognormal(size=(5, 5)), np.ones(5) / 5)
    m = M.diagonalize_rate_matrix(Q, np.ones(5) / 5)
    np.testing.assert_allclose(O.reconstruct_rate_matrix(O.DiagModel(*m)), Q, atol=1e-13)


def test_log_kappa_weights_match_oracle_gradient():
    rng = np.random.default_rng(1)
    tree = S.random_tree(6, rng, 0.2)
    pi = rng.dirichlet(np.ones(4) * 5)
    aln = S.simulate(tree, M.hky85_diag(2.0, pi), 30, rng, gap_frac=0.1, ambig_frac=0.05)
    W = M.hky85_log_kappa_weights(2.0, pi)
    N = O.Counts(aln, O.Tree(*tree), O.hky85_diag(2.0, pi), branch_mask=None)
    got = np.einsum('ij,ijc->', W, N)
    np.testing.assert_allclose(got, O.hky85_loglike_kappa_grad(aln, O.Tree(*tree), 2.0, pi), rtol=1e-12)


def test_random_tree_is_binary_preorder():
    rng = np.random.default_rng(2)
    for n in [1, 2, 3, 17, 64, 100]:
        t = S.random_tree(n, rng, 0.1)
        R = len(t.parentIndex)
        assert R == 2 * n - 1
        O.validate_binary_tree(t.parentIndex)
        assert t.distanceToParent[0] == 0 and np.all(t.distanceToParent[1:] > 0)


def test_simulation_statistics():
    """Leaf frequencies of a star-like tree approach pi; gap/ambiguous rates hold."""
    rng = np.random.default_rng(3)
    t = S.random_tree(8, rng, 0.05)
    pi = np.array([0.1, 0.2, 0.3, 0.4])
    aln = S.simulate(t, M.f81_model(pi), 40000, rng, gap_frac=0.05, ambig_frac=0.02)
    assert aln.dtype == np.int8
    internal = np.zeros(len(t.parentIndex), bool)
    internal[t.parentIndex[1:]] = True
    assert np.all(aln[internal] == 4)
    leaves = aln[~internal]
    assert abs((leaves == -1).mean() - 0.05) < 0.005
    assert abs((leaves == 4).mean() - 0.02) < 0.004
    freq = np.array([(leaves == a).sum() for a in range(4)], float)
    np.testing.assert_allclose(freq / freq.sum(), pi, atol=0.01)
