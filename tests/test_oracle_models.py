"""Pin the oracle's rate categories, mixtures, per-column model lists and
per-branch model grids.

* MixturePosterior and its per-component log-likelihoods against the
  reference's own golden fixture (tests/golden/7node_mixture_jc4: K = 3
  rate-scaled JC models);
* gamma rate categories against scipy.stats.gamma.ppf (the quantile
  function the reference's bisection approximates);
* per-column model lists: [m] * C equals the single model and column c equals
  that column run alone (tests/test_api.py:397-430 of the reference); the
  root-distribution convention (models[0].pi) against enumeration;
* per-branch models: LogLike, RootProb, BranchCounts and Counts against
  enumeration of node states with expm(Q_n t_n) per branch and quadrature of
  the endpoint-conditioned integral with each branch's own Q_n;
* model grids (C, R): the broadcast (1, R) form equals the explicit one, and
  column c equals that column run alone with grid[c].
"""
import itertools

import numpy as np
import pytest
from scipy.linalg import expm
from scipy.stats import gamma

from oracle import subby_oracle as O
from tests.golden_cases import load
from tests.test_oracle_invariants import _I_integral, _rand_tree, _rev_Q


def test_mixture_posterior_matches_golden():
    g = load('7node_mixture_jc4')
    tree = O.Tree(g['inputs.parentIndex'], g['inputs.distanceToParent'])
    base = O.DiagModel(g['inputs.base_model.eigenvalues'], g['inputs.base_model.eigenvectors'],
                       g['inputs.base_model.pi'])
    models = [O.scale_model(base, r) for r in g['inputs.rates']]
    aln = g['inputs.alignment']
    ll = np.stack([O.LogLike(aln, tree, m) for m in models])
    np.testing.assert_allclose(ll, g['outputs.log_likes'], rtol=1e-12, atol=1e-12)
    post = O.MixturePosterior(aln, tree, models, g['inputs.log_weights'])
    np.testing.assert_allclose(post, g['outputs.posteriors'], atol=1e-13)
    np.testing.assert_allclose(post.sum(0), 1.0, atol=1e-14)


@pytest.mark.parametrize('alpha,K', [(0.5, 4), (1.0, 4), (2.3, 6), (0.2, 8)])
def test_gamma_rate_categories_are_normalised_quantile_medians(alpha, K):
    rates, w = O.gamma_rate_categories(alpha, K)
    q = gamma.ppf((np.arange(K) + 0.5) / K, a=alpha, scale=1.0 / alpha)
    np.testing.assert_allclose(rates, q * K / q.sum(), rtol=1e-10)
    np.testing.assert_allclose(w, np.full(K, 1.0 / K))
    np.testing.assert_allclose(rates.mean(), 1.0, rtol=1e-14)


def _setup(seed, R=7, C=5, A=4):
    rng = np.random.default_rng(seed)
    tree = _rand_tree(R, rng)
    aln = rng.integers(-1, A + 2, size=(R, C))
    aln[:, 0] = rng.integers(0, A, size=R)
    return rng, tree, aln


def test_model_list_of_one_model_is_the_single_model():
    rng, tree, aln = _setup(3, R=9, C=6)
    m = O.hky85_diag(2.0, rng.dirichlet(np.ones(4) * 4))
    ms = [m] * 6
    np.testing.assert_allclose(O.LogLike_columns(aln, tree, ms), O.LogLike(aln, tree, m), rtol=1e-14)
    np.testing.assert_allclose(O.RootProb_columns(aln, tree, ms), O.RootProb(aln, tree, m), atol=1e-14)
    for mask in ('auto', None):
        np.testing.assert_allclose(O.Counts_columns(aln, tree, ms, branch_mask=mask),
                                   O.Counts(aln, tree, m, branch_mask=mask), atol=1e-13)
    np.testing.assert_allclose(O.BranchCounts_columns(aln, tree, ms), O.BranchCounts(aln, tree, m), atol=1e-13)


def _enumerate_rows(aln, tree, Qs, pi0):
    """Exact P(x), root and pair posteriors with branch n under expm(Qs[n] t_n)."""
    parent = tree.parentIndex
    R, C = aln.shape
    A = len(pi0)
    M = np.array([expm(Qs[n] * tree.distanceToParent[n]) for n in range(R)])
    L = O.token_to_likelihood(aln, A)
    P = np.zeros(C)
    root = np.zeros((A, C))
    pair = np.zeros((R, A, A, C))
    for states in itertools.product(range(A), repeat=R):
        w = pi0[states[0]] * np.ones(C)
        for n in range(R):
            w = w * L[n, :, states[n]]
        for n in range(1, R):
            w = w * M[n][states[parent[n]], states[n]]
        P += w
        root[states[0]] += w
        for n in range(1, R):
            pair[n, states[parent[n]], states[n]] += w
    return P, root / P, pair / P, M


def test_model_list_root_distribution_is_the_first_models():
    """Different pi per column: LogLike uses pi of models[0] at the root,
    RootProb = pi_c U0 / P_0(x) (the reference's stacked-model convention)."""
    rng, tree, aln = _setup(4, R=5, C=3)
    pis = [rng.dirichlet(np.ones(4) * 2) for _ in range(3)]
    ms = [O.diagonalize_rate_matrix(_rev_Q(rng.lognormal(0, 1, (4, 4)), p), p) for p in pis]
    ll = O.LogLike_columns(aln, tree, ms)
    rp = O.RootProb_columns(aln, tree, ms)
    for c in range(3):
        Q = O.reconstruct_rate_matrix(ms[c])
        P, root, _, _ = _enumerate_rows(aln[:, c:c + 1], tree, [Q] * 5, pis[0])
        np.testing.assert_allclose(ll[c], np.log(P[0]), rtol=1e-12)
        np.testing.assert_allclose(rp[:, c], root[:, 0] * pis[c] / pis[0], atol=1e-12)
        # Counts use the column's own model throughout
        np.testing.assert_allclose(O.Counts_columns(aln, tree, ms, branch_mask=None)[..., c],
                                   O.Counts(aln[:, c:c + 1], tree, ms[c], branch_mask=None)[..., 0],
                                   atol=1e-13)


def test_model_list_posteriors_by_enumeration():
    """InsideOutside over a model list: node and branch posteriors of the
    non-root nodes of column c are those of exhaustive enumeration under Q_c
    with the root distribution of models[0]; the root row is the root
    marginal of the enumeration under models[c]'s own pi."""
    rng, tree, aln = _setup(6, R=5, C=3)
    pis = [rng.dirichlet(np.ones(4) * 2) for _ in range(3)]
    ms = [O.diagonalize_rate_matrix(_rev_Q(rng.lognormal(0, 1, (4, 4)), p), p) for p in pis]
    npost = O.node_posterior_columns(aln, tree, ms)
    bpost = O.branch_posterior_columns(aln, tree, ms)
    assert np.all(bpost[0] == 0)
    for c in range(3):
        Q = O.reconstruct_rate_matrix(ms[c])
        P, root, pair, _ = _enumerate_rows(aln[:, c:c + 1], tree, [Q] * 5, pis[0])
        np.testing.assert_allclose(bpost[1:, :, :, c], pair[1:, :, :, 0], atol=1e-11)
        for n in range(1, 5):
            np.testing.assert_allclose(npost[n, :, c], pair[n, :, :, 0].sum(0), atol=1e-11)
        _, root_c, _, _ = _enumerate_rows(aln[:, c:c + 1], tree, [Q] * 5, pis[c])
        np.testing.assert_allclose(npost[0, :, c], root_c[:, 0], atol=1e-11)
    # the first column's own root distribution: the plain posteriors
    np.testing.assert_allclose(npost[..., 0], O.node_posterior(aln[:, :1], tree, ms[0])[..., 0], atol=1e-13)


@pytest.mark.parametrize('seed', [0, 1])
def test_per_branch_models_by_enumeration_and_quadrature(seed):
    rng, tree, aln = _setup(10 + seed, R=5, C=4, A=3)
    A = 3
    pis = [rng.dirichlet(np.ones(A) * 3) for _ in range(5)]
    Qs = [_rev_Q(rng.lognormal(0, 1, (A, A)), p) for p in pis]
    rows = [O.diagonalize_rate_matrix(Q, p) for Q, p in zip(Qs, pis)]
    P, root, pair, M = _enumerate_rows(aln, tree, Qs, pis[0])
    np.testing.assert_allclose(O.LogLike_rows(aln, tree, rows), np.log(P), rtol=1e-12)
    np.testing.assert_allclose(O.RootProb_rows(aln, tree, rows), root, atol=1e-12)
    bc = O.BranchCounts_rows(aln, tree, rows)
    off = ~np.eye(A, dtype=bool)
    for n in range(1, 5):
        I = _I_integral(Qs[n], tree.distanceToParent[n])
        E = np.einsum('abc,abij->ijc', pair[n] / M[n][:, :, None], I)
        want = np.where(off[..., None], Qs[n][..., None] * E, E)
        np.testing.assert_allclose(bc[n], want, atol=1e-10)
    assert np.all(bc[0] == 0)
    for mask in ('auto', None):
        np.testing.assert_allclose(O.Counts_rows(aln, tree, rows, mask),
                                   O.BranchCounts_rows(aln, tree, rows, mask).sum(0), atol=1e-14)
    # irreversible branches: the reference convention integrates with Q^T
    Qi = rng.uniform(0.05, 1.0, (A, A))
    np.fill_diagonal(Qi, 0.0)
    np.fill_diagonal(Qi, -Qi.sum(1))
    rows_i = list(rows)
    rows_i[2] = O.diagonalize_irreversible(Qi, pis[2])
    Qs_i = list(Qs)
    Qs_i[2] = Qi
    P, root, pair, M = _enumerate_rows(aln, tree, Qs_i, pis[0])
    np.testing.assert_allclose(O.LogLike_rows(aln, tree, rows_i), np.log(P), rtol=1e-12)
    bci = O.BranchCounts_rows(aln, tree, rows_i)
    E = np.einsum('abc,abij->ijc', pair[2] / M[2][:, :, None], _I_integral(Qi.T, tree.distanceToParent[2]))
    np.testing.assert_allclose(bci[2], np.where(off[..., None], Qi[..., None] * E, E), atol=1e-10)


def test_per_branch_models_all_equal_is_the_single_model():
    rng, tree, aln = _setup(20, R=9, C=7)
    m = O.hky85_diag(3.0, rng.dirichlet(np.ones(4) * 4))
    rows = [m] * 9
    np.testing.assert_allclose(O.LogLike_rows(aln, tree, rows), O.LogLike(aln, tree, m), rtol=1e-14)
    np.testing.assert_allclose(O.RootProb_rows(aln, tree, rows), O.RootProb(aln, tree, m), atol=1e-14)
    np.testing.assert_allclose(O.Counts_rows(aln, tree, rows), O.Counts(aln, tree, m), atol=1e-12)
    np.testing.assert_allclose(O.BranchCounts_rows(aln, tree, rows), O.BranchCounts(aln, tree, m), atol=1e-12)


def test_model_grid_columns_and_broadcast():
    rng, tree, aln = _setup(30, R=7, C=4)
    base = O.hky85_diag(2.0, rng.dirichlet(np.ones(4) * 4))
    row = [O.scale_model(base, r) for r in rng.uniform(0.5, 2.0, 7)]
    grid = [[O.scale_model(m, r) for m in row] for r in (0.7, 1.0, 1.3, 2.0)]
    ll = O.LogLike_grid(aln, tree, grid)
    rp = O.RootProb_grid(aln, tree, grid)
    cnt = O.Counts_grid(aln, tree, grid)
    for c in range(4):
        np.testing.assert_allclose(ll[c], O.LogLike_rows(aln[:, c:c + 1], tree, grid[c])[0], rtol=1e-14)
        np.testing.assert_allclose(rp[:, c], O.RootProb_rows(aln[:, c:c + 1], tree, grid[c])[:, 0], atol=1e-14)
        np.testing.assert_allclose(cnt[..., c], O.Counts_rows(aln[:, c:c + 1], tree, grid[c])[..., 0], atol=1e-13)
    same = [row] * 4
    np.testing.assert_allclose(O.LogLike_grid(aln, tree, same), O.LogLike_rows(aln, tree, row), rtol=1e-14)
    np.testing.assert_allclose(O.Counts_grid(aln, tree, same), O.Counts_rows(aln, tree, row), atol=1e-13)
