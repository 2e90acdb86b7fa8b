"""Pin the oracle's irreversible-model path with what the mathematics fixes.

The reference's own irreversible tests (tests/test_oracle.py:612-760) compare
its JAX backend with its oracle, which share one formula; there is no golden
fixture.  So here:

* M(t) = expm(Q t) (as tests/test_oracle.py:638-649), rows sum to one;
* complex J(t) against quadrature of int_0^t e^{mu_k s} e^{mu_l (t-s)} ds;
* LogLike, RootProb, node and branch posteriors against exhaustive
  enumeration of node states with expm(Q t);
* counts under both conventions against quadrature: the exact convention is
  sum_ab P(a,b|x)/M_ab int M_ai(s) Q_ij M_jb(t-s) ds (the expected counts
  under Q); the reference convention is the same expression with Q^T in the
  integrand (what eigensub.py:271-385 computes: projection by V^{-1} on the
  outside and V on the inside, back-transform V C V^{-1});
* the exact convention on a reversible Q equals the reversible path;
* Counts = sum of BranchCounts under every mask; counts stay non-negative on
  the reference's own random irreversible matrices (test_oracle.py:615-625,
  721-734).
"""
import itertools

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import subby_oracle as O
from tests.test_oracle_invariants import _I_integral, _rand_tree, _rev_Q


def _irrev_Q(A, rng):
    Q = rng.uniform(0.01, 1.0, (A, A))
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(1))
    return Q


def _reference_style_Q(A, seed):
    """The reference's random irreversible matrix (test_oracle.py:615-625)."""
    rs = np.random.RandomState(seed)
    Q = rs.uniform(0.01, 1.0, (A, A))
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(axis=1))
    pi = np.ones(A) / A
    Q /= -np.sum(pi * np.diag(Q))
    return Q, pi


def _enumerate(aln, tree, Q, pi):
    parent = tree.parentIndex
    R, C = aln.shape
    A = len(pi)
    M = np.array([expm(Q * t) for t in tree.distanceToParent])
    L = O.token_to_likelihood(aln, A)
    P = np.zeros(C)
    root = np.zeros((A, C))
    pair = np.zeros((R, A, A, C))
    for states in itertools.product(range(A), repeat=R):
        w = pi[states[0]] * np.ones(C)
        for n in range(R):
            w = w * L[n, :, states[n]]
        for n in range(1, R):
            w = w * M[n][states[parent[n]], states[n]]
        P += w
        root[states[0]] += w
        for n in range(1, R):
            pair[n, states[parent[n]], states[n]] += w
    return P, root / P, pair / P, M


def _expected(pair, M, Q, tree, integrand_Q):
    """Per-branch sum_ab P(a,b|x)/M_ab * I^{ab}_{ij} (x Q_ij off the diagonal),
    I built from expm(integrand_Q s)."""
    R, A = pair.shape[0], pair.shape[1]
    C = pair.shape[-1]
    out = np.zeros((R, A, A, C))
    off = ~np.eye(A, dtype=bool)
    for n in range(1, R):
        I = _I_integral(integrand_Q, tree.distanceToParent[n])
        E = np.einsum('abc,abij->ijc', pair[n] / M[n][:, :, None], I)
        out[n] = np.where(off[..., None], Q[..., None] * E, E)
    return out


def test_irrev_sub_matrices_are_expm():
    for seed in (200, 201):
        Q, pi = _reference_style_Q(4, seed)
        m = O.diagonalize_irreversible(Q, pi)
        t = np.array([0.0, 0.01, 0.1, 0.5, 1.0, 2.0])
        M = O.compute_sub_matrices(m, t)
        for r, tt in enumerate(t):
            np.testing.assert_allclose(M[r], expm(Q * tt), atol=1e-12)
            np.testing.assert_allclose(M[r].sum(1), 1.0, atol=1e-12)


def test_irrev_J_against_quadrature():
    rng = np.random.default_rng(7)
    # a matrix with a complex-conjugate pair of eigenvalues
    Q = np.array([[-1.0, 1.0, 0.0], [0.0, -1.0, 1.0], [1.0, 0.0, -1.0]]) + 0.05 * _irrev_Q(3, rng)
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(1))
    mu = np.linalg.eigvals(Q)
    assert np.max(np.abs(mu.imag)) > 0.1
    t = np.array([0.0, 0.3, 1.7])
    J = O.compute_J_complex(mu, t)
    x, w = np.polynomial.legendre.leggauss(40)
    for r, tt in enumerate(t):
        s = 0.5 * tt * (x + 1)
        ww = 0.5 * tt * w
        ref = np.einsum('q,qk,ql->kl', ww, np.exp(np.outer(s, mu)), np.exp(np.outer(tt - s, mu)))
        np.testing.assert_allclose(J[r], ref, atol=1e-13)
    # degenerate pair: t e^{mu t}
    Jd = O.compute_J_complex(np.array([-1.0 + 0.5j, -1.0 + 0.5j]), np.array([0.8]))
    np.testing.assert_allclose(Jd[0], 0.8 * np.exp((-1.0 + 0.5j) * 0.8) * np.ones((2, 2)), atol=1e-15)


@pytest.mark.parametrize('seed', [0, 1])
@pytest.mark.parametrize('A', [3, 4])
def test_irrev_loglike_root_posteriors_counts_by_enumeration(seed, A):
    rng = np.random.default_rng(500 + seed + 10 * A)
    tree = _rand_tree(5, rng)
    R, C = 5, 4
    aln = rng.integers(-1, A + 2, size=(R, C))
    aln[:, 0] = rng.integers(0, A, size=R)
    Q = _irrev_Q(A, rng)
    pi = rng.dirichlet(np.ones(A) * 3)
    P, root, pair, M = _enumerate(aln, tree, Q, pi)
    for exact in (False, True):
        m = O.diagonalize_irreversible(Q, pi, exact=exact)
        np.testing.assert_allclose(O.LogLike(aln, tree, m), np.log(P), rtol=1e-11)
        np.testing.assert_allclose(O.RootProb(aln, tree, m), root, atol=1e-11)
        np.testing.assert_allclose(O.branch_posterior(aln, tree, m)[1:], pair[1:], atol=1e-11)
        npost = O.node_posterior(aln, tree, m)
        np.testing.assert_allclose(npost[0], root, atol=1e-11)
        for n in range(1, R):
            np.testing.assert_allclose(npost[n], pair[n].sum(0), atol=1e-11)
        # the reference convention integrates with Q^T, the exact one with Q
        want = _expected(pair, M, Q, tree, Q if exact else Q.T)
        bc = O.BranchCounts(aln, tree, m)
        assert np.all(bc[0] == 0)
        np.testing.assert_allclose(bc, want, atol=1e-9, err_msg=f'exact={exact}')
        np.testing.assert_allclose(O.Counts(aln, tree, m, branch_mask=None), want.sum(0), atol=1e-9)
        for mask in (None, 'auto'):
            np.testing.assert_allclose(O.BranchCounts(aln, tree, m, branch_mask=mask).sum(0),
                                       O.Counts(aln, tree, m, branch_mask=mask), atol=1e-13)
    # the two conventions really differ on an irreversible Q
    d = (O.Counts(aln, tree, O.diagonalize_irreversible(Q, pi), branch_mask=None)
         - O.Counts(aln, tree, O.diagonalize_irreversible(Q, pi, exact=True), branch_mask=None))
    assert np.max(np.abs(d)) > 1e-4


@pytest.mark.parametrize('seed', [0, 1])
def test_irrev_exact_on_reversible_Q_equals_reversible_path(seed):
    rng = np.random.default_rng(900 + seed)
    A = 4
    tree = _rand_tree(9, rng)
    aln = rng.integers(-1, A + 2, size=(9, 30))
    pi = rng.dirichlet(np.ones(A) * 3)
    Q = _rev_Q(rng.lognormal(0, 1, (A, A)), pi)
    rev = O.diagonalize_rate_matrix(Q, pi)
    irr = O.diagonalize_irreversible(Q, pi, exact=True)
    np.testing.assert_allclose(O.LogLike(aln, tree, irr), O.LogLike(aln, tree, rev), rtol=1e-12)
    for mask in (None, 'auto'):
        np.testing.assert_allclose(O.Counts(aln, tree, irr, branch_mask=mask),
                                   O.Counts(aln, tree, rev, branch_mask=mask), atol=1e-11)
    np.testing.assert_allclose(O.BranchCounts(aln, tree, irr), O.BranchCounts(aln, tree, rev), atol=1e-11)


@pytest.mark.parametrize('seed', [205, 211, 214])
def test_irrev_counts_real_and_nonnegative(seed):
    Q, pi = _reference_style_Q(4, seed)
    rng = np.random.default_rng(seed)
    tree = _rand_tree(7, rng)
    aln = rng.integers(0, 4, size=(7, 6))
    for exact in (False, True):
        m = O.diagonalize_irreversible(Q, pi, exact=exact)
        J = O.compute_J_complex(m.eigenvalues, tree.distanceToParent)
        Mt, U, lnU, ll, D, lnD = O.inside_outside(aln, tree, m)
        Ut, Dt = O.eigenbasis_project_irrev(U, D, m)
        Ce = O.accumulate_C_complex(Dt, Ut, J, lnU, lnD, ll, tree.parentIndex)
        VCV = O._vcv_irrev(Ce, m, False)
        full = np.where(np.eye(4, dtype=bool)[..., None], VCV, O.rate_matrix_irrev(m)[..., None] * VCV)
        assert np.max(np.abs(full.imag)) < 1e-12 * max(1.0, np.max(np.abs(full.real)))
        assert O.Counts(aln, tree, m, branch_mask=None).min() >= -1e-10


@pytest.mark.parametrize('seed', [0, 1])
def test_irrev_branch_gradient_vs_finite_differences(seed):
    """d logL / d t_n of an irreversible model (vjp.py:80-131: D projected with
    V, U with V^{-1}, real part) against central differences of LogLike, whose
    M(t) = expm(Q t) is pinned above."""
    rng = np.random.default_rng(40 + seed)
    A = 4
    tree = _rand_tree(9, rng)
    aln = rng.integers(-1, A + 2, size=(9, 20))
    Q = _irrev_Q(A, rng)
    pi = rng.dirichlet(np.ones(A) * 3)
    m = O.diagonalize_irreversible(Q, pi)
    assert np.max(np.abs(m.eigenvalues.imag)) > 0 or seed == 1
    g = rng.normal(size=20)
    grad, _ = O.loglike_branch_grad(aln, tree, m, g)
    assert grad[0] == 0.0
    h = 1e-5
    for n in range(1, 9):
        tp = tree.distanceToParent.copy()
        tm = tp.copy()
        tp[n] += h
        tm[n] -= h
        fd = (O.LogLike(aln, O.Tree(tree.parentIndex, tp), m)
              - O.LogLike(aln, O.Tree(tree.parentIndex, tm), m)) @ g / (2 * h)
        np.testing.assert_allclose(grad[n], fd, rtol=1e-7, atol=1e-9)


def test_irrev_scaling_identity_links_gradient_and_exact_counts():
    """sum_n t_n dlogL/dt_n = sum_{i!=j} N_ij + sum_i T_i Q_ii with the exact
    (Q-integrand) unmasked counts; the reference convention does not satisfy it."""
    rng = np.random.default_rng(77)
    A = 4
    tree = _rand_tree(11, rng)
    aln = rng.integers(-1, A + 2, size=(11, 15))
    Q = _irrev_Q(A, rng)
    pi = rng.dirichlet(np.ones(A) * 3)
    off = ~np.eye(A, dtype=bool)
    grad, _ = O.loglike_branch_grad(aln, tree, O.diagonalize_irreversible(Q, pi), np.ones(15))
    lhs = tree.distanceToParent @ grad
    for exact, holds in ((True, True), (False, False)):
        N = O.Counts(aln, tree, O.diagonalize_irreversible(Q, pi, exact=exact), branch_mask=None)
        rhs = (np.where(off[..., None], N, 0).sum((0, 1)) + np.einsum('i,iic->c', np.diag(Q), N)).sum()
        if holds:
            np.testing.assert_allclose(lhs, rhs, rtol=1e-10)
        else:
            assert abs(lhs - rhs) > 1e-6 * abs(lhs)
