"""Pin the oracle with what the mathematics fixes, independently of its own formulas.

* exhaustive enumeration of all node states on small trees -> logLike, root
  posterior and joint branch-endpoint posteriors;
* counts from those posteriors times numerical quadrature of
  I^{ab}_{ij}(T) = int_0^T M_ai(s) M_jb(T-s) ds with scipy.linalg.expm
  (the reference's brute force, tests/test_eigensub.py:20-46, made exact);
* Steiner mask vs the graph definition (branch on a path between two
  ungapped leaves);
* closed forms: J(t), HKY85 eigensystem, M(0)=I, cherry likelihood;
* gradient vs central finite differences.
Each check is chosen so that a dropped term, wrong index or transposed
operand in the oracle fails.
"""
import itertools

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import subby_oracle as O


def _rand_tree(R, rng):
    """Random binary preorder tree with R (odd) nodes."""
    n_leaves = (R + 1) // 2
    children = {0: []}
    leaves = [0]
    nxt = 1
    while len(leaves) < n_leaves:
        k = rng.integers(len(leaves))
        v = leaves[k]
        children[v] = [nxt, nxt + 1]
        children[nxt] = []
        children[nxt + 1] = []
        leaves[k] = nxt
        leaves.append(nxt + 1)
        nxt += 2
    order, par = [], {0: -1}
    st = [0]
    while st:
        v = st.pop()
        order.append(v)
        for c in reversed(children[v]):
            par[c] = v
            st.append(c)
    new = {v: i for i, v in enumerate(order)}
    parent = np.array([-1] + [new[par[v]] for v in order[1:]], dtype=np.int32)
    t = rng.uniform(0.02, 0.6, size=R)
    t[0] = 0.0
    return O.Tree(parent, t)


def _models(A, rng):
    pi = rng.dirichlet(np.ones(A) * 3)
    s = rng.lognormal(0, 1, (A, A))
    gtr = O.diagonalize_rate_matrix(_rev_Q(s, pi), pi)
    out = [('jc', O.jukes_cantor_model(A)), ('f81', O.f81_model(pi)), ('gtr', gtr)]
    if A == 4:
        out.append(('hky', O.hky85_diag(2.5, pi)))
    return out


def _rev_Q(s, pi):
    s = 0.5 * (s + s.T)
    Q = s * pi[None, :]
    np.fill_diagonal(Q, 0)
    np.fill_diagonal(Q, -Q.sum(1))
    return Q / -(pi @ np.diag(Q))


def _enumerate(aln, tree, model):
    """Exact joint posteriors by summing over every assignment of node states."""
    parent = tree.parentIndex
    R, C = aln.shape
    A = len(model.pi)
    Q = O.reconstruct_rate_matrix(model)
    M = np.array([expm(Q * t) for t in tree.distanceToParent])
    L = O.token_to_likelihood(aln, A)            # (R, C, A)
    P = np.zeros(C)
    root = np.zeros((A, C))
    pair = np.zeros((R, A, A, C))                # P(x_p=a, x_n=b, data)
    for states in itertools.product(range(A), repeat=R):
        w = model.pi[states[0]] * np.ones(C)
        for n in range(R):
            w = w * L[n, :, states[n]]
        for n in range(1, R):
            w = w * M[n][states[parent[n]], states[n]]
        P += w
        root[states[0]] += w
        for n in range(1, R):
            pair[n, states[parent[n]], states[n]] += w
    return P, root / P, pair / P, M, Q


def _I_integral(Q, T, nodes=40):
    """I[a,b,i,j] = int_0^T M_ai(s) M_jb(T-s) ds by Gauss-Legendre quadrature."""
    x, w = np.polynomial.legendre.leggauss(nodes)
    s = 0.5 * T * (x + 1)
    w = 0.5 * T * w
    A = Q.shape[0]
    I = np.zeros((A, A, A, A))
    for sk, wk in zip(s, w):
        I += wk * np.einsum('ai,jb->abij', expm(Q * sk), expm(Q * (T - sk)))
    return I


@pytest.mark.parametrize('seed', [0, 1])
@pytest.mark.parametrize('A', [3, 4])
def test_loglike_root_counts_by_enumeration(seed, A):
    rng = np.random.default_rng(100 + seed + 10 * A)
    tree = _rand_tree(5, rng)
    R = 5
    C = 4
    aln = rng.integers(-1, A + 2, size=(R, C))
    aln[:, 0] = rng.integers(0, A, size=R)        # a fully observed column
    for name, m in _models(A, rng):
        P, root, pair, M, Q = _enumerate(aln, tree, m)
        np.testing.assert_allclose(O.LogLike(aln, tree, m), np.log(P), rtol=1e-11, err_msg=name)
        np.testing.assert_allclose(O.RootProb(aln, tree, m), root, atol=1e-11, err_msg=name)
        # counts_ij = sum_n sum_ab P(a,b|x) I^{ab}_{ij}(t_n) Q_ij / M_ab(t_n)  (i != j)
        # dwell_i  = sum_n sum_ab P(a,b|x) I^{ab}_{ii}(t_n) / M_ab(t_n)
        exp_cnt = np.zeros((A, A, C))
        for n in range(1, R):
            I = _I_integral(Q, tree.distanceToParent[n])
            W = pair[n] / M[n][:, :, None]
            E = np.einsum('abc,abij->ijc', W, I)
            off = ~np.eye(A, dtype=bool)
            exp_cnt += np.where(off[..., None], Q[..., None] * E, E)
        got = O.Counts(aln, tree, m, branch_mask=None)
        np.testing.assert_allclose(got, exp_cnt, atol=1e-9, err_msg=name)
        if name in ('jc', 'f81'):
            np.testing.assert_allclose(O.Counts(aln, tree, m, f81_fast=True, branch_mask=None),
                                       exp_cnt, atol=1e-9, err_msg=name)


def test_internal_observed_tokens_by_enumeration():
    """Observed tokens at internal nodes (present in the goldens) are honoured."""
    rng = np.random.default_rng(7)
    tree = _rand_tree(5, rng)
    aln = rng.integers(0, 4, size=(5, 3))
    m = O.hky85_diag(1.7, rng.dirichlet(np.ones(4) * 4))
    P, root, pair, M, Q = _enumerate(aln, tree, m)
    np.testing.assert_allclose(O.LogLike(aln, tree, m), np.log(P), rtol=1e-11)
    np.testing.assert_allclose(O.RootProb(aln, tree, m), root, atol=1e-11)


def _brute_mask(aln, parent, A):
    R, C = aln.shape
    cnt = np.zeros(R, int)
    for n in range(1, R):
        cnt[parent[n]] += 1
    leaf = cnt == 0
    under = [set() for _ in range(R)]
    for n in range(R - 1, -1, -1):
        if leaf[n]:
            under[n].add(n)
        if n > 0:
            under[parent[n]] |= under[n]
    mask = np.zeros((R, C), bool)
    for c in range(C):
        ung = {n for n in range(R) if leaf[n] and 0 <= aln[n, c] <= A}
        for n in range(1, R):
            k = len(under[n] & ung)
            mask[n, c] = k >= 1 and len(ung) - k >= 1
    return mask


@pytest.mark.parametrize('seed', range(4))
def test_branch_mask_is_steiner_tree(seed):
    rng = np.random.default_rng(seed)
    tree = _rand_tree(2 * int(rng.integers(2, 30)) + 1, rng)
    R = len(tree.parentIndex)
    aln = rng.integers(-1, 6, size=(R, 40))
    aln[:, rng.random(40) < 0.3] = -1
    np.testing.assert_array_equal(O.compute_branch_mask(aln, tree.parentIndex, 4),
                                  _brute_mask(aln, tree.parentIndex, 4))


def test_branch_mask_deep_caterpillar():
    """A caterpillar deeper than ceil(log2 R)+1 (the JAX bound) is exact here."""
    L = 30
    parent = [-1]
    spine = 0
    for k in range(L - 1):
        parent += [spine, spine]
        spine = len(parent) - 1
    parent = np.array(parent)
    # parent list above is not preorder-sorted when leaves are appended; relabel
    R = len(parent)
    assert np.all(parent[1:] < np.arange(1, R))
    aln = np.full((R, 2), -1)
    aln[R - 1, :] = 0           # deepest leaf
    aln[1, 0] = 2               # the leaf hanging from the root
    mask = O.compute_branch_mask(aln, parent, 4)
    np.testing.assert_array_equal(mask, _brute_mask(aln, parent, 4))
    assert mask[:, 0].sum() > 2 * (int(np.ceil(np.log2(R))) + 1)


def test_J_closed_forms():
    mu = np.array([0.0, -1.0, -1.0, -2.0])
    J = O.compute_J(mu, np.array([0.5]))[0]
    for k in range(4):
        assert np.isclose(J[k, k], 0.5 * np.exp(mu[k] * 0.5), rtol=1e-14)
    assert np.isclose(J[1, 2], 0.5 * np.exp(-0.5), rtol=1e-14)
    assert np.isclose(J[0, 1], 1 - np.exp(-0.5), rtol=1e-14)
    assert np.isclose(J[0, 3], (1 - np.exp(-1.0)) / 2.0, rtol=1e-14)
    assert np.isclose(J[3, 0], J[0, 3], rtol=1e-14)


def test_hky85_eigensystem():
    pi = np.array([0.3, 0.2, 0.25, 0.25])
    m = O.hky85_diag(2.0, pi)
    V = m.eigenvectors
    np.testing.assert_allclose(V.T @ V, np.eye(4), atol=1e-14)
    Q = O.hky85_rate_matrix(2.0, pi)
    np.testing.assert_allclose(O.reconstruct_rate_matrix(m), Q, atol=1e-14)
    assert np.isclose(-(pi @ np.diag(Q)), 1.0)
    M = O.compute_sub_matrices(m, np.array([0.0, 0.3, 50.0]))
    np.testing.assert_allclose(M[0], np.eye(4), atol=1e-14)
    np.testing.assert_allclose(M[1], expm(Q * 0.3), atol=1e-13)
    np.testing.assert_allclose(M[1].sum(1), 1.0, atol=1e-14)
    np.testing.assert_allclose(M[2], np.tile(pi, (4, 1)), atol=1e-12)
    # kappa = 1 is F81
    np.testing.assert_allclose(O.reconstruct_rate_matrix(O.hky85_diag(1.0, pi)),
                               O.reconstruct_rate_matrix(O.f81_model(pi)), atol=1e-14)


def test_cherry_loglike_pulley():
    """Two observed leaves under a gapped root: logL = log(pi_x M_xy(t1+t2))."""
    m = O.hky85_diag(3.0, np.array([0.1, 0.4, 0.3, 0.2]))
    tree = O.Tree(np.array([-1, 0, 0]), np.array([0.0, 0.13, 0.41]))
    aln = np.array([[-1] * 16, np.repeat(np.arange(4), 4), np.tile(np.arange(4), 4)])
    M = expm(O.reconstruct_rate_matrix(m) * 0.54)
    exp_ll = np.log(m.pi[aln[1]] * M[aln[1], aln[2]])
    np.testing.assert_allclose(O.LogLike(aln, tree, m), exp_ll, rtol=1e-12)


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_sum_product_identity_and_dwell(seed):
    """sum_ab D_a M_ab U_b scale = 1 per branch; dwell sums to the active branch lengths."""
    rng = np.random.default_rng(seed)
    tree = _rand_tree(21, rng)
    aln = rng.integers(-1, 6, size=(21, 30))
    m = O.hky85_diag(2.0, rng.dirichlet(np.ones(4) * 5))
    M, U, lnU, ll, D, lnD = O.inside_outside(aln, tree, m)
    sc = O.branch_scale(lnU, lnD, ll)
    ident = np.einsum('nca,nab,ncb->nc', D[1:], M[1:], U[1:]) * sc[1:]
    np.testing.assert_allclose(ident, 1.0, rtol=1e-12)
    for mask in (None, 'auto'):
        cnt = O.Counts(aln, tree, m, branch_mask=mask)
        act = np.ones((21, 30), bool) if mask is None else O.compute_branch_mask(aln, tree.parentIndex, 4)
        act[0] = False
        np.testing.assert_allclose(np.einsum('iic->c', cnt),
                                   (tree.distanceToParent[:, None] * act).sum(0), rtol=1e-11, atol=1e-14)
        assert cnt.min() > -1e-12


def test_masked_equals_unmasked_without_gaps_and_single_leaf_zero():
    rng = np.random.default_rng(3)
    tree = _rand_tree(9, rng)
    m = O.jukes_cantor_model(4)
    aln = rng.integers(0, 4, size=(9, 10))
    np.testing.assert_allclose(O.Counts(aln, tree, m, branch_mask='auto'),
                               O.Counts(aln, tree, m, branch_mask=None), atol=1e-12)
    t3 = O.Tree(np.array([-1, 0, 0]), np.array([0.0, 0.1, 0.2]))
    a3 = np.array([[0, 0, 0], [0, 1, 2], [5, 5, 5]])
    np.testing.assert_allclose(O.Counts(a3, t3, m, branch_mask='auto'), 0.0, atol=1e-15)


def test_chunked_dataflow_equals_counts():
    rng = np.random.default_rng(4)
    tree = _rand_tree(15, rng)
    aln = rng.integers(-1, 5, size=(15, 300))
    m = O.hky85_diag(2.0, rng.dirichlet(np.ones(4) * 5))
    np.testing.assert_allclose(O.counts_chunked(aln, tree, m, chunk=128), O.Counts(aln, tree, m), atol=1e-13)


def test_branch_gradient_vs_finite_differences():
    rng = np.random.default_rng(5)
    tree = _rand_tree(11, rng)
    aln = rng.integers(-1, 5, size=(11, 25))
    m = O.hky85_diag(2.0, rng.dirichlet(np.ones(4) * 5))
    g = rng.normal(size=25)
    grad, _ = O.loglike_branch_grad(aln, tree, m, g)
    h = 1e-5
    fd = np.zeros(11)
    for n in range(1, 11):
        tp = tree.distanceToParent.copy()
        tm = tp.copy()
        tp[n] += h
        tm[n] -= h
        fd[n] = (O.LogLike(aln, O.Tree(tree.parentIndex, tp), m)
                 - O.LogLike(aln, O.Tree(tree.parentIndex, tm), m)) @ g / (2 * h)
    np.testing.assert_allclose(grad, fd, rtol=1e-7, atol=1e-9)
    assert grad[0] == 0.0


def test_log_kappa_gradient_fisher_vs_finite_differences():
    rng = np.random.default_rng(6)
    tree = _rand_tree(13, rng)
    aln = rng.integers(-1, 5, size=(13, 20))
    pi = rng.dirichlet(np.ones(4) * 5)
    g = rng.normal(size=20)
    kappa = 2.3
    got = O.hky85_loglike_kappa_grad(aln, tree, kappa, pi, g)
    h = 1e-5
    f = lambda lk: O.LogLike(aln, tree, O.hky85_diag(np.exp(lk), pi)) @ g
    fd = (f(np.log(kappa) + h) - f(np.log(kappa) - h)) / (2 * h)
    np.testing.assert_allclose(got, fd, rtol=1e-7)


def test_scaling_identity_links_gradient_and_counts():
    """sum_n t_n dlogL/dt_n = sum_{i!=j} N_ij + sum_i T_i Q_ii (unmasked counts)."""
    rng = np.random.default_rng(8)
    tree = _rand_tree(15, rng)
    aln = rng.integers(-1, 5, size=(15, 12))
    m = O.hky85_diag(2.0, rng.dirichlet(np.ones(4) * 5))
    grad, _ = O.loglike_branch_grad(aln, tree, m, np.ones(12))
    N = O.Counts(aln, tree, m, branch_mask=None)
    Q = O.reconstruct_rate_matrix(m)
    off = ~np.eye(4, dtype=bool)
    rhs = (np.where(off[..., None], N, 0).sum((0, 1)) + np.einsum('i,iic->c', np.diag(Q), N)).sum()
    np.testing.assert_allclose(tree.distanceToParent @ grad, rhs, rtol=1e-10)


def test_J_expm1_form_is_the_accurate_one():
    """For close eigenvalues the reference's J = (e^{mu_k t} - e^{mu_l t})/(mu_k - mu_l)
    cancels; e^{mu_l t} expm1((mu_k - mu_l) t)/(mu_k - mu_l) (the form the kernels
    evaluate in FP64) agrees with an 80-bit extended-precision evaluation to a few ulp."""
    mu = np.array([-1.4474406912345678, -1.4474406912345678 + 3e-6, -1.4266377, -0.7])
    t = np.array([0.0, 0.05, 0.3])
    ref = O.compute_J(mu, t)
    mul = mu.astype(np.longdouble)
    tl = t.astype(np.longdouble)[:, None, None]
    ek = np.exp(mul[None, :, None] * tl)
    el = np.exp(mul[None, None, :] * tl)
    d = (mul[:, None] - mul[None, :])[None]
    deg = np.abs(d) < 1e-8
    with np.errstate(all='ignore'):
        exact = np.where(deg, tl * ek, el * np.expm1(d * tl) / np.where(deg, 1, d))
    with np.errstate(all='ignore'):
        d64 = (mu[:, None] - mu[None, :])[None]
        alt = np.where(np.abs(d64) < 1e-8, t[:, None, None] * np.exp(mu[None, :, None] * t[:, None, None]),
                       np.exp(mu[None, None, :] * t[:, None, None]) * np.expm1(d64 * t[:, None, None])
                       / np.where(np.abs(d64) < 1e-8, 1, d64))
    err_alt = float(np.max(np.abs(alt - exact) / np.maximum(np.abs(exact), 1e-300)))
    err_ref = float(np.max(np.abs(ref - exact) / np.maximum(np.abs(exact), 1e-300)))
    assert err_alt < 1e-14
    assert err_ref > 10 * err_alt


@pytest.mark.parametrize('seed', [0, 1])
@pytest.mark.parametrize('A', [3, 4])
def test_branch_counts_and_posteriors_by_enumeration(seed, A):
    """Per-branch counts, node posteriors and branch joint posteriors against
    exhaustive enumeration (+ quadrature of I^{ab}_{ij}); Counts = sum of
    BranchCounts under every mask."""
    rng = np.random.default_rng(300 + seed + 10 * A)
    tree = _rand_tree(5, rng)
    R, C = 5, 4
    aln = rng.integers(-1, A + 2, size=(R, C))
    aln[:, 0] = rng.integers(0, A, size=R)
    for name, m in _models(A, rng):
        P, root, pair, M, Q = _enumerate(aln, tree, m)
        bp = O.branch_posterior(aln, tree, m)
        np.testing.assert_allclose(bp[1:], pair[1:], atol=1e-11, err_msg=name)
        assert np.all(bp[0] == 0)
        npost = O.node_posterior(aln, tree, m)
        np.testing.assert_allclose(npost[0], root, atol=1e-11, err_msg=name)
        for n in range(1, R):
            np.testing.assert_allclose(npost[n], pair[n].sum(0), atol=1e-11, err_msg=name)
        bc = O.BranchCounts(aln, tree, m)
        assert np.all(bc[0] == 0)
        off = ~np.eye(A, dtype=bool)
        for n in range(1, R):
            I = _I_integral(Q, tree.distanceToParent[n])
            E = np.einsum('abc,abij->ijc', pair[n] / M[n][:, :, None], I)
            np.testing.assert_allclose(bc[n], np.where(off[..., None], Q[..., None] * E, E),
                                       atol=1e-9, err_msg=name)
        for mask in (None, 'auto'):
            np.testing.assert_allclose(O.BranchCounts(aln, tree, m, branch_mask=mask).sum(0),
                                       O.Counts(aln, tree, m, branch_mask=mask), atol=1e-13)
            if name in ('jc', 'f81'):
                np.testing.assert_allclose(
                    O.BranchCounts(aln, tree, m, f81_fast=True, branch_mask=mask),
                    O.BranchCounts(aln, tree, m, branch_mask=mask), atol=1e-12, err_msg=name)
