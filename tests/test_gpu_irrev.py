"""GPU parity for irreversible (complex-eigensystem) models against the FP64 oracle.

The oracle's irreversible path is pinned in tests/test_oracle_irrev.py (expm,
quadrature of J, enumeration, quadrature of both counts conventions, finite
differences of the gradient).  Tolerances, per output:
  FP32: logLike max_c |d|/max(|ref|,1) <= 1e-5; counts normwise <= 1e-4;
        root max |d| <= 1e-5; gradient max|d|/max|ref| <= 1e-4.
  FP64: logLike, root <= 1e-12; counts normwise <= 1e-11 (the kernels form
        J with a complex expm1, the oracle with the reference's difference
        quotient, which differ by rounding of order 1e-16 / |mu_k - mu_l|);
        gradient <= 1e-10.
"""
import numpy as np
import pytest
import torch

from oracle import subby_oracle as O
from tests.test_gpu_parity import err_counts, err_ll, err_root, otree, say

pytestmark = pytest.mark.gpu

TOL = {'fp32': dict(ll=1e-5, counts=1e-4, root=1e-5, grad=1e-4),
       'fp64': dict(ll=1e-12, counts=1e-11, root=1e-12, grad=1e-10)}


@pytest.fixture(scope='module')
def sb():
    import subby_amd
    assert torch.cuda.is_available(), 'GPU tests need a ROCm GPU'
    return subby_amd


@pytest.fixture(scope='module')
def eng(sb):
    return sb.Engine(0)


def _irrev_Q(A, rng):
    """Random irreversible Q with a rotational part (complex eigenvalues)."""
    Q = rng.uniform(0.05, 1.0, (A, A))
    for a in range(A):
        Q[a, (a + 1) % A] += 1.5
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(1))
    return Q


def _case(sb, A, seed, leaves=20, C=700, reference_style=False):
    rng = np.random.default_rng(seed)
    if reference_style:   # the reference's random irreversible matrix (test_oracle.py:615-625)
        rs = np.random.RandomState(seed)
        Q = rs.uniform(0.01, 1.0, (A, A))
        np.fill_diagonal(Q, 0.0)
        np.fill_diagonal(Q, -Q.sum(axis=1))
        pi = np.ones(A) / A
        Q /= -np.sum(pi * np.diag(Q))
    else:
        Q = _irrev_Q(A, rng)
        pi = rng.dirichlet(np.full(A, 3.0))
    m = sb.diagonalize_irreversible(Q, pi)
    t = sb.synthetic.random_tree(leaves, rng, 0.15)
    aln = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.05, ambig_frac=0.02)
    return Q, pi, m, t, aln


def oirr(m, exact):
    return O.IrrevModel(np.asarray(m.eigenvalues), np.asarray(m.eigenvectors),
                        np.asarray(m.eigenvectors_inv), np.asarray(m.pi), exact)


def check_irrev(sb, eng, aln, t, m, precision, mask, exact):
    tol = TOL[precision]
    conv = 'exact' if exact else 'reference'
    ref_cnt, ref_ll, ref_root = O.counts_loglike_root(np.asarray(aln), otree(t), oirr(m, exact),
                                                      branch_mask=mask)
    cnt, ll, root = sb.Counts(aln, t, m, branch_mask=mask, precision=precision, engine=eng,
                              return_loglike=True, return_root=True, irrev_convention=conv)
    ll2 = sb.LogLike(aln, t, m, precision=precision, engine=eng)
    root2 = sb.RootProb(aln, t, m, precision=precision, engine=eng)
    e = dict(ll=err_ll(ll, ref_ll), counts=err_counts(cnt, ref_cnt), root=err_root(root, ref_root),
             ll2=err_ll(ll2, ref_ll), root2=err_root(root2, ref_root))
    say('   irrev', precision, conv, 'mask', 'explicit' if not isinstance(mask, (str, type(None))) else mask,
        {k: f'{v:.2e}' for k, v in e.items()})
    assert max(e['ll'], e['ll2']) <= tol['ll']
    assert e['counts'] <= tol['counts']
    assert max(e['root'], e['root2']) <= tol['root']


@pytest.mark.parametrize('A', [3, 4, 6])
def test_irrev_counts_loglike_root(sb, eng, A):
    say('irreversible A =', A)
    Q, pi, m, t, aln = _case(sb, A, 300 + A)
    assert np.max(np.abs(np.asarray(m.eigenvalues).imag)) > 0.1
    mk = np.random.default_rng(A).random(aln.shape) < 0.7
    for precision in ('fp64', 'fp32'):
        for exact in (False, True):
            for mask in ('auto', None, mk):
                check_irrev(sb, eng, aln, t, m, precision, mask, exact)


@pytest.mark.parametrize('seed', [205, 211])
def test_irrev_reference_style_matrices(sb, eng, seed):
    say('reference-style irreversible Q, seed', seed)
    Q, pi, m, t, aln = _case(sb, 4, seed, leaves=40, C=1500, reference_style=True)
    for precision in ('fp64', 'fp32'):
        for exact in (False, True):
            check_irrev(sb, eng, aln, t, m, precision, 'auto', exact)


def test_irrev_edge_shapes(sb, eng):
    say('irreversible edge shapes')
    rng = np.random.default_rng(31)
    Q = _irrev_Q(4, rng)
    m = sb.diagonalize_irreversible(Q, rng.dirichlet(np.full(4, 3.0)))
    for R, C in [(1, 5), (3, 1), (3, 65), (31, 129)]:
        t = sb.Tree(np.array([-1]), np.array([0.0])) if R == 1 else \
            sb.synthetic.random_tree((R + 1) // 2, rng, 0.2)
        aln = rng.integers(-1, 6, size=(R, C)).astype(np.int8)
        aln[:, 0] = -1
        for precision in ('fp32', 'fp64'):
            check_irrev(sb, eng, aln, t, m, precision, 'auto', False)


def test_irrev_gradient(sb, eng):
    """Branch-length gradient (vjp.py:80-131) vs the oracle, and a rate parameter
    through the Fisher identity: Q(theta) = e^theta Q has W_ij = 1 (i != j),
    W_ii = Q_ii, so dlogL/dtheta = sum_n t_n dlogL/dt_n (exact-convention counts)."""
    say('irreversible gradient')
    Q, pi, m, t, aln = _case(sb, 4, 41, leaves=15, C=3000)
    g = np.random.default_rng(42).normal(size=3000)
    ref, _ = O.loglike_branch_grad(aln, otree(t), oirr(m, False), g)
    W = np.where(np.eye(4, dtype=bool), np.diag(Q)[:, None] * np.eye(4), 1.0)
    for prec in ('fp64', 'fp32'):
        d, r, _ = sb.loglike_grad(aln, t, m, g=g, rate_weights=W, precision=prec, engine=eng)
        d = d.cpu().numpy()
        err = np.abs(d - ref).max() / np.abs(ref).max()
        want = float(np.asarray(t.distanceToParent) @ ref)
        errk = abs(float(r[0]) - want) / max(abs(want), 1.0)
        say('   grad err', prec, f'{err:.2e}', 'rate', f'{errk:.2e}')
        assert err <= TOL[prec]['grad'] and errk <= TOL[prec]['grad']
    # the autograd wrapper accepts an IrrevDiagModel
    dist = torch.tensor(np.asarray(t.distanceToParent), dtype=torch.float64, requires_grad=True)
    ll = sb.LogLikeCustomGrad(aln, sb.Tree(t.parentIndex, dist), m, precision='fp64', engine=eng)
    (ll.cpu() * torch.as_tensor(g)).sum().backward()
    assert np.abs(dist.grad.numpy() - ref).max() <= 1e-10 * np.abs(ref).max()
    # method='pade' gives the same values (same adjoint)
    dist2 = dist.detach().clone().requires_grad_(True)
    ll2 = sb.LogLikeCustomGrad(aln, sb.Tree(t.parentIndex, dist2), m, precision='fp64', engine=eng, method='pade')
    (ll2.cpu() * torch.as_tensor(g)).sum().backward()
    assert torch.equal(ll2, ll) and torch.equal(dist2.grad, dist.grad)


@pytest.mark.parametrize('A', [12, 40])
def test_irrev_gradient_large_alphabets(sb, eng, A):
    """The same gradient checks above A = 8 (the complex counts variant with
    the exact projections): branch lengths vs the oracle, the rate-scaling
    parameter through the identity dlogL/dtheta = sum_n t_n dlogL/dt_n."""
    say('irreversible gradient, A =', A)
    Q, pi, m, t, aln = _case(sb, A, 43 + A, leaves=9, C=200)
    g = np.random.default_rng(44).normal(size=200)
    ref, _ = O.loglike_branch_grad(aln, otree(t), oirr(m, False), g)
    W = np.where(np.eye(A, dtype=bool), np.diag(Q)[:, None] * np.eye(A), 1.0)
    for prec in ('fp64', 'fp32'):
        d, r, _ = sb.loglike_grad(aln, t, m, g=g, rate_weights=W, precision=prec, engine=eng)
        d = d.cpu().numpy()
        err = np.abs(d - ref).max() / np.abs(ref).max()
        want = float(np.asarray(t.distanceToParent) @ ref)
        errk = abs(float(r[0]) - want) / max(abs(want), 1.0)
        say('   grad err', prec, f'{err:.2e}', 'rate', f'{errk:.2e}')
        assert err <= TOL[prec]['grad'] and errk <= TOL[prec]['grad']


def test_irrev_batch_and_graph_equal_single(sb, eng):
    """Several irreversible models and trees in one launch, eager and as a
    replayed hipGraph (with eager work between replays), equal single calls bitwise."""
    say('irreversible batch / graph')
    from subby_amd.batch import Launcher, Block
    rng = np.random.default_rng(51)
    L = Launcher(eng)
    for rep in range(3):
        blocks = []
        for i in range(int(rng.integers(2, 7))):
            m = sb.diagonalize_irreversible(_irrev_Q(4, rng), rng.dirichlet(np.full(4, 3.0)))
            t = sb.synthetic.random_tree(int(rng.integers(2, 60)), rng, 0.1)
            blocks.append(Block(sb.synthetic.simulate(t, m, int(rng.integers(1, 700)), rng, 0.05), t, m))
        L.submit(blocks, use_graph=rep > 0)
        L.sync()
        for b in blocks:
            cnt, ll, root = sb.Counts(b.alignment, b.tree, b.model, engine=eng, return_loglike=True,
                                      return_root=True)
            torch.cuda.synchronize()
            assert torch.equal(cnt, b.counts) and torch.equal(ll, b.loglike) and torch.equal(root, b.root)
        b = blocks[0]
        ref_cnt, ref_ll, _ = O.counts_loglike_root(b.alignment, otree(b.tree), oirr(b.model, False))
        assert err_counts(b.counts, ref_cnt) <= 1e-4 and err_ll(b.loglike, ref_ll) <= 1e-5


@pytest.mark.parametrize('A', [4, 6])
def test_irrev_per_branch_outputs(sb, eng, A):
    """BranchCounts (both conventions, all masks) and node / branch posteriors of
    an irreversible model against the oracle (pinned by enumeration and quadrature)."""
    say('irreversible per-branch outputs A =', A)
    Q, pi, m, t, aln = _case(sb, A, 700 + A, leaves=16, C=300)
    mk = np.random.default_rng(A).random(aln.shape) < 0.7
    for precision in ('fp64', 'fp32'):
        tb = 1e-11 if precision == 'fp64' else 1e-4
        tp = 1e-12 if precision == 'fp64' else 1e-5
        for exact in (False, True):
            conv = 'exact' if exact else 'reference'
            o = oirr(m, exact)
            for mask in (None, 'auto', mk):
                ref = O.BranchCounts(aln, otree(t), o, branch_mask=mask)
                got = sb.BranchCounts(aln, t, m, branch_mask=mask, precision=precision, engine=eng,
                                      irrev_convention=conv).detach().cpu().double().numpy()
                num = np.abs(got - ref).max(axis=(1, 2))
                den = np.maximum(np.abs(ref).max(axis=(1, 2)), 1e-300)
                e = float(np.max(num[1:] / den[1:]))
                assert np.all(got[0] == 0) and e <= tb, (precision, conv, e)
        io = sb.InsideOutside(aln, t, m, precision=precision, engine=eng)
        e_np = float(np.abs(io.node_posterior().cpu().double().numpy() - O.node_posterior(aln, otree(t), oirr(m, False))).max())
        e_bp = float(np.abs(io.branch_posterior().cpu().double().numpy() - O.branch_posterior(aln, otree(t), oirr(m, False))).max())
        say('   ', precision, f'node {e_np:.2e} branch {e_bp:.2e}')
        assert e_np <= tp and e_bp <= tp


def test_irrev_per_branch_models_a6(sb, eng):
    """Per-branch irreversible models at A = 6 (the AP = 8 kernels): Counts and
    BranchCounts with each branch in its own complex eigenbasis vs the oracle."""
    say('irreversible per-branch models, A = 6')
    Q, pi, m, t, aln = _case(sb, 6, 71, leaves=10, C=300)
    R = len(t.parentIndex)
    rng = np.random.default_rng(72)
    rows = [sb.diagonalize_irreversible(_irrev_Q(6, rng), pi) for _ in range(R)]
    orows = [oirr(r, False) for r in rows]
    for precision in ('fp64', 'fp32'):
        tol = TOL[precision]
        e = dict(ll=err_ll(sb.LogLike(aln, t, [rows], precision=precision, engine=eng),
                           O.LogLike_rows(aln, otree(t), orows)),
                 counts=err_counts(sb.Counts(aln, t, [rows], precision=precision, engine=eng),
                                   O.Counts_rows(aln, otree(t), orows, 'auto')))
        bc = sb.BranchCounts(aln, t, [rows], precision=precision, engine=eng).cpu().double().numpy()
        ref = O.BranchCounts_rows(aln, otree(t), orows)
        e['bcounts'] = float(np.max(np.abs(bc - ref).max(axis=(1, 2))[1:] / np.abs(ref).max(axis=(1, 2))[1:]))
        say('   ', precision, {k: f'{v:.2e}' for k, v in e.items()})
        assert e['ll'] <= tol['ll'] and max(e['counts'], e['bcounts']) <= tol['counts']


@pytest.mark.parametrize('A,C', [(9, 60), (20, 150), (64, 40)])
def test_irrev_large_alphabets(sb, eng, A, C):
    """Irreversible models above A = 8 on the large-alphabet kernel (real M
    from the complex eigensystem, the irreversible record layouts): LogLike,
    RootProb and the InsideOutside posteriors vs the oracle, FP64 and FP32;
    Counts (both conventions, masks auto/none) and BranchCounts through the
    complex counts variant; gradients stay at A <= 8."""
    say('irreversible, A =', A, 'C =', C)
    Q, pi, m, t, aln = _case(sb, A, 300 + A, leaves=9, C=C)
    ot, om = otree(t), oirr(m, False)
    ref_ll, ref_root = O.LogLike(aln, ot, om), O.RootProb(aln, ot, om)
    rn, rb = O.node_posterior(aln, ot, om), O.branch_posterior(aln, ot, om)
    for precision in ('fp64', 'fp32'):
        tol = TOL[precision]
        tp = 1e-11 if precision == 'fp64' else 1e-5
        ll = sb.LogLike(aln, t, m, precision=precision, engine=eng)
        root = sb.RootProb(aln, t, m, precision=precision, engine=eng)
        io = sb.InsideOutside(aln, t, m, precision=precision, engine=eng)
        e = dict(ll=err_ll(ll, ref_ll), root=err_root(root, ref_root),
                 node=float(np.abs(io.node_posterior().cpu().double().numpy() - rn).max()),
                 branch=float(np.abs(io.branch_posterior().cpu().double().numpy() - rb).max()))
        say('   ', precision, {k: f'{v:.2e}' for k, v in e.items()})
        assert e['ll'] <= tol['ll'] and e['root'] <= tol['root']
        assert e['node'] <= tp and e['branch'] <= tp
        # counts (both conventions) and branch counts: the complex variant
        tc = 1e-10 if precision == 'fp64' else tol['counts']
        for exact in (False, True):
            conv = 'exact' if exact else 'reference'
            for mask in ('auto', None):
                cnt = sb.Counts(aln, t, m, branch_mask=mask, precision=precision, engine=eng, irrev_convention=conv)
                rc, _, _ = O.counts_loglike_root(np.asarray(aln), ot, oirr(m, exact), branch_mask=mask)
                ec = err_counts(cnt, rc)
                assert ec <= tc, (precision, conv, mask, ec)
        bc = sb.BranchCounts(aln, t, m, precision=precision, engine=eng).cpu().double().numpy()
        ref = O.BranchCounts(np.asarray(aln), ot, om)
        eb = float(np.max(np.abs(bc - ref).max(axis=(1, 2))[1:] / np.abs(ref).max(axis=(1, 2))[1:]))
        say('   ', precision, f'counts {ec:.2e} branch counts {eb:.2e}')
        assert eb <= tc and float(np.abs(bc[0]).max()) == 0.0


def test_irrev_errors(sb, eng):
    say('irreversible errors')
    rng = np.random.default_rng(61)
    Q = _irrev_Q(4, rng)
    m = sb.diagonalize_irreversible(Q, np.full(4, 0.25))
    t = sb.synthetic.random_tree(5, rng, 0.1)
    aln = sb.synthetic.simulate(t, m, 50, rng)
    with pytest.raises(sb.SubbyError, match='reversible-only'):
        sb.Counts(aln, t, m, f81_fast_flag=True, engine=eng)
    bad = sb.IrrevDiagModel(m.eigenvalues, m.eigenvectors, 2.0 * np.asarray(m.eigenvectors_inv), m.pi)
    with pytest.raises(sb.SubbyError, match='inverse'):
        eng.model_id(bad)
    from subby_amd.batch import Launcher, Block, counts_plain
    with pytest.raises(sb.SubbyError, match='reversible'):
        counts_plain(aln, t, m, engine=eng)
    rev = sb.hky85_diag(2.0, np.full(4, 0.25))
    with pytest.raises(sb.SubbyError, match='share a batch'):
        Launcher(eng).submit([Block(aln, t, m), Block(aln, t, rev)])
    # the engine still works after errors
    ref_ll = O.LogLike(aln, otree(t), oirr(m, False))
    assert err_ll(sb.LogLike(aln, t, m, engine=eng), ref_ll) <= 1e-5
