"""GPU parity on trees whose node order is breadth-first, not depth-first.

The reference accepts any parents-before-children numbering
(docs/architecture.md:39-48), and its own balanced trees are breadth-first:
parentIndex[i] = (i - 1) // 2 (tests/test_oracle.py:59,
benchmarks/run_benchmarks.py:119-128).  Every kernel family runs on such trees
(15 to 199 nodes) against the FP64 oracle, whose results on these trees are
pinned by enumeration and by relabelling invariance
(tests/test_oracle_tree_order.py).  Tolerances as in test_gpu_parity.
"""
import numpy as np
import pytest
import torch

from oracle import subby_oracle as O
from tests.test_gpu_parity import check_all, err_counts, err_ll, err_root, omodel, otree, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def sb():
    import subby_amd
    from subby_amd import _lib
    assert torch.cuda.is_available(), 'GPU tests need a ROCm GPU'
    _lib.load()
    return subby_amd


@pytest.fixture(scope='module')
def eng(sb):
    return sb.Engine(0)


def heap_tree(sb, R, rng, lo=0.01, hi=0.5):
    parent = np.array([-1] + [(i - 1) // 2 for i in range(1, R)], dtype=np.int32)
    t = rng.uniform(lo, hi, size=R)
    t[0] = 0.0
    return sb.Tree(parent, t)


def per_branch_err(got, ref):
    got = got.detach().cpu().double().numpy()
    num = np.abs(got - ref).max(axis=(1, 2))
    den = np.abs(ref).max(axis=(1, 2))
    ok = den > 0
    assert np.all(num[~ok] == 0)
    return float(np.max(num[ok] / den[ok])) if ok.any() else 0.0


@pytest.mark.parametrize('R', [15, 63, 127])
def test_heap_trees_four_states(sb, eng, R):
    """HKY85 / F81 / general A = 6 on breadth-first balanced trees: fused kernels, every
    mask mode, FP32 and FP64, the plain path and the gradient."""
    say('breadth-first tree, four states, R =', R)
    rng = np.random.default_rng(500 + R)
    t = heap_tree(sb, R, rng)
    pi = rng.dirichlet(np.full(4, 5.0))
    m = sb.hky85_diag(2.0, pi)
    aln = sb.synthetic.simulate(t, m, 700, rng, gap_frac=0.1, ambig_frac=0.02)
    for precision in ('fp32', 'fp64'):
        for mask in ('auto', None):
            check_all(sb, eng, aln, t, m, precision, mask=mask)
    mk = rng.random(aln.shape) < 0.6
    ref = O.Counts(aln, otree(t), omodel(m), branch_mask=mk)
    assert err_counts(sb.Counts(aln, t, m, branch_mask=mk, engine=eng), ref) <= 1e-4
    check_all(sb, eng, aln, t, sb.f81_model(pi), 'fp32', f81=True)
    from subby_amd.batch import counts_plain
    ref_cnt, ref_ll, ref_root = O.counts_loglike_root(aln, otree(t), omodel(m))
    cnt, ll, root = counts_plain(aln, t, m, engine=eng)
    assert err_ll(ll, ref_ll) <= 1e-5 and err_counts(cnt, ref_cnt) <= 1e-4 and err_root(root, ref_root) <= 1e-5
    d, _, _ = sb.loglike_grad(aln, t, m, precision='fp64', engine=eng)
    ref_g, _ = O.loglike_branch_grad(aln, otree(t), omodel(m))
    assert np.abs(d.cpu().numpy() - ref_g).max() <= 1e-9 * np.abs(ref_g).max()
    A6 = 6
    p6 = rng.dirichlet(np.full(A6, 3.0))
    m6 = sb.diagonalize_rate_matrix(sb.reversible_rate_matrix(rng.lognormal(0, 1, (A6, A6)), p6), p6)
    aln6 = sb.synthetic.simulate(t, m6, 300, rng, gap_frac=0.05, ambig_frac=0.02)
    check_all(sb, eng, aln6, t, m6, 'fp32')


def test_heap_tree_posteriors_and_irreversible(sb, eng):
    """BranchCounts, node and branch posteriors (A = 4) and an irreversible model on a
    breadth-first 63-node tree."""
    say('breadth-first tree: posteriors, irreversible')
    rng = np.random.default_rng(77)
    t = heap_tree(sb, 63, rng)
    m = sb.hky85_diag(3.0, rng.dirichlet(np.full(4, 5.0)))
    aln = sb.synthetic.simulate(t, m, 400, rng, gap_frac=0.08, ambig_frac=0.02)
    ot, om = otree(t), omodel(m)
    for mask in (None, 'auto'):
        assert per_branch_err(sb.BranchCounts(aln, t, m, branch_mask=mask, engine=eng),
                              O.BranchCounts(aln, ot, om, branch_mask=mask)) <= 1e-4
    io = sb.InsideOutside(aln, t, m, engine=eng)
    assert np.abs(io.node_posterior().cpu().double().numpy() - O.node_posterior(aln, ot, om)).max() <= 1e-5
    assert np.abs(io.branch_posterior().cpu().double().numpy() - O.branch_posterior(aln, ot, om)).max() <= 1e-5
    Q = rng.uniform(0.05, 1.0, (4, 4))
    for a in range(4):
        Q[a, (a + 1) % 4] += 1.5
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(1))
    pi = rng.dirichlet(np.full(4, 3.0))
    mi = sb.diagonalize_irreversible(Q, pi)
    alni = sb.synthetic.simulate(t, mi, 300, rng, gap_frac=0.05)
    oi = O.IrrevModel(np.asarray(mi.eigenvalues), np.asarray(mi.eigenvectors), np.asarray(mi.eigenvectors_inv),
                      np.asarray(mi.pi), False)
    ref_cnt, ref_ll, ref_root = O.counts_loglike_root(alni, ot, oi)
    cnt, ll, root = sb.Counts(alni, t, mi, engine=eng, return_loglike=True, return_root=True)
    assert err_ll(ll, ref_ll) <= 1e-5 and err_counts(cnt, ref_cnt) <= 1e-4 and err_root(root, ref_root) <= 1e-5


@pytest.mark.parametrize('A,C', [(20, 150), (64, 96)])
def test_heap_trees_large_alphabets(sb, eng, A, C):
    """Large alphabets on a breadth-first 127-node tree: the two-pass matrix-core path
    (FP32 general), the VALU kernel (FP64), the F81 matrix-core kernel and BranchCounts."""
    say('breadth-first tree, A =', A)
    rng = np.random.default_rng(900 + A)
    t = heap_tree(sb, 127, rng, 0.01, 0.3)
    pi = rng.dirichlet(np.full(A, 2.0))
    m = sb.diagonalize_rate_matrix(sb.reversible_rate_matrix(rng.lognormal(0, 1, (A, A)), pi), pi)
    aln = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.05, ambig_frac=0.02)
    check_all(sb, eng, aln, t, m, 'fp32')
    check_all(sb, eng, aln, t, m, 'fp64', tol=dict(ll=1e-12, counts=1e-10, root=1e-12))
    f = sb.f81_model(pi)
    check_all(sb, eng, aln, t, f, 'fp32', f81=True)
    assert per_branch_err(sb.BranchCounts(aln, t, m, branch_mask='auto', engine=eng),
                          O.BranchCounts(aln, otree(t), omodel(m), branch_mask='auto')) <= 1e-4


def test_heap_and_depth_first_trees_in_one_launch(sb, eng):
    """A batch mixing breadth-first and depth-first trees equals one call per block, bitwise."""
    say('breadth-first and depth-first trees in one launch')
    from subby_amd.batch import Block, Launcher
    rng = np.random.default_rng(31)
    blocks = []
    for i in range(8):
        m = sb.hky85_diag(float(rng.uniform(1, 5)), rng.dirichlet(np.full(4, 5.0)))
        t = heap_tree(sb, int(rng.choice([15, 31, 63, 127])), rng) if i % 2 == 0 else \
            sb.synthetic.random_tree(int(rng.integers(3, 60)), rng, 0.1)
        blocks.append(Block(sb.synthetic.simulate(t, m, int(rng.integers(10, 600)), rng, gap_frac=0.05), t, m))
    L = Launcher(eng)
    L.submit(blocks)
    L.sync()
    for b in blocks:
        cnt, ll, root = sb.Counts(b.alignment, b.tree, b.model, engine=eng, return_loglike=True, return_root=True)
        torch.cuda.synchronize()
        assert torch.equal(cnt, b.counts) and torch.equal(ll, b.loglike) and torch.equal(root, b.root)
        ref_cnt, ref_ll, _ = O.counts_loglike_root(b.alignment, otree(b.tree), omodel(b.model))
        assert err_counts(b.counts, ref_cnt) <= 1e-4 and err_ll(b.loglike, ref_ll) <= 1e-5
