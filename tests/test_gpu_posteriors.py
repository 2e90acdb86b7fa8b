"""BranchCounts and InsideOutside posteriors on the GPU vs the FP64 oracle.

The oracle functions are pinned by exhaustive enumeration
(tests/test_oracle_invariants.py); here the GPU path is compared with them,
and the per-branch counts are tied to the reference's own golden counts
(Counts = sum over branches of BranchCounts, unmasked)."""
import numpy as np
import pytest
import torch

from oracle import subby_oracle as O
from tests.golden_cases import CASES, load

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope='module')
def sb():
    import subby_amd
    return subby_amd


@pytest.fixture(scope='module')
def eng(sb):
    return sb.Engine(0)


def otree(t):
    return O.Tree(np.asarray(t.parentIndex), np.asarray(t.distanceToParent, dtype=np.float64))


def omodel(m):
    return O.DiagModel(*(np.asarray(x, dtype=np.float64) for x in m))


def per_branch_err(got, ref):
    """max over branches and columns of max_ij |d| / max_ij |ref| (branches with ref == 0 skipped)."""
    got = got.detach().cpu().double().numpy()
    num = np.abs(got - ref).max(axis=(1, 2))
    den = np.abs(ref).max(axis=(1, 2))
    ok = den > 0
    assert np.all(num[~ok] == 0)
    return float(np.max(num[ok] / den[ok])) if ok.any() else 0.0


def abs_err(got, ref):
    return float(np.max(np.abs(got.detach().cpu().double().numpy() - ref)))


def _block(sb, A, C, seed, leaves=12):
    rng = np.random.default_rng(seed)
    pi = rng.dirichlet(np.full(A, 3.0))
    if A == 4:
        m = sb.hky85_diag(2.3, pi)
    else:
        Q = sb.reversible_rate_matrix(rng.lognormal(0, 1, (A, A)), pi)
        m = sb.diagonalize_rate_matrix(Q, pi)
    t = sb.synthetic.random_tree(leaves, rng, 0.15)
    aln = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.08, ambig_frac=0.03)
    return t, m, aln, pi


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
@pytest.mark.parametrize('A', [3, 4, 6])
def test_branch_counts_and_posteriors(sb, eng, A, precision):
    say('posteriors', A, precision)
    t, m, aln, pi = _block(sb, A, 333, seed=60 + A)
    ot, om = otree(t), omodel(m)
    tol_c = 1e-11 if precision == 'fp64' else 1e-4
    tol_p = 1e-12 if precision == 'fp64' else 1e-5
    for mask in (None, 'auto'):
        ref = O.BranchCounts(aln, ot, om, branch_mask=mask)
        got = sb.BranchCounts(aln, t, m, branch_mask=mask, precision=precision, engine=eng)
        e = per_branch_err(got, ref)
        say('   branch counts', mask, f'{e:.2e}')
        assert e <= tol_c
    mk = np.random.default_rng(5).random((len(t.parentIndex), aln.shape[1])) < 0.5
    ref = O.BranchCounts(aln, ot, om, branch_mask=mk)
    got = sb.BranchCounts(aln, t, m, branch_mask=mk, precision=precision, engine=eng)
    assert per_branch_err(got, ref) <= tol_c
    io = sb.InsideOutside(aln, t, m, precision=precision, engine=eng)
    en = abs_err(io.node_posterior(), O.node_posterior(aln, ot, om))
    eb = abs_err(io.branch_posterior(), O.branch_posterior(aln, ot, om))
    say('   node / branch posterior', f'{en:.2e} {eb:.2e}')
    assert en <= tol_p and eb <= tol_p
    assert abs_err(io.node_posterior(0), O.RootProb(aln, ot, om)) <= tol_p
    ll = io.log_likelihood.cpu().numpy()
    assert np.max(np.abs(ll - O.LogLike(aln, ot, om)) / np.maximum(np.abs(ll), 1)) <= (1e-12 if precision == 'fp64' else 1e-5)
    # F81 closed form per branch (F81 model)
    f = sb.f81_model(pi)
    ref = O.BranchCounts(aln, ot, omodel(f), f81_fast=True)
    got = sb.BranchCounts(aln, t, f, f81_fast_flag=True, precision=precision, engine=eng)
    assert per_branch_err(got, ref) <= tol_c


@pytest.mark.parametrize('case', [c for c in CASES if c != '7node_jc64'])
def test_branch_counts_sum_to_golden_counts(sb, eng, case):
    """sum_n BranchCounts = the reference's golden (unmasked) counts."""
    g = load(case)
    m = sb.DiagModel(g['inputs.model.eigenvalues'], g['inputs.model.eigenvectors'], g['inputs.model.pi'])
    t = sb.Tree(g['inputs.parentIndex'], g['inputs.distanceToParent'])
    aln = g['inputs.alignment']
    bc = sb.BranchCounts(aln, t, m, precision='fp64', engine=eng).sum(0).cpu().numpy()
    ref = g['outputs.counts']
    num = np.abs(bc - ref).max(axis=(0, 1))
    den = np.maximum(np.abs(ref).max(axis=(0, 1)), 1e-300)
    assert np.max(num / den) <= 1e-12


@pytest.mark.parametrize('A,C', [(12, 70), (20, 5), (20, 131), (64, 40)])
def test_branch_counts_large_alphabets(sb, eng, A, C):
    """Per-branch counts above A = 8 (FP32 general models: the two-pass matrix-core
    path; FP64 and the F81 closed form: the large-alphabet kernel's per-branch
    counts variant): all mask modes, and sum over branches = Counts (the
    reference's test_large_alphabet, tests/test_api.py:1406-1419, at A = 20)."""
    say('branch counts, A =', A, 'C =', C)
    t, m, aln, pi = _block(sb, A, C, seed=80 + A + C, leaves=7)
    ot, om = otree(t), omodel(m)
    mk = np.random.default_rng(A).random((len(t.parentIndex), C)) < 0.6
    for precision in ('fp64', 'fp32'):
        tol = 1e-10 if precision == 'fp64' else 1e-4
        for mask in (None, 'auto', mk):
            ref = O.BranchCounts(aln, ot, om, branch_mask=mask)
            got = sb.BranchCounts(aln, t, m, branch_mask=mask, precision=precision, engine=eng)
            e = per_branch_err(got, ref)
            assert float(got[0].abs().max()) == 0.0
            assert e <= tol, (precision, mask if isinstance(mask, (str, type(None))) else 'explicit', e)
        jc = sb.jukes_cantor_model(A)
        bc = sb.BranchCounts(aln, t, jc, precision=precision, engine=eng)
        c = sb.Counts(aln, t, jc, branch_mask=None, precision=precision, engine=eng)
        assert float((bc.sum(0) - c).abs().max()) <= (1e-10 if precision == 'fp64' else 1e-4) * float(c.abs().max())
        f = sb.f81_model(pi)
        ref = O.BranchCounts(aln, ot, omodel(f), f81_fast=True)
        got = sb.BranchCounts(aln, t, f, f81_fast_flag=True, precision=precision, engine=eng)
        e = per_branch_err(got, ref)
        say('   ', precision, 'f81', f'{e:.2e}')
        assert e <= tol


@pytest.mark.parametrize('A,C', [(20, 300), (64, 70)])
def test_branch_counts_on_199_node_tree(sb, eng, A, C):
    """BranchCounts of a general model on the large-alphabet configurations' 100-leaf
    (R = 199) tree, FP32 (two-pass matrix-core path: per-branch GEMMs over 16-column
    groups) against the oracle, masks none and auto, ragged last tile."""
    say('branch counts on the 199-node tree, A =', A, 'C =', C)
    rng = np.random.default_rng(900 + A)
    pi = rng.dirichlet(np.full(A, 2.0))
    m = sb.diagonalize_rate_matrix(sb.reversible_rate_matrix(rng.lognormal(0, 1, (A, A)), pi), pi)
    t = sb.synthetic.random_tree(100, rng, 0.1)
    aln = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.03, ambig_frac=0.01)
    ot, om = otree(t), omodel(m)
    for mask in (None, 'auto'):
        ref = O.BranchCounts(aln, ot, om, branch_mask=mask)
        got = sb.BranchCounts(aln, t, m, branch_mask=mask, precision='fp32', engine=eng)
        e = per_branch_err(got, ref)
        say(f'    mask {mask}: {e:.2e}')
        assert float(got[0].abs().max()) == 0.0
        assert e <= 1e-4


@pytest.mark.parametrize('A,C', [(12, 70), (20, 131), (33, 45), (64, 40)])
def test_posteriors_large_alphabets(sb, eng, A, C):
    """Node and branch posteriors above A = 8 (the large-alphabet kernel's
    posterior variant): general model and Jukes-Cantor (F81-structured M,
    applied in O(A)), FP64 / FP32, vs the oracle; row 0 of the node
    posteriors is RootProb; every posterior sums to 1; branch counts and
    posteriors asked for together give what each gives alone."""
    say('posteriors, A =', A, 'C =', C)
    t, m, aln, pi = _block(sb, A, C, seed=90 + A + C, leaves=7)
    ot = otree(t)
    for name, mod in (('general', m), ('jc', sb.jukes_cantor_model(A))):
        om = omodel(mod)
        rn, rb = O.node_posterior(aln, ot, om), O.branch_posterior(aln, ot, om)
        for precision in ('fp64', 'fp32'):
            tol = 1e-12 if precision == 'fp64' else 1e-5
            io = sb.InsideOutside(aln, t, mod, precision=precision, engine=eng)
            npost, bpost = io.node_posterior(), io.branch_posterior()
            en, eb = abs_err(npost, rn), abs_err(bpost, rb)
            say('   ', name, precision, f'node {en:.2e} branch {eb:.2e}')
            assert en <= tol and eb <= tol, (name, precision, en, eb)
            assert float(bpost[0].abs().max()) == 0.0
            assert abs_err(io.node_posterior(0), O.RootProb(aln, ot, om)) <= tol
            assert float((npost.sum(1) - 1).abs().max()) <= 10 * tol
            assert float((bpost[1:].sum((1, 2)) - 1).abs().max()) <= 10 * tol
    from subby_amd.api import _inside_outside
    both = _inside_outside(aln, t, m, {'bcounts', 'npost', 'bpost', 'll'}, precision='fp64', engine=eng)
    assert per_branch_err(both['bcounts'], O.BranchCounts(aln, ot, omodel(m))) <= 1e-10
    assert abs_err(both['npost'], O.node_posterior(aln, ot, omodel(m))) <= 1e-12
    assert abs_err(both['bpost'], O.branch_posterior(aln, ot, omodel(m))) <= 1e-12
    ll = O.LogLike(aln, ot, omodel(m))
    assert np.max(np.abs(both['ll'].cpu().numpy() - ll) / np.maximum(np.abs(ll), 1)) <= 1e-12


def test_large_alphabet_per_branch_edge_shapes(sb, eng):
    """Edge shapes above A = 8 for the posterior, per-branch counts and
    irreversible counts variants: a single-node tree, one column, ragged
    column tiles, all-gap and ambiguous columns; FP64 and FP32 vs the oracle."""
    say('large-alphabet per-branch edge shapes')
    rng = np.random.default_rng(77)
    A = 12
    pi = rng.dirichlet(np.full(A, 3.0))
    m = sb.diagonalize_rate_matrix(sb.reversible_rate_matrix(rng.lognormal(0, 1, (A, A)), pi), pi)
    Q = rng.uniform(0.05, 1.0, (A, A))
    for a in range(A):
        Q[a, (a + 1) % A] += 1.5
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(1))
    mi = sb.diagonalize_irreversible(Q, pi)
    omi = O.IrrevModel(np.asarray(mi.eigenvalues), np.asarray(mi.eigenvectors),
                       np.asarray(mi.eigenvectors_inv), np.asarray(mi.pi), False)
    for R, C in [(1, 5), (3, 1), (3, 7), (9, 66)]:
        t = sb.Tree(np.array([-1]), np.array([0.0])) if R == 1 else sb.synthetic.random_tree((R + 1) // 2, rng, 0.2)
        aln = rng.integers(-1, A + 2, size=(R, C)).astype(np.int8)
        aln[:, 0] = -1                                   # an all-gap column
        ot = otree(t)
        for precision in ('fp64', 'fp32'):
            tol = 1e-10 if precision == 'fp64' else 1e-4
            tp = 1e-11 if precision == 'fp64' else 1e-5
            io = sb.InsideOutside(aln, t, m, precision=precision, engine=eng)
            assert abs_err(io.node_posterior(), O.node_posterior(aln, ot, omodel(m))) <= tp, (R, C, precision)
            assert abs_err(io.branch_posterior(), O.branch_posterior(aln, ot, omodel(m))) <= tp, (R, C, precision)
            bc = sb.BranchCounts(aln, t, m, precision=precision, engine=eng)
            rbc = O.BranchCounts(aln, ot, omodel(m))
            assert abs_err(bc, rbc) <= tol * max(1.0, float(np.abs(rbc).max())), (R, C, precision)
            ci = sb.Counts(aln, t, mi, precision=precision, engine=eng)
            rci, _, _ = O.counts_loglike_root(aln, ot, omi)
            assert abs_err(ci, rci) <= tol * max(1.0, float(np.abs(rci).max())), (R, C, precision)
