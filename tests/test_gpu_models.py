"""GPU parity for mixtures, rate categories, stacked models, per-column model
lists and per-branch model grids against the FP64 oracle.

The oracle's functions are pinned in tests/test_oracle_models.py (the
reference's mixture golden, scipy gamma quantiles, enumeration, quadrature).
Tolerances, per output:
  FP32: logLike max_c |d|/max(|ref|,1) <= 1e-5; counts normwise <= 1e-4;
        root and mixture posteriors max |d| <= 1e-5.
  FP64: logLike <= 1e-12; root <= 1e-11 and counts <= 1e-10: slow gamma
        categories (rate 0.03) scale the eigenvalues down, so M(t) = I + O(1e-3)
        comes out of a cancelling eigen-sum in both implementations (root
        differences near 2e-12), and the eigenvalue gaps near 1e-3 cost the
        oracle's J, the reference's difference quotient
        (E_k - E_l) / (mu_k - mu_l), digits that the kernels' expm1 form keeps.
"""
import numpy as np
import pytest
import torch

from oracle import subby_oracle as O
from tests.golden_cases import load
from tests.test_gpu_parity import err_counts, err_ll, err_root, otree, say

pytestmark = pytest.mark.gpu

TOL = {'fp32': dict(ll=1e-5, counts=1e-4, root=1e-5),
       'fp64': dict(ll=1e-12, counts=1e-10, root=1e-11)}


@pytest.fixture(scope='module')
def sb():
    import subby_amd
    assert torch.cuda.is_available(), 'GPU tests need a ROCm GPU'
    return subby_amd


@pytest.fixture(scope='module')
def eng(sb):
    return sb.Engine(0)


def om(m):
    return O.DiagModel(*(np.asarray(x, dtype=np.float64) for x in m))


def err_bc(got, ref):
    got = got.detach().cpu().double().numpy()
    num = np.abs(got - ref).max(axis=(1, 2))
    den = np.maximum(np.abs(ref).max(axis=(1, 2)), 1e-300)
    return float(np.max(num[1:] / den[1:])) if ref.shape[0] > 1 else 0.0


def test_mixture_posterior_golden(sb, eng):
    say('mixture golden')
    g = load('7node_mixture_jc4')
    tree = sb.Tree(g['inputs.parentIndex'], g['inputs.distanceToParent'])
    base = sb.DiagModel(g['inputs.base_model.eigenvalues'], g['inputs.base_model.eigenvectors'],
                        g['inputs.base_model.pi'])
    aln = g['inputs.alignment']
    for prec in ('fp64', 'fp32'):
        for models in ([sb.scale_model(base, r) for r in g['inputs.rates']],
                       sb.scale_model(base, g['inputs.rates'])):
            post, ll = sb.MixturePosterior(aln, tree, models, g['inputs.log_weights'], precision=prec,
                                           engine=eng, return_loglikes=True)
            e_ll = float(np.max(np.abs(ll.cpu().numpy() - g['outputs.log_likes'])
                                / np.maximum(np.abs(g['outputs.log_likes']), 1)))
            e_p = float(np.abs(post.cpu().numpy() - g['outputs.posteriors']).max())
            say('   ', prec, f'loglikes {e_ll:.2e} posteriors {e_p:.2e}')
            assert e_ll <= TOL[prec]['ll'] and e_p <= TOL[prec]['root']
        # the stacked model through LogLike gives (K, C)
        ll2 = sb.LogLike(aln, tree, sb.scale_model(base, g['inputs.rates']), precision=prec, engine=eng)
        assert ll2.shape == (3, 8)
        assert err_ll(ll2.reshape(-1), g['outputs.log_likes'].reshape(-1)) <= TOL[prec]['ll']


def _hky_case(sb, seed, leaves=30, C=900):
    rng = np.random.default_rng(seed)
    pi = rng.dirichlet(np.full(4, 5.0))
    m = sb.hky85_diag(2.0, pi)
    t = sb.synthetic.random_tree(leaves, rng, 0.1)
    aln = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.05, ambig_frac=0.02)
    return rng, m, t, aln


def test_gamma_mixture_and_stacked_models(sb, eng):
    say('gamma rate categories, stacked models')
    rng, m, t, aln = _hky_case(sb, 71)
    rates, w = sb.gamma_rate_categories(0.5, 4)
    ref_r, _ = O.gamma_rate_categories(0.5, 4)
    np.testing.assert_allclose(rates, ref_r, rtol=1e-14)
    stack = sb.scale_model(m, rates)
    oms = [O.scale_model(om(m), r) for r in rates]
    ref_post = O.MixturePosterior(aln, otree(t), oms, np.log(w))
    for prec in ('fp64', 'fp32'):
        post = sb.MixturePosterior(aln, t, stack, np.log(w), precision=prec, engine=eng)
        assert float(np.abs(post.cpu().numpy() - ref_post).max()) <= TOL[prec]['root']
        cnt, ll, root = sb.Counts(aln, t, stack, precision=prec, engine=eng, return_loglike=True,
                                  return_root=True)
        assert cnt.shape == (4, 4, 4, aln.shape[1]) and ll.shape == (4, aln.shape[1])
        for k in range(4):
            rc, rl, rr = O.counts_loglike_root(aln, otree(t), oms[k])
            assert err_counts(cnt[k], rc) <= TOL[prec]['counts']
            assert err_ll(ll[k], rl) <= TOL[prec]['ll'] and err_root(root[k], rr) <= TOL[prec]['root']
        rp = sb.RootProb(aln, t, stack, precision=prec, engine=eng)
        assert err_root(rp[2], O.RootProb(aln, otree(t), oms[2])) <= TOL[prec]['root']
    # InsideOutside and BranchCounts over the stack: leading K axis, each slice the
    # single category's result (bitwise) and the oracle's
    io = sb.InsideOutside(aln, t, stack, precision='fp64', engine=eng)
    npost, bpost = io.node_posterior(), io.branch_posterior()
    bc = sb.BranchCounts(aln, t, stack, precision='fp64', engine=eng)
    R, C = aln.shape
    assert npost.shape == (4, R, 4, C) and bpost.shape == (4, R, 4, 4, C) and bc.shape == (4, R, 4, 4, C)
    assert io.node_posterior(3).shape == (4, 4, C)
    for k in (0, 3):
        one = sb.InsideOutside(aln, t, sb.scale_model(m, rates[k]), precision='fp64', engine=eng)
        assert torch.equal(npost[k], one.node_posterior()) and torch.equal(bpost[k], one.branch_posterior())
        assert float(np.abs(npost[k].cpu().numpy() - O.node_posterior(aln, otree(t), oms[k])).max()) \
            <= TOL['fp64']['root']
        assert torch.equal(bc[k], sb.BranchCounts(aln, t, sb.scale_model(m, rates[k]), precision='fp64', engine=eng))


def test_stacked_models_per_model_masks(sb, eng):
    """Stacked models with an explicit (*H, R, C) mask: model h uses slice h
    (subby/jax/__init__.py:355-356); an (R, C) mask is shared; wrong shapes raise
    on the host before anything is launched."""
    say('stacked models, per-model explicit masks')
    rng, m, t, aln = _hky_case(sb, 72, leaves=20, C=500)
    R, C = aln.shape
    rates = np.array([0.5, 1.0, 2.0])
    stack = sb.scale_model(m, rates)
    oms = [O.scale_model(om(m), r) for r in rates]
    masks = rng.random((3, R, C)) < np.array([0.3, 0.6, 0.9])[:, None, None]
    for prec, tol in (('fp64', 1e-12), ('fp32', 1e-4)):
        cnt = sb.Counts(aln, t, stack, branch_mask=masks, precision=prec, engine=eng)
        bc = sb.BranchCounts(aln, t, stack, branch_mask=torch.as_tensor(masks), precision=prec, engine=eng)
        shared = sb.Counts(aln, t, stack, branch_mask=masks[1], precision=prec, engine=eng)
        for h in range(3):
            ref = O.Counts(aln, otree(t), oms[h], branch_mask=masks[h])
            assert err_counts(cnt[h], ref) <= tol, (prec, h)
            ref_bc = O.BranchCounts(aln, otree(t), oms[h], branch_mask=masks[h])
            assert err_bc(bc[h], ref_bc) <= tol, (prec, h)
            ref = O.Counts(aln, otree(t), oms[h], branch_mask=masks[1])
            assert err_counts(shared[h], ref) <= tol, (prec, h)
    for bad in (masks[:2], masks[:, :-1], masks[0, :, :-1], masks[None]):
        with pytest.raises(ValueError, match='branch_mask'):
            sb.Counts(aln, t, stack, branch_mask=bad, engine=eng)
    from subby_amd.batch import Block
    with pytest.raises(ValueError, match='branch_mask'):
        eng.launcher().submit([Block(aln, t, m, branch_mask=masks[0, :-1])])
    with pytest.raises(ValueError, match='branch_mask'):
        eng.launcher().submit([Block(aln, t, m, branch_mask='bogus')])


def test_model_list(sb, eng):
    """Per-column models from three distinct models (different pi): counts per
    column under its own model; loglike / root with the first column's root
    distribution (the reference's stacked-model convention)."""
    say('per-column model list')
    rng, m, t, aln = _hky_case(sb, 81, C=700)
    C = aln.shape[1]
    base = [sb.hky85_diag(k, rng.dirichlet(np.full(4, 4.0))) for k in (1.5, 2.5, 4.0)]
    which = rng.integers(0, 3, C)
    models = [base[i] for i in which]
    oms = [om(x) for x in models]
    mk = rng.random(aln.shape) < 0.7
    ref_ll = O.LogLike_columns(aln, otree(t), oms)
    ref_root = O.RootProb_columns(aln, otree(t), oms)
    for prec in ('fp64', 'fp32'):
        tol = TOL[prec]
        ll = sb.LogLike(aln, t, models, precision=prec, engine=eng)
        rp = sb.RootProb(aln, t, models, precision=prec, engine=eng)
        e = dict(ll=err_ll(ll, ref_ll), root=err_root(rp, ref_root))
        for mask in ('auto', None, mk):
            cnt, ll2, rp2 = sb.Counts(aln, t, models, branch_mask=mask, precision=prec, engine=eng,
                                      return_loglike=True, return_root=True)
            ref_c = O.Counts_columns(aln, otree(t), oms, branch_mask=mask)
            e['counts'] = max(e.get('counts', 0.0), err_counts(cnt, ref_c))
            e['ll'] = max(e['ll'], err_ll(ll2, ref_ll))
            e['root'] = max(e['root'], err_root(rp2, ref_root))
        bc = sb.BranchCounts(aln, t, models, precision=prec, engine=eng)
        e['bcounts'] = err_bc(bc, O.BranchCounts_columns(aln, otree(t), oms))
        f = sb.Counts(aln, t, [sb.f81_model(x.pi) for x in models], f81_fast_flag=True, precision=prec,
                      engine=eng)
        e['f81'] = err_counts(f, O.Counts_columns(aln, otree(t), [om(sb.f81_model(x.pi)) for x in models],
                                                  f81_fast=True))
        say('   ', prec, {k: f'{v:.2e}' for k, v in e.items()})
        assert e['ll'] <= tol['ll'] and e['root'] <= tol['root']
        assert max(e['counts'], e['bcounts'], e['f81']) <= tol['counts']
    # [m] * C is the single model, bitwise
    same = sb.Counts(aln, t, [m] * C, engine=eng)
    assert torch.equal(same, sb.Counts(aln, t, m, engine=eng))


@pytest.mark.parametrize('A', [4, 12])
def test_inside_outside_model_list(sb, eng, A):
    """InsideOutside over a per-column model list: node and branch posteriors
    with the passes under the first model's root distribution and the root row
    under each column's own (the reference's convention), log-likelihood of
    the list, counts and branch counts of the list; A = 4 (per-branch kernel)
    and A = 12 (large-alphabet posterior variant)."""
    say('InsideOutside over a model list, A =', A)
    rng = np.random.default_rng(95 + A)
    t = sb.synthetic.random_tree(9, rng, 0.15)
    C = 300 if A == 4 else 90
    base = []
    for k in range(3):
        pi = rng.dirichlet(np.full(A, 3.0))
        Q = sb.reversible_rate_matrix(rng.lognormal(0, 1, (A, A)), pi)
        base.append(sb.diagonalize_rate_matrix(Q, pi))
    aln = sb.synthetic.simulate(t, base[0], C, rng, gap_frac=0.05, ambig_frac=0.02)
    which = rng.integers(0, 3, C)
    which[0] = 1                                   # the first column's model is not base[0]
    models = [base[i] for i in which]
    oms = [om(x) for x in models]
    ot = otree(t)
    rn = O.node_posterior_columns(aln, ot, oms)
    rb = O.branch_posterior_columns(aln, ot, oms)
    rl = O.LogLike_columns(aln, ot, oms)
    for prec in ('fp64', 'fp32'):
        tp = 1e-12 if prec == 'fp64' else 1e-5
        io = sb.InsideOutside(aln, t, models, precision=prec, engine=eng)
        en = float(np.abs(io.node_posterior().double().cpu().numpy() - rn).max())
        eb = float(np.abs(io.branch_posterior().double().cpu().numpy() - rb).max())
        el = err_ll(io.log_likelihood, rl)
        say('   ', prec, f'node {en:.2e} branch {eb:.2e} ll {el:.2e}')
        assert en <= tp and eb <= tp and el <= TOL[prec]['ll']
        assert float(io.branch_posterior(0).abs().max()) == 0.0
        assert err_counts(io.counts(branch_mask=None), O.Counts_columns(aln, ot, oms, branch_mask=None)) \
            <= TOL[prec]['counts']
        assert err_bc(io.branch_counts(), O.BranchCounts_columns(aln, ot, oms)) <= TOL[prec]['counts']
    # a list of one repeated model is that model, bitwise
    io1 = sb.InsideOutside(aln, t, [base[0]] * C, precision='fp64', engine=eng)
    io2 = sb.InsideOutside(aln, t, base[0], precision='fp64', engine=eng)
    assert torch.equal(io1.node_posterior(), io2.node_posterior())
    assert torch.equal(io1.branch_posterior(), io2.branch_posterior())


def test_inside_outside_irreversible_model_list(sb, eng):
    """The same over irreversible models (A = 4, complex eigensystems)."""
    say('InsideOutside over an irreversible model list')
    rng = np.random.default_rng(97)
    t = sb.synthetic.random_tree(8, rng, 0.15)
    C = 200
    base = []
    for k in range(2):
        Q = rng.uniform(0.05, 1.0, (4, 4))
        for a in range(4):
            Q[a, (a + 1) % 4] += 1.5                # a rotational part: complex eigenvalues
        np.fill_diagonal(Q, 0.0)
        np.fill_diagonal(Q, -Q.sum(1))
        base.append(sb.diagonalize_irreversible(Q, rng.dirichlet(np.full(4, 3.0))))
    aln = sb.synthetic.simulate(t, base[0], C, rng, gap_frac=0.05, ambig_frac=0.02)
    which = rng.integers(0, 2, C)
    which[0] = 1
    models = [base[i] for i in which]
    oms = [O.IrrevModel(np.asarray(m.eigenvalues), np.asarray(m.eigenvectors), np.asarray(m.eigenvectors_inv),
                        np.asarray(m.pi), False) for m in models]
    ot = otree(t)
    rn = O.node_posterior_columns(aln, ot, oms)
    rb = O.branch_posterior_columns(aln, ot, oms)
    for prec in ('fp64', 'fp32'):
        tp = 1e-11 if prec == 'fp64' else 1e-5
        io = sb.InsideOutside(aln, t, models, precision=prec, engine=eng)
        en = float(np.abs(io.node_posterior().double().cpu().numpy() - rn).max())
        eb = float(np.abs(io.branch_posterior().double().cpu().numpy() - rb).max())
        say('   ', prec, f'node {en:.2e} branch {eb:.2e}')
        assert en <= tp and eb <= tp


def test_per_branch_models(sb, eng):
    """Model grid (1, R): branch n under its own model, row 0 the root distribution."""
    say('per-branch models')
    rng, m, t, aln = _hky_case(sb, 91, leaves=25, C=500)
    R = len(t.parentIndex)
    rows = [sb.hky85_diag(float(rng.uniform(1, 5)), rng.dirichlet(np.full(4, 4.0))) for _ in range(R)]
    orows = [om(x) for x in rows]
    mk = rng.random(aln.shape) < 0.6
    for prec in ('fp64', 'fp32'):
        tol = TOL[prec]
        e = dict(ll=err_ll(sb.LogLike(aln, t, [rows], precision=prec, engine=eng),
                           O.LogLike_rows(aln, otree(t), orows)),
                 root=err_root(sb.RootProb(aln, t, [rows], precision=prec, engine=eng),
                               O.RootProb_rows(aln, otree(t), orows)))
        for mask in ('auto', None, mk):
            cnt = sb.Counts(aln, t, [rows], branch_mask=mask, precision=prec, engine=eng)
            e['counts'] = max(e.get('counts', 0.0), err_counts(cnt, O.Counts_rows(aln, otree(t), orows, mask)))
        e['bcounts'] = err_bc(sb.BranchCounts(aln, t, [rows], precision=prec, engine=eng),
                              O.BranchCounts_rows(aln, otree(t), orows))
        say('   ', prec, {k: f'{v:.2e}' for k, v in e.items()})
        assert e['ll'] <= tol['ll'] and e['root'] <= tol['root']
        assert max(e['counts'], e['bcounts']) <= tol['counts']
    # all branches on one model: the single-model result
    got = sb.Counts(aln, t, [[m] * R], precision='fp64', engine=eng)
    ref = O.Counts(aln, otree(t), om(m))
    assert err_counts(got, ref) <= 1e-11
    # irreversible per-branch models: log-likelihood, root, counts (each branch
    # in its own complex eigenbasis, the reference's convention) and branch counts
    Q = rng.uniform(0.05, 1.0, (4, 4))
    for a in range(4):
        Q[a, (a + 1) % 4] += 1.5                     # complex eigenvalues
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(1))
    irows = [sb.diagonalize_irreversible(Q * float(rng.uniform(0.5, 2)), rows[0].pi) for _ in range(R)]
    oir = [O.IrrevModel(*(np.asarray(x) for x in r), False) for r in irows]
    for prec in ('fp64', 'fp32'):
        tol = TOL[prec]
        e = dict(ll=err_ll(sb.LogLike(aln, t, [irows], precision=prec, engine=eng), O.LogLike_rows(aln, otree(t), oir)),
                 root=err_root(sb.RootProb(aln, t, [irows], precision=prec, engine=eng),
                               O.RootProb_rows(aln, otree(t), oir)))
        for mask in ('auto', None, mk):
            cnt = sb.Counts(aln, t, [irows], branch_mask=mask, precision=prec, engine=eng)
            e['counts'] = max(e.get('counts', 0.0), err_counts(cnt, O.Counts_rows(aln, otree(t), oir, mask)))
        e['bcounts'] = err_bc(sb.BranchCounts(aln, t, [irows], precision=prec, engine=eng),
                              O.BranchCounts_rows(aln, otree(t), oir))
        say('    irreversible', prec, {k: f'{v:.2e}' for k, v in e.items()})
        assert e['ll'] <= tol['ll'] and e['root'] <= tol['root']
        assert max(e['counts'], e['bcounts']) <= tol['counts']


@pytest.mark.parametrize('A', [12, 20])
def test_per_branch_models_large_alphabets(sb, eng, A):
    """Model grid (1, R) above A = 8: LogLike and RootProb through the large-
    alphabet kernel with per-branch M; Counts (every mask) and BranchCounts
    through its per-branch counts variant with each branch's own eigenbasis
    and back-transform; irreversible per-branch LogLike; a grid of one model is
    that model."""
    say('per-branch models, A =', A)
    rng = np.random.default_rng(190 + A)
    t = sb.synthetic.random_tree(8, rng, 0.15)
    R = len(t.parentIndex)
    C = 120

    def rev(scale):
        pi = rng.dirichlet(np.full(A, 3.0))
        Q = sb.reversible_rate_matrix(rng.lognormal(0, 1, (A, A)), pi) * scale
        return sb.diagonalize_rate_matrix(Q, pi)
    rows = [rev(float(rng.uniform(0.5, 2.0))) for _ in range(R)]
    aln = sb.synthetic.simulate(t, rows[0], C, rng, gap_frac=0.05, ambig_frac=0.02)
    orows = [om(x) for x in rows]
    ot = otree(t)
    mk = rng.random(aln.shape) < 0.6
    for prec in ('fp64', 'fp32'):
        tol = TOL[prec]
        e = dict(ll=err_ll(sb.LogLike(aln, t, [rows], precision=prec, engine=eng), O.LogLike_rows(aln, ot, orows)),
                 root=err_root(sb.RootProb(aln, t, [rows], precision=prec, engine=eng),
                               O.RootProb_rows(aln, ot, orows)))
        for mask in ('auto', None, mk):
            cnt = sb.Counts(aln, t, [rows], branch_mask=mask, precision=prec, engine=eng)
            e['counts'] = max(e.get('counts', 0.0), err_counts(cnt, O.Counts_rows(aln, ot, orows, mask)))
        e['bcounts'] = err_bc(sb.BranchCounts(aln, t, [rows], precision=prec, engine=eng),
                              O.BranchCounts_rows(aln, ot, orows))
        say('   ', prec, {k: f'{v:.2e}' for k, v in e.items()})
        assert e['ll'] <= tol['ll'] and e['root'] <= tol['root']
        assert max(e['counts'], e['bcounts']) <= tol['counts']
    got = sb.Counts(aln, t, [[rows[3]] * R], branch_mask=None, precision='fp64', engine=eng)
    assert err_counts(got, O.Counts(aln, ot, om(rows[3]), branch_mask=None)) <= 1e-10
    Q = rng.uniform(0.05, 1.0, (A, A))
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(1))
    irows = [sb.diagonalize_irreversible(Q * float(rng.uniform(0.5, 2)), rows[0].pi) for _ in range(R)]
    oir = [O.IrrevModel(*(np.asarray(x) for x in r), False) for r in irows]
    assert err_ll(sb.LogLike(aln, t, [irows], precision='fp64', engine=eng), O.LogLike_rows(aln, ot, oir)) <= 1e-12
    for prec in ('fp64', 'fp32'):
        tc = 1e-10 if prec == 'fp64' else TOL[prec]['counts']
        ec = err_counts(sb.Counts(aln, t, [irows], branch_mask=mk, precision=prec, engine=eng),
                        O.Counts_rows(aln, ot, oir, mk))
        eb = err_bc(sb.BranchCounts(aln, t, [irows], precision=prec, engine=eng), O.BranchCounts_rows(aln, ot, oir))
        say('    irreversible', prec, f'counts {ec:.2e} branch counts {eb:.2e}')
        assert ec <= tc and eb <= tc


def test_model_grid(sb, eng):
    """Model grid (C, R): a few distinct rows spread over the columns."""
    say('model grid (C, R)')
    rng, m, t, aln = _hky_case(sb, 101, leaves=12, C=300)
    R, C = aln.shape
    rowsets = [[sb.hky85_diag(float(rng.uniform(1, 5)), rng.dirichlet(np.full(4, 4.0))) for _ in range(R)]
               for _ in range(3)]
    grid = [rowsets[i] for i in rng.integers(0, 3, C)]
    ogrid = [[om(x) for x in row] for row in grid]
    for prec in ('fp64', 'fp32'):
        tol = TOL[prec]
        cnt, ll, rp = sb.Counts(aln, t, grid, precision=prec, engine=eng, return_loglike=True, return_root=True)
        e = dict(counts=err_counts(cnt, O.Counts_grid(aln, otree(t), ogrid)),
                 ll=err_ll(ll, O.LogLike_grid(aln, otree(t), ogrid)),
                 root=err_root(rp, O.RootProb_grid(aln, otree(t), ogrid)))
        say('   ', prec, {k: f'{v:.2e}' for k, v in e.items()})
        assert e['ll'] <= tol['ll'] and e['root'] <= tol['root'] and e['counts'] <= tol['counts']


def test_model_table_errors(sb, eng):
    say('model table errors')
    rng, m, t, aln = _hky_case(sb, 111, leaves=6, C=40)
    R, C = aln.shape
    with pytest.raises(ValueError):
        sb.LogLike(aln, t, [m] * (C - 1), engine=eng)
    with pytest.raises(sb.SubbyError, match='F81'):
        sb.Counts(aln, t, [[sb.f81_model(m.pi)] * R], f81_fast_flag=True, engine=eng)
    with pytest.raises(NotImplementedError):
        sb.InsideOutside(aln, t, [[m] * R], engine=eng)
    with pytest.raises(sb.SubbyError, match='alphabet'):
        sb.LogLike(aln, t, [m] + [sb.jukes_cantor_model(5)] * (C - 1), engine=eng)
    # the engine still works after errors
    assert err_ll(sb.LogLike(aln, t, [m] * C, engine=eng), O.LogLike(aln, otree(t), om(m))) <= 1e-5


def test_padding_round_trips(sb, eng):
    """Gap-padded columns and dummy leaves change nothing (the reference's
    TestPadAlignment, tests/test_api.py:758-969): round trips, padded columns
    with logL = 0, root posterior = pi and zero masked counts."""
    say('padding round trips')
    rng, m, t, aln = _hky_case(sb, 121, leaves=9, C=37)
    pt, pa, R0, C0 = sb.pad_tree_and_alignment(t, aln)
    assert pa.shape[0] > R0 and pa.shape[1] > C0
    for prec in ('fp64', 'fp32'):
        tol = TOL[prec]
        ll = sb.LogLike(aln, t, m, precision=prec, engine=eng)
        llp = sb.LogLike(pa, pt, m, precision=prec, engine=eng)
        assert float((sb.unpad_columns(llp, C0) - ll).abs().max()) <= tol['ll'] * max(1.0, float(ll.abs().max()))
        # an all-gap column has logL = 0 exactly in FP64; FP32 rows of M sum to 1 only to rounding
        assert float(llp[C0:].abs().max()) <= (1e-12 if prec == 'fp64' else tol['ll'])
        rp = sb.RootProb(pa, pt, m, precision=prec, engine=eng)
        assert float((sb.unpad_columns(rp, C0) - sb.RootProb(aln, t, m, precision=prec, engine=eng)).abs().max()) <= tol['root']
        pi = torch.as_tensor(np.asarray(m.pi), dtype=rp.dtype, device=rp.device)
        assert float((rp[:, C0:] - pi[:, None]).abs().max()) <= tol['root']
        cp = sb.Counts(pa, pt, m, precision=prec, engine=eng)
        c = sb.Counts(aln, t, m, precision=prec, engine=eng)
        assert err_counts(sb.unpad_columns(cp, C0), c.double().cpu().numpy()) <= tol['counts']
        assert float(cp[..., C0:].abs().max()) == 0.0
        d, _, _ = sb.loglike_grad(pa, pt, m, precision=prec, engine=eng)
        d0, _, _ = sb.loglike_grad(aln, t, m, precision=prec, engine=eng)
        assert d.shape[0] == pa.shape[0] and float(d[R0:].abs().max()) == 0.0
        assert float((d[:R0] - d0).abs().max()) <= (1e-10 if prec == 'fp64' else 1e-4) * float(d0.abs().max())
    padded, C1 = sb.pad_alignment(aln, bin_size=16)
    post = sb.MixturePosterior(padded, t, [sb.scale_model(m, r) for r in (0.5, 1.0, 2.0)], np.log(np.ones(3) / 3),
                               engine=eng, precision='fp64')
    ref = sb.MixturePosterior(aln, t, [sb.scale_model(m, r) for r in (0.5, 1.0, 2.0)], np.log(np.ones(3) / 3),
                              engine=eng, precision='fp64')
    assert float((sb.unpad_columns(post, C1) - ref).abs().max()) <= 1e-12


def test_gy94_codon_model(sb, eng):
    """GY94 on the 61 sense codons through the large-alphabet paths.  FP32 counts
    are held to 1e-3 here, not 1e-4: the dwell times of this sparse 61-state
    model come out of the eigenbasis with heavy cancellation, and the same
    algorithm run in FP32 on the CPU (FP32 partials and projections) is off by
    3e-4 on this block (measured on the GPU: 1.2e-4).  FP64 is the precision for
    codon models."""
    say('GY94 codon model')
    rng = np.random.default_rng(131)
    pi = rng.dirichlet(np.full(61, 10.0))
    m = sb.gy94_model(0.4, 2.2, pi)
    t = sb.synthetic.random_tree(10, rng, 0.2)
    aln = sb.synthetic.simulate(t, m, 200, rng, gap_frac=0.05, ambig_frac=0.02)
    for prec in ('fp64', 'fp32'):
        tol = TOL[prec]
        rc, rl, rr = O.counts_loglike_root(aln, otree(t), om(m))
        cnt, ll, root = sb.Counts(aln, t, m, precision=prec, engine=eng, return_loglike=True, return_root=True)
        e = dict(counts=err_counts(cnt, rc), ll=err_ll(ll, rl), root=err_root(root, rr))
        say('   ', prec, {k: f'{v:.2e}' for k, v in e.items()})
        assert e['counts'] <= (tol['counts'] if prec == 'fp64' else 1e-3)
        assert e['ll'] <= tol['ll'] and e['root'] <= tol['root']
