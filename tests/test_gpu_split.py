"""Split walks: two waves walk each 64-column tile of the small-alphabet
single-model kinds (four states, reversible or F81), one a subtree the planner
names and the other the rest of the tree.  Log-likelihoods and root
posteriors equal the one-wave walk's bitwise; counts (the subtree's branch
terms summed apart) agree with it to rounding, and everything with the FP64
oracle within the FP32 / FP64 tolerances of the parity tests."""
import numpy as np
import pytest
import torch

from oracle import subby_oracle as O
from tests.test_gpu_parity import TOL32, TOL64, err_counts, err_ll, err_root, omodel, otree, say

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


def _trees(sb, rng):
    """name -> (tree, whether the planner gives a second wave a subtree)."""
    R = 127   # balanced, numbered breadth-first (parentIndex[i] = (i - 1) // 2)
    heap = sb.Tree(np.array([-1] + [(i - 1) // 2 for i in range(1, R)], dtype=np.int32),
                   np.r_[0.0, rng.uniform(0.01, 0.2, R - 1)])
    parent, spine = [-1], 0   # caterpillar: every subtree is a leaf away from the whole
    for _ in range(39):
        parent += [spine, spine]
        spine = len(parent) - 1
    cat = sb.Tree(np.array(parent, dtype=np.int32), np.r_[0.0, rng.uniform(0.01, 0.2, len(parent) - 1)])
    small = sb.synthetic.random_tree(5, rng, 0.1)   # 4 internal nodes: too few to split
    return dict(heap=(heap, True), random=(sb.synthetic.random_tree(64, rng, 0.08), None),
                caterpillar=(cat, False), small=(small, False))


ROUND = dict(fp32=1e-5, fp64=1e-13)   # counts, split vs one wave (normwise)


def _same(x, y, precision):
    """(counts, loglike, root) of a split and a one-wave launch."""
    assert torch.equal(x[1], y[1]) and torch.equal(x[2], y[2])
    assert err_counts(x[0], y[0].double().cpu().numpy()) <= ROUND[precision]


def _run(sb, L, blocks, precision, split, use_graph=False):
    from subby_amd import _lib
    L.set_tune(_lib.SB_TUNE_SPLIT, split)
    L.submit(blocks, precision=precision, use_graph=use_graph)
    L.sync()
    return L.get_tune(_lib.SB_TUNE_LAST_SPLIT), [(b.counts.clone(), b.loglike.clone(), b.root.clone())
                                                 for b in blocks]


@pytest.mark.parametrize('precision', ['fp32', 'fp64'])
@pytest.mark.parametrize('kind', ['hky85', 'hky85-nomask', 'f81'])
def test_split_equals_one_wave_and_oracle(sb, eng, precision, kind):
    from subby_amd.batch import Block, Launcher
    rng = np.random.default_rng(401)
    f81 = kind == 'f81'
    mask = None if kind == 'hky85-nomask' else 'auto'
    tol = TOL32 if precision == 'fp32' else TOL64
    L = Launcher(eng)
    for name, (t, splits) in _trees(sb, rng).items():
        pi = rng.dirichlet(np.full(4, 4.0))
        m = sb.f81_model(pi) if f81 else sb.hky85_diag(2.5, pi)
        C = 1000 if name != 'small' else 130
        aln = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.04, ambig_frac=0.01)
        say('split', precision, kind, name, 'R =', len(t.parentIndex), 'C =', C)
        runs = {}
        for split in (0, 1):
            runs[split] = _run(sb, L, [Block(aln, t, m, f81_fast=f81, branch_mask=mask)], precision, split)
        assert runs[0][0] == 0
        if splits is not None:
            assert runs[1][0] == int(splits), name
        _same(runs[1][1][0], runs[0][1][0], precision)
        if not runs[1][0]:   # (no subtree given: the one-wave walk itself)
            for a, b in zip(runs[0][1][0], runs[1][1][0]):
                assert torch.equal(a, b), name
        cnt, ll, root = runs[1][1][0]
        ref_cnt, ref_ll, ref_root = O.counts_loglike_root(aln, otree(t), omodel(m), f81_fast=f81,
                                                          branch_mask=mask)
        e = dict(ll=err_ll(ll, ref_ll), counts=err_counts(cnt, ref_cnt), root=err_root(root, ref_root))
        say('   errors', {k: f'{v:.2e}' for k, v in e.items()})
        assert e['ll'] <= tol['ll'] and e['counts'] <= tol['counts'] and e['root'] <= tol['root']


def test_split_mixed_batch_and_graph(sb, eng):
    """Blocks whose trees split and blocks whose trees do not, partial tiles
    and a one-column block in one launch: the one-wave launch's results (to
    rounding in counts), and replays of a captured graph (eager work between
    replays) equal the eager split launch bitwise."""
    from subby_amd.batch import Block, Launcher
    rng = np.random.default_rng(402)
    trees = _trees(sb, rng)
    order = ['heap', 'caterpillar', 'random', 'small', 'heap', 'random']
    blocks = []
    for name, C in zip(order, [1, 65, 200, 64, 777, 129]):
        t = trees[name][0]
        m = sb.hky85_diag(float(rng.uniform(1, 5)), rng.dirichlet(np.full(4, 5.0)))
        blocks.append(Block(sb.synthetic.simulate(t, m, C, rng, gap_frac=0.05), t, m))
    L = Launcher(eng)
    say('split mixed batch')
    s0, one = _run(sb, L, blocks, 'fp32', 0)
    s2, two = _run(sb, L, blocks, 'fp32', 1)
    assert s0 == 0 and s2 == 1
    for x, y in zip(two, one):
        _same(x, y, 'fp32')
    for rep in range(3):
        s, g = _run(sb, L, blocks, 'fp32', 1, use_graph=True)
        assert s == 1
        torch.ones(1024, device=eng.device).cumsum(0)
        for x, y in zip(two, g):
            for a, b in zip(x, y):
                assert torch.equal(a, b), rep
    # no tree here has 64 internal nodes: none splits, the one-wave results exactly
    from subby_amd import _lib
    L.set_tune(_lib.SB_TUNE_SPLIT_MIN, 64)
    s64, same = _run(sb, L, blocks, 'fp32', 1)
    assert s64 == 0 and L.get_tune(_lib.SB_TUNE_SPLIT_MIN) == 64
    for x, y in zip(one, same):
        for a, b in zip(x, y):
            assert torch.equal(a, b)
    b = blocks[4]
    ref_cnt, ref_ll, _ = O.counts_loglike_root(b.alignment, otree(b.tree), omodel(b.model))
    assert err_counts(two[4][0], ref_cnt) <= TOL32['counts'] and err_ll(two[4][1], ref_ll) <= TOL32['ll']


def test_split_gradient(sb, eng):
    """Branch-length and log-kappa gradients on the balanced tree, walked with
    two waves per tile, against the oracle; branch-length gradients equal the
    one-wave walk's bitwise."""
    from subby_amd import _lib
    rng = np.random.default_rng(403)
    t = _trees(sb, rng)['heap'][0]
    pi = rng.dirichlet(np.full(4, 5.0))
    m = sb.hky85_diag(2.0, pi)
    C = 3000
    aln = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.03, ambig_frac=0.01)
    g = rng.normal(size=C)
    W = sb.hky85_log_kappa_weights(2.0, pi)
    say('split gradient')
    ref, _ = O.loglike_branch_grad(aln, otree(t), omodel(m), g)
    refk = O.hky85_loglike_kappa_grad(aln, otree(t), 2.0, pi, g)
    for prec, tol in (('fp64', 1e-10), ('fp32', 1e-4)):
        out = {}
        for split in (0, 1):
            _lib.check(eng.lib.sb_context_set_tune(eng.ctx, _lib.SB_TUNE_SPLIT, split))
            out[split] = sb.loglike_grad(aln, t, m, g=g, rate_weights=W, precision=prec, engine=eng)
        d, r, ll = out[1]
        assert torch.equal(d, out[0][0]) and torch.equal(ll, out[0][2])
        err = np.abs(d.cpu().numpy() - ref).max() / np.abs(ref).max()
        errk = abs(float(r[0]) - refk) / max(1.0, abs(refk))
        say('   grad err', prec, f'{err:.2e}', 'log kappa', f'{errk:.2e}')
        assert err <= tol and errk <= (1e-9 if prec == 'fp64' else 1e-4)
    # a fresh context tuned before its first call, then called on a side stream
    e2 = sb.Engine(0)
    _lib.check(e2.lib.sb_context_set_tune(e2.ctx, _lib.SB_TUNE_SPLIT, 0))
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        d2, _, ll2 = sb.loglike_grad(aln, t, m, g=g, rate_weights=W, precision='fp32', engine=e2)
    st.synchronize()
    assert torch.equal(d2, d) and torch.equal(ll2, ll)
