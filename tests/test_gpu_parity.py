"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Tolerances (stated per output, FP32 path vs FP64 oracle):
  logLike       max_c |d| / max(|ref|, 1)                  <= 1e-5
  counts        max_c max_ij |d| / max_ij |ref|   (normwise) <= 1e-4
  root          max |d|                                       <= 1e-5
FP64 path: the same measures <= 1e-12.
"""
import sys

import numpy as np
import pytest
import torch

from oracle import subby_oracle as O
from tests.golden_cases import CASES, load

pytestmark = pytest.mark.gpu

TOL32 = dict(ll=1e-5, counts=1e-4, root=1e-5)
TOL64 = dict(ll=1e-12, counts=1e-12, root=1e-12)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


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


def otree(t):
    return O.Tree(np.asarray(t.parentIndex), np.asarray(t.distanceToParent))


def omodel(m):
    return O.DiagModel(*(np.asarray(x, dtype=np.float64) for x in m))


def err_ll(got, ref):
    got = got.detach().cpu().numpy()
    return float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0))) if len(ref) else 0.0


def err_counts(got, ref):
    got = got.detach().cpu().double().numpy()
    if ref.shape[-1] == 0:
        return 0.0
    num = np.abs(got - ref).max(axis=(0, 1))
    den = np.maximum(np.abs(ref).max(axis=(0, 1)), 1e-300)
    return float(np.max(num / den))


def err_root(got, ref):
    got = got.detach().cpu().double().numpy()
    return float(np.max(np.abs(got - ref))) if ref.size else 0.0


def check_all(sb, eng, aln, tree, model, precision='fp32', mask='auto', f81=False, tol=None):
    tol = tol or (TOL32 if precision == 'fp32' else TOL64)
    ot, om = otree(tree), omodel(model)
    ref_cnt, ref_ll, ref_root = O.counts_loglike_root(np.asarray(aln), ot, om, f81_fast=f81, branch_mask=mask)
    cnt, ll, root = sb.Counts(aln, tree, model, f81_fast_flag=f81, branch_mask=mask, precision=precision,
                              engine=eng, return_loglike=True, return_root=True)
    ll2 = sb.LogLike(aln, tree, model, precision=precision, engine=eng)
    root2 = sb.RootProb(aln, tree, model, precision=precision, engine=eng)
    e = dict(ll=err_ll(ll, ref_ll), counts=err_counts(cnt, ref_cnt), root=err_root(root, ref_root),
             ll2=err_ll(ll2, ref_ll), root2=err_root(root2, ref_root))
    say('   errors', precision, {k: f'{v:.2e}' for k, v in e.items()})
    assert e['ll'] <= tol['ll'] and e['ll2'] <= tol['ll']
    assert e['counts'] <= tol['counts']
    assert e['root'] <= tol['root'] and e['root2'] <= tol['root']
    return cnt, ll, root


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
@pytest.mark.parametrize('case', [c for c in CASES if c != '7node_jc64'])
def test_goldens(sb, eng, case, precision):
    say('golden', case, precision)
    g = load(case)
    m = sb.DiagModel(g['inputs.model.eigenvalues'], g['inputs.model.eigenvectors'], g['inputs.model.pi'])
    t = sb.Tree(g['inputs.parentIndex'], g['inputs.distanceToParent'])
    aln = g['inputs.alignment']
    tol = TOL32 if precision == 'fp32' else TOL64
    cnt, ll, root = sb.Counts(aln, t, m, branch_mask=None, precision=precision, engine=eng,
                              return_loglike=True, return_root=True)
    assert err_ll(ll, g['outputs.logLike']) <= tol['ll']
    assert err_counts(cnt, g['outputs.counts']) <= tol['counts']
    assert err_root(root, g['outputs.root_prob']) <= tol['root']
    check_all(sb, eng, aln, t, m, precision, mask='auto')
    check_all(sb, eng, aln, t, m, precision, mask='auto', f81=True)


def _c1(sb, seed=1):
    rng = np.random.default_rng(seed)
    t = sb.synthetic.random_tree(8, rng, 0.1)
    m = sb.jukes_cantor_model(4)
    aln = sb.synthetic.simulate(t, m, 1000, rng, gap_frac=0.05)
    return t, m, aln


def test_config_small_jc_fp64(sb, eng):
    """JC A=4, 8 leaves, C=1000, 5% gaps: FP64 end to end at 1e-12."""
    say('config small JC fp64')
    t, m, aln = _c1(sb)
    for mask in ('auto', None):
        check_all(sb, eng, aln, t, m, 'fp64', mask=mask)
        check_all(sb, eng, aln, t, m, 'fp64', mask=mask, f81=True)
        check_all(sb, eng, aln, t, m, 'fp32', mask=mask)


def _hky(sb, C, seed=2, leaves=64, mean=0.05):
    rng = np.random.default_rng(seed)
    pi = rng.dirichlet(np.full(4, 5.0))
    m = sb.hky85_diag(2.0, pi)
    t = sb.synthetic.random_tree(leaves, rng, mean)
    aln = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.03, ambig_frac=0.01)
    return t, m, aln


def test_hky85_127_nodes_fp32(sb, eng):
    say('HKY85 127 nodes C=20000 fp32')
    t, m, aln = _hky(sb, 20000)
    check_all(sb, eng, aln, t, m, 'fp32', mask='auto')
    check_all(sb, eng, aln, t, m, 'fp32', mask=None)
    check_all(sb, eng, aln, t, m, 'fp64', mask='auto')


@pytest.mark.parametrize('A', [2, 3, 5, 8])
def test_padded_alphabets(sb, eng, A):
    say('padded alphabet', A)
    rng = np.random.default_rng(10 + A)
    pi = rng.dirichlet(np.full(A, 3.0))
    Q = sb.reversible_rate_matrix(rng.lognormal(0, 1, (A, A)), pi)
    m = sb.diagonalize_rate_matrix(Q, pi)
    t = sb.synthetic.random_tree(20, rng, 0.1)
    aln = sb.synthetic.simulate(t, m, 777, rng, gap_frac=0.05, ambig_frac=0.02)
    check_all(sb, eng, aln, t, m, 'fp32')
    check_all(sb, eng, aln, t, m, 'fp64')
    f = sb.f81_model(pi)
    check_all(sb, eng, aln, t, f, 'fp32', f81=True)
    check_all(sb, eng, aln, t, f, 'fp64', f81=True)


def test_edge_shapes(sb, eng):
    say('edge shapes')
    rng = np.random.default_rng(5)
    m = sb.hky85_diag(3.0, rng.dirichlet(np.full(4, 5.0)))
    # single-node tree, 3-node tree, C = 1, 63, 65, all-gap and observed internal nodes
    for R, C in [(1, 5), (3, 1), (3, 65), (9, 63), (31, 129)]:
        if R == 1:
            t = sb.Tree(np.array([-1]), np.array([0.0]))
        else:
            t = sb.synthetic.random_tree((R + 1) // 2, rng, 0.2)
        aln = rng.integers(-1, 6, size=(R, C)).astype(np.int8)
        aln[:, 0] = -1
        for prec in ('fp32', 'fp64'):
            check_all(sb, eng, aln, t, m, prec, mask='auto')
            check_all(sb, eng, aln, t, m, prec, mask=None)


def test_deep_caterpillar_mask(sb, eng):
    """Deep unbalanced tree: the exact (oracle) Steiner mask, not a depth-bounded one."""
    say('caterpillar')
    L = 40
    parent = [-1]
    spine = 0
    for _ in range(L - 1):
        parent += [spine, spine]
        spine = len(parent) - 1
    parent = np.array(parent, dtype=np.int32)
    R = len(parent)
    rng = np.random.default_rng(6)
    t = sb.Tree(parent, np.r_[0.0, rng.uniform(0.01, 0.3, R - 1)])
    aln = np.full((R, 64), -1, dtype=np.int8)
    aln[R - 1] = rng.integers(0, 4, 64)
    aln[1] = rng.integers(0, 4, 64)
    aln[5, ::3] = 2
    m = sb.jukes_cantor_model(4)
    check_all(sb, eng, aln, t, m, 'fp64', mask='auto')
    check_all(sb, eng, aln, t, m, 'fp32', mask='auto')


def test_explicit_mask(sb, eng):
    say('explicit mask')
    t, m, aln = _hky(sb, 300, seed=8, leaves=10)
    mk = np.random.default_rng(9).random((len(t.parentIndex), 300)) < 0.6
    ref = O.Counts(aln, otree(t), omodel(m), branch_mask=mk)
    got = sb.Counts(aln, t, m, branch_mask=mk, precision='fp64', engine=eng)
    assert err_counts(got, ref) <= 1e-12
    got = sb.Counts(aln, t, m, branch_mask=torch.as_tensor(mk), precision='fp32', engine=eng)
    assert err_counts(got, ref) <= 1e-4


def test_errors(sb, eng):
    say('errors')
    m = sb.jukes_cantor_model(4)
    aln = np.zeros((4, 3), dtype=np.int8)
    with pytest.raises(sb.SubbyError, match='children'):
        sb.LogLike(aln, sb.Tree(np.array([-1, 0, 0, 0]), np.zeros(4)), m, engine=eng)
    with pytest.raises(sb.SubbyError, match='preorder'):
        sb.LogLike(aln[:3], sb.Tree(np.array([-1, 2, 0]), np.zeros(3)), m, engine=eng)
    with pytest.raises(sb.SubbyError, match='1 child'):
        sb.LogLike(aln[:2], sb.Tree(np.array([-1, 0]), np.zeros(2)), m, engine=eng)
    with pytest.raises(sb.SubbyError, match='alphabet'):
        eng.model_id(sb.jukes_cantor_model(65))
    with pytest.raises(ValueError):
        sb.Counts(aln[:3], sb.Tree(np.array([-1, 0, 0]), np.zeros(3)), m, branch_mask='bogus', engine=eng)
    # the engine still works after errors
    ll = sb.LogLike(aln[:3], sb.Tree(np.array([-1, 0, 0]), np.array([0, .1, .2])), m, engine=eng)
    assert torch.isfinite(ll).all()


def test_batch_equals_single_bitwise(sb, eng):
    """Uneven trees, column counts and models in one launch == one block per call, bitwise."""
    say('batch vs single')
    from subby_amd.batch import Launcher, Block
    rng = np.random.default_rng(11)
    blocks = []
    for i in range(12):
        leaves = int(rng.integers(2, 70))
        C = int(rng.integers(1, 900))
        m = sb.hky85_diag(float(rng.uniform(1, 5)), rng.dirichlet(np.full(4, 5.0)))
        t = sb.synthetic.random_tree(leaves, rng, 0.08)
        aln = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.05)
        blocks.append(Block(aln, t, m))
    L = Launcher(eng)
    L.submit(blocks)
    L.sync()
    for b in blocks:
        cnt, ll, root = sb.Counts(b.alignment, b.tree, b.model, engine=eng, return_loglike=True,
                                  return_root=True)
        torch.cuda.synchronize()
        assert torch.equal(cnt, b.counts) and torch.equal(ll, b.loglike) and torch.equal(root, b.root)
    # and against the oracle
    for b in blocks[:4]:
        ref_cnt, ref_ll, ref_root = O.counts_loglike_root(b.alignment, otree(b.tree), omodel(b.model))
        assert err_counts(b.counts, ref_cnt) <= 1e-4 and err_ll(b.loglike, ref_ll) <= 1e-5


def test_graph_replays_match_eager(sb, eng):
    """hipGraph replays, with eager work and syncs in between, equal the eager launch."""
    say('graph replay')
    from subby_amd.batch import Launcher, Block
    rng = np.random.default_rng(12)
    L = Launcher(eng)
    for rep in range(4):
        blocks = []
        for i in range(int(rng.integers(1, 6))):
            m = sb.hky85_diag(float(rng.uniform(1, 5)), rng.dirichlet(np.full(4, 5.0)))
            t = sb.synthetic.random_tree(int(rng.integers(2, 50)), rng, 0.08)
            blocks.append(Block(sb.synthetic.simulate(t, m, int(rng.integers(1, 500)), rng), t, m))
        L.submit(blocks, use_graph=True)
        L.sync()
        for b in blocks:
            cnt, ll, root = sb.Counts(b.alignment, b.tree, b.model, engine=eng, return_loglike=True,
                                      return_root=True)
            torch.cuda.synchronize()
            assert torch.equal(cnt, b.counts) and torch.equal(ll, b.loglike) and torch.equal(root, b.root)


def test_gradient_vs_oracle_and_fd(sb, eng):
    say('gradient')
    t, m, aln = _hky(sb, 3000, seed=13, leaves=20, mean=0.1)
    rng = np.random.default_rng(14)
    g = rng.normal(size=3000)
    ref, _ = O.loglike_branch_grad(aln, otree(t), omodel(m), g)
    for prec, tol in (('fp64', 1e-10), ('fp32', 1e-4)):
        d, _, ll = sb.loglike_grad(aln, t, m, g=g, precision=prec, engine=eng)
        d = d.cpu().numpy()
        err = np.abs(d - ref).max() / np.abs(ref).max()
        say('   grad err', prec, err)
        assert err <= tol
    # log kappa through the Fisher identity
    pi = np.asarray(m.pi)
    W = sb.hky85_log_kappa_weights(2.0, pi)
    _, r, _ = sb.loglike_grad(aln, t, m, g=g, rate_weights=W, precision='fp64', engine=eng)
    refk = O.hky85_loglike_kappa_grad(aln, otree(t), 2.0, pi, g)
    assert abs(float(r[0]) - refk) <= 1e-9 * max(1.0, abs(refk))


def test_plain_path(sb, eng):
    say('plain path')
    from subby_amd.batch import counts_plain
    t, m, aln = _hky(sb, 5000, seed=15)
    for mask in ('auto', None):
        ref_cnt, ref_ll, ref_root = O.counts_loglike_root(aln, otree(t), omodel(m), branch_mask=mask)
        cnt, ll, root = counts_plain(aln, t, m, branch_mask=mask, engine=eng)
        assert err_ll(ll, ref_ll) <= 1e-5
        assert err_counts(cnt, ref_cnt) <= 1e-4
        assert err_root(root, ref_root) <= 1e-5


def test_full_size_hky85_properties(sb, eng):
    """HKY85, 127 nodes, C = 200,000: size-independent checks plus a column sample vs the oracle."""
    say('full size HKY85 C=200000')
    t, m, aln = _hky(sb, 200000, seed=16)
    cnt, ll, root = sb.Counts(aln, t, m, engine=eng, return_loglike=True, return_root=True)
    torch.cuda.synchronize()
    mask = O.compute_branch_mask(aln, np.asarray(t.parentIndex), 4)
    mask[0] = False
    dwell = torch.einsum('iic->c', cnt.double()).cpu().numpy()
    expect = (np.asarray(t.distanceToParent)[:, None] * mask).sum(0)
    assert np.max(np.abs(dwell - expect) / np.maximum(expect, 1e-3)) <= 1e-4
    assert np.abs(root.double().sum(0).cpu().numpy() - 1).max() <= 1e-5
    assert cnt.min().item() > -1e-4
    cols = np.random.default_rng(17).choice(200000, 3000, replace=False)
    sub = aln[:, cols]
    rc, rl, rr = O.counts_loglike_root(sub, otree(t), omodel(m))
    assert err_counts(cnt[:, :, cols], rc) <= 1e-4
    assert err_ll(ll[cols], rl) <= 1e-5
    assert err_root(root[:, cols], rr) <= 1e-5


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_golden_jc64(sb, eng, precision):
    say('golden 7node_jc64', precision)
    g = load('7node_jc64')
    m = sb.DiagModel(g['inputs.model.eigenvalues'], g['inputs.model.eigenvectors'], g['inputs.model.pi'])
    t = sb.Tree(g['inputs.parentIndex'], g['inputs.distanceToParent'])
    aln = g['inputs.alignment']
    tol = TOL32 if precision == 'fp32' else TOL64
    for f81 in (False, True):
        cnt, ll, root = sb.Counts(aln, t, m, branch_mask=None, f81_fast_flag=f81, precision=precision,
                                  engine=eng, return_loglike=True, return_root=True)
        assert err_ll(ll, g['outputs.logLike']) <= tol['ll']
        assert err_counts(cnt, g['outputs.counts']) <= tol['counts']
        assert err_root(root, g['outputs.root_prob']) <= tol['root']


@pytest.mark.parametrize('A', [12, 20, 61, 64])
def test_large_alphabets(sb, eng, A):
    say('large alphabet', A)
    rng = np.random.default_rng(40 + A)
    pi = rng.dirichlet(np.full(A, 2.0))
    Q = sb.reversible_rate_matrix(rng.lognormal(0, 1, (A, A)), pi)
    m = sb.diagonalize_rate_matrix(Q, pi)
    t = sb.synthetic.random_tree(15, rng, 0.1)
    aln = sb.synthetic.simulate(t, m, 203, rng, gap_frac=0.05, ambig_frac=0.02)
    check_all(sb, eng, aln, t, m, 'fp32')
    # FP64 bound 1e-10 for a general 64-state model: its closest eigenvalues are
    # 0.002 apart, and the oracle's (reference's) J = (e^{mu_k t} - e^{mu_l t})/(mu_k - mu_l)
    # loses digits there; the kernel's e^{mu_l t} expm1((mu_k - mu_l) t)/(mu_k - mu_l)
    # does not (tests/test_oracle_invariants.py::test_J_expm1_form_is_the_accurate_one).
    check_all(sb, eng, aln, t, m, 'fp64', tol=dict(ll=1e-12, counts=1e-10, root=1e-12))
    f = sb.f81_model(pi)
    check_all(sb, eng, aln, t, f, 'fp32', f81=True)
    check_all(sb, eng, aln, t, f, 'fp64', f81=True)
    check_all(sb, eng, aln, t, f, 'fp32', mask=None)
    d, _, _ = sb.loglike_grad(aln, t, m, precision='fp64', engine=eng)
    ref, _ = O.loglike_branch_grad(aln, otree(t), omodel(m))
    assert np.abs(d.cpu().numpy() - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize('A,C', [(64, 400), (64, 203), (32, 256), (20, 129), (16, 64)])
def test_f81_structured_large_alphabets(sb, eng, A, C):
    """F81 fast path on F81-structured models in FP32 (the matrix-core kernel): widths
    with C % 4 == 0 (vector stores) and ragged tails, every mask mode, and a
    non-F81-structured model with f81_fast (the general large-alphabet kernel)."""
    say('F81-structured', A, C)
    rng = np.random.default_rng(700 + A + C)
    pi = rng.dirichlet(np.full(A, 2.0))
    f = sb.f81_model(pi)
    t = sb.synthetic.random_tree(40, rng, 0.1)
    aln = sb.synthetic.simulate(t, f, C, rng, gap_frac=0.05, ambig_frac=0.02)
    aln[3, : C // 3] = -1          # a ragged run of gaps on one internal row
    for mask in ('auto', None):
        check_all(sb, eng, aln, t, f, 'fp32', mask=mask, f81=True)
    mk = rng.random((len(t.parentIndex), C)) < 0.6
    ref = O.Counts(aln, otree(t), omodel(f), f81_fast=True, branch_mask=mk)
    got = sb.Counts(aln, t, f, f81_fast_flag=True, branch_mask=mk, precision='fp32', engine=eng)
    assert err_counts(got, ref) <= 1e-4
    # JC is F81-structured too
    jc = sb.jukes_cantor_model(A)
    check_all(sb, eng, aln, t, jc, 'fp32', f81=True)
    # f81_fast on a general reversible model (not F81-structured)
    Q = sb.reversible_rate_matrix(rng.lognormal(0, 1, (A, A)), pi)
    check_all(sb, eng, aln[:, :37], t, sb.diagonalize_rate_matrix(Q, pi), 'fp32', f81=True)


def _c3_tree(sb, seed):
    """The large-alphabet configurations' tree: 100 leaves (R = 199), t ~ Exp(mean 0.1)."""
    rng = np.random.default_rng(seed)
    return rng, sb.synthetic.random_tree(100, rng, 0.1)


def test_f81_a64_on_199_node_tree(sb, eng):
    """F81 A = 64 (pi ~ Dirichlet(2 * 1_64)) on a 100-leaf tree at C = 400 (C % 4 == 0) and
    C = 397 (ragged tail): the matrix-core F81 kernel against the oracle, masks auto and none."""
    say('F81 A=64 on the 199-node tree')
    rng, t = _c3_tree(sb, 901)
    assert len(t.parentIndex) == 199
    f = sb.f81_model(rng.dirichlet(np.full(64, 2.0)))
    aln = sb.synthetic.simulate(t, f, 400, rng, gap_frac=0.03, ambig_frac=0.01)
    for mask in ('auto', None):
        check_all(sb, eng, aln, t, f, 'fp32', mask=mask, f81=True)
    check_all(sb, eng, aln[:, :397], t, f, 'fp32', mask='auto', f81=True)


def test_general_a64_on_199_node_tree(sb, eng):
    """General reversible A = 64 (exchangeabilities ~ LogNormal(0, 1) times pi) on a 100-leaf
    tree at C = 300: the two-pass matrix-core path (FP32) and the FP64 kernel against the
    oracle (FP64 bound 1e-10: near-degenerate eigenvalues cost the oracle's difference
    quotient J digits, see test_large_alphabets)."""
    say('general A=64 on the 199-node tree')
    rng, t = _c3_tree(sb, 902)
    pi = rng.dirichlet(np.full(64, 2.0))
    m = sb.diagonalize_rate_matrix(sb.reversible_rate_matrix(rng.lognormal(0, 1, (64, 64)), pi), pi)
    aln = sb.synthetic.simulate(t, m, 300, rng, gap_frac=0.03, ambig_frac=0.01)
    check_all(sb, eng, aln, t, m, 'fp32', mask='auto')
    check_all(sb, eng, aln, t, m, 'fp64', mask='auto', tol=dict(ll=1e-12, counts=1e-10, root=1e-12))


def test_full_size_f81_a64_properties(sb, eng):
    """F81 A = 64, 100-leaf tree, C = 1,000,000 (a 16.4 GB count output whose element offsets
    pass 2^31): dwell times sum to the active branch lengths of each column, root posteriors
    sum to 1, and 400 sampled columns (the last ones included) match the oracle."""
    say('full size F81 A=64 C=1e6')
    rng, t = _c3_tree(sb, 903)
    f = sb.f81_model(rng.dirichlet(np.full(64, 2.0)))
    C = 1_000_000
    aln = sb.synthetic.simulate(t, f, C, rng, gap_frac=0.03, ambig_frac=0.01)
    say('   simulated')
    aln_d = torch.from_numpy(aln).to(eng.device)
    cnt, ll, root = sb.Counts(aln_d, t, f, f81_fast_flag=True, engine=eng, return_loglike=True,
                              return_root=True)
    torch.cuda.synchronize()
    say('   counted')
    dwell = cnt.diagonal(dim1=0, dim2=1).double().sum(1).cpu().numpy()
    rsum = root.double().sum(0).cpu().numpy()
    cmin = float(cnt.min())
    mask = O.compute_branch_mask(aln, np.asarray(t.parentIndex), 64)
    mask[0] = False
    expect = (np.asarray(t.distanceToParent)[:, None] * mask).sum(0)
    e_dwell = float(np.max(np.abs(dwell - expect) / np.maximum(expect, 1e-3)))
    say(f'   dwell {e_dwell:.2e} root sum {np.abs(rsum - 1).max():.2e} min {cmin:.2e}')
    assert e_dwell <= 1e-4
    assert np.abs(rsum - 1).max() <= 1e-5
    assert cmin > -1e-4
    cols = np.concatenate([rng.choice(C - 64, 336, replace=False), np.arange(C - 64, C)])
    idx = torch.as_tensor(cols, device=eng.device)
    sub = aln[:, cols]
    rc, rl, rr = O.counts_loglike_root(sub, otree(t), omodel(f), f81_fast=True)
    assert err_counts(cnt.index_select(2, idx), rc) <= 1e-4
    assert err_ll(ll[idx], rl) <= 1e-5
    assert err_root(root.index_select(1, idx), rr) <= 1e-5
    del cnt


def test_autograd_functions(sb, eng):
    """torch.autograd.Function wrappers: branch lengths and log kappa."""
    say('autograd functions')
    t, m, aln = _hky(sb, 4000, seed=21, leaves=30, mean=0.1)
    d = torch.tensor(np.asarray(t.distanceToParent), dtype=torch.float64, requires_grad=True)
    ll = sb.LogLikeCustomGrad(aln, sb.Tree(t.parentIndex, d), m, engine=eng)
    g = torch.as_tensor(np.random.default_rng(22).normal(size=4000), device=ll.device)
    (ll * g).sum().backward()
    ref, _ = O.loglike_branch_grad(aln, otree(t), omodel(m), g.cpu().numpy())
    err = np.abs(d.grad.numpy() - ref).max() / np.abs(ref).max()
    say('   branch grad err', err)
    assert err <= 1e-4
    pi = np.asarray(m.pi)
    lk = torch.tensor(np.log(2.0), dtype=torch.float64, requires_grad=True)
    d2 = torch.tensor(np.asarray(t.distanceToParent), dtype=torch.float64, requires_grad=True)
    ll = sb.hky85_loglike(aln, sb.Tree(t.parentIndex, d2), lk, pi, engine=eng)
    ll.sum().backward()
    refk = O.hky85_loglike_kappa_grad(aln, otree(t), 2.0, pi)
    say('   log kappa grad', float(lk.grad), refk)
    assert abs(float(lk.grad) - refk) <= 1e-4 * abs(refk)
    ref2, _ = O.loglike_branch_grad(aln, otree(t), omodel(m))
    assert np.abs(d2.grad.numpy() - ref2).max() <= 1e-4 * np.abs(ref2).max()


def test_gradient_config_fd(sb, eng):
    """HKY85, 127 nodes, C = 50,000: gradient of sum_c logL_c w.r.t. the 126 branch lengths
    and log kappa, stated error <= 1e-4 (relative).

    References:
      * the oracle's FP64 adjoint on the whole block (per-column derivatives x_nc, their sum
        and the cancellation ratio k_n = sum_c |x_nc| / |sum_c x_nc| of every branch);
      * central finite differences with h = 1e-4, Richardson-extrapolated with h/2
        ((4 D(h/2) - D(h)) / 3, truncation O(h^4)), of the GPU's FP64 LogLike on the whole
        block and of the oracle's FP64 LogLike on a 2,000-column subsample.
    Bounds: FP64 adjoint per branch <= 1e-10 vs the oracle, <= 1e-4 vs FD.  FP32 adjoint per
    branch relative <= 1e-4 vs the oracle's FP64 adjoint on every branch (cancellation ratios
    here reach 3.6e3; measured worst 4.8e-5 on that branch, median 9e-7).  log kappa FP32
    <= 1e-4 relative.  Two FP32 calls give bitwise equal results (fixed reduction order)."""
    say('gradient config C=50000')
    t, m, aln = _hky(sb, 50000, seed=23)
    R = len(t.parentIndex)
    dev = eng.device
    aln_d = torch.from_numpy(aln).to(dev)
    pi = np.asarray(m.pi)
    W = sb.hky85_log_kappa_weights(2.0, pi)
    d64, r64, _ = sb.loglike_grad(aln_d, t, m, rate_weights=W, precision='fp64', engine=eng)
    d32, r32, _ = sb.loglike_grad(aln_d, t, m, rate_weights=W, precision='fp32', engine=eng)
    d32b, r32b, _ = sb.loglike_grad(aln_d, t, m, rate_weights=W, precision='fp32', engine=eng)
    assert torch.equal(d32, d32b) and torch.equal(r32, r32b)
    d64, d32 = d64.cpu().numpy(), d32.cpu().numpy()

    per, _ = O.loglike_branch_grad_columns(aln, otree(t), omodel(m))
    ref = per.sum(1)
    mag = np.abs(per).sum(1)
    kap = mag[1:] / np.maximum(np.abs(ref[1:]), 1e-300)
    e64 = np.abs(d64[1:] - ref[1:]) / np.abs(ref[1:])
    e32 = np.abs(d32[1:] - ref[1:]) / np.abs(ref[1:])
    say(f'   oracle adjoint: fp64 max rel {e64.max():.2e}; fp32 max rel {e32.max():.2e}, '
        f'median {np.median(e32):.2e}; cancellation ratio median {np.median(kap):.1f} max {kap.max():.1f}')
    say('   fp32 rel / cancellation ratio, worst 5 branches:',
        [(int(n) + 1, f'{e32[n]:.2e}', f'{kap[n]:.0f}') for n in np.argsort(e32)[-5:]])
    assert e64.max() <= 1e-10
    assert e32.max() <= 1e-4

    def total(tree, model):
        return sb.LogLike(aln_d, tree, model, precision='fp64', engine=eng).sum().item()

    def fd(f, n, h):
        tp = np.array(t.distanceToParent, dtype=np.float64)
        tm = tp.copy()
        tp[n] += h
        tm[n] -= h
        return (f(sb.Tree(t.parentIndex, tp)) - f(sb.Tree(t.parentIndex, tm))) / (2 * h)

    h = 1e-4
    for n in np.linspace(1, R - 1, 8).astype(int):
        g = lambda tr: total(tr, m)  # noqa: E731
        rich = (4 * fd(g, n, h / 2) - fd(g, n, h)) / 3
        assert abs(d64[n] - rich) <= 1e-4 * abs(rich), (n, d64[n], rich)

    def fd_k(h, f):
        return (f(sb.hky85_diag(2.0 * np.exp(h), pi)) - f(sb.hky85_diag(2.0 * np.exp(-h), pi))) / (2 * h)

    richk = (4 * fd_k(h / 2, lambda mm: total(t, mm)) - fd_k(h, lambda mm: total(t, mm))) / 3
    say('   log kappa: fp64', float(r64[0]), 'fp32', float(r32[0]), 'FD', richk)
    assert abs(float(r64[0]) - richk) <= 1e-4 * abs(richk)
    assert abs(float(r32[0]) - float(r64[0])) <= 1e-4 * abs(float(r64[0]))

    # the oracle's own FP64 LogLike on a column subsample, differentiated numerically
    cols = np.sort(np.random.default_rng(24).choice(aln.shape[1], 2000, replace=False))
    sub = np.ascontiguousarray(aln[:, cols])
    s64, sr64, _ = sb.loglike_grad(sub, t, m, rate_weights=W, precision='fp64', engine=eng)
    s32, sr32, _ = sb.loglike_grad(sub, t, m, rate_weights=W, precision='fp32', engine=eng)
    s64, s32 = s64.cpu().numpy(), s32.cpu().numpy()
    otot = lambda tr, mm=m: float(O.LogLike(sub, otree(tr), omodel(mm)).sum())  # noqa: E731
    for n in np.linspace(1, R - 1, 6).astype(int):
        rich = (4 * fd(otot, n, h / 2) - fd(otot, n, h)) / 3
        say(f'   subsample branch {n}: FD {rich:.6e} fp64 {s64[n]:.6e} fp32 {s32[n]:.6e}')
        assert abs(s64[n] - rich) <= 1e-4 * abs(rich)
        assert abs(s32[n] - rich) <= 1e-4 * abs(rich)
    ok = lambda mm: float(O.LogLike(sub, otree(t), omodel(mm)).sum())  # noqa: E731
    richk = (4 * fd_k(h / 2, ok) - fd_k(h, ok)) / 3
    assert abs(float(sr64[0]) - richk) <= 1e-4 * abs(richk)
    assert abs(float(sr32[0]) - richk) <= 1e-4 * abs(richk)


def test_models_uploaded_into_reused_records(sb):
    """Many new models whose arena records are scattered over reused ones go up
    in one staged copy plus a scatter; every model must arrive intact (results
    bitwise those of a fresh engine, a sample against the oracle)."""
    from subby_amd.batch import Launcher, Block
    say('model upload into scattered reused records')
    rng = np.random.default_rng(21)
    t = sb.synthetic.random_tree(20, rng, 0.1)
    base = sb.hky85_diag(2.0, np.full(4, 0.25))
    aln = sb.synthetic.simulate(t, base, 300, rng, gap_frac=0.03)

    def models(n):
        return [sb.hky85_diag(float(rng.uniform(1, 5)), rng.dirichlet(np.full(4, 5.0))) for _ in range(n)]

    def run(e, ms):
        bl = [Block(aln, t, m) for m in ms]
        L = Launcher(e)
        L.submit(bl)
        L.sync()
        return bl

    eng = sb.Engine(0, max_models=24)
    first = models(24)
    run(eng, first)
    run(eng, first[1::2])            # the odd ones become the recently used
    fresh = models(12)               # evicts the even ones: 12 scattered records
    got = run(eng, fresh)
    r0, n0, q0 = eng.arena_stats()
    ref_eng = sb.Engine(0, max_models=None)
    ref = run(ref_eng, fresh)
    for a, b in zip(got, ref):
        assert torch.equal(a.counts, b.counts) and torch.equal(a.loglike, b.loglike)
        assert torch.equal(a.root, b.root)
    for b in got[:3]:
        rc, rl, _ = O.counts_loglike_root(b.alignment, otree(b.tree), omodel(b.model))
        assert err_counts(b.counts, rc) <= 1e-4 and err_ll(b.loglike, rl) <= 1e-5
    # the new models took over the dropped records: the arena did not grow
    run(eng, models(12))
    assert eng.arena_stats()[0] == r0
    ref_eng.close()
    eng.close()


@pytest.mark.parametrize('A', [20, 64])
def test_f81_partial_tiles(sb, eng, A):
    """Blocks narrower than one 16-column group tile of the F81 matrix-core kernel, and
    ragged tails: waves whose columns all lie past the block's last column read the
    clamped last column through offsets from the tile's first column (a regression test
    for offsets taken from a wave's own first column, which went negative there)."""
    say('F81 partial tiles, A =', A)
    rng = np.random.default_rng(780 + A)
    f = sb.f81_model(rng.dirichlet(np.full(A, 2.0)))
    t = sb.synthetic.random_tree(9, rng, 0.1)
    for C in (1, 2, 3, 6, 15, 17, 33):
        aln = sb.synthetic.simulate(t, f, C, rng, gap_frac=0.1, ambig_frac=0.05)
        for mask in ('auto', None):
            check_all(sb, eng, aln, t, f, 'fp32', mask=mask, f81=True)
        mk = rng.random(aln.shape) < 0.5
        ref = O.Counts(aln, otree(t), omodel(f), f81_fast=True, branch_mask=mk)
        got = sb.Counts(aln, t, f, f81_fast_flag=True, branch_mask=mk, precision='fp32', engine=eng)
        assert err_counts(got, ref) <= 1e-4
