"""Launch-shape and tile-shape knobs: every candidate of the autotuner gives
bitwise the same results (only the time may differ), and tune() times the whole
search space of the batch's kernel family and applies the fastest."""
import numpy as np
import pytest
import torch

from tests.test_gpu_parity import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def eng():
    import subby_amd
    return subby_amd.Engine(0)


def _blocks(sb, A, C, seed, leaves):
    rng = np.random.default_rng(seed)
    pi = rng.dirichlet(np.full(A, 3.0))
    if A == 4:
        m = sb.hky85_diag(2.0, pi)
    else:
        m = sb.diagonalize_rate_matrix(sb.reversible_rate_matrix(rng.lognormal(0, 1, (A, A)), pi), pi)
    t = sb.synthetic.random_tree(leaves, rng, 0.1)
    aln = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.03, ambig_frac=0.01)
    return t, m, aln


@pytest.mark.parametrize('A,C,leaves', [(4, 3000, 64), (20, 700, 30)])
def test_every_candidate_gives_the_same_bits(eng, A, C, leaves):
    import subby_amd as sb
    from subby_amd.autotune import apply, candidates, family, tune
    from subby_amd.batch import Block, Launcher
    say('tuning knobs, A =', A)
    t, m, aln = _blocks(sb, A, C, 300 + A, leaves)
    aln_d = torch.from_numpy(aln).to(eng.device)
    L = Launcher(eng)
    fam = family([Block(aln_d, t, m)])
    assert fam == ('small' if A == 4 else 'gen2')
    cands = candidates(fam)
    if fam == 'small':
        cands = cands[::5]
    cands = cands + [dict(gen2_round=1)] if fam == 'gen2' else cands
    ref = None
    for cand in cands:
        apply(L, cand)
        blk = Block(aln_d, t, m)
        L.submit([blk])
        L.sync()
        got = (blk.counts.clone(), blk.loglike.clone(), blk.root.clone())
        if ref is None:
            ref = got
        else:
            for a, b in zip(ref, got):
                assert torch.equal(a, b), cand
    res = tune(L, [Block(aln_d, t, m)], reps=1)
    assert res['family'] == fam and len(res['table']) == len(candidates(fam, res['split']))
    assert res['split'] == (fam == 'small')   # (the 64-leaf tree gives a second wave a subtree)
    assert res['ms'] == min(r['ms'] for r in res['table'])
    say('   best', res['best'], f"{res['ms']:.3f} ms, slowest {res['worst_ms']:.3f} ms")
    blk = Block(aln_d, t, m)
    L.submit([blk])
    L.sync()
    assert torch.equal(blk.counts, ref[0])


@pytest.mark.parametrize('A,C', [(64, 20000), (20, 36000)])
def test_two_pass_tail_overlap(eng, A, C):
    """More pass-1 tiles than the CUs hold at once (A = 64: 313 tiles of 64 columns on 256
    CUs; A = 20: 563 tiles on 512 slots): the last pass-1 workgroups run beside pass 2 of
    the first tiles on a second stream.  The results equal, bitwise, the same call in
    rounds too small to split (gen2_round below one CU-wave of tiles), and the oracle on
    columns taken from the first tiles and from the tail."""
    import subby_amd as sb
    from oracle import subby_oracle as O
    from subby_amd.autotune import apply
    from subby_amd.batch import Block, Launcher
    say('two-pass tail overlap, A =', A, 'C =', C)
    t, m, aln = _blocks(sb, A, C, 910 + A, 7)
    aln_d = torch.from_numpy(aln).to(eng.device)
    L = Launcher(eng)
    out = []
    for rnd in (0, 100):
        apply(L, dict(gen2_round=rnd))
        blk = Block(aln_d, t, m)
        L.submit([blk])
        L.sync()
        out.append((blk.counts.clone(), blk.loglike.clone(), blk.root.clone()))
    for a, b in zip(*out):
        assert torch.equal(a, b)
    cols = np.r_[0:40, 16380:16420, C - 60:C]
    ot = O.Tree(np.asarray(t.parentIndex), np.asarray(t.distanceToParent))
    om = O.DiagModel(*(np.asarray(x, dtype=np.float64) for x in m))
    rc, rl, _ = O.counts_loglike_root(aln[:, cols], ot, om)
    got = out[0][0][:, :, torch.from_numpy(cols).to(eng.device)].double().cpu().numpy()
    num = np.abs(got - rc).max(axis=(0, 1))
    den = np.abs(rc).max(axis=(0, 1))
    assert np.max(num / den) <= 1e-4
    ll = out[0][1].cpu().numpy()[cols]
    assert np.max(np.abs(ll - rl) / np.maximum(np.abs(rl), 1)) <= 1e-5


def test_repeated_and_changed_plans(eng):
    """A launcher that skips re-sending a plan equal to the one the device holds:
    the same prepared batch launched again (direct and replayed) gives the same
    bits, and a batch that differs only in its branch lengths or its output
    buffers (a different plan of the same size) gets its own results, against
    the oracle; two launchers alternate on one stream."""
    import subby_amd as sb
    from oracle import subby_oracle as O
    from subby_amd.batch import Block, Launcher
    from tests.test_gpu_parity import err_counts, err_ll, omodel, otree
    say('repeated and changed plans')
    t, m, aln = _blocks(sb, 4, 2000, 911, 40)
    aln_d = torch.from_numpy(aln).to(eng.device)
    L = Launcher(eng)
    blk = Block(aln_d, t, m)
    prep = L.prepare([blk])
    L.launch(prep, use_graph=False)
    L.sync()
    ref = (blk.counts.clone(), blk.loglike.clone())
    for g in (False, True, False, True):
        blk.counts.zero_()
        blk.loglike.zero_()
        torch.cuda.synchronize()
        L.launch(prep, use_graph=g)
        L.sync()
        assert torch.equal(blk.counts, ref[0]) and torch.equal(blk.loglike, ref[1]), g
    oc, ol, _ = O.counts_loglike_root(aln, otree(t), omodel(m))
    assert err_counts(ref[0], oc) <= 1e-4 and err_ll(ref[1], ol) <= 1e-5
    # same tree and sizes, other branch lengths: the plan differs only in its jobs
    t2 = sb.Tree(np.asarray(t.parentIndex), np.asarray(t.distanceToParent) * 1.7)
    blk2 = Block(aln_d, t2, m)
    L2 = Launcher(eng)
    for _ in range(2):
        L.submit([blk2])
        L2.submit([blk])
        L.sync()
        L2.sync()
        oc2, ol2, _ = O.counts_loglike_root(aln, otree(t2), omodel(m))
        assert err_counts(blk2.counts, oc2) <= 1e-4 and err_ll(blk2.loglike, ol2) <= 1e-5
        assert torch.equal(blk.counts, ref[0]) and torch.equal(blk.loglike, ref[1])
        L.launch(prep, use_graph=True)
        L.sync()
        assert torch.equal(blk.counts, ref[0])
