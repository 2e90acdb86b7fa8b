"""Serving loop: continuous batching over several lanes, HBM-budget admission.

Each served result must equal the same request computed alone through the
single-block API (bitwise: columns are computed independently of the batch),
and a sample is checked against the FP64 oracle at the FP32 tolerances.
"""
import sys

import numpy as np
import pytest
import torch

from oracle import subby_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope='module')
def eng():
    import subby_amd
    return subby_amd.Engine(0)


def _check_alone(sb, eng, reqs, results, oracle_every=25):
    for r, res in zip(reqs, results):
        assert res.rid == r.rid
        cnt, ll, root = sb.Counts(r.alignment, r.tree, r.model, engine=eng, return_loglike=True, return_root=True)
        torch.cuda.synchronize()
        assert np.array_equal(cnt.cpu().numpy(), res.counts), r.rid
        assert np.array_equal(ll.cpu().numpy(), res.loglike), r.rid
        assert np.array_equal(root.cpu().numpy(), res.root), r.rid
        if r.rid % oracle_every == 0:
            rc, rl, rr = O.counts_loglike_root(r.alignment, O.Tree(*r.tree), O.DiagModel(*r.model))
            e_c = np.max(np.abs(res.counts - rc).max((0, 1)) / np.maximum(np.abs(rc).max((0, 1)), 1e-300))
            e_l = np.max(np.abs(res.loglike - rl) / np.maximum(np.abs(rl), 1))
            assert e_c <= 1e-4 and e_l <= 1e-5 and np.abs(res.root - rr).max() <= 1e-5


def test_served_equals_alone(eng):
    import subby_amd as sb
    from subby_amd.serve import Server, make_stream
    say('serve: 400 requests, Poisson arrivals')
    reqs = make_stream(400, qps=4000.0, seed=11, leaves=(4, 60), mean_cols=300)
    srv = Server(eng, hbm_budget=1 << 30, lanes=3)
    results, stats = srv.run(reqs)
    say('   stats', stats)
    assert stats['requests'] == 400 and stats['batches'] >= 2
    _check_alone(sb, eng, reqs, results)


def test_tight_budget_and_burst(eng):
    """A burst with a budget that admits only a few requests at a time."""
    import subby_amd as sb
    from subby_amd.serve import Server, make_stream
    say('serve: tight budget burst')
    reqs = make_stream(120, qps=1e9, seed=12, leaves=(4, 40), mean_cols=200)
    srv = Server(eng, hbm_budget=200_000, lanes=2, max_blocks=16)
    results, stats = srv.run(reqs, realtime=False)
    say('   stats', stats)
    assert stats['batches'] >= 120 // 16
    _check_alone(sb, eng, reqs, results, oracle_every=40)
