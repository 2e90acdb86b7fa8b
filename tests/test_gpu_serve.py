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
    reqs = make_stream(400, qps=4000.0, seed=11, leaves=(4, 100), mean_cols=300)
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


def test_native_served_equals_alone(eng):
    """The native loop (sb_server_*), Poisson arrivals replayed from native code."""
    import subby_amd as sb
    from subby_amd.serve import NativeServer, make_stream
    say('native serve: 400 requests, Poisson arrivals')
    reqs = make_stream(400, qps=4000.0, seed=13, leaves=(4, 100), mean_cols=300)
    srv = NativeServer(eng, hbm_budget=1 << 30, lanes=3)
    try:
        results, stats = srv.run(reqs)
    finally:
        srv.close()
    say('   stats', stats)
    assert stats['requests'] == 400
    assert all(res.done >= res.arrival for res in results)
    _check_alone(sb, eng, reqs, results)


def test_native_timed_replay(eng):
    """sb_server_replay_timed: same results as the untimed replay (bitwise), and
    a consistent per-request breakdown (ordered host times, GPU spans, batch
    sizes) whose spans add up to the latency."""
    from subby_amd.serve import NativeServer, make_stream
    say('native serve: timed replay, 300 requests')
    reqs = make_stream(300, qps=3000.0, seed=15, leaves=(4, 100), mean_cols=300)
    srv = NativeServer(eng, hbm_budget=1 << 30, lanes=3)
    try:
        plain, _ = srv.run(reqs)
        timed, stats = srv.run(reqs, timed=True)
    finally:
        srv.close()
    for a, b in zip(plain, timed):
        assert np.array_equal(a.counts, b.counts) and np.array_equal(a.loglike, b.loglike)
        assert np.array_equal(a.root, b.root)
    t = stats['timing']
    arrival = np.array([r.arrival for r in reqs])
    eps = 1e-6
    assert np.all(t['admitted_s'] >= arrival - eps)
    assert np.all(t['enqueued_s'] >= t['admitted_s'] - eps)
    assert np.all(t['gpu_done_s'] >= t['enqueued_s'] - eps)
    assert np.all(t['done_s'] >= t['gpu_done_s'] - eps)
    for k in ('upload_ms', 'kernel_ms', 'download_ms', 'plan_ms', 'models_ms'):
        assert np.all(t[k] >= 0) and np.all(np.isfinite(t[k])), k
    assert np.all(t['kernel_ms'] > 0)
    assert np.all((t['lane'] >= 0) & (t['lane'] < 3))
    assert np.all(t['batch_requests'] >= 1)
    # requests of one batch share its record; the batches' sizes add up
    key = np.stack([t['lane'], np.round(t['enqueued_s'] * 1e7).astype(np.int64)], 1)
    _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    assert np.array_equal(cnt[inv.ravel()], t['batch_requests'])
    bd = stats['breakdown']['all']
    spans = bd['queue_ms'] + bd['models_ms'] + bd['pack_ms'] + bd['plan_ms'] + bd['gpu_ms'] + bd['unpack_ms']
    assert abs(spans - bd['latency_ms']) <= 1e-6 * max(1.0, bd['latency_ms'])
    assert abs(bd['latency_ms'] - stats['mean_ms']) <= 1e-6 * max(1.0, stats['mean_ms'])
    say('   breakdown', stats['breakdown'])


def test_native_tight_budget_burst_and_poll(eng):
    """Burst under a budget that admits few requests at a time; then the
    submit / drain / poll interface with caller tags."""
    import ctypes
    import subby_amd as sb
    from subby_amd import _lib
    from subby_amd.serve import NativeServer, make_stream
    say('native serve: tight budget burst, submit/poll')
    reqs = make_stream(120, qps=1e9, seed=14, leaves=(4, 40), mean_cols=200)
    srv = NativeServer(eng, hbm_budget=200_000, lanes=2, max_blocks=16)
    try:
        results, stats = srv.run(reqs, realtime=False)
        say('   stats', stats)
        _check_alone(sb, eng, reqs, results, oracle_every=40)
        sub = reqs[:10]
        arr, keep, outs = srv.prepare(sub)
        for k in range(len(sub)):
            arr[k].tag = 1000 + k
        _lib.check(srv.lib.sb_server_submit(srv.h, len(sub), ctypes.cast(arr, ctypes.c_void_p)))
        _lib.check(srv.lib.sb_server_drain(srv.h))
        tags = np.zeros(64, dtype=np.int64)
        times = np.zeros(64, dtype=np.float64)
        nd = ctypes.c_int()
        _lib.check(srv.lib.sb_server_poll(srv.h, 64, tags.ctypes.data, times.ctypes.data, ctypes.byref(nd)))
        assert sorted(tags[:nd.value].tolist()) == list(range(1000, 1010))
        for (cnt, ll, root), res in zip(outs, results[:10]):
            assert np.array_equal(cnt, res.counts) and np.array_equal(ll, res.loglike)
        # a malformed request is refused at submission
        bad = (_lib.SbRequest * 1)()
        bad[0].model_id = -1
        bad[0].A = 70
        with pytest.raises(sb.SubbyError):
            _lib.check(srv.lib.sb_server_submit(srv.h, 1, ctypes.cast(bad, ctypes.c_void_p)))
    finally:
        srv.close()


def _mixed_stream(sb, n, seed, irreversible):
    """HKY85 stream with every third request on a general A = 20 model and (if
    asked) every fifth on an irreversible A = 4 model."""
    from subby_amd.serve import make_stream
    reqs = make_stream(n, qps=1e9, seed=seed, leaves=(4, 30), mean_cols=150)
    rng = np.random.default_rng(seed + 1)
    pi20 = rng.dirichlet(np.full(20, 3.0))
    m20 = sb.diagonalize_rate_matrix(sb.reversible_rate_matrix(rng.lognormal(0, 1, (20, 20)), pi20), pi20)
    Q = rng.uniform(0.05, 1.0, (4, 4))
    for a in range(4):
        Q[a, (a + 1) % 4] += 1.5
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(1))
    mirr = sb.diagonalize_irreversible(Q, rng.dirichlet(np.full(4, 3.0)))
    for r in reqs:
        if r.rid % 3 == 1:
            r.model = m20
        elif irreversible and r.rid % 5 == 2:
            r.model = mirr
        else:
            continue
        r.alignment = sb.synthetic.simulate(r.tree, r.model, r.alignment.shape[1], rng, gap_frac=0.03)
    return reqs


def _check_alone_any(sb, eng, reqs, results):
    for r, res in zip(reqs, results):
        assert res.rid == r.rid
        cnt, ll, root = sb.Counts(r.alignment, r.tree, r.model, engine=eng, return_loglike=True, return_root=True)
        torch.cuda.synchronize()
        assert np.array_equal(cnt.cpu().numpy(), res.counts), r.rid
        assert np.array_equal(ll.cpu().numpy(), res.loglike), r.rid
        assert np.array_equal(root.cpu().numpy(), res.root), r.rid


def test_served_mixed_alphabets_and_models(eng):
    """Requests over A = 4 and A = 20 and reversible / irreversible models in
    one stream: each launch takes one kind (padded alphabet, reversibility);
    every result equals the request computed alone."""
    import subby_amd as sb
    from subby_amd.serve import Server
    say('serve: mixed alphabets and irreversible models')
    reqs = _mixed_stream(sb, 90, 21, irreversible=True)
    srv = Server(eng, hbm_budget=1 << 30, lanes=2, max_blocks=16)
    results, stats = srv.run(reqs, realtime=False)
    say('   stats', stats)
    assert stats['requests'] == 90
    _check_alone_any(sb, eng, reqs, results)


def test_native_served_mixed_alphabets(eng):
    """The native loop with A = 4 and A = 20 requests interleaved, and
    irreversible models as resident model ids."""
    import subby_amd as sb
    from subby_amd.serve import NativeServer
    say('native serve: mixed alphabets and irreversible models')
    reqs = _mixed_stream(sb, 90, 22, irreversible=True)
    srv = NativeServer(eng, hbm_budget=1 << 30, lanes=2, max_blocks=16)
    try:
        results, stats = srv.run(reqs, realtime=False)
    finally:
        srv.close()
    say('   stats', stats)
    assert stats['requests'] == 90
    _check_alone_any(sb, eng, reqs, results)


def test_native_arena_flat_with_distinct_models(eng):
    """A stream whose every request carries its own inline model: the server keeps at most
    4096 of them resident and reuses the arena records of evicted ones, so a second stream
    of distinct models does not grow the model arena; equal models share one record."""
    import subby_amd as sb
    from subby_amd.serve import NativeServer, make_stream
    say('native serve: arena over distinct inline models')
    srv = NativeServer(eng, hbm_budget=1 << 30, lanes=3, max_blocks=256)
    try:
        _, n0, _ = eng.arena_stats()         # models of the engine itself (other calls)
        reqs = make_stream(6000, qps=1e9, seed=31, leaves=(4, 12), mean_cols=40, pool=40)
        res1, st1 = srv.run(reqs, realtime=False)
        r1, n1, q1 = eng.arena_stats()
        reqs2 = make_stream(6000, qps=1e9, seed=32, leaves=(4, 12), mean_cols=40, pool=40)
        res2, st2 = srv.run(reqs2, realtime=False)
        r2, n2, q2 = eng.arena_stats()
        say(f'   arena after 6000: {r1} B, {n1} live, {q1} retired; after 12000: {r2} B, {n2} live, {q2} retired')
        rec = 1536   # one A = 4 model record (FP64 master + float + double records)
        assert n2 - n0 <= 4096 + 3 * 256
        assert r2 - r1 <= 3 * 256 * rec
        # equal models share one record
        same = make_stream(300, qps=1e9, seed=33, leaves=(4, 12), mean_cols=40, pool=10)
        for r in same:
            r.model = same[0].model
        res3, _ = srv.run(same, realtime=False)
        r3, n3, _ = eng.arena_stats()
        assert r3 <= r2 + rec and n3 <= n2 + 1
    finally:
        srv.close()
    _check_alone(sb, eng, reqs[:100], res1[:100], oracle_every=50)
    _check_alone(sb, eng, reqs2[-100:], res2[-100:], oracle_every=50)
    _check_alone(sb, eng, same[:20], res3[:20], oracle_every=10)


def test_engine_model_cache_bounded():
    """An Engine keeps at most max_models eigensystems; evicted records are reused, and a
    call that needs more distinct models than that (a model list) still runs."""
    import subby_amd as sb
    say('engine model cache bound')
    e = sb.Engine(0, max_models=32)
    try:
        rng = np.random.default_rng(41)
        t = sb.synthetic.random_tree(10, rng, 0.1)
        aln = sb.synthetic.simulate(t, sb.hky85_diag(2.0, np.full(4, 0.25)), 64, rng)
        models = [sb.hky85_diag(float(rng.uniform(1, 5)), rng.dirichlet(np.full(4, 5.0))) for _ in range(300)]
        for k, m in enumerate(models):
            ll = sb.LogLike(aln, t, m, engine=e)
            if k % 97 == 0:
                assert err_ll_np(ll, O.LogLike(aln, O.Tree(*t), O.DiagModel(*m))) <= 1e-5
        r, n, q = e.arena_stats()
        say(f'   arena {r} B, {n} live, {q} retired')
        assert n <= 32 and r <= 64 * 1536
        lst = models[:64]
        ll = sb.LogLike(aln, t, lst, engine=e)       # 64 per-column models > max_models
        ref = O.LogLike_columns(aln, O.Tree(*t), [O.DiagModel(*m) for m in lst])
        assert err_ll_np(ll, ref) <= 1e-5
    finally:
        e.close()


def err_ll_np(got, ref):
    got = got.detach().cpu().numpy()
    return float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)))
