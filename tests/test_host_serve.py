"""Host side of the serving benchmark: the synthetic request stream and the
latency breakdown (no GPU)."""
import numpy as np

from subby_amd import _lib
from subby_amd.serve import breakdown, make_stream


def test_stream_distinct_and_independent_of_workers():
    a = make_stream(40, qps=1000.0, seed=7, leaves=(4, 30), mean_cols=60)
    b = make_stream(40, qps=1000.0, seed=7, leaves=(4, 30), mean_cols=60, workers=2)
    assert len(a) == len(b) == 40
    for x, y in zip(a, b):
        assert x.rid == y.rid and x.arrival == y.arrival
        assert np.array_equal(x.alignment, y.alignment)
        assert np.array_equal(x.tree.parentIndex, y.tree.parentIndex)
        assert np.array_equal(x.model.pi, y.model.pi)
    assert len({x.alignment.tobytes() for x in a}) == 40
    arr = np.array([x.arrival for x in a])
    assert np.all(np.diff(arr) > 0)
    for x in a:
        R, C = x.alignment.shape
        assert 7 <= R <= 59 and 20 <= C <= 10000 and R == len(x.tree.parentIndex)


def test_stream_pool_draws_shapes():
    a = make_stream(30, qps=1e9, seed=8, leaves=(4, 10), mean_cols=40, pool=3)
    assert len({x.alignment.tobytes() for x in a}) <= 3
    assert len({x.model.pi.tobytes() for x in a}) == 30


def test_breakdown_spans_add_up():
    n = 200
    rng = np.random.default_rng(1)
    t = np.zeros(n, dtype=_lib.TIMING_DTYPE)
    arrival = np.sort(rng.uniform(0, 1, n))
    t['admitted_s'] = arrival + rng.uniform(0, 1e-3, n)
    t['plan_ms'] = rng.uniform(0, 0.2, n)
    t['models_ms'] = rng.uniform(0, 0.1, n)
    t['enqueued_s'] = t['admitted_s'] + (t['plan_ms'] + t['models_ms']) * 1e-3 + rng.uniform(0, 1e-4, n)
    t['gpu_done_s'] = t['enqueued_s'] + rng.uniform(1e-4, 2e-3, n)
    t['done_s'] = t['gpu_done_s'] + rng.uniform(0, 1e-4, n)
    t['batch_requests'] = rng.integers(1, 50, n)
    t[-1]['gpu_done_s'] += 0.5          # one slow request
    t[-1]['done_s'] += 0.5
    bd = breakdown(t, arrival)
    for band in ('all', 'p45_p55', 'slowest_1pct'):
        b = bd[band]
        spans = b['queue_ms'] + b['models_ms'] + b['pack_ms'] + b['plan_ms'] + b['gpu_ms'] + b['unpack_ms']
        assert b['pack_ms'] >= 0
        assert abs(spans - b['latency_ms']) < 1e-9 * max(1, b['latency_ms'])
    lat = (t['done_s'] - arrival) * 1e3
    assert abs(bd['all']['latency_ms'] - lat.mean()) < 1e-9
    assert bd['slowest_1pct']['gpu_ms'] > 200      # the slow request is in the tail band
    assert bd['p45_p55']['latency_ms'] < bd['slowest_1pct']['latency_ms']
