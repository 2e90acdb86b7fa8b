"""Continuous-batching server for (tree, alignment-block) requests.

Requests arrive over time (Poisson in the benchmark).  The loop keeps several
launcher lanes, each a persistent-kernel launcher on its own HIP stream with
its own pinned staging and device buffers.  Whenever a lane is free, the
requests that have arrived are admitted into it, first come first served,
while the in-flight bytes (alignments + outputs of every batch still on the
GPU) stay within the HBM budget; the batch is uploaded with one copy, run as
one persistent launch (replayed hipGraph), and its outputs come back with one
copy.  Batches on different lanes overlap on the device, and new requests are
admitted as soon as a lane and memory free up; nothing waits for a full batch.

Models stay resident: each distinct eigensystem is uploaded once into the
engine's model arena.  Every column is computed independently of what else is
in the batch, so a request's result is bitwise the same served alone or in
any mix.

The reference's batched driver (scripts/batch_eigensub_maf.py:423-698) buckets
blocks by tree size, pads columns to 128 and sizes jax.vmap batches by a 4 GiB
budget; here no padding or bucketing is needed.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .api import Engine, default_engine, _PREC
from .models import IrrevDiagModel, ensure_diag

_BLOCK_DTYPE = np.dtype({
    'names': ['model_id', 'R', 'C', 'aln', 'parent', 'dist', 'want', 'f81', 'mask_mode',
              'mask', 'counts', 'root', 'loglike'],
    'formats': ['<i4', '<i4', '<i4', '<u8', '<u8', '<u8', '<i4', '<i4', '<i4', '<u8', '<u8', '<u8', '<u8'],
    'offsets': [0, 4, 8, 16, 24, 32, 40, 44, 48, 56, 64, 72, 80],
    'itemsize': 88})
assert _BLOCK_DTYPE.itemsize == ctypes.sizeof(_lib.SbBlock)


def _align(x, a=256):
    return (x + a - 1) // a * a


@dataclass
class Request:
    alignment: np.ndarray          # (R, C) int8 tokens
    tree: object                   # Tree(parentIndex, distanceToParent)
    model: object                  # DiagModel / RateModel
    arrival: float = 0.0           # seconds after the start of the run
    rid: int = 0
    branch_mask: str | None = 'auto'


@dataclass
class Result:
    rid: int
    counts: np.ndarray
    loglike: np.ndarray
    root: np.ndarray
    arrival: float
    done: float

    @property
    def latency(self):
        return self.done - self.arrival


@dataclass
class _Lane:
    launcher: object
    stream: torch.cuda.Stream
    host_in: torch.Tensor = None
    dev_in: torch.Tensor = None
    dev_out: torch.Tensor = None
    host_out: torch.Tensor = None
    event: torch.cuda.Event = None
    batch: list = field(default_factory=list)
    bytes: int = 0
    busy: bool = False


class Server:
    def __init__(self, engine: Engine | None = None, hbm_budget: int = 8 << 30, lanes: int = 3,
                 max_blocks: int = 2048, max_cols: int = 1 << 19, precision: str = 'fp32',
                 use_graph: bool = True):
        from .batch import Launcher
        self.eng = engine or default_engine()
        self.dev = self.eng.device
        self.budget = int(hbm_budget)
        self.max_blocks = max_blocks
        self.max_cols = max_cols
        self.precision = precision
        self.dt, self.tdt = _PREC[precision]
        self.esz = 4 if precision == 'fp32' else 8
        self.use_graph = use_graph
        self.lanes = []
        for _ in range(lanes):
            st = torch.cuda.Stream(self.dev)
            self.lanes.append(_Lane(launcher=Launcher(self.eng, stream=st), stream=st))
        self.in_flight = 0

    # ---- per-request footprint
    def _bytes(self, r: Request):
        R, C = r.alignment.shape
        A = len(r.model.pi)
        return _align(R * C), _align(A * A * C * self.esz) + _align(C * 8) + _align(A * C * self.esz)

    def _ensure(self, lane, nin, nout):
        if lane.host_in is None or lane.host_in.numel() < nin:
            n = max(nin, 1 << 22)
            lane.host_in = torch.empty(n, dtype=torch.uint8, pin_memory=True)
            lane.dev_in = torch.empty(n, dtype=torch.uint8, device=self.dev)
        if lane.dev_out is None or lane.dev_out.numel() < nout:
            n = max(nout, 1 << 22)
            lane.dev_out = torch.empty(n, dtype=torch.uint8, device=self.dev)
            lane.host_out = torch.empty(n, dtype=torch.uint8, pin_memory=True)

    def _launch(self, lane, reqs):
        self.eng.new_call()
        sizes = [self._bytes(r) for r in reqs]
        nin = sum(s[0] for s in sizes)
        nout = sum(s[1] for s in sizes)
        self._ensure(lane, nin, nout)
        hin = lane.host_in.numpy()
        blocks = np.zeros(len(reqs), dtype=_BLOCK_DTYPE)
        keep = []
        din = lane.dev_in.data_ptr()
        dout = lane.dev_out.data_ptr()
        oi = oo = 0
        layout = []
        for k, (r, (bi, bo)) in enumerate(zip(reqs, sizes)):
            R, C = r.alignment.shape
            m = ensure_diag(r.model)
            A = len(m.pi)
            hin[oi:oi + R * C] = r.alignment.reshape(-1).view(np.uint8)
            parent = np.ascontiguousarray(r.tree.parentIndex, dtype=np.int32)
            dist = np.ascontiguousarray(r.tree.distanceToParent, dtype=np.float64)
            keep += [parent, dist]
            oc = oo
            ol = oc + _align(A * A * C * self.esz)
            orr = ol + _align(C * 8)
            b = blocks[k]
            b['model_id'] = self.eng.model_id(m)
            b['R'], b['C'] = R, C
            b['aln'] = din + oi
            b['parent'] = parent.ctypes.data
            b['dist'] = dist.ctypes.data
            b['want'] = _lib.SB_WANT_LOGLIKE | _lib.SB_WANT_ROOT | _lib.SB_WANT_COUNTS
            b['mask_mode'] = _lib.SB_MASK_AUTO if r.branch_mask == 'auto' else _lib.SB_MASK_NONE
            b['counts'] = dout + oc
            b['loglike'] = dout + ol
            b['root'] = dout + orr
            layout.append((r, A, C, oc, ol, orr))
            oi += bi
            oo += bo
        with torch.cuda.stream(lane.stream):
            lane.dev_in[:oi].copy_(lane.host_in[:oi], non_blocking=True)
        _lib.check(self.eng.lib.sb_batch_submit(lane.launcher.h, len(reqs), blocks.ctypes.data, self.dt,
                                                1 if self.use_graph else 0))
        with torch.cuda.stream(lane.stream):
            lane.host_out[:oo].copy_(lane.dev_out[:oo], non_blocking=True)
            lane.event = torch.cuda.Event()
            lane.event.record(lane.stream)
        lane.batch = layout
        lane.bytes = nin + nout
        lane.busy = True
        self.in_flight += lane.bytes
        del keep  # the planner copied the tree arrays during submit

    def _complete(self, lane, now, results):
        hout = lane.host_out.numpy()
        np_t = np.float32 if self.esz == 4 else np.float64
        for r, A, C, oc, ol, orr in lane.batch:
            cnt = hout[oc:oc + A * A * C * self.esz].view(np_t).reshape(A, A, C).copy()
            ll = hout[ol:ol + C * 8].view(np.float64).copy()
            root = hout[orr:orr + A * C * self.esz].view(np_t).reshape(A, C).copy()
            results.append(Result(r.rid, cnt, ll, root, r.arrival, now))
        self.in_flight -= lane.bytes
        lane.busy = False
        lane.batch = []

    def run(self, requests, realtime: bool = True):
        """Serve ``requests`` (sorted by arrival).  Returns (results, stats)."""
        pending = list(requests)
        admitted = [False] * len(pending)
        results = []
        head = 0                        # first request not yet admitted
        t0 = time.perf_counter()
        batches = 0
        while head < len(pending) or any(l.busy for l in self.lanes):
            now = time.perf_counter() - t0
            progressed = False
            for lane in self.lanes:
                if lane.busy and lane.event.query():
                    self._complete(lane, time.perf_counter() - t0, results)
                    progressed = True
            for lane in self.lanes:
                if lane.busy or head >= len(pending):
                    continue
                if realtime and pending[head].arrival > now:
                    break
                # one launch runs one kind: the oldest waiting request sets it, and
                # requests of that kind are taken in order from a bounded window
                take, cols, nbytes = [], 0, 0
                kind0, j, scanned = None, head, 0
                while j < len(pending) and len(take) < self.max_blocks and scanned < 4 * self.max_blocks:
                    if admitted[j]:
                        j += 1
                        continue
                    r = pending[j]
                    if realtime and r.arrival > now:
                        break
                    scanned += 1
                    k = _kind(r)
                    if take and k != kind0:
                        j += 1
                        continue
                    bi, bo = self._bytes(r)
                    if take and (cols + r.alignment.shape[1] > self.max_cols or
                                 self.in_flight + nbytes + bi + bo > self.budget):
                        break
                    if not take and self.in_flight > 0 and self.in_flight + bi + bo > self.budget:
                        break
                    take.append(r)
                    kind0 = k
                    admitted[j] = True
                    cols += r.alignment.shape[1]
                    nbytes += bi + bo
                    j += 1
                while head < len(pending) and admitted[head]:
                    head += 1
                if take:
                    self._launch(lane, take)
                    batches += 1
                    progressed = True
            if not progressed:
                if head < len(pending) and realtime:
                    wait = pending[head].arrival - (time.perf_counter() - t0)
                    if wait > 2e-4 and not any(l.busy for l in self.lanes):
                        time.sleep(min(wait, 1e-3))
                        continue
                time.sleep(2e-5)
        makespan = time.perf_counter() - t0
        lat = np.array([r.latency for r in results]) if results else np.zeros(1)
        units = sum(int(r.alignment.shape[1]) * (len(r.tree.parentIndex) - 1) for r in requests)
        stats = {'requests': len(results), 'batches': batches, 'makespan_s': makespan,
                 'throughput': units / makespan, 'p50_ms': float(np.percentile(lat, 50) * 1e3),
                 'p99_ms': float(np.percentile(lat, 99) * 1e3), 'mean_ms': float(lat.mean() * 1e3)}
        results.sort(key=lambda r: r.rid)
        return results, stats


class NativeServer:
    """The same continuous batching in the library's native loop (sb_server_*).

    A C++ worker thread owns the lanes (launcher + HIP stream + pinned and
    device buffers), admits queued requests first come first served within
    the HBM budget, packs and unpacks batches with a small thread pool and
    copies results straight into each request's host arrays; `run` replays a
    request stream at its arrival times from native code, so no Python runs
    per request while the clock is running.  Each request carries its model
    inline (uploaded into the resident arena when admitted).  FP32 outputs.
    """

    def __init__(self, engine: Engine | None = None, hbm_budget: int = 8 << 30, lanes: int = 4,
                 max_blocks: int = 2048, max_cols: int = 1 << 17):
        self.eng = engine or default_engine()
        self.lib = self.eng.lib
        h = ctypes.c_void_p()
        _lib.check(self.lib.sb_server_create(self.eng.ctx, lanes, int(hbm_budget), max_blocks,
                                             int(max_cols), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            h, self.h = self.h, None
            _lib.check(self.lib.sb_server_destroy(h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self, requests):
        """Host-side request records and output arrays (not timed)."""
        n = len(requests)
        self.eng.new_call()
        arr = (_lib.SbRequest * n)()
        keep, outs = [], []
        for k, r in enumerate(requests):
            m = ensure_diag(r.model)
            irr = isinstance(m, IrrevDiagModel)
            ev = None if irr else np.ascontiguousarray(m.eigenvalues, dtype=np.float64)
            V = None if irr else np.ascontiguousarray(m.eigenvectors, dtype=np.float64)
            pi = np.ascontiguousarray(m.pi, dtype=np.float64)
            aln = np.ascontiguousarray(r.alignment, dtype=np.int8)
            parent = np.ascontiguousarray(r.tree.parentIndex, dtype=np.int32)
            dist = np.ascontiguousarray(r.tree.distanceToParent, dtype=np.float64)
            R, C = aln.shape
            A = len(pi)
            # output arrays are touched here, so no page is first faulted in
            # while the server copies results into them
            cnt = np.full((A, A, C), np.nan, dtype=np.float32)
            root = np.full((A, C), np.nan, dtype=np.float32)
            ll = np.full(C, np.nan, dtype=np.float64)
            q = arr[k]
            q.A = A
            if irr:   # resident model (inline models are reversible)
                q.model_id = self.eng.model_id(m)
            else:
                q.model_id = -1
                q.eigenvalues, q.eigenvectors, q.pi = ev.ctypes.data, V.ctypes.data, pi.ctypes.data
            q.R, q.C = R, C
            q.alignment, q.parent_index, q.distances = aln.ctypes.data, parent.ctypes.data, dist.ctypes.data
            q.mask_mode = _lib.SB_MASK_AUTO if r.branch_mask == 'auto' else _lib.SB_MASK_NONE
            q.want = _lib.SB_WANT_LOGLIKE | _lib.SB_WANT_ROOT | _lib.SB_WANT_COUNTS
            q.counts, q.root, q.loglike = cnt.ctypes.data, root.ctypes.data, ll.ctypes.data
            keep += [ev, V, pi, aln, parent, dist]
            outs.append((cnt, ll, root))
        return arr, keep, outs

    def run(self, requests, realtime: bool = True, timed: bool = False):
        """Serve ``requests`` (sorted by arrival).  Returns (results, stats) like
        Server.run; with ``timed`` the stats also hold ``breakdown`` (where the
        latency went, see `breakdown`) and ``timing`` (the per-request records)."""
        n = len(requests)
        arr, keep, outs = self.prepare(requests)
        arrival = np.array([r.arrival if realtime else 0.0 for r in requests], dtype=np.float64)
        timing = None
        if timed:
            timing = np.zeros(n, dtype=_lib.TIMING_DTYPE)
            _lib.check(self.lib.sb_server_replay_timed(self.h, n, ctypes.cast(arr, ctypes.c_void_p),
                                                       arrival.ctypes.data, timing.ctypes.data))
            done = timing['done_s'].copy()
        else:
            done = np.zeros(n, dtype=np.float64)
            _lib.check(self.lib.sb_server_replay(self.h, n, ctypes.cast(arr, ctypes.c_void_p),
                                                 arrival.ctypes.data, done.ctypes.data))
        del keep
        results = [Result(r.rid, cnt, ll, root, float(arrival[k]), float(done[k]))
                   for k, (r, (cnt, ll, root)) in enumerate(zip(requests, outs))]
        lat = done - arrival if n else np.zeros(1)
        makespan = float(done.max()) if n else 0.0
        units = sum(int(r.alignment.shape[1]) * (len(r.tree.parentIndex) - 1) for r in requests)
        stats = {'requests': n, 'makespan_s': makespan,
                 'throughput': units / makespan if makespan > 0 else 0.0,
                 'p50_ms': float(np.percentile(lat, 50) * 1e3), 'p99_ms': float(np.percentile(lat, 99) * 1e3),
                 'mean_ms': float(lat.mean() * 1e3)}
        if timed and n:
            stats['breakdown'] = breakdown(timing, arrival)
            stats['timing'] = timing
        results.sort(key=lambda r: r.rid)
        return results, stats


def breakdown(timing, arrival):
    """Split each request's latency (arrival -> results in the caller's arrays)
    into consecutive host-clock spans:

      queue      arrival -> a lane admits it (waiting for a free lane / budget)
      models     resolving the batch's inline models (sharing by content,
                 evictions, puts into the resident arena)
      pack       admission -> its batch is enqueued, less ``models`` and
                 ``plan`` (alignment copies into pinned staging, upload)
      plan       host time of the batch submission (tree planning, launch)
      gpu        enqueued -> the lane sees the batch's download complete
      unpack     results copied from pinned staging into the caller's arrays

    plus the batch's GPU spans from HIP events (upload, kernel, download) and its
    size.  Returns the mean of every field over all requests, over the median
    band (45th-55th percentile of latency) and over the slowest 1%."""
    t = timing
    lat = (t['done_s'] - arrival) * 1e3
    f = {
        'latency_ms': lat,
        'queue_ms': (t['admitted_s'] - arrival) * 1e3,
        'models_ms': t['models_ms'],
        'pack_ms': (t['enqueued_s'] - t['admitted_s']) * 1e3 - t['plan_ms'] - t['models_ms'],
        'plan_ms': t['plan_ms'],
        'gpu_ms': (t['gpu_done_s'] - t['enqueued_s']) * 1e3,
        'unpack_ms': (t['done_s'] - t['gpu_done_s']) * 1e3,
        'gpu_upload_ms': t['upload_ms'],
        'gpu_kernel_ms': t['kernel_ms'],
        'gpu_download_ms': t['download_ms'],
        'batch_requests': t['batch_requests'].astype(np.float64),
        'batch_columns': t['batch_columns'].astype(np.float64),
    }
    lo, mid, hi = np.percentile(lat, [45, 55, 99])
    bands = {'all': np.ones(len(lat), bool), 'p45_p55': (lat >= lo) & (lat <= mid), 'slowest_1pct': lat >= hi}
    return {b: {k: float(v[m].mean()) if m.any() else None for k, v in f.items()} for b, m in bands.items()}


_PADDED = (4, 8, 16, 20, 32, 64)


def _kind(r):
    """Requests that can share a persistent launch: same padded alphabet and
    reversibility (the kernels are specialised on both)."""
    m = ensure_diag(r.model)
    A = len(m.pi)
    return next(p for p in _PADDED if A <= p), isinstance(m, IrrevDiagModel)


def _simulate(args):
    """Requests ``lo .. hi-1`` of a stream (each from its own seeded generator)."""
    seed, lo, hi, leaves, mu, sd, clip, shapes = args
    from . import synthetic
    from .models import hky85_diag
    out = []
    for i in range(lo, hi):
        rng = np.random.default_rng([seed, i])
        m = hky85_diag(float(rng.uniform(1, 5)), rng.dirichlet(np.full(4, 5.0)))
        if shapes is not None:
            out.append((int(rng.integers(shapes)), m))
            continue
        nl = int(rng.integers(leaves[0], leaves[1] + 1))
        C = int(np.clip(round(rng.lognormal(mu, sd)), clip[0], clip[1]))
        tr = synthetic.random_tree(nl, rng, 0.1)
        aln = synthetic.simulate(tr, m, C, rng, gap_frac=0.03, ambig_frac=0.01)
        out.append((tr, m, aln))
    return out


def make_stream(n, qps, seed=4, leaves=(4, 100), mean_cols=500.0, sd=1.0, clip=(20, 10000),
                pool=None, workers=None):
    """Synthetic request stream: Poisson arrivals at ``qps``; each request has a
    random tree (leaves uniform in ``leaves``), LogNormal block length with
    arithmetic mean ``mean_cols`` (log-sd ``sd``) clipped to ``clip``, and its own
    HKY85 model with kappa ~ U(1,5), pi ~ Dirichlet(5,5,5,5).  Request i is
    simulated from its own generator (seeded by (seed, i)), so every request is
    distinct and the stream does not depend on ``workers``: with ``workers`` > 1
    the requests are simulated by that many forked processes (call it before
    anything initialises the GPU).  With ``pool`` set, only that many (tree,
    alignment) pairs are simulated and requests draw from them (each request
    still gets its own model), to keep generation fast in tests."""
    mu = np.log(mean_cols) - 0.5 * sd * sd
    shapes = None
    if pool:
        shapes = [(tr, aln) for tr, _, aln in _simulate((seed + 1, 0, pool, leaves, mu, sd, clip, None))]
    args = (seed, leaves, mu, sd, clip, len(shapes) if shapes else None)
    if workers and workers > 1 and n > workers:
        import multiprocessing as mp
        step = -(-n // (4 * workers))
        parts = [(args[0], lo, min(n, lo + step)) + args[1:] for lo in range(0, n, step)]
        with mp.get_context('fork').Pool(workers) as p:
            sims = [x for part in p.map(_simulate, parts) for x in part]
    else:
        sims = _simulate((args[0], 0, n) + args[1:])
    arrivals = np.cumsum(np.random.default_rng([seed, n]).exponential(1.0 / qps, size=n))
    out = []
    for i, x in enumerate(sims):
        if shapes is not None:
            (tr, aln), m = shapes[x[0]], x[1]
        else:
            tr, m, aln = x
        out.append(Request(aln, tr, m, arrival=float(arrivals[i]), rid=i))
    return out
