"""Benchmark: FP32 Counts column*branches/s on the HKY85 127-node workload.

Workload (one GPU): HKY85, kappa = 2, pi ~ Dirichlet(5,5,5,5); random binary
tree with 64 leaves (127 nodes), branch lengths ~ Exp(mean 0.05); C = 200,000
columns simulated under the model, 3% gaps, 1% ambiguous leaf cells; int8
alignment resident in HBM; FP32 partials, Steiner branch mask "auto"; each
step returns Counts (4,4,C), logLike (C,) and root posteriors (4,C).

A step is one Counts call through the persistent fused path, launched as a
captured hipGraph.  value = C*(R-1) / (seconds per step).

Extra keys on the JSON line:
  roofline      HBM roofline of the fused kernel: algorithmic bytes per launch
                (every inside/outside partial written once and read once,
                int8 tokens, FP32 partials, FP64 log-scales) / kernel time
                measured with HIP events on the launch stream.
  cpu_baseline  the FP64 NumPy oracle (subby.jax data flow, 128-column inside
                chunks) on the whole block, one forked worker per core of the
                CPU affinity (at most 16), one BLAS thread each; median of 5.
  plain         the plain one-block path (per-level launches, partials in HBM)
                on the same block.
  batched_vs_plain  a stream of HKY85 127-node blocks (lengths LogNormal,
                mean 500, clipped to [20, 10000]): one persistent launch for all
                blocks vs the plain path one block per call.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILE = os.path.join(ROOT, 'profiles', 'fused_small_hky85_r127_c2e5.md')


def profiled_traffic(path=PROFILE):
    """HBM bytes per fused launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by a 1 GiB copy), or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith('traffic_bytes_per_launch:'):
                    return int(float(line.split(':', 1)[1]))
    except OSError:
        pass
    return None


def b_alg_grad(R, A, C):
    """Algorithmic bytes of one gradient step: tokens, U and D written and read
    (as b_alg_counts) and the FP64 logLike out; the per-branch partials are
    R doubles per 64-column tile."""
    return C * (R * 1 + 2 * R * (4 * A + 8) + 2 * (R - 1) * (4 * A + 8) + 8) + (C + 63) // 64 * R * 8


def b_alg_counts(R, A, C):
    """Algorithmic bytes of one Counts call: tokens (1 B), U written+read for R
    nodes, D written+read for R-1 nodes (A FP32 + one FP64 log-scale each),
    counts + root + FP64 logLike out."""
    return C * (R * 1 + 2 * R * (4 * A + 8) + 2 * (R - 1) * (4 * A + 8) + 4 * A * A + 4 * A + 8)


def make_c2(sb, C=200000, seed=20260):
    rng = np.random.default_rng(seed)
    pi = rng.dirichlet(np.full(4, 5.0))
    m = sb.hky85_diag(2.0, pi)
    t = sb.synthetic.random_tree(64, rng, 0.05)
    aln = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.03, ambig_frac=0.01)
    return t, m, aln


def make_stream(sb, n_blocks, seed=777):
    rng = np.random.default_rng(seed)
    blocks = []
    sigma = 1.0
    mu = np.log(500.0) - 0.5 * sigma ** 2
    for _ in range(n_blocks):
        C = int(np.clip(round(rng.lognormal(mu, sigma)), 20, 10000))
        m = sb.hky85_diag(float(rng.uniform(1, 5)), rng.dirichlet(np.full(4, 5.0)))
        t = sb.synthetic.random_tree(64, rng, 0.05)
        aln = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.03, ambig_frac=0.01)
        blocks.append((t, m, aln))
    return blocks


_CPU = {}


def _cpu_chunk(bounds):
    """One worker's share of the CPU baseline: Counts of columns [c0, c1)."""
    from threadpoolctl import threadpool_limits
    from oracle import subby_oracle as O
    c0, c1 = bounds
    with threadpool_limits(limits=1):
        O.counts_chunked(_CPU['aln'][:, c0:c1], _CPU['tree'], _CPU['model'], chunk=128)
    return c1 - c0


def _host_cores(max_workers=16):
    return max(1, min(len(os.sched_getaffinity(0)), max_workers))


def _round(d, nd=4):
    return {k: (round(v, nd) if isinstance(v, float) else v) for k, v in d.items()}


def cpu_baseline(t, m, aln, reps=5, max_workers=16):
    """The FP64 oracle (subby.jax data flow: 128-column inside chunks, einsum
    contractions, Counts with mask auto) over the whole block on every core of
    this process's CPU affinity (at most ``max_workers``): one forked worker
    process per core, each with one BLAS thread, on contiguous column ranges.
    Must run before anything initialises the GPU (the workers are forked).
    1 warm-up + ``reps`` timed reps; median and sd."""
    import multiprocessing as mp
    from oracle import subby_oracle as O
    cores = len(os.sched_getaffinity(0))
    n = max(1, min(cores, max_workers))
    R, C = aln.shape
    _CPU.update(aln=np.ascontiguousarray(aln), tree=O.Tree(np.asarray(t.parentIndex), np.asarray(t.distanceToParent)),
                model=O.DiagModel(*m))
    edges = np.linspace(0, C, 4 * n + 1).astype(int)
    work = [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:]) if b > a]
    times = []
    with mp.get_context('fork').Pool(n) as pool:
        pool.map(_cpu_chunk, work)                      # warm-up
        for _ in range(reps):
            s = time.perf_counter()
            done = sum(pool.map(_cpu_chunk, work))
            times.append(time.perf_counter() - s)
    assert done == C
    _CPU.clear()
    med, sd = float(np.median(times)), float(np.std(times))
    return {"value": C * (R - 1) / med, "unit": "column*branches/s", "cores": n, "kind": "port",
            "affinity_cores": cores, "blas_threads_per_worker": 1, "reps": reps,
            "seconds_median": med, "seconds_sd": sd,
            "sample": f"the whole block ({C} columns, R = {R}): Counts (mask auto) in FP64 NumPy "
                      f"(subby.jax data flow, 128-column inside chunks), {n} forked worker processes "
                      f"x 1 BLAS thread over {len(work)} column ranges; 1 warm-up + {reps} reps, "
                      f"median {med:.3f} s, sd {sd:.3f} s; reported baseline only"}


FP32_PEAK_TFLOPS = 157.3


def flops_counts_per_column(R, A):
    """Separable algorithmic flops of one column: inside 2A^2(R-1), outside 4A^2(R-1),
    projections 4A^2(R-1), accumulation 3A^2(R-1), back-transform 4A^3."""
    return 13 * A * A * (R - 1) + 4 * A ** 3


def other_workload(args, log):
    """Measurement lines for the other configurations (not the headline metric)."""
    import torch
    import subby_amd as sb
    from subby_amd.batch import Launcher, Block
    from oracle import subby_oracle as O
    eng = sb.Engine(0)
    dev = eng.device
    w = args.workload
    if w in ('f81-a64', 'general-a64'):
        A = 64
        rng = np.random.default_rng(3000 if w == 'f81-a64' else 3001)
        t = sb.synthetic.random_tree(100, rng, 0.1)
        pi = rng.dirichlet(np.full(A, 2.0))
        if w == 'f81-a64':
            m, C, f81 = sb.f81_model(pi), 1_000_000, True
        else:
            Q = sb.reversible_rate_matrix(rng.lognormal(0.0, 1.0, (A, A)), pi)
            m, C, f81 = sb.diagonalize_rate_matrix(Q, pi), 20_000, False
        log(f'bench {w}: simulating A=64, R={len(t.parentIndex)}, C={C}')
        aln_np = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.03, ambig_frac=0.01)
        R = len(t.parentIndex)
        aln = torch.from_numpy(aln_np).to(dev)
        L = Launcher(eng)
        blk = Block(aln, t, m, f81_fast=f81)
        tuned = None
        if not args.no_tune:
            from subby_amd.autotune import tune
            tuned = tune(L, [blk], reps=2)
            log(f"bench {w}: autotuned {tuned['family']} {tuned['best']} ({tuned['ms']:.3f} ms; "
                f"slowest candidate {tuned['worst_ms']:.3f} ms)")
        prep = L.prepare([blk])
        L.launch(prep, use_graph=False)
        torch.cuda.synchronize()
        steps = max(3, args.steps // 10)
        t0 = time.perf_counter()
        for _ in range(steps):
            L.launch(prep, use_graph=True)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        L.set_timing(True)
        for _ in range(3):
            L.launch(prep, use_graph=False)
        kms, nl = L.kernel_ms()
        kavg = kms / nl
        cols = np.arange(0, C, max(1, C // 400))[:400]
        rc, rl, rr = O.counts_loglike_root(aln_np[:, cols], O.Tree(*t), O.DiagModel(*m), f81_fast=f81)
        gc = blk.counts[:, :, cols].double().cpu().numpy()
        err_c = float(np.max(np.abs(gc - rc).max((0, 1)) / np.abs(rc).max((0, 1))))
        err_l = float(np.max(np.abs(blk.loglike[cols].cpu().numpy() - rl) / np.maximum(np.abs(rl), 1)))
        units = C * (R - 1)
        if w == 'f81-a64':
            balg = b_alg_counts(R, A, C)
            roof = {"bound": "hbm", "achieved": balg / (kavg * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": balg / (kavg * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                    "alg_bytes_per_launch": balg}
        else:
            fl = flops_counts_per_column(R, A) * C
            ach = fl / (kavg * 1e-3) / 1e12
            roof = {"bound": "mfma", "achieved": ach, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": ach / FP32_PEAK_TFLOPS, "traffic": None, "alg_flops_per_launch": fl,
                    "note": "GEMMs on MFMA (pass 1), J-weighted accumulation on VALU (pass 2); "
                            "kernel_ms spans both passes"}
        roof.update({"kernel": "sb_f81w_kernel<64>" if f81 else "sb_gwalk_kernel<64> + sb_gacc_kernel<64>",
                     "kernel_ms": kavg})
        out = {"metric": f"Counts column*branches/sec ({'F81 fast path' if f81 else 'general reversible'}, A=64, FP32)",
               "value": units / wall, "unit": "column*branches/s", "n_gpus": 1, "steps": steps,
               "ms_per_step": wall * 1e3, "higher_is_better": True, "dtype": "f32",
               "config": {"workload": w, "A": A, "R": R, "C": C}, "roofline": roof,
               "parity": {"counts_normwise": err_c, "loglike_rel": err_l, "columns_checked": len(cols)}}
        if tuned is not None:
            out["launch_shape"] = {"family": tuned['family'], "best": tuned['best'], "best_ms": tuned['ms'],
                                   "slowest_ms": tuned['worst_ms'], "autotune_table": tuned['table']}
    elif w == 'latency-hky85':
        # one call at a time (submit, then wait), the headline's tree and model, a
        # block of 64, 6,400 and 64,000 columns, split walks on and off
        from subby_amd import _lib
        rng = np.random.default_rng(20260)
        pi = rng.dirichlet(np.full(4, 5.0))
        m = sb.hky85_diag(2.0, pi)
        t = sb.synthetic.random_tree(64, rng, 0.05)
        R = len(t.parentIndex)
        rows = []
        for C in (64, 6400, 64000):
            aln = torch.from_numpy(sb.synthetic.simulate(t, m, C, rng, gap_frac=0.03, ambig_frac=0.01)).to(dev)
            for split in (1, 0):
                L = Launcher(eng)
                L.set_tune(_lib.SB_TUNE_SPLIT, split)
                blk = [Block(aln, t, m)]
                for _ in range(5):
                    L.submit(blk)
                    L.sync()
                L.set_timing(True)
                L.kernel_ms()
                lat = []
                for _ in range(max(10, args.steps)):
                    t0 = time.perf_counter()
                    L.submit(blk)
                    L.sync()
                    lat.append(time.perf_counter() - t0)
                kms, nl = L.kernel_ms()
                rows.append({"C": C, "tiles": (C + 63) // 64, "split": L.get_tune(_lib.SB_TUNE_LAST_SPLIT),
                             "kernel_ms": kms / nl, "call_p50_ms": 1e3 * float(np.median(lat))})
                L.set_timing(False)
                L.close()
        first = rows[0]
        out = {"metric": "call latency, one 64-column block (FP32 Counts+logLike+root, HKY85 R=127)",
               "value": first["call_p50_ms"], "unit": "ms", "n_gpus": 1, "steps": max(10, args.steps),
               "ms_per_step": first["call_p50_ms"], "higher_is_better": False, "dtype": "f32",
               "config": {"workload": w, "A": 4, "R": R, "C": [64, 6400, 64000]},
               "kernel": "sb_fused_small_kernel<float,4,false,false,false,false,true|false>",
               "calls": rows}
    elif w == 'irrev':
        # irreversible A=4 model (complex eigensystem) on a block of the headline's shape
        rng = np.random.default_rng(6000)
        A = 4
        Q = rng.uniform(0.05, 1.0, (A, A))
        for a in range(A):
            Q[a, (a + 1) % A] += 1.5           # rotational part: complex eigenvalue pair
        np.fill_diagonal(Q, 0.0)
        np.fill_diagonal(Q, -Q.sum(1))
        pi = rng.dirichlet(np.full(A, 5.0))
        Q /= -np.sum(pi * np.diag(Q))
        m = sb.diagonalize_irreversible(Q, pi)
        t = sb.synthetic.random_tree(64, rng, 0.05)
        C = 200_000
        aln_np = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.03, ambig_frac=0.01)
        R = len(t.parentIndex)
        aln = torch.from_numpy(aln_np).to(dev)
        L = Launcher(eng)
        blk = Block(aln, t, m)
        prep = L.prepare([blk])
        L.launch(prep, use_graph=False)
        torch.cuda.synchronize()
        steps = max(5, args.steps)
        for _ in range(3):
            L.launch(prep, use_graph=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            L.launch(prep, use_graph=True)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        L.set_timing(True)
        for _ in range(5):
            L.launch(prep, use_graph=False)
        kms, nl = L.kernel_ms()
        kavg = kms / nl
        cols = np.arange(0, C, max(1, C // 1000))[:1000]
        om = O.IrrevModel(*(np.asarray(x) for x in m), False)
        rc, rl, rr = O.counts_loglike_root(aln_np[:, cols], O.Tree(*t), om)
        gc = blk.counts[:, :, cols].double().cpu().numpy()
        err_c = float(np.max(np.abs(gc - rc).max((0, 1)) / np.abs(rc).max((0, 1))))
        err_l = float(np.max(np.abs(blk.loglike[cols].cpu().numpy() - rl) / np.maximum(np.abs(rl), 1)))
        err_r = float(np.abs(blk.root[:, cols].double().cpu().numpy() - rr).max())
        balg = b_alg_counts(R, A, C)
        units = C * (R - 1)
        out = {"metric": "Counts column*branches/sec (FP32, irreversible A=4 model, complex eigenbasis, "
                         "127-node tree, batched persistent path)",
               "value": units / wall, "unit": "column*branches/s", "n_gpus": 1, "steps": steps,
               "ms_per_step": wall * 1e3, "higher_is_better": True, "dtype": "f32",
               "config": {"workload": "irrev", "A": A, "R": R, "C": C,
                          "model": "random irreversible Q with a complex eigenvalue pair, "
                                   "pi~Dir(5,5,5,5); counts in the reference convention"},
               "roofline": {"bound": "hbm", "achieved": balg / (kavg * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": balg / (kavg * 1e-3) / 1e9 / HBM_PEAK_GBS,
                            "traffic": None, "alg_bytes_per_launch": balg,
                            "kernel": "sb_fused_small_kernel<float,4,false,false,true>",
                            "kernel_ms": kavg},
               "parity": {"counts_normwise": err_c, "loglike_rel": err_l, "root_abs": err_r,
                          "columns_checked": len(cols)}}
    elif w in ('mixture', 'rows'):
        # 'mixture': MixturePosterior over 4 gamma rate categories of HKY85 (K blocks in one
        #            launch over the shared alignment, then the softmax kernel);
        # 'rows':    Counts with one HKY85 model per branch (model grid (1, R)): branch-by-branch
        #            back-transform in each branch's eigenbasis.
        rng = np.random.default_rng(7000)
        pi = rng.dirichlet(np.full(4, 5.0))
        m = sb.hky85_diag(2.0, pi)
        t = sb.synthetic.random_tree(64, rng, 0.05)
        C = 200_000
        aln_np = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.03, ambig_frac=0.01)
        R = len(t.parentIndex)
        A = 4
        aln = torch.from_numpy(aln_np).to(dev)
        cols = np.arange(0, C, max(1, C // 1000))[:1000]
        sub = aln_np[:, cols]
        if w == 'mixture':
            K = 4
            rates, wts = sb.gamma_rate_categories(0.5, K)
            stack = sb.scale_model(m, rates)
            lw = np.log(wts)
            fn = lambda: sb.MixturePosterior(aln, t, stack, lw, engine=eng)  # noqa: E731
            units = C * (R - 1) * K
            metric = ("MixturePosterior column*branches/sec (FP32, HKY85 x 4 gamma categories, R=127; "
                      "units = columns x branches x categories)")
            kname = "sb_fused_small_kernel<float,4,false,false,false,false,true> (LogLike only, split walks) + sb_mixture_softmax_kernel"
        else:
            rows = [sb.hky85_diag(float(rng.uniform(1, 5)), rng.dirichlet(np.full(4, 5.0))) for _ in range(R)]
            fn = lambda: sb.Counts(aln, t, [rows], engine=eng)  # noqa: E731
            units = C * (R - 1)
            metric = "Counts column*branches/sec (FP32, one HKY85 model per branch, R=127, mask auto)"
            kname = "sb_fused_small_kernel<float,4,false,false,false,true>"
        out_t = fn()
        torch.cuda.synchronize()
        steps = max(5, args.steps)
        t0 = time.perf_counter()
        for _ in range(steps):
            out_t = fn()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        ot = O.Tree(*t)
        if w == 'mixture':
            ref = O.MixturePosterior(sub, ot, [O.scale_model(O.DiagModel(*m), r) for r in rates], lw)
            par = {"posterior_abs": float(np.abs(out_t[:, cols].cpu().numpy() - ref).max())}
        else:
            ref = O.Counts_rows(sub, ot, [O.DiagModel(*x) for x in rows])
            g = out_t[:, :, cols].double().cpu().numpy()
            par = {"counts_normwise": float(np.max(np.abs(g - ref).max((0, 1)) / np.abs(ref).max((0, 1))))}
        par["columns_checked"] = len(cols)
        out = {"metric": metric, "value": units / wall, "unit": "column*branches/s", "n_gpus": 1,
               "steps": steps, "ms_per_step": wall * 1e3, "higher_is_better": True, "dtype": "f32",
               "config": {"workload": w, "A": A, "R": R, "C": C}, "kernel": kname, "parity": par}
    elif w == 'maf':
        # MAF ingestion end to end: a synthetic MAF of 5,000 blocks on a 100-species guide
        # tree (species subsets of 3-100, lengths LogNormal mean 500 clipped to [20, 2048]),
        # parsed, pruned and expanded per block, all blocks in persistent launches, 20-feature
        # output formed on the GPU and downloaded (files not written in the timed run).
        import tempfile
        from subby_amd import ingest
        rng = np.random.default_rng(8000)
        names = [f'sp{i}' for i in range(100)]
        gt = sb.synthetic.random_tree(100, rng, 0.05)
        nwk_names = {}
        par = np.asarray(gt.parentIndex)
        is_par = np.zeros(len(par), dtype=bool)
        is_par[par[1:]] = True
        leaf_ids = [n for n in range(len(par)) if not is_par[n]]
        guide = {'parentIndex': par, 'distanceToParent': np.asarray(gt.distanceToParent),
                 'node_names': [None] * len(par), 'leaf_names': names}
        for k, n in enumerate(leaf_ids):
            guide['node_names'][n] = names[k]
            nwk_names[n] = names[k]
        m = sb.hky85_diag(2.0, rng.dirichlet(np.full(4, 5.0)))
        nblk = 5000
        sigma = 1.0
        lens = np.clip(np.round(rng.lognormal(np.log(500.0) - 0.5 * sigma ** 2, sigma, nblk)), 20, 2048).astype(int)
        full = sb.synthetic.simulate(gt, m, int(lens.sum()), rng, gap_frac=0.03, ambig_frac=0.01)
        letters = np.frombuffer(b'ACGTN-', dtype=np.uint8)
        tmpd = tempfile.mkdtemp(prefix='sbmaf')
        path = os.path.join(tmpd, 'synthetic.maf')
        log('bench maf: writing a synthetic MAF of 5,000 blocks')
        with open(path, 'w') as f:
            f.write('##maf version=1\n')
            off = 0
            for b in range(nblk):
                L = int(lens[b])
                k = int(rng.integers(3, 101))
                sub = rng.choice(len(leaf_ids), k, replace=False)
                f.write(f'a score={b}\n')
                for q in sub:
                    row = full[leaf_ids[q], off:off + L].astype(np.int64)
                    seq = letters[np.where(row < 0, 5, np.minimum(row, 4))].tobytes().decode()
                    f.write(f's {names[q]}.chr1 {off} {L} + 10000000 {seq}\n')
                f.write('\n')
                off += L
        ingest.precompute_maf(path, None, m, guide, engine=eng, write=False, max_cols=2048)   # warm-up
        torch.cuda.synchronize()
        st = {}
        t0 = time.perf_counter()
        manifest, feats = ingest.precompute_maf(path, None, m, guide, engine=eng, write=False, max_cols=2048,
                                                stats=st)
        wall = time.perf_counter() - t0
        # parity on a sample of blocks
        om = O.DiagModel(*m)
        blocks = [bk for bk in ingest.iter_maf_blocks(path) if len(bk['species']) >= 3]
        worst = 0.0
        for bk, f in list(zip(blocks, feats))[::250]:
            pr = ingest.prune_tree(guide, set(bk['species']))
            aln = ingest.expand_alignment_to_tree(ingest.seqs_to_alignment(bk['seqs']), bk['species'], pr)
            rf = ingest.counts_to_features(O.Counts(aln, O.Tree(pr['parentIndex'], pr['distanceToParent']), om))
            num = np.abs(f[:, :16] - rf[:, :16]).max(1)
            worst = max(worst, float((num / np.abs(rf[:, :16]).max(1)).max()))
        os.remove(path)
        out = {"metric": "MAF to features column*branches/sec (FP32 Counts, HKY85, per-block pruned guide tree, "
                         "mask auto; parse + prune + expand + GPU + feature download)",
               "value": st['column_branches'] / wall, "unit": "column*branches/s", "n_gpus": 1, "steps": 1,
               "ms_per_step": wall * 1e3, "higher_is_better": True, "dtype": "f32",
               "config": {"workload": "maf", "blocks": st['blocks'], "columns": st['columns'],
                          "guide_tree_leaves": 100},
               "stages_s": {"scan_prune_expand": st['scan_s'], "gpu_and_download": st['gpu_s']},
               "gpu_stage_column_branches_per_s": st['column_branches'] / st['gpu_s'],
               "parity": {"counts_features_normwise": worst, "blocks_checked": len(list(zip(blocks, feats))[::250])}}
    elif w == 'branch-hky85':
        # per-branch outputs on the headline block: BranchCounts, then node + branch posteriors
        rng = np.random.default_rng(2000)
        pi = rng.dirichlet(np.full(4, 5.0))
        m = sb.hky85_diag(2.0, pi)
        t = sb.synthetic.random_tree(64, rng, 0.05)
        C = 200_000
        aln_np = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.03, ambig_frac=0.01)
        aln = torch.from_numpy(aln_np).to(dev)
        R = len(t.parentIndex)
        A = 4

        def timed(fn, reps):
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / reps

        reps = max(5, args.steps // 5)
        bc_holder = {}
        t_bc = timed(lambda: bc_holder.__setitem__('bc', sb.BranchCounts(aln, t, m, engine=eng)), reps)
        io = sb.InsideOutside(aln, t, m, engine=eng)

        def post():   # node and branch posteriors in one launch
            io._npost = io._bpost = None
            io._posteriors(('npost', 'bpost'))
        t_post = timed(post, reps)
        cols = np.arange(0, C, max(1, C // 300))[:300]
        ot, om = O.Tree(*t), O.DiagModel(*m)
        sub = aln_np[:, cols]
        rbc = O.BranchCounts(sub, ot, om)
        gbc = bc_holder['bc'][:, :, :, cols].double().cpu().numpy()
        num = np.abs(gbc - rbc).max(axis=(1, 2))
        den = np.abs(rbc).max(axis=(1, 2))
        err_bc = float(np.max(num[1:] / den[1:]))
        err_np = float(np.abs(io.node_posterior()[:, :, cols].double().cpu().numpy() - O.node_posterior(sub, ot, om)).max())
        err_bp = float(np.abs(io.branch_posterior()[:, :, :, cols].double().cpu().numpy() - O.branch_posterior(sub, ot, om)).max())
        units = C * (R - 1)
        out_bc = R * A * A * C * 4
        out_post = R * A * C * 4 + R * A * A * C * 4
        out = {"metric": "BranchCounts column*branches/sec (FP32, HKY85 R=127, unmasked, (R,A,A,C) output)",
               "value": units / t_bc, "unit": "column*branches/s", "n_gpus": 1, "steps": reps,
               "ms_per_step": t_bc * 1e3, "higher_is_better": True, "dtype": "f32",
               "config": {"workload": "branch-hky85", "A": A, "R": R, "C": C},
               "roofline": {"bound": "hbm", "note": "write-dominated: output bytes / call time",
                            "achieved": out_bc / t_bc / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": out_bc / t_bc / 1e9 / HBM_PEAK_GBS, "traffic": None,
                            "output_bytes_per_call": out_bc},
               "posteriors": {"ms_per_call": t_post * 1e3, "column*branches/s": units / t_post,
                              "output_GB_per_s": out_post / t_post / 1e9},
               "parity": {"branch_counts_per_branch_normwise": err_bc, "node_posterior_abs": err_np,
                          "branch_posterior_abs": err_bp, "columns_checked": len(cols)}}
    elif w == 'post-a20':
        # node + branch posteriors above A = 8: general A = 20 model, 100-leaf tree
        A = 20
        rng = np.random.default_rng(2020)
        pi = rng.dirichlet(np.full(A, 2.0))
        Q = sb.reversible_rate_matrix(rng.lognormal(0.0, 1.0, (A, A)), pi)
        m = sb.diagonalize_rate_matrix(Q, pi)
        t = sb.synthetic.random_tree(100, rng, 0.1)
        C = 20_000
        log(f'bench post-a20: simulating A={A}, R={len(t.parentIndex)}, C={C}')
        aln_np = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.03, ambig_frac=0.01)
        aln = torch.from_numpy(aln_np).to(dev)
        R = len(t.parentIndex)
        io = sb.InsideOutside(aln, t, m, engine=eng)

        def post():   # node and branch posteriors in one launch
            io._npost = io._bpost = None
            io._posteriors(('npost', 'bpost'))
        post()
        torch.cuda.synchronize()
        reps = max(5, args.steps // 5)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record()
        for _ in range(reps):
            post()
        ev1.record()
        torch.cuda.synchronize()
        t_post = (time.perf_counter() - t0) / reps
        ev_ms = ev0.elapsed_time(ev1) / reps
        cols = np.arange(0, C, max(1, C // 200))[:200]
        ot, om = O.Tree(*t), O.DiagModel(*m)
        sub = aln_np[:, cols]
        err_np = float(np.abs(io.node_posterior()[:, :, cols].double().cpu().numpy() - O.node_posterior(sub, ot, om)).max())
        err_bp = float(np.abs(io.branch_posterior()[:, :, :, cols].double().cpu().numpy() - O.branch_posterior(sub, ot, om)).max())
        io._npost = io._bpost = None
        # per-branch counts on the same block (per-branch epilogue of the large-alphabet kernel)
        bc_holder = {}

        def bcounts():
            bc_holder['bc'] = None
            bc_holder['bc'] = sb.BranchCounts(aln, t, m, engine=eng)
        bcounts()
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(reps):
            bcounts()
        ev1.record()
        torch.cuda.synchronize()
        bc_ms = ev0.elapsed_time(ev1) / reps
        rbc = O.BranchCounts(sub, ot, om)
        gbc = bc_holder['bc'][:, :, :, cols].double().cpu().numpy()
        num = np.abs(gbc - rbc).max(axis=(1, 2))
        den = np.abs(rbc).max(axis=(1, 2))
        err_bc = float(np.max(num[1:] / den[1:]))
        bc_holder.clear()
        # Counts (+ logLike, root) on the same block, for scale
        cnt_holder = {}

        def counts():
            cnt_holder['c'] = sb.Counts(aln, t, m, engine=eng)
        counts()
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(reps):
            counts()
        ev1.record()
        torch.cuda.synchronize()
        counts_ms = ev0.elapsed_time(ev1) / reps
        cnt_holder.clear()
        units = C * (R - 1)
        out_b = R * (A + A * A) * C * 4
        out = {"metric": "InsideOutside node + branch posteriors column*branches/sec (FP32, general A=20, R=199)",
               "value": units / t_post, "unit": "column*branches/s", "n_gpus": 1, "steps": reps,
               "ms_per_step": t_post * 1e3, "higher_is_better": True, "dtype": "f32",
               "config": {"workload": "post-a20", "A": A, "R": R, "C": C},
               "roofline": {"bound": "hbm", "note": "write-dominated: (R,A,C) + (R,A,A,C) output bytes / stream time",
                            "achieved": out_b / (ev_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": out_b / (ev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                            "output_bytes_per_call": out_b, "stream_ms_per_call": ev_ms,
                            "kernel": "sb_large_kernel<float, 20, false, false, LOUT_POST>"},
               "branch_counts": {"stream_ms_per_call": bc_ms, "column*branches/s": units / (bc_ms * 1e-3),
                                 "output_GB_per_s": R * A * A * C * 4 / (bc_ms * 1e-3) / 1e9,
                                 "kernel": "sb_gwalk_kernel<32> + sb_gbranch_kernel<32>"},
               "counts": {"stream_ms_per_call": counts_ms, "column*branches/s": units / (counts_ms * 1e-3)},
               "parity": {"node_posterior_abs": err_np, "branch_posterior_abs": err_bp,
                          "branch_counts_per_branch_normwise": err_bc, "columns_checked": len(cols)}}
    else:  # grad-hky85: the gradient
        rng = np.random.default_rng(5000)
        pi = rng.dirichlet(np.full(4, 5.0))
        kappa = 2.0
        m = sb.hky85_diag(kappa, pi)
        t = sb.synthetic.random_tree(64, rng, 0.05)
        C = 50_000
        aln_np = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.03, ambig_frac=0.01)
        aln = torch.from_numpy(aln_np).to(dev)
        R = len(t.parentIndex)
        W = sb.hky85_log_kappa_weights(kappa, pi)
        timing = {}
        grads = {}
        for prec in ('fp32', 'fp64'):
            for _ in range(3):
                sb.loglike_grad(aln, t, m, rate_weights=W, precision=prec, engine=eng)
            torch.cuda.synchronize()
            steps = max(5, args.steps // 5)
            t0 = time.perf_counter()
            for _ in range(steps):
                dg, rg, _ = sb.loglike_grad(aln, t, m, rate_weights=W, precision=prec, engine=eng)
            torch.cuda.synchronize()
            timing[prec] = (time.perf_counter() - t0) / steps
            grads[prec] = (dg.cpu().numpy(), float(rg.cpu().numpy()[0]))
        wall = timing['fp32']
        # central finite differences of the FP64 GPU log-likelihood (h = 1e-4) on 12 branches and
        # log kappa, plain and Richardson-extrapolated with h/2 (truncation O(h^4))
        h = 1e-4
        idx = np.linspace(1, R - 1, 12).astype(int)

        def total(tree, model):
            return sb.LogLike(aln, tree, model, precision='fp64', engine=eng).sum().item()

        def fd_b(n, hh):
            tp = np.array(t.distanceToParent, dtype=np.float64)
            tm = tp.copy()
            tp[n] += hh
            tm[n] -= hh
            return (total(sb.Tree(t.parentIndex, tp), m) - total(sb.Tree(t.parentIndex, tm), m)) / (2 * hh)

        def fd_k(hh):
            return (total(t, sb.hky85_diag(kappa * np.exp(hh), pi))
                    - total(t, sb.hky85_diag(kappa * np.exp(-hh), pi))) / (2 * hh)

        fd = np.array([fd_b(n, h) for n in idx])
        rich = np.array([(4 * fd_b(n, h / 2) - fd_b(n, h)) / 3 for n in idx])
        fdk = fd_k(h)
        richk = (4 * fd_k(h / 2) - fdk) / 3
        par = {"fd_step": h, "branches_checked": len(idx)}
        for prec in ('fp32', 'fp64'):
            dg, rg = grads[prec]
            par[f"{prec}_branch_rel_vs_fd"] = float(np.max(np.abs(dg[idx] - fd) / np.abs(fd)))
            par[f"{prec}_branch_rel_vs_richardson_fd"] = float(np.max(np.abs(dg[idx] - rich) / np.abs(rich)))
            par[f"{prec}_log_kappa_rel_vs_fd"] = abs(rg - fdk) / abs(fdk)
            par[f"{prec}_log_kappa_rel_vs_richardson_fd"] = abs(rg - richk) / abs(richk)
        par["fp32_normwise_vs_fp64"] = float(np.abs(grads['fp32'][0] - grads['fp64'][0]).max()
                                             / np.abs(grads['fp64'][0]).max())
        par["fp64_ms_per_step"] = timing['fp64'] * 1e3
        units = C * (R - 1)
        out = {"metric": "gradient column*branches/sec (FP32 forward + adjoint, HKY85 R=127, "
                         "d/d branch lengths and d/d log kappa)",
               "value": units / wall, "unit": "column*branches/s", "n_gpus": 1, "steps": steps,
               "ms_per_step": wall * 1e3, "higher_is_better": True, "dtype": "f32",
               "config": {"workload": "grad-hky85", "A": 4, "R": R, "C": C},
               "roofline": {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None,
                            "achieved": b_alg_grad(R, 4, C) / wall / 1e9,
                            "frac": b_alg_grad(R, 4, C) / wall / 1e9 / HBM_PEAK_GBS,
                            "alg_bytes_per_step": b_alg_grad(R, 4, C),
                            "note": "whole step (forward + adjoint kernel, per-tile reduction, host glue) "
                                    "against the fused pass's algorithmic bytes; 782 column tiles, "
                                    "two waves each (split walks), fill under two fifths of the "
                                    "wave slots",
                            "kernel": "sb_fused_small_kernel<float,4,false,true,false,false,true> + "
                                      "sb_grad_reduce_kernel"},
               "parity": par}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-plain', action='store_true')
    ap.add_argument('--stream-blocks', type=int, default=300)
    ap.add_argument('--no-tune', action='store_true', help='skip the launch-shape autotuner')
    ap.add_argument('--serve-requests', type=int, default=20000)
    ap.add_argument('--serve-pool', type=int, default=0,
                    help='if set, the request stream draws (tree, alignment) pairs from this many '
                         'simulated ones (default: every request distinct)')
    ap.add_argument('--serve-plain-sample', type=int, default=1000,
                    help='served blocks timed through the plain path')
    ap.add_argument('--workload', default='hky85',
                    choices=['hky85', 'f81-a64', 'general-a64', 'grad-hky85', 'branch-hky85', 'post-a20',
                             'irrev', 'mixture', 'rows', 'maf', 'latency-hky85'],
                    help='hky85 (default, the headline line) or one of the other configurations: '
                         'f81-a64 F81 A=64 R=199 C=1e6; general-a64 general A=64 R=199 C=2e4; '
                         'grad-hky85 HKY85 gradient R=127 C=5e4; branch-hky85 per-branch outputs (BranchCounts, '
                         'node and branch posteriors) HKY85 R=127 C=2e5; post-a20 node and branch '
                         'posteriors, general A=20 R=199 C=2e4; latency-hky85 one call at a time, '
                         'split walks on and off')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='first run a 1 GiB device copy (reference kernel for HBM byte counters)')
    ap.add_argument('--graph', action='store_true',
                    help='time the headline steps as hipGraph replays (default: direct launches, which '
                         'queue fewer packets between steps on ROCm)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU only; --gpus {args.gpus} is not supported in this build '
              '(columns shard across GPUs with no exchange; see the design notes)', file=sys.stderr)
        sys.exit(2)

    import torch
    import subby_amd as sb
    from subby_amd.batch import Launcher, Block, counts_plain

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    if args.workload != 'hky85':
        return other_workload(args, log)

    if args.copy_calibration:
        src = torch.empty(1 << 28, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        src.fill_(1.0)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log('bench: copy calibration done (3 x 1 GiB read + 1 GiB write)')

    log('bench: building the HKY85 127-node workload (C = 200,000)')
    t, m, aln_np = make_c2(sb)
    R, C, A = len(t.parentIndex), aln_np.shape[1], 4
    cpu = None
    if not args.no_cpu:   # first: its workers are forked before anything touches the GPU
        log('bench: CPU baseline (FP64 oracle, whole block, all cores)')
        cpu = cpu_baseline(t, m, aln_np)
        log(f"bench: CPU baseline {cpu['value']:.3e} col*br/s on {cpu['cores']} cores")
    serve_reqs = None
    if args.serve_requests > 0:   # simulated by forked workers, before the GPU is touched
        from subby_amd.serve import make_stream as make_requests
        nw = 1 if torch.cuda.is_initialized() else _host_cores()
        log(f'bench: simulating {args.serve_requests} serving requests ({nw} workers)')
        serve_reqs = make_requests(args.serve_requests, qps=1e9, seed=99, pool=args.serve_pool or None,
                                   workers=nw)
    eng = sb.Engine(0)
    dev = eng.device
    aln = torch.from_numpy(aln_np).to(dev)
    L = Launcher(eng)
    blk = Block(aln, t, m)
    from subby_amd.autotune import tune
    if args.no_tune:
        tuned = {'best': {}, 'ms': None, 'worst_ms': None, 'table': [], 'family': 'small'}
    else:
        tuned = tune(L, [blk])
        log(f"bench: autotuned launch shape {tuned['best']} ({tuned['ms']:.3f} ms; "
            f"slowest candidate {tuned['worst_ms']:.3f} ms)")
    prep = L.prepare([blk])
    torch.cuda.synchronize()

    log(f'bench: warmup {args.warmup}, timed {args.steps} steps')
    for _ in range(args.warmup):
        L.launch(prep, use_graph=args.graph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        L.launch(prep, use_graph=args.graph)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    units = C * (R - 1)
    value = units / (wall / args.steps)

    # the other launch mode, for comparison (same prepared batch)
    for _ in range(3):
        L.launch(prep, use_graph=not args.graph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        L.launch(prep, use_graph=not args.graph)
    torch.cuda.synchronize()
    ms_other = (time.perf_counter() - t0) / args.steps * 1e3

    # kernel duration with HIP events on the launcher's stream (eager launches)
    L.set_timing(True)
    for _ in range(max(5, args.steps // 2)):
        L.launch(prep, use_graph=False)
    kms, nl = L.kernel_ms()
    L.set_timing(False)
    kavg = kms / nl
    balg = b_alg_counts(R, A, C)
    achieved = balg / (kavg * 1e-3) / 1e9
    log(f'bench: {ms_step:.3f} ms/step, kernel {kavg:.3f} ms, {value:.3e} col*br/s')

    # parity spot check of the timed outputs against the oracle on a column sample
    from oracle import subby_oracle as O
    cols = np.arange(0, C, 97)[:1500]
    rc, rl, rr = O.counts_loglike_root(aln_np[:, cols], O.Tree(*t), O.DiagModel(*m))
    gc = blk.counts[:, :, cols].double().cpu().numpy()
    err_c = float(np.max(np.abs(gc - rc).max((0, 1)) / np.abs(rc).max((0, 1))))
    gl = blk.loglike[cols].cpu().numpy()
    err_l = float(np.max(np.abs(gl - rl) / np.maximum(np.abs(rl), 1)))

    out = {
        "metric": "Counts column*branches/sec (FP32, HKY85 127-node tree, batched persistent path)",
        "value": value, "unit": "column*branches/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded simulation under the model)",
        "config": {"workload": "HKY85 kappa=2 pi~Dir(5,5,5,5); 64-leaf random tree (R=127), "
                               "t~Exp(0.05); C=200000; 3% gaps, 1% ambiguous; int8 alignment; "
                               "Counts+logLike+root, mask auto", "R": R, "C": C, "A": A,
                   "parallelism": "single GPU", "launch": "hipGraph replay" if args.graph else "direct"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(),
                     # the counters' bytes (FETCH_SIZE + WRITE_SIZE) over this run's launch time
                     "counter_frac": (profiled_traffic() / (kavg * 1e-3) / 1e9 / HBM_PEAK_GBS
                                      if profiled_traffic() else None),
                     "kernel": "sb_fused_small_kernel<float,4,false,false,false,false,true> (two waves per tile)",
                     "kernel_ms": kavg,
                     "alg_bytes_per_launch": balg},
        "parity": {"counts_normwise": err_c, "loglike_rel": err_l, "columns_checked": int(len(cols))},
        ("ms_per_step_graph" if not args.graph else "ms_per_step_direct"): ms_other,
        "launch_shape": {"family": tuned['family'], "best": tuned['best'], "best_ms": tuned['ms'],
                         "slowest_ms": tuned['worst_ms'], "autotune_table": tuned['table']},
    }

    if not args.no_plain:
        log('bench: plain path on the same block')
        counts_plain(aln, t, m, engine=eng)
        torch.cuda.synchronize()
        reps = 5
        s = time.perf_counter()
        for _ in range(reps):
            counts_plain(aln, t, m, engine=eng)
        torch.cuda.synchronize()
        pt = (time.perf_counter() - s) / reps
        out["plain"] = {"value": units / pt, "ms_per_call": pt * 1e3, "unit": "column*branches/s",
                        "speedup_fused_vs_plain": value / (units / pt)}

        log(f'bench: block stream ({args.stream_blocks} blocks, 127-node trees)')
        stream = make_stream(sb, args.stream_blocks)
        dblocks = [Block(torch.from_numpy(a).to(dev), tt, mm) for tt, mm, a in stream]
        sunits = sum(a.shape[1] * (len(tt.parentIndex) - 1) for tt, mm, a in stream)
        L2 = Launcher(eng)
        if not args.no_tune:
            tune(L2, dblocks)
        sprep = L2.prepare(dblocks)
        L2.launch(sprep, use_graph=True)
        torch.cuda.synchronize()
        s = time.perf_counter()
        for _ in range(5):
            L2.launch(sprep, use_graph=True)
        torch.cuda.synchronize()
        bt = (time.perf_counter() - s) / 5
        for tt, mm, a in stream[:3]:
            counts_plain(a, tt, mm, engine=eng)
        torch.cuda.synchronize()
        s = time.perf_counter()
        for b in dblocks:
            counts_plain(b.alignment, b.tree, b.model, engine=eng)
        torch.cuda.synchronize()
        ptot = time.perf_counter() - s
        out["batched_vs_plain"] = {
            "blocks": len(stream), "columns": int(sum(a.shape[1] for _, _, a in stream)),
            "batched": sunits / bt, "plain": sunits / ptot, "ratio": (sunits / bt) / (sunits / ptot),
            "unit": "column*branches/s"}

    if args.serve_requests > 0:
        from subby_amd.serve import NativeServer, Server
        reqs = serve_reqs
        units = sum(int(r.alignment.shape[1]) * (len(r.tree.parentIndex) - 1) for r in reqs)
        ncols = sum(int(r.alignment.shape[1]) for r in reqs)
        from subby_amd.batch import Block
        # GPU-only reference: the same blocks as one persistent batch, inputs resident
        log('bench: serving stream as one resident batch')
        blocks = [Block(torch.from_numpy(r.alignment).to(dev), r.tree, r.model) for r in reqs]
        Lg = Launcher(eng)
        prep = Lg.prepare(blocks)
        Lg.launch(prep, use_graph=False)
        torch.cuda.synchronize()
        s_ = time.perf_counter()
        for _ in range(3):
            Lg.launch(prep, use_graph=False)
        torch.cuda.synchronize()
        gpu_only = units / ((time.perf_counter() - s_) / 3)
        del prep, Lg
        # plain path (one block per call, per-level launches) on the same blocks:
        # a sample of them, inputs resident, outputs left on the device
        nplain = min(len(blocks), args.serve_plain_sample)
        log(f'bench: plain path on {nplain} of the served blocks')
        for b_ in blocks[:20]:
            counts_plain(b_.alignment, b_.tree, b_.model, engine=eng)
        torch.cuda.synchronize()
        s_ = time.perf_counter()
        for b_ in blocks[:nplain]:
            counts_plain(b_.alignment, b_.tree, b_.model, engine=eng)
        torch.cuda.synchronize()
        pt = time.perf_counter() - s_
        plain_units = sum(int(b_.alignment.shape[1]) * (len(b_.tree.parentIndex) - 1) for b_ in blocks[:nplain])
        plain = plain_units / pt
        del blocks
        eng.trim()              # the resident batch resolved every request's model
        torch.cuda.empty_cache()

        log(f'bench: serving {len(reqs)} distinct requests (native continuous batching)')
        nsrv = NativeServer(eng, hbm_budget=16 << 30, lanes=4, max_cols=1 << 17)
        nsrv.run(reqs[:2000], realtime=False)                   # warm-up (graphs, buffers)
        arena = [eng.arena_stats()]
        _, sat = nsrv.run(reqs, realtime=False, timed=True)
        arena.append(eng.arena_stats())
        rate = sat['requests'] / sat['makespan_s']
        log(f"bench: saturated {rate:.0f} requests/s, {sat['throughput']:.3e} col*br/s")
        sweep = []
        for k, frac in enumerate((0.25, 0.5, 0.75, 0.9)):
            qps = frac * rate
            arr = np.cumsum(np.random.default_rng(5 + k).exponential(1.0 / qps, size=len(reqs)))
            for r, a_ in zip(reqs, arr):
                r.arrival = float(a_)
            _, st = nsrv.run(reqs, realtime=True, timed=True)
            arena.append(eng.arena_stats())
            bd = st['breakdown']
            sweep.append({"load": frac, "qps": qps, "achieved_qps": st['requests'] / st['makespan_s'],
                          "throughput": st['throughput'], "p50_ms": st['p50_ms'], "p99_ms": st['p99_ms'],
                          "mean_ms": st['mean_ms'], "breakdown_p45_p55": _round(bd['p45_p55']),
                          "breakdown_slowest_1pct": _round(bd['slowest_1pct'])})
            log(f"bench: load {frac:.2f} ({qps:.0f} req/s): p50 {st['p50_ms']:.2f} ms, p99 {st['p99_ms']:.2f} ms")
        nsrv.close()
        psrv = Server(eng, hbm_budget=16 << 30, lanes=3)       # the Python loop, for comparison
        npy = min(len(reqs), 5000)
        psrv.run(reqs[:200], realtime=False)
        _, psat = psrv.run(reqs[:npy], realtime=False)
        del psrv
        out["served"] = {
            "workload": "stream of distinct HKY85 blocks: random trees with 4-100 leaves, block length LogNormal "
                        "(mean 500, sd 1.0) clipped to [20, 10000], 3% gaps and 1% ambiguous tokens, "
                        "kappa~U(1,5), pi~Dir(5,5,5,5) per request, each request simulated from its own seed; "
                        "FP32 Counts+logLike+root, alignments from host memory, results copied back to host",
            "server": "native loop (sb_server_*), 4 lanes, at most 131072 columns per launch, "
                      "arrivals replayed from native code",
            "requests": len(reqs), "distinct_blocks": len(reqs) if not args.serve_pool else args.serve_pool,
            "columns": ncols,
            "saturated": {"throughput": sat['throughput'], "requests_per_s": rate,
                          "p50_ms": sat['p50_ms'], "p99_ms": sat['p99_ms'],
                          "breakdown_all": _round(sat['breakdown']['all'])},
            "sweep": sweep,
            "arena": {"note": "model arena after the warm-up and after each of the 5 runs above "
                              "(saturated, then the sweep); inline models are capped at 4096 per server",
                      "reserved_bytes": [a_[0] for a_ in arena], "live_models": [a_[1] for a_ in arena]},
            "plain": {"throughput": plain, "blocks": nplain,
                      "note": "sb_counts_plain per block on the first blocks of the same stream, "
                              "alignments resident in HBM, outputs left in HBM"},
            "served_vs_plain": sat['throughput'] / plain,
            "gpu_only_batch": gpu_only,
            "gpu_only_batch_vs_plain": gpu_only / plain,
            "python_loop_saturated": {"throughput": psat['throughput'], "requests": npy,
                                      "requests_per_s": psat['requests'] / psat['makespan_s']},
            "unit": "column*branches/s"}

    if cpu is not None:
        out["cpu_baseline"] = cpu

    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
