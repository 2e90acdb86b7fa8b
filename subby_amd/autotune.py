"""Autotuner for the persistent launch shape and tile shapes.

For a representative batch it times every candidate of the kernel family the
batch runs on (HIP events around the main kernels on the launcher's stream),
keeps the fastest, and applies it to the launcher.  The search space per
family:

  small  (padded alphabet <= 8): waves per workgroup x resident waves per CU;
  gen2   (general models above 8 states, FP32 Counts / BranchCounts): 64-column
         tiles per round of the two-pass path (the stage footprint: small rounds
         keep it in cache, large ones expose more parallelism);
  large  (FP64, F81 and the other large-alphabet kernels): resident waves per CU.

Results do not depend on any of these (bitwise), only the time does.  Results
can be cached in a JSON file keyed by device, family, alphabet, precision and
batch size bucket, so serving processes reuse them without re-timing.  The
reference sizes its batches by a memory budget only
(scripts/batch_eigensub_maf.py:296-304); it has no tuner.
"""
from __future__ import annotations

import itertools
import json
import math
import os

import torch

from . import _lib
from .models import IrrevDiagModel, ensure_diag

WAVES_PER_GROUP = (1, 2, 4)
WAVES_PER_CU = (0, 8, 12, 16, 20, 24)          # 0 = from the kernel's occupancy
GEN2_ROUNDS = (0, 4, 8, 16, 32, 64, 128)        # 0 = as many tiles as the stage budget holds
LARGE_WAVES_PER_CU = (0, 1, 2, 4, 8)

KNOBS = {'waves_per_group': _lib.SB_TUNE_WAVES_PER_GROUP, 'waves_per_cu': _lib.SB_TUNE_WAVES_PER_CU,
         'gen2_round': _lib.SB_TUNE_GEN2_ROUND}


def family(blocks, precision='fp32'):
    """Kernel family the batch runs on: 'small', 'gen2' or 'large'."""
    m = ensure_diag(blocks[0].model)
    A = len(m.pi)
    if A <= 8:
        return 'small'
    f81 = any(b.f81_fast for b in blocks)
    if precision == 'fp32' and not f81 and not isinstance(m, IrrevDiagModel):
        return 'gen2'
    return 'large'


def candidates(fam, split=False):
    """Launch shapes to time.  ``split``: the batch's tiles are walked by two
    waves each (split walks), whose workgroups always hold two waves, so only
    the resident waves per CU are searched."""
    if fam == 'small' and split:
        return [dict(waves_per_cu=c) for c in WAVES_PER_CU]
    if fam == 'small':
        return [dict(waves_per_group=g, waves_per_cu=c)
                for g, c in itertools.product(WAVES_PER_GROUP, WAVES_PER_CU)]
    if fam == 'gen2':
        return [dict(gen2_round=r) for r in GEN2_ROUNDS]
    return [dict(waves_per_cu=c) for c in LARGE_WAVES_PER_CU]


def apply(launcher, cand):
    for k, v in cand.items():
        launcher.set_tune(KNOBS[k], v)


def _key(launcher, blocks, precision, fam):
    dev = torch.cuda.get_device_name(launcher.eng.device)
    A = max(len(ensure_diag(b.model).pi) for b in blocks)
    cols = sum(int(b.alignment.shape[1]) for b in blocks)
    R = max(len(b.tree.parentIndex) for b in blocks)
    f81 = any(b.f81_fast for b in blocks)
    return (f'{dev}|{fam}|A{A}|{precision}|f81={int(f81)}|cols2^{int(math.log2(max(cols, 1)))}'
            f'|R2^{int(math.log2(max(R, 1)))}')


def tune(launcher, blocks, precision: str = 'fp32', reps: int = 5, cache: str | None = None,
         cands=None):
    """Time the candidates on ``blocks`` and apply the best to ``launcher``.

    Returns {"best": {knob: value}, "ms": ..., "table": [...], "family": ...}."""
    fam = family(blocks, precision)
    key = _key(launcher, blocks, precision, fam)
    if cache and os.path.exists(cache):
        with open(cache) as f:
            db = json.load(f)
        if key in db:
            apply(launcher, db[key]['best'])
            return dict(db[key], cached=True, family=fam)
    prep = launcher.prepare(blocks, precision)
    split = False
    if fam == 'small':
        launcher.launch(prep, use_graph=False)
        split = launcher.get_tune(_lib.SB_TUNE_LAST_SPLIT) == 1
    cands = cands or candidates(fam, split)
    table = []
    launcher.set_timing(True)
    try:
        for cand in cands:
            apply(launcher, cand)
            launcher.launch(prep, use_graph=False)       # warm-up (and buffer growth)
            launcher.kernel_ms()
            for _ in range(reps):
                launcher.launch(prep, use_graph=False)
            ms, n = launcher.kernel_ms()
            table.append(dict(cand, ms=ms / n))
    finally:
        launcher.set_timing(False)
    best = min(table, key=lambda r: r['ms'])
    worst = max(table, key=lambda r: r['ms'])
    best_cand = {k: v for k, v in best.items() if k != 'ms'}
    apply(launcher, best_cand)
    res = {'best': best_cand, 'ms': best['ms'], 'worst_ms': worst['ms'], 'table': table, 'key': key,
           'family': fam, 'split': split}
    if cache:
        db = {}
        if os.path.exists(cache):
            with open(cache) as f:
                db = json.load(f)
        db[key] = {'best': best_cand, 'ms': best['ms']}
        with open(cache, 'w') as f:
            json.dump(db, f, indent=1)
    return res
