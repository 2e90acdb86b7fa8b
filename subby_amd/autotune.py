"""Autotuner for the persistent launch shape.

For a representative batch it times every candidate (waves per workgroup x
resident waves per CU) with HIP events around the main kernel on the
launcher's stream, keeps the fastest, and applies it to the launcher.
Results can be cached in a JSON file keyed by device, alphabet, precision
and batch size bucket, so serving processes reuse them without re-timing.
"""
from __future__ import annotations

import json
import math
import os

import torch

WAVES_PER_GROUP = (1, 2, 4)
WAVES_PER_CU = (0, 4, 8, 12, 16, 20, 24, 32)   # 0 = from the kernel's occupancy


def _key(launcher, blocks, precision):
    dev = torch.cuda.get_device_name(launcher.eng.device)
    A = max(len(b.model.pi) for b in blocks)
    cols = sum(int(b.alignment.shape[1]) for b in blocks)
    f81 = any(b.f81_fast for b in blocks)
    return f'{dev}|A{A}|{precision}|f81={int(f81)}|cols2^{int(math.log2(max(cols, 1)))}'


def tune(launcher, blocks, precision: str = 'fp32', reps: int = 5, cache: str | None = None,
         candidates=None):
    """Time the candidates on ``blocks`` and apply the best to ``launcher``.

    Returns {"best": (waves_per_group, waves_per_cu), "ms": ..., "table": [...]}."""
    key = _key(launcher, blocks, precision)
    if cache and os.path.exists(cache):
        with open(cache) as f:
            db = json.load(f)
        if key in db:
            wpg, wpc = db[key]['best']
            launcher.set_tiling(wpg, wpc)
            return dict(db[key], cached=True)
    prep = launcher.prepare(blocks, precision)
    cands = candidates or [(g, c) for g in WAVES_PER_GROUP for c in WAVES_PER_CU]
    table = []
    launcher.set_timing(True)
    try:
        for g, c in cands:
            launcher.set_tiling(g, c)
            launcher.launch(prep, use_graph=False)       # warm-up (and scratch growth)
            launcher.kernel_ms()
            for _ in range(reps):
                launcher.launch(prep, use_graph=False)
            ms, n = launcher.kernel_ms()
            table.append({'waves_per_group': g, 'waves_per_cu': c, 'ms': ms / n})
    finally:
        launcher.set_timing(False)
    best = min(table, key=lambda r: r['ms'])
    launcher.set_tiling(best['waves_per_group'], best['waves_per_cu'])
    res = {'best': (best['waves_per_group'], best['waves_per_cu']), 'ms': best['ms'], 'table': table,
           'key': key}
    if cache:
        db = {}
        if os.path.exists(cache):
            with open(cache) as f:
                db = json.load(f)
        db[key] = {'best': res['best'], 'ms': res['ms']}
        with open(cache, 'w') as f:
            json.dump(db, f, indent=1)
    return res
