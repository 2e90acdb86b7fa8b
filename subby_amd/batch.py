"""Batched launches: many (tree, alignment-block, model) triples per persistent kernel.

The reference batches with ``jax.vmap(Counts)`` over blocks of equal padded
shape (scripts/batch_eigensub_maf.py:519-641, bucketed by node count, columns
padded to COL_BIN=128).  Here blocks of any tree size and column count share
one launch: the planner turns every block into 64-column tiles, orders the
tiles longest-tree-first, and the persistent kernel pulls them from a device
work counter.  No padding of trees or columns is needed, and each column's
arithmetic does not depend on what else is in the batch, so a block's
results are bitwise the same alone or in any mix.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .api import Engine, default_engine, _PREC
from .models import ensure_diag

WANT = {'loglike': _lib.SB_WANT_LOGLIKE, 'root': _lib.SB_WANT_ROOT, 'counts': _lib.SB_WANT_COUNTS}
MASK = {None: _lib.SB_MASK_NONE, 'auto': _lib.SB_MASK_AUTO}


@dataclass
class Block:
    """One request: alignment (R, C) int8 tokens, tree and model."""
    alignment: object
    tree: object
    model: object
    want: tuple = ('loglike', 'counts', 'root')
    f81_fast: bool = False
    branch_mask: object = 'auto'
    # filled by submit()
    counts: torch.Tensor | None = None
    loglike: torch.Tensor | None = None
    root: torch.Tensor | None = None
    _keep: list = field(default_factory=list)


class Launcher:
    """A sb_batch bound to one HIP stream (several may run concurrently)."""

    def __init__(self, engine: Engine | None = None, stream: torch.cuda.Stream | None = None,
                 waves_per_group: int = 0, waves_per_cu: int = -1):
        self.eng = engine or default_engine()
        self.stream = stream or torch.cuda.current_stream(self.eng.device)
        self.h = ctypes.c_void_p()
        _lib.check(self.eng.lib.sb_batch_create(self.eng.ctx, ctypes.c_void_p(self.stream.cuda_stream),
                                                ctypes.byref(self.h)))
        if waves_per_group or waves_per_cu >= 0:
            self.set_tiling(waves_per_group, waves_per_cu)

    def set_tiling(self, waves_per_group=0, waves_per_cu=-1):
        _lib.check(self.eng.lib.sb_batch_set_tiling(self.h, waves_per_group, waves_per_cu))

    def set_tune(self, knob, value):
        """Set a launch-shape knob (SB_TUNE_*); results do not depend on it."""
        _lib.check(self.eng.lib.sb_batch_set_tune(self.h, int(knob), int(value)))

    def get_tune(self, knob):
        v = ctypes.c_int()
        _lib.check(self.eng.lib.sb_batch_get_tune(self.h, int(knob), ctypes.byref(v)))
        return v.value

    def close(self):
        if self.h:
            self.eng.lib.sb_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, blocks, precision: str = 'fp32', use_graph: bool = False, allocate: bool = True):
        """Enqueue all blocks in one launch on this launcher's stream (asynchronous).

        Output tensors are allocated (on this stream) unless the blocks already
        carry them.  Returns the blocks."""
        arr = self.submit_arrays(blocks, precision, allocate)
        dt, _ = _PREC[precision]
        _lib.check(self.eng.lib.sb_batch_submit(self.h, len(blocks), arr, dt, 1 if use_graph else 0))
        return blocks

    def submit_arrays(self, blocks, precision: str = 'fp32', allocate: bool = True):
        """Build the sb_block array for ``blocks`` (outputs allocated here).  Host
        tree arrays stay referenced by the blocks for replays."""
        dt, tdt = _PREC[precision]
        dev = self.eng.new_call().device
        arr = (_lib.SbBlock * len(blocks))()
        with torch.cuda.stream(self.stream):
            for i, b in enumerate(blocks):
                m = ensure_diag(b.model)
                A = len(m.pi)
                aln = self.eng.alignment(b.alignment, A)
                R, C = aln.shape
                parent, dist = self.eng.tree_arrays(b.tree, R)
                want = 0
                for w in b.want:
                    want |= WANT[w]
                if allocate:
                    if 'counts' in b.want and b.counts is None:
                        b.counts = torch.empty((A, A, C), dtype=tdt, device=dev)
                    if 'loglike' in b.want and b.loglike is None:
                        b.loglike = torch.empty(C, dtype=torch.float64, device=dev)
                    if 'root' in b.want and b.root is None:
                        b.root = torch.empty((A, C), dtype=tdt, device=dev)
                if isinstance(b.branch_mask, str) or b.branch_mask is None:
                    if b.branch_mask not in MASK:
                        raise ValueError(f'block {i}: branch_mask must be "auto", None or an (R, C) array, '
                                         f'got {b.branch_mask!r}')
                    mode, mptr = MASK[b.branch_mask], None
                else:
                    mk = torch.as_tensor(np.asarray(b.branch_mask) if not isinstance(b.branch_mask, torch.Tensor)
                                         else b.branch_mask)
                    if tuple(mk.shape) != (R, C):
                        raise ValueError(f'block {i}: branch_mask must have shape ({R}, {C}), '
                                         f'got {tuple(mk.shape)}')
                    mk = (mk.to(dev) != 0).to(torch.uint8).contiguous()
                    b._keep.append(mk)
                    mode, mptr = _lib.SB_MASK_EXPLICIT, mk.data_ptr()
                b._keep += [aln, parent, dist]
                s = arr[i]
                s.model_id = self.eng.model_id(m)
                s.R, s.C = R, C
                s.d_alignment = aln.data_ptr()
                s.parent_index = parent.ctypes.data
                s.distances = dist.ctypes.data
                s.want = want
                s.f81_fast = 1 if b.f81_fast else 0
                s.mask_mode = mode
                s.d_mask = mptr
                s.d_counts = b.counts.data_ptr() if b.counts is not None else None
                s.d_root = b.root.data_ptr() if b.root is not None else None
                s.d_loglike = b.loglike.data_ptr() if b.loglike is not None else None
        return arr

    def sync(self):
        _lib.check(self.eng.lib.sb_batch_sync(self.h))

    def set_timing(self, enable=True):
        _lib.check(self.eng.lib.sb_batch_set_timing(self.h, 1 if enable else 0))

    def kernel_ms(self):
        """(summed main-kernel milliseconds, launches) since the last call."""
        tot, n = ctypes.c_double(), ctypes.c_int()
        _lib.check(self.eng.lib.sb_batch_kernel_ms(self.h, ctypes.byref(tot), ctypes.byref(n)))
        return tot.value, n.value

    def prepare(self, blocks, precision='fp32'):
        """Allocate outputs and build the sb_block array once; replay with launch()."""
        self._prepared = (self.submit_arrays(blocks, precision), precision)
        return self._prepared

    def launch(self, prepared=None, use_graph=True):
        arr, precision = prepared or self._prepared
        dt, _ = _PREC[precision]
        _lib.check(self.eng.lib.sb_batch_submit(self.h, len(arr), arr, dt, 1 if use_graph else 0))


def counts_plain(alignment, tree, model, branch_mask='auto', engine: Engine | None = None):
    """Plain one-block path (A = 4, FP32): per-level launches, partials in HBM.
    Returns (counts (4,4,C) f32, logLike (C,) f64, root (4,C) f32); synchronous."""
    eng = (engine or default_engine()).new_call()
    m = ensure_diag(model)
    A = len(m.pi)
    aln = eng.alignment(alignment, A)
    R, C = aln.shape
    parent, dist = eng.tree_arrays(tree, R)
    dev = eng.device
    cnt = torch.empty((A, A, C), dtype=torch.float32, device=dev)
    ll = torch.empty(C, dtype=torch.float64, device=dev)
    root = torch.empty((A, C), dtype=torch.float32, device=dev)
    mode = MASK[branch_mask] if (branch_mask is None or isinstance(branch_mask, str)) else None
    if mode is None:
        raise ValueError('plain path supports branch_mask "auto" or None')
    _lib.check(eng.lib.sb_counts_plain(eng.ctx, eng.model_id(m), aln.data_ptr(), parent.ctypes.data,
                                       dist.ctypes.data, R, C, mode, cnt.data_ptr(), root.data_ptr(),
                                       ll.data_ptr(), ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)))
    return cnt, ll, root
