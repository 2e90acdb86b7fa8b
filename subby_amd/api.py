"""Drop-in host API: LogLike / Counts / RootProb / LogLikeCustomGrad.

Same call shape as the reference's subby.jax wrappers
(subby/jax/__init__.py:207-291 LogLike, :338-447 Counts, :543-650 RootProb,
:294-335 LogLikeCustomGrad) for a single DiagModel or RateModel:

    alignment : (R, C) int tokens (numpy, torch CPU or torch CUDA)
    tree      : Tree(parentIndex (R,), distanceToParent (R,))
    model     : DiagModel(eigenvalues (A,), eigenvectors (A,A) V[a,k], pi (A,))

Results are torch tensors on the engine's GPU: LogLike (C,) float64,
Counts (A, A, C) and RootProb (A, C) in the compute precision ('fp32' by
default, 'fp64' on request).  ``maxChunkSize`` is accepted for signature
compatibility; results do not depend on it.  Errors the reference raises as
Python assertions (non-binary tree, F81 with an unsupported model) raise
``SubbyError`` / ``ValueError`` here.
"""
from __future__ import annotations

import ctypes
from collections import OrderedDict
import threading

import numpy as np
import torch

from . import _lib
from .models import Tree, DiagModel, IrrevDiagModel, RateModel, ensure_diag
from .padding import strip_dummy_leaves

_PREC = {'fp32': (_lib.SB_F32, torch.float32), 'fp64': (_lib.SB_F64, torch.float64)}


def _stream_ptr(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class Engine:
    """One sb_context on one GPU plus the resident-model cache.

    The cache holds up to ``max_models`` eigensystems by content (None: no
    limit).  Beyond that, the least recently used model that no call since
    the last ``new_call()`` has asked for is dropped; the library reuses its
    device record once the launches that read it have finished.  Every API
    call starts a new generation, so the ids a call resolves stay valid until
    it has submitted its launch (one calling thread per Engine)."""

    def __init__(self, device=0, max_models: int | None = 4096):
        if not torch.cuda.is_available():
            raise RuntimeError('subby_amd needs a ROCm GPU (torch.cuda.is_available() is False)')
        self.lib = _lib.load()
        self.device = torch.device('cuda', device)
        self.ctx = ctypes.c_void_p()
        _lib.check(self.lib.sb_create(ctypes.byref(self.ctx), device))
        self._models = OrderedDict()      # key -> [model id, generation of last use]
        self._gen = 0
        self.max_models = max_models
        self._launchers = {}
        self._lock = threading.Lock()

    def new_call(self):
        """Start a call: models used by earlier calls become evictable."""
        self._gen += 1
        return self

    def arena_stats(self):
        """(reserved model-arena bytes, live models, retired records)."""
        r, n, q = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_int()
        _lib.check(self.lib.sb_arena_stats(self.ctx, ctypes.byref(r), ctypes.byref(n), ctypes.byref(q)))
        return r.value, n.value, q.value

    def close(self):
        for L in getattr(self, '_launchers', {}).values():
            L.close()
        self._launchers = {}
        if self.ctx:
            self.lib.sb_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- resident models
    def model_id(self, model, exact: bool = False) -> int:
        """Resident id of a model (uploaded once per engine).  For an
        IrrevDiagModel, ``exact`` selects the counts convention (False: the
        reference's formulas, True: expected counts under Q itself)."""
        m = ensure_diag(model)
        pi = np.ascontiguousarray(m.pi, dtype=np.float64)
        A = len(pi)
        irr = isinstance(m, IrrevDiagModel)
        cdt = np.complex128 if irr else np.float64
        ev = np.ascontiguousarray(m.eigenvalues, dtype=cdt)
        V = np.ascontiguousarray(m.eigenvectors, dtype=cdt)
        Vi = np.ascontiguousarray(m.eigenvectors_inv, dtype=cdt) if irr else None
        if ev.shape != (A,) or V.shape != (A, A) or (irr and Vi.shape != (A, A)):
            raise ValueError(f'model shapes {ev.shape}, {V.shape}, {pi.shape} are inconsistent')
        key = (A, irr, bool(exact) and irr, ev.tobytes(), V.tobytes(),
               Vi.tobytes() if irr else b'', pi.tobytes())
        with self._lock:
            ent = self._models.get(key)
            if ent is not None:
                ent[1] = self._gen
                self._models.move_to_end(key)
                return ent[0]
            self._evict()
            if True:
                out = ctypes.c_int()
                if irr:
                    _lib.check(self.lib.sb_model_put_irrev(self.ctx, A, ev.ctypes.data, V.ctypes.data,
                                                           Vi.ctypes.data, pi.ctypes.data,
                                                           1 if exact else 0, ctypes.byref(out)))
                else:
                    _lib.check(self.lib.sb_model_put(self.ctx, A, ev.ctypes.data, V.ctypes.data,
                                                     pi.ctypes.data, ctypes.byref(out)))
                mid = out.value
                self._models[key] = [mid, self._gen]
        return mid

    def trim(self):
        """Start a call and drop cached models beyond ``max_models`` (least
        recently used first), e.g. after one call resolved many models."""
        self.new_call()
        with self._lock:
            self._evict(0)

    def _evict(self, slack: int = 1):
        if self.max_models is None:
            return
        while len(self._models) > self.max_models - slack:
            key, (mid, gen) = next(iter(self._models.items()))
            if gen >= self._gen:        # everything left was asked for by the current call
                return
            del self._models[key]
            _lib.check(self.lib.sb_model_drop(self.ctx, mid))

    def launcher(self):
        """A batch launcher on the current stream (created once per stream)."""
        from .batch import Launcher
        st = torch.cuda.current_stream(self.device)
        key = st.cuda_stream
        with self._lock:
            L = self._launchers.get(key)
            if L is None:
                L = Launcher(self, stream=st)
                self._launchers[key] = L
        return L

    def drop_models(self):
        with self._lock:
            for mid, _ in self._models.values():
                self.lib.sb_model_drop(self.ctx, mid)
            self._models.clear()

    # ---- inputs
    def alignment(self, alignment, A) -> torch.Tensor:
        """(R, C) int8 tokens on the device; tokens outside -1..A+1 read as gaps."""
        if isinstance(alignment, torch.Tensor) and alignment.device == self.device \
                and alignment.dtype == torch.int8 and alignment.is_contiguous():
            return alignment
        a = alignment.detach().cpu().numpy() if isinstance(alignment, torch.Tensor) else np.asarray(alignment)
        if a.ndim != 2:
            raise ValueError(f'alignment must be (R, C), got shape {a.shape}')
        a = np.where((a >= -1) & (a <= A + 1), a, -1).astype(np.int8)
        return torch.from_numpy(np.ascontiguousarray(a)).to(self.device)

    @staticmethod
    def tree_arrays(tree, R):
        parent = np.ascontiguousarray(_to_np(tree.parentIndex), dtype=np.int32)
        dist = np.ascontiguousarray(_to_np(tree.distanceToParent), dtype=np.float64)
        if parent.shape != (R,) or dist.shape != (R,):
            raise ValueError(f'tree arrays must have shape ({R},), got {parent.shape} and {dist.shape}')
        return parent, dist


def _to_np(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


_default = None
_default_lock = threading.Lock()


def default_engine() -> Engine:
    global _default
    with _default_lock:
        if _default is None:
            _default = Engine(torch.cuda.current_device() if torch.cuda.is_available() else 0)
        return _default


def _prep(alignment, tree, model, engine, strip=False):
    """Device alignment, host tree arrays and the model.  ``strip``: drop the
    trailing dummy leaves of a padded tree (pad_tree), which leave LogLike,
    RootProb and Counts unchanged."""
    eng = (engine or default_engine()).new_call()
    m = ensure_diag(model)
    A = len(m.pi)
    aln = eng.alignment(alignment, A)
    R, C = aln.shape
    parent, dist = eng.tree_arrays(tree, R)
    if strip:
        keep = strip_dummy_leaves(parent, dist, aln)
        if keep < R:
            aln, parent, dist, R = aln[:keep], parent[:keep].copy(), dist[:keep].copy(), keep
    return eng, m, A, aln, R, C, parent, dist


_MODEL_TYPES = (Tree, DiagModel, IrrevDiagModel, RateModel)


def _is_model_list(model) -> bool:
    """A list of per-column models (subby/jax/__init__.py:52-55)."""
    return isinstance(model, (list, tuple)) and len(model) > 0 and not isinstance(model, _MODEL_TYPES)


def _is_model_grid(model) -> bool:
    """A list of lists: grid[c][n] (subby/jax/__init__.py:58-70)."""
    return _is_model_list(model) and isinstance(model[0], (list, tuple)) \
        and not isinstance(model[0], _MODEL_TYPES)


def _stacked(model):
    """Models stacked along leading *H axes (e.g. scale_model with a vector of
    rates) as (H shape, list of models), or None for a single model."""
    if _is_model_list(model) or isinstance(model, dict):
        return None
    m = ensure_diag(model)
    H = np.shape(m.eigenvalues)[:-1]
    if not H:
        return None
    flat = [type(m)(*(np.asarray(x).reshape((-1,) + np.shape(x)[len(H):])[h] for x in m))
            for h in range(int(np.prod(H)))]
    return H, flat


def _multi(alignment, tree, model, want, precision, engine, f81=False, branch_mask='auto',
           irrev_convention='reference'):
    """One sb_counts_multi / sb_branch_counts_multi / sb_inside_outside_multi call
    for a model list (C, 1) or a model grid (1 or C, R).  ``want``: subset of
    {'counts', 'root', 'll', 'bcounts'} or of {'npost', 'bpost', 'll'}."""
    eng = (engine or default_engine()).new_call()
    exact = _exact(irrev_convention)
    if _is_model_grid(model):
        table = [list(row) for row in model]
    else:
        table = [[m] for m in model]
    cache = {}

    def mid(m):
        k = id(m)
        if k not in cache:
            cache[k] = (m, eng.model_id(m, exact))
        return cache[k][1]

    ids = np.ascontiguousarray([[mid(m) for m in row] for row in table], dtype=np.int32)
    m0 = ensure_diag(table[0][0])
    A = len(m0.pi)
    aln = eng.alignment(alignment, A)
    R, C = aln.shape
    n_cols, n_rows = ids.shape
    if _is_model_grid(model):
        if n_cols not in (1, C) or n_rows != R:
            raise ValueError(f'model grid must be (1, {R}) or ({C}, {R}), got ({n_cols}, {n_rows})')
    elif n_cols != C:
        raise ValueError(f'model list has {n_cols} models for {C} columns')
    parent, dist = eng.tree_arrays(tree, R)
    dt, tdt = _PREC[precision]
    mode, mk, mptr = _mask_args(branch_mask, eng, R, C)
    dev = eng.device
    out = {}
    if 'counts' in want:
        out['counts'] = torch.empty((A, A, C), dtype=tdt, device=dev)
    if 'root' in want:
        out['root'] = torch.empty((A, C), dtype=tdt, device=dev)
    if 'll' in want:
        out['ll'] = torch.empty(C, dtype=torch.float64, device=dev)
    if 'bcounts' in want:
        out['bcounts'] = torch.empty((R, A, A, C), dtype=tdt, device=dev)
    if 'npost' in want:
        out['npost'] = torch.empty((R, A, C), dtype=tdt, device=dev)
    if 'bpost' in want:
        out['bpost'] = torch.empty((R, A, A, C), dtype=tdt, device=dev)
    ptr = lambda k: out[k].data_ptr() if k in out else None  # noqa: E731
    if 'npost' in want or 'bpost' in want:
        if n_rows != 1:
            raise NotImplementedError('InsideOutside does not support model grids (2D model arrays); '
                                      'use LogLike, Counts, BranchCounts and RootProb directly')
        _lib.check(eng.lib.sb_inside_outside_multi(
            eng.ctx, n_cols, ids.ctypes.data, aln.data_ptr(), parent.ctypes.data, dist.ctypes.data,
            R, C, dt, ptr('npost'), ptr('bpost'), ptr('ll'), _stream_ptr(dev)))
    elif 'bcounts' in want:
        _lib.check(eng.lib.sb_branch_counts_multi(
            eng.ctx, n_cols, n_rows, ids.ctypes.data, aln.data_ptr(), parent.ctypes.data,
            dist.ctypes.data, R, C, dt, 1 if f81 else 0, mode, mptr, ptr('bcounts'), ptr('ll'),
            _stream_ptr(dev)))
    else:
        _lib.check(eng.lib.sb_counts_multi(
            eng.ctx, n_cols, n_rows, ids.ctypes.data, aln.data_ptr(), parent.ctypes.data,
            dist.ctypes.data, R, C, dt, 1 if f81 else 0, mode, mptr, ptr('counts'), ptr('root'),
            ptr('ll'), _stream_ptr(dev)))
    return out


def _stack_masks(branch_mask, H, R, C):
    """Per-model branch masks for H stacked models: "auto" / None apply to every
    model; an explicit mask is (R, C) (shared) or (*H, R, C) (one slice per model,
    subby/jax/__init__.py:355-356), or any shape that broadcasts to (*H, R, C)."""
    n = int(np.prod(H))
    if branch_mask is None or isinstance(branch_mask, str):
        return [branch_mask] * n
    mk = branch_mask if isinstance(branch_mask, torch.Tensor) else torch.from_numpy(np.asarray(branch_mask))
    full = tuple(H) + (R, C)
    if mk.dim() < 2 or tuple(mk.shape[-2:]) != (R, C) or mk.dim() > len(full):
        raise ValueError(f'branch_mask for {tuple(H)} stacked models must have shape ({R}, {C}) or '
                         f'{full}, got {tuple(mk.shape)}')
    try:
        mk = torch.broadcast_to(mk, full)
    except RuntimeError:
        raise ValueError(f'branch_mask of shape {tuple(mk.shape)} does not broadcast to {full}') from None
    mk = mk.reshape((n, R, C))
    return [mk[h] for h in range(n)]


def _stack_call(alignment, tree, stack, want, precision, engine, f81=False, branch_mask='auto'):
    """(*H, ...) outputs of H stacked models over one alignment: one persistent
    launch with one block per model, all reading the same device alignment;
    model h uses slice h of a (*H, R, C) explicit branch mask."""
    from .batch import Block
    H, models = stack
    eng = (engine or default_engine()).new_call()
    A = len(ensure_diag(models[0]).pi)
    aln = eng.alignment(alignment, A)
    R, C = aln.shape
    masks = _stack_masks(branch_mask, H, R, C)
    blocks = [Block(aln, tree, m, want=want, f81_fast=f81, branch_mask=mk) for m, mk in zip(models, masks)]
    eng.launcher().submit(blocks, precision=precision)
    return {w: torch.stack([getattr(b, w) for b in blocks]).reshape(H + getattr(blocks[0], w).shape)
            for w in want}


def LogLike(alignment, tree, model, maxChunkSize: int = 128, precision: str = 'fp32',
            engine: Engine | None = None) -> torch.Tensor:
    """Per-column log-likelihoods, (C,) float64 (subby/jax/__init__.py:207-291).
    ``model``: one model, a list of C per-column models, a model grid
    ([[m_0 .. m_{R-1}]] per branch, or C such rows), or models stacked along
    leading axes (result (*H, C))."""
    if _is_model_list(model):
        return _multi(alignment, tree, model, {'ll'}, precision, engine)['ll']
    st = _stacked(model)
    if st is not None:
        return _stack_call(alignment, tree, st, ('loglike',), precision, engine)['loglike']
    eng, m, A, aln, R, C, parent, dist = _prep(alignment, tree, model, engine, strip=True)
    dt, _ = _PREC[precision]
    out = torch.empty(C, dtype=torch.float64, device=eng.device)
    _lib.check(eng.lib.sb_loglike(eng.ctx, eng.model_id(m), aln.data_ptr(), parent.ctypes.data,
                                  dist.ctypes.data, R, C, dt, out.data_ptr(), _stream_ptr(eng.device)))
    return out


def RootProb(alignment, tree, model, maxChunkSize: int = 128, precision: str = 'fp32',
             engine: Engine | None = None) -> torch.Tensor:
    """Posterior root distribution, (A, C) (subby/jax/__init__.py:543-650); model
    lists, grids and stacked models as in LogLike."""
    if _is_model_list(model):
        return _multi(alignment, tree, model, {'root'}, precision, engine)['root']
    st = _stacked(model)
    if st is not None:
        return _stack_call(alignment, tree, st, ('root',), precision, engine)['root']
    eng, m, A, aln, R, C, parent, dist = _prep(alignment, tree, model, engine, strip=True)
    dt, tdt = _PREC[precision]
    out = torch.empty((A, C), dtype=tdt, device=eng.device)
    _lib.check(eng.lib.sb_root_prob(eng.ctx, eng.model_id(m), aln.data_ptr(), parent.ctypes.data,
                                    dist.ctypes.data, R, C, dt, out.data_ptr(), None,
                                    _stream_ptr(eng.device)))
    return out


def _mask_args(branch_mask, eng, R, C):
    if branch_mask is None:
        return _lib.SB_MASK_NONE, None, None
    if isinstance(branch_mask, str):
        if branch_mask != 'auto':
            raise ValueError(f'branch_mask must be "auto", None or an (R, C) bool array, got {branch_mask!r}')
        return _lib.SB_MASK_AUTO, None, None
    mk = branch_mask
    if isinstance(mk, torch.Tensor):
        mk = mk.to(eng.device)
    else:
        mk = torch.from_numpy(np.asarray(mk)).to(eng.device)
    if tuple(mk.shape) != (R, C):
        raise ValueError(f'branch_mask must have shape ({R}, {C}), got {tuple(mk.shape)}')
    mk = (mk != 0).to(torch.uint8).contiguous()
    return _lib.SB_MASK_EXPLICIT, mk, mk.data_ptr()


def Counts(alignment, tree, model, maxChunkSize: int = 128, f81_fast_flag: bool = False,
           branch_mask='auto', precision: str = 'fp32', engine: Engine | None = None,
           return_loglike: bool = False, return_root: bool = False,
           irrev_convention: str = 'reference'):
    """Expected substitution counts (off-diagonal) and dwell times (diagonal), (A, A, C)
    (subby/jax/__init__.py:338-447).  ``branch_mask``: "auto" (Steiner mask of
    the ungapped leaves), None, or an (R, C) bool array.  For an IrrevDiagModel
    ``irrev_convention`` is "reference" (the
    reference's formulas, eigensub.py:271-434) or "exact" (expected counts
    under Q itself).  Model lists, grids and stacked models as in LogLike."""
    if _is_model_list(model):
        want = {'counts'} | ({'ll'} if return_loglike else set()) | ({'root'} if return_root else set())
        out = _multi(alignment, tree, model, want, precision, engine, f81_fast_flag, branch_mask,
                     irrev_convention)
        res = [out['counts']] + ([out['ll']] if return_loglike else []) + ([out['root']] if return_root else [])
        return tuple(res) if len(res) > 1 else res[0]
    st = _stacked(model)
    if st is not None:
        want = ('counts',) + (('loglike',) if return_loglike else ()) + (('root',) if return_root else ())
        out = _stack_call(alignment, tree, st, want, precision, engine, f81_fast_flag, branch_mask)
        res = [out[w] for w in want]
        return tuple(res) if len(res) > 1 else res[0]
    eng, m, A, aln, R, C, parent, dist = _prep(alignment, tree, model, engine, strip=True)
    dt, tdt = _PREC[precision]
    mode, mk, mptr = _mask_args(branch_mask, eng, R, C)
    counts = torch.empty((A, A, C), dtype=tdt, device=eng.device)
    ll = torch.empty(C, dtype=torch.float64, device=eng.device) if return_loglike else None
    root = torch.empty((A, C), dtype=tdt, device=eng.device) if return_root else None
    _lib.check(eng.lib.sb_counts(eng.ctx, eng.model_id(m, _exact(irrev_convention)), aln.data_ptr(),
                                 parent.ctypes.data,
                                 dist.ctypes.data, R, C, dt, 1 if f81_fast_flag else 0, mode, mptr,
                                 counts.data_ptr(), root.data_ptr() if root is not None else None,
                                 ll.data_ptr() if ll is not None else None, _stream_ptr(eng.device)))
    if return_loglike or return_root:
        res = [counts]
        if return_loglike:
            res.append(ll)
        if return_root:
            res.append(root)
        return tuple(res)
    return counts


def _exact(irrev_convention):
    if irrev_convention not in ('reference', 'exact'):
        raise ValueError(f'irrev_convention must be "reference" or "exact", got {irrev_convention!r}')
    return irrev_convention == 'exact'


def _inside_outside(alignment, tree, model, want, f81_fast=False, branch_mask=None, precision='fp32',
                    engine=None, irrev_convention='reference'):
    """One sb_inside_outside call; ``want`` is a set of 'bcounts', 'npost', 'bpost', 'll'."""
    eng, m, A, aln, R, C, parent, dist = _prep(alignment, tree, model, engine)
    dt, tdt = _PREC[precision]
    mode, mk, mptr = _mask_args(branch_mask, eng, R, C)
    dev = eng.device
    out = {}
    if 'bcounts' in want:
        out['bcounts'] = torch.empty((R, A, A, C), dtype=tdt, device=dev)
    if 'npost' in want:
        out['npost'] = torch.empty((R, A, C), dtype=tdt, device=dev)
    if 'bpost' in want:
        out['bpost'] = torch.empty((R, A, A, C), dtype=tdt, device=dev)
    if 'll' in want:
        out['ll'] = torch.empty(C, dtype=torch.float64, device=dev)
    ptr = lambda k: out[k].data_ptr() if k in out else None  # noqa: E731
    _lib.check(eng.lib.sb_inside_outside(
        eng.ctx, eng.model_id(m, _exact(irrev_convention)), aln.data_ptr(), parent.ctypes.data, dist.ctypes.data, R, C, dt,
        1 if f81_fast else 0, mode, mptr, ptr('bcounts'), ptr('npost'), ptr('bpost'), ptr('ll'),
        _stream_ptr(dev)))
    return out


def BranchCounts(alignment, tree, model, maxChunkSize: int = 128, f81_fast_flag: bool = False,
                 branch_mask=None, precision: str = 'fp32', engine: Engine | None = None,
                 irrev_convention: str = 'reference'):
    """Per-branch counts and dwell times, (R, A, A, C), branch 0 zeros
    (subby/jax/__init__.py:450-540).  subby.jax applies no branch mask here; the
    default ``branch_mask=None`` does the same ("auto" or an (R, C) array as in
    Counts).  Any alphabet up to 64 (above 8 states: the large-alphabet kernel's
    per-branch counts variant).  Model lists and grids as in LogLike; stacked
    models give (*H, R, A, A, C)."""
    if _is_model_list(model):
        return _multi(alignment, tree, model, {'bcounts'}, precision, engine, f81_fast_flag,
                      branch_mask, irrev_convention)['bcounts']
    st = _stacked(model)
    if st is not None:
        H, flat = st
        R, C = np.shape(alignment)
        masks = _stack_masks(branch_mask, H, R, C)
        outs = [_inside_outside(alignment, tree, m, {'bcounts'}, f81_fast_flag, mk, precision,
                                engine, irrev_convention)['bcounts'] for m, mk in zip(flat, masks)]
        return torch.stack(outs).reshape(tuple(H) + tuple(outs[0].shape))
    return _inside_outside(alignment, tree, model, {'bcounts'}, f81_fast_flag, branch_mask,
                           precision, engine, irrev_convention)['bcounts']


class InsideOutside:
    """Posterior queries on the inside-outside tables (subby/jax/__init__.py:681-1040).

    The reference stores U and D once and answers queries from them; here each
    query runs the fused GPU passes again (the tables never leave the chip),
    and results are torch tensors on the GPU.  Every query takes any
    alphabet up to 64 states.  ``model`` is one model or a per-column model
    list (C models); as in the reference, a list's passes run under each
    column's matrices with the root distribution of the first model, the root
    row of ``node_posterior`` uses each column's own, and ``counts`` /
    ``branch_counts`` are those of Counts / BranchCounts of the list.  Stacked
    models (leading *H axes) give every query with those leading axes.  Model
    grids raise NotImplementedError, as in the reference.
    """

    def __init__(self, alignment, tree, model, maxChunkSize: int = 128, precision: str = 'fp32',
                 engine: Engine | None = None):
        self._eng = engine or default_engine()
        if _is_model_grid(model):
            raise NotImplementedError('InsideOutside does not support model grids (2D model arrays); '
                                      'use LogLike, Counts, BranchCounts and RootProb directly')
        self._list = _is_model_list(model)
        self._stack = _stacked(model)
        if self._stack is not None:
            A = len(self._stack[1][0].pi)
            self._args = (self._eng.alignment(alignment, A), tree, model)
        elif self._list:
            A = len(ensure_diag(model[0]).pi)
            self._args = (self._eng.alignment(alignment, A), tree, list(model))
        else:
            m = ensure_diag(model)
            self._args = (self._eng.alignment(alignment, len(m.pi)), tree, m)
        self._precision = precision
        self._ll = None
        self._npost = None
        self._bpost = None

    @property
    def log_likelihood(self):
        """Per-column log-likelihoods, (C,) float64."""
        if self._ll is None:
            self._ll = LogLike(*self._args, precision=self._precision, engine=self._eng)
        return self._ll

    def counts(self, f81_fast_flag: bool = False, branch_mask='auto',
               irrev_convention: str = 'reference'):
        """(A, A, C) counts (as Counts)."""
        return Counts(*self._args, f81_fast_flag=f81_fast_flag, branch_mask=branch_mask,
                      precision=self._precision, engine=self._eng,
                      irrev_convention=irrev_convention)

    def branch_counts(self, f81_fast_flag: bool = False):
        """(R, A, A, C) per-branch counts, unmasked (as subby.jax)."""
        return BranchCounts(*self._args, f81_fast_flag=f81_fast_flag, branch_mask=None,
                            precision=self._precision, engine=self._eng)

    def _posteriors(self, want=('npost', 'bpost')):
        """Compute the requested posteriors ('npost', 'bpost') that are not cached
        yet, in one launch (each query asks only for its own output: the branch
        joints are R A^2 C values)."""
        want = {w for w in want if getattr(self, '_' + w) is None}
        if not want:
            return
        if self._stack is not None:
            H, flat = self._stack
            outs = [_inside_outside(self._args[0], self._args[1], m, want,
                                    precision=self._precision, engine=self._eng) for m in flat]
            out = {w: torch.stack([o[w] for o in outs]).reshape(tuple(H) + tuple(outs[0][w].shape))
                   for w in want}
        elif self._list:
            out = _multi(*self._args, want, self._precision, self._eng)
        else:
            out = _inside_outside(*self._args, want, precision=self._precision, engine=self._eng)
        for w in want:
            setattr(self, '_' + w, out[w])

    def node_posterior(self, node: int | None = None):
        """(A, C) posterior of one node, or (R, A, C) for all nodes (row 0: root)."""
        self._posteriors(('npost',))
        return self._npost if node is None else self._npost[..., node, :, :]

    def branch_posterior(self, node: int | None = None):
        """(A, A, C) joint posterior P(x_parent = i, x_node = j) of branch ``node``
        (> 0), or (R, A, A, C) for all branches (branch 0 zeros)."""
        self._posteriors(('bpost',))
        return self._bpost if node is None else self._bpost[..., node, :, :, :]


def MixturePosterior(alignment, tree, models, log_weights, maxChunkSize: int = 128,
                     precision: str = 'fp32', engine: Engine | None = None,
                     return_loglikes: bool = False):
    """(K, C) float64 posterior over mixture components per column,
    softmax_k(LogLike_k + log_weights[k]) (subby/jax/__init__.py:653-678,
    mixture.py:6-30).  ``models``: a list of K models or models stacked along
    one leading axis (e.g. scale_model(base, gamma rates)); K <= 64."""
    eng = (engine or default_engine()).new_call()
    st = None if _is_model_list(models) else _stacked(models)
    ms = st[1] if st is not None else list(models)
    K = len(ms)
    lw = np.ascontiguousarray(np.asarray(log_weights, dtype=np.float64).reshape(-1))
    if lw.shape != (K,):
        raise ValueError(f'log_weights must have shape ({K},), got {lw.shape}')
    ids = np.ascontiguousarray([eng.model_id(m) for m in ms], dtype=np.int32)
    A = len(ensure_diag(ms[0]).pi)
    aln = eng.alignment(alignment, A)
    R, C = aln.shape
    parent, dist = eng.tree_arrays(tree, R)
    dt, _ = _PREC[precision]
    post = torch.empty((K, C), dtype=torch.float64, device=eng.device)
    ll = torch.empty((K, C), dtype=torch.float64, device=eng.device) if return_loglikes else None
    _lib.check(eng.lib.sb_mixture_posterior(
        eng.ctx, K, ids.ctypes.data, aln.data_ptr(), parent.ctypes.data, dist.ctypes.data, R, C, dt,
        lw.ctypes.data, post.data_ptr(), ll.data_ptr() if ll is not None else None,
        _stream_ptr(eng.device)))
    return (post, ll) if return_loglikes else post


def loglike_grad(alignment, tree, model, g=None, rate_weights=None, precision: str = 'fp32',
                 engine: Engine | None = None):
    """d(sum_c g_c logL_c)/d(distanceToParent) (R,) and, for each Fisher weight
    matrix W in ``rate_weights``, sum_ij W_ij Counts_ij (unmasked): the
    derivative w.r.t. that rate parameter (subby/jax/vjp.py:62-133 + HR2002).

    Returns (dist_grad (R,) f64, rate_grad (n_rate,) f64, logLike (C,) f64)."""
    R_full = len(np.asarray(_to_np(tree.parentIndex)))
    eng, m, A, aln, R, C, parent, dist = _prep(alignment, tree, model, engine, strip=True)
    dt, _ = _PREC[precision]
    dev = eng.device
    gt = None
    if g is not None:
        gt = torch.as_tensor(g, dtype=torch.float64, device=dev).contiguous()
        if gt.shape != (C,):
            raise ValueError(f'g must have shape ({C},)')
    W = np.zeros((0, A, A)) if rate_weights is None else np.asarray(rate_weights, dtype=np.float64)
    if W.ndim == 2:
        W = W[None]
    W = np.ascontiguousarray(W)
    nr = W.shape[0]
    dgrad = torch.empty(R, dtype=torch.float64, device=dev)
    rgrad = torch.empty(max(nr, 1), dtype=torch.float64, device=dev)
    ll = torch.empty(C, dtype=torch.float64, device=dev)
    _lib.check(eng.lib.sb_loglike_grad(
        eng.ctx, eng.model_id(m), aln.data_ptr(), parent.ctypes.data, dist.ctypes.data, R, C, dt,
        gt.data_ptr() if gt is not None else None, dgrad.data_ptr(), nr,
        W.ctypes.data if nr else None, rgrad.data_ptr() if nr else None, ll.data_ptr(),
        _stream_ptr(dev)))
    if R < R_full:   # dummy leaves of a padded tree: M(t) 1 = 1, derivative 0
        dgrad = torch.cat([dgrad, torch.zeros(R_full - R, dtype=dgrad.dtype, device=dev)])
    return dgrad, rgrad[:nr], ll
