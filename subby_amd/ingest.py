"""Ingestion for the served pipeline: Newick guide trees, MAF blocks, per-block
tree pruning and alignment expansion, and the 20-feature output format.

Mirrors the data path of the reference's batch driver
(scripts/batch_eigensub_maf.py:67-321, 423-698) and its Newick parser
(subby/formats.py:63-225).  The driver there buckets blocks by pruned-tree
size, pads columns to 128 and runs jax.vmap(Counts) per bucket; here every
block, whatever its tree, goes into one persistent launch (no bucketing, no
padding), one download per launch brings the counts back, and the 20 features
of each block are formed from them on the host.
"""
from __future__ import annotations

import ctypes
import gzip
import os
import time

import numpy as np
import torch

from . import _lib
from .models import Tree

FEATURE_DIM = 20


# ---------------------------------------------------------------------------
# Newick  (subby/formats.py:63-225)
# ---------------------------------------------------------------------------

def _strip_comments(s: str) -> str:
    """Drop [...] comments outside quoted labels."""
    out, depth, quoted = [], 0, False
    for ch in s:
        if ch == "'" and depth == 0:
            quoted = not quoted
            out.append(ch)
        elif quoted:
            out.append(ch)
        elif ch == '[':
            depth += 1
        elif ch == ']' and depth > 0:
            depth -= 1
        elif depth == 0:
            out.append(ch)
    return ''.join(out)


class _Reader:
    def __init__(self, s):
        self.s, self.i = s, 0

    def skip_ws(self):
        while self.i < len(self.s) and self.s[self.i].isspace():
            self.i += 1

    def peek(self):
        self.skip_ws()
        return self.s[self.i] if self.i < len(self.s) else None

    def take(self, ch):
        if self.peek() != ch:
            raise ValueError(f"Newick: expected {ch!r} at position {self.i}")
        self.i += 1

    def label(self):
        if self.peek() == "'":
            self.i += 1
            parts = []
            while self.i < len(self.s):
                ch = self.s[self.i]
                self.i += 1
                if ch == "'":
                    if self.i < len(self.s) and self.s[self.i] == "'":   # '' is a literal quote
                        parts.append("'")
                        self.i += 1
                    else:
                        return ''.join(parts)
                else:
                    parts.append(ch)
            raise ValueError('Newick: unterminated quoted label')
        j = self.i
        while self.i < len(self.s) and self.s[self.i] not in '(),:;':
            self.i += 1
        text = self.s[j:self.i].strip()
        return text or None

    def length(self):
        if self.peek() != ':':
            return 0.0
        self.i += 1
        j = self.i
        while self.i < len(self.s) and self.s[self.i] not in '(),;':
            self.i += 1
        text = self.s[j:self.i].strip()
        return float(text) if text else 0.0


def parse_newick(text: str) -> dict:
    """Newick string -> {'parentIndex' (R,) int32 preorder with root -1,
    'distanceToParent' (R,) float64 with root 0, 'leaf_names', 'node_names'}.
    Children keep their order in the string; comments are ignored; quoted
    labels may contain any character ('' for a quote)."""
    s = _strip_comments(text).strip()
    while s.endswith(';'):
        s = s[:-1].rstrip()
    if not s:
        raise ValueError('Empty Newick string')
    rd = _Reader(s)
    parent, dist, names, leaves = [], [], [], []

    def node(p):
        me = len(parent)
        parent.append(p)
        dist.append(0.0)
        names.append(None)
        kids = 0
        if rd.peek() == '(':
            rd.take('(')
            node(me)
            kids = 1
            while rd.peek() == ',':
                rd.take(',')
                node(me)
                kids += 1
            rd.take(')')
        names[me] = rd.label()
        d = rd.length()
        dist[me] = d if p >= 0 else 0.0
        if kids == 0:
            leaves.append(names[me])

    node(-1)
    if rd.peek() is not None:
        raise ValueError(f'Newick: unexpected text at position {rd.i}')
    return {'parentIndex': np.asarray(parent, dtype=np.int32),
            'distanceToParent': np.asarray(dist, dtype=np.float64),
            'leaf_names': leaves, 'node_names': names}


# ---------------------------------------------------------------------------
# MAF blocks and DNA tokens  (scripts/batch_eigensub_maf.py:67-157)
# ---------------------------------------------------------------------------

def iter_maf_blocks(source):
    """Alignment blocks of a MAF file (path, .gz path, or an iterable of lines):
    dicts with species, seqs, srcs, starts, sizes, strands, src_sizes; species is
    the source name up to its first '.'."""
    if isinstance(source, (str, os.PathLike)):
        path = str(source)
        f = gzip.open(path, 'rt') if path.endswith('.gz') else open(path)
        close = True
    else:
        f, close = source, False
    block = None
    try:
        for line in f:
            line = line.rstrip('\n\r')
            if line.startswith('a'):
                if block is not None and block['species']:
                    yield block
                block = {k: [] for k in ('species', 'seqs', 'srcs', 'starts', 'sizes', 'strands', 'src_sizes')}
            elif line.startswith('s') and block is not None:
                f_ = line.split()
                block['species'].append(f_[1].split('.')[0])
                block['srcs'].append(f_[1])
                block['starts'].append(int(f_[2]))
                block['sizes'].append(int(f_[3]))
                block['strands'].append(f_[4])
                block['src_sizes'].append(int(f_[5]))
                block['seqs'].append(f_[6])
        if block is not None and block['species']:
            yield block
    finally:
        if close:
            f.close()


_DNA = np.full(256, 4, dtype=np.int8)          # anything else: ungapped unobserved (N)
for _c, _v in (('A', 0), ('C', 1), ('G', 2), ('T', 3)):
    _DNA[ord(_c)] = _v
    _DNA[ord(_c.lower())] = _v
_DNA[ord('-')] = 5
_DNA[ord('.')] = 5


def seqs_to_alignment(seqs) -> np.ndarray:
    """Aligned DNA strings -> (R, C) int8 tokens: A,C,G,T = 0..3, '-' and '.' =
    5 (gap), anything else = 4 (unobserved) (batch_eigensub_maf.py:124-157)."""
    if not seqs:
        return np.zeros((0, 0), dtype=np.int8)
    raw = np.frombuffer(''.join(seqs).encode('latin-1'), dtype=np.uint8).reshape(len(seqs), -1)
    return _DNA[raw]


# ---------------------------------------------------------------------------
# Guide-tree pruning and alignment expansion  (batch_eigensub_maf.py:164-291)
# ---------------------------------------------------------------------------

def prune_tree(tree: dict, keep_leaves) -> dict | None:
    """Restrict a parsed tree to the leaves in keep_leaves: absent leaves are
    removed, internal nodes left with one kept child are spliced out (their
    branch lengths added to the child's), and the result is renumbered in
    preorder.  None if fewer than two leaves remain."""
    parent = np.asarray(tree['parentIndex'])
    dist = np.asarray(tree['distanceToParent'], dtype=np.float64)
    names = tree['node_names']
    R = len(parent)
    children = [[] for _ in range(R)]
    for n in range(1, R):
        children[parent[n]].append(n)
    keep_set = set(keep_leaves)
    kept = np.zeros(R, dtype=bool)
    for n in range(R - 1, -1, -1):       # children have larger preorder indices
        if children[n]:
            kept[n] = any(kept[c] for c in children[n])
        else:
            kept[n] = names[n] in keep_set
    if not kept[0]:
        return None
    new_parent, new_dist, new_names, new_leaves = [], [], [], []
    stack = [(0, -1, 0.0)]                # (old node, new parent, length carried down)
    while stack:
        n, p, extra = stack.pop()
        kids = [c for c in children[n] if kept[c]]
        d = dist[n] + extra
        if kids and len(kids) == 1:
            stack.append((kids[0], p, d))
            continue
        me = len(new_parent)
        new_parent.append(p)
        new_dist.append(d if p >= 0 else 0.0)
        new_names.append(names[n])
        if not kids:
            new_leaves.append(names[n])
        for c in reversed(kids):
            stack.append((c, me, 0.0))
    if len(new_leaves) < 2:
        return None
    return {'parentIndex': np.asarray(new_parent, dtype=np.int32),
            'distanceToParent': np.asarray(new_dist, dtype=np.float64),
            'leaf_names': new_leaves, 'node_names': new_names}


def expand_alignment_to_tree(alignment: np.ndarray, species, tree: dict) -> np.ndarray:
    """(n_species, L) tokens -> (R, L): each leaf of the tree gets its species'
    row, every other node the gap token 5 (batch_eigensub_maf.py:270-291)."""
    parent = np.asarray(tree['parentIndex'])
    R = len(parent)
    out = np.full((R, alignment.shape[1]), 5, dtype=np.int8)
    is_parent = np.zeros(R, dtype=bool)
    is_parent[parent[1:]] = True
    row = {sp: i for i, sp in enumerate(species)}
    for n, name in enumerate(tree['node_names']):
        if not is_parent[n] and name in row:
            out[n] = alignment[row[name]]
    return out


# ---------------------------------------------------------------------------
# Features  (batch_eigensub_maf.py:311-321)
# ---------------------------------------------------------------------------

def counts_to_features(counts):
    """(4, 4, L) counts -> (L, 20): the 16 entries of each column's 4 x 4 matrix
    (row-major, dwell times on its diagonal) then the 4 dwell times again."""
    if isinstance(counts, torch.Tensor):
        per = counts.permute(2, 0, 1)
        return torch.cat([per.reshape(per.shape[0], 16), torch.diagonal(per, dim1=1, dim2=2)], dim=1)
    per = np.transpose(counts, (2, 0, 1))
    return np.concatenate([per.reshape(per.shape[0], 16), np.diagonal(per, axis1=1, axis2=2)], axis=1)


# ---------------------------------------------------------------------------
# Driver  (batch_eigensub_maf.py:423-698)
# ---------------------------------------------------------------------------

def precompute_maf(maf_path, output_dir, model, guide_tree: dict, exclude_species=None,
                   min_species: int = 3, min_cols: int = 10, max_cols: int = 2048,
                   launch_cols: int = 1 << 20, engine=None, write: bool = True, stats=None):
    """Counts features of every MAF block on its pruned guide tree.

    Blocks pass the reference's filters (length in [min_cols, max_cols], at
    least min_species species after exclusion, a prunable tree); each keeps its
    own pruned tree, and blocks are launched together, up to launch_cols columns
    per persistent launch, with the Steiner mask "auto".  With ``write`` each
    block goes to ``features_{index:06d}.npz`` (features (L, 20) float32,
    species, n_species, n_species_original, length) as the reference writes
    them.  Returns (manifest, features list); ``stats`` (a dict) receives the
    seconds spent scanning / pruning / expanding, on the GPU launches with the
    feature download, and writing, with the block, column and column·branch
    totals."""
    from .api import default_engine
    eng = (engine or default_engine()).new_call()
    if len(np.asarray(model.pi)) != 4:
        raise ValueError('the MAF driver reads DNA blocks: the model needs A = 4')
    exclude = set(exclude_species or ())
    if write:
        os.makedirs(output_dir, exist_ok=True)
    cache = {}
    recs = []                               # (species, expanded tokens, tree, n_species_original)
    t_scan = time.perf_counter()
    for block in iter_maf_blocks(maf_path):
        seqs = block['seqs']
        L = len(seqs[0]) if seqs else 0
        if L < min_cols or L > max_cols:
            continue
        keep = [i for i, sp in enumerate(block['species']) if sp not in exclude]
        if len(keep) < min_species:
            continue
        species = [block['species'][i] for i in keep]
        key = tuple(sorted(species))
        if key not in cache:
            cache[key] = prune_tree(guide_tree, set(species))
        pruned = cache[key]
        if pruned is None:
            continue
        aln = expand_alignment_to_tree(seqs_to_alignment([seqs[i] for i in keep]), species, pruned)
        recs.append((species, aln, pruned, len(block['species'])))

    manifest, feats = [], []
    launcher = eng.launcher()
    mid = eng.model_id(model)
    A = 4
    i = 0
    t0 = time.perf_counter()
    t_write = 0.0
    while i < len(recs):
        j, cols = i, 0
        while j < len(recs) and (j == i or cols + recs[j][1].shape[1] <= launch_cols):
            cols += recs[j][1].shape[1]
            j += 1
        group = recs[i:j]
        # one upload of all alignments, one counts buffer, one sb_block per MAF block
        flat = np.concatenate([r[1].reshape(-1) for r in group])
        with torch.cuda.stream(launcher.stream):
            dev = torch.from_numpy(flat).to(eng.device, non_blocking=False)
            out = torch.empty(A * A * cols, dtype=torch.float32, device=eng.device)
        arr = (_lib.SbBlock * len(group))()
        keep = []
        off = coff = 0
        for k, (sp, aln, tr, _) in enumerate(group):
            R, L = aln.shape
            par = np.ascontiguousarray(tr['parentIndex'], dtype=np.int32)
            dst = np.ascontiguousarray(tr['distanceToParent'], dtype=np.float64)
            keep += [par, dst]
            b = arr[k]
            b.model_id, b.R, b.C = mid, R, L
            b.d_alignment = dev.data_ptr() + off
            b.parent_index, b.distances = par.ctypes.data, dst.ctypes.data
            b.want, b.mask_mode = _lib.SB_WANT_COUNTS, _lib.SB_MASK_AUTO
            b.d_counts = out.data_ptr() + 4 * coff
            off += R * L
            coff += A * A * L
        _lib.check(eng.lib.sb_batch_submit(launcher.h, len(group), arr, _lib.SB_F32, 0))
        host = out.cpu().numpy()            # one download (stream-ordered after the launch)
        coff = 0
        for k, (sp, aln, tr, n0) in enumerate(group):
            L = aln.shape[1]
            f = counts_to_features(host[coff:coff + A * A * L].reshape(A, A, L))
            coff += A * A * L
            idx = i + k
            entry = {'file': f'features_{idx:06d}.npz', 'n_species': len(sp), 'n_species_original': n0,
                     'length': L}
            if write:
                tw = time.perf_counter()
                np.savez_compressed(os.path.join(output_dir, entry['file']), features=f,
                                    species=np.array(sp), n_species=len(sp), n_species_original=n0, length=L)
                t_write += time.perf_counter() - tw
            manifest.append(entry)
            feats.append(f)
        del keep
        i = j
    if stats is not None:
        stats.update({'scan_s': t0 - t_scan, 'gpu_s': time.perf_counter() - t0 - t_write,
                      'write_s': t_write, 'blocks': len(recs),
                      'columns': int(sum(r[1].shape[1] for r in recs)),
                      'column_branches': int(sum(r[1].shape[1] * (r[1].shape[0] - 1) for r in recs))})
    return manifest, feats
