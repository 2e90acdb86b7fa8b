"""End to end on the GPU: a synthetic MAF on a guide tree through
ingest.precompute_maf (scan, prune, expand, one persistent launch for all
blocks, features on the device) against the FP64 oracle's Counts of each block
on its pruned tree (FP32 counts normwise <= 1e-4 per column)."""
import numpy as np
import pytest
import torch

from oracle import subby_oracle as O
from subby_amd import ingest
from tests.test_host_ingest import _random_newick

pytestmark = pytest.mark.gpu


def test_precompute_maf_matches_oracle(tmp_path):
    import subby_amd as sb
    print('precompute_maf end to end', flush=True)
    eng = sb.Engine(0)
    rng = np.random.default_rng(2024)
    names = [f'sp{i}' for i in range(12)]
    guide = ingest.parse_newick(_random_newick(rng, names))
    m = sb.hky85_diag(2.0, rng.dirichlet(np.full(4, 5.0)))
    lines = ['##maf version=1']
    want = []
    for b in range(25):
        k = int(rng.integers(2, 13))
        sp = list(rng.choice(names, k, replace=False))
        L = int(rng.integers(4, 400))
        pr = ingest.prune_tree(guide, set(sp)) if k >= 2 else None
        t = sb.Tree(pr['parentIndex'], pr['distanceToParent'])
        sim = sb.synthetic.simulate(t, m, L, rng, gap_frac=0.05, ambig_frac=0.02)
        leaf_row = {n: i for i, n in enumerate(pr['node_names'])}
        lines.append(f'a score={b}')
        for s in sp:
            seq = ''.join('ACGTN-'[min(int(x), 5)] if x >= 0 else '-' for x in sim[leaf_row[s]])
            lines.append(f's {s}.chr1 {b * 1000} {L} + 100000 {seq}')
        lines.append('')
        if k >= 3 and 10 <= L <= 2048:
            want.append((sp, L))
    path = tmp_path / 'synthetic.maf'
    path.write_text('\n'.join(lines))
    manifest, feats = ingest.precompute_maf(str(path), str(tmp_path / 'out'), m, guide, engine=eng,
                                            launch_cols=2000)
    assert [e['length'] for e in manifest] == [L for _, L in want]
    om = O.DiagModel(*(np.asarray(x) for x in m))
    worst = 0.0
    blocks = [b for b in ingest.iter_maf_blocks(str(path))
              if len(b['species']) >= 3 and 10 <= len(b['seqs'][0]) <= 2048]
    for b, f, e in zip(blocks, feats, manifest):
        pr = ingest.prune_tree(guide, set(b['species']))
        aln = ingest.expand_alignment_to_tree(ingest.seqs_to_alignment(b['seqs']), b['species'], pr)
        ref = O.Counts(aln, O.Tree(pr['parentIndex'], pr['distanceToParent']), om, branch_mask='auto')
        rf = ingest.counts_to_features(ref)
        num = np.abs(f[:, :16] - rf[:, :16]).max(1)
        den = np.maximum(np.abs(rf[:, :16]).max(1), 1e-300)
        worst = max(worst, float((num / den).max()))
        assert f.shape == (e['length'], 20) and np.array_equal(f[:, 16:], f[:, [0, 5, 10, 15]])
        saved = np.load(tmp_path / 'out' / e['file'])
        assert np.array_equal(saved['features'], f) and int(saved['length']) == e['length']
    print('   worst per-column counts error', f'{worst:.2e}', 'blocks', len(manifest), flush=True)
    assert worst <= 1e-4
