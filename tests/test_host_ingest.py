"""Ingestion helpers on the host: Newick parsing (the reference's own
known-answer strings, tests/test_formats.py:63-150), MAF block iteration,
DNA tokens, guide-tree pruning (checked by the invariant that fixes it: the
path length between any two kept leaves is unchanged) and the 20-feature
layout."""
import gzip
import io
import itertools

import numpy as np
import pytest
import torch

from subby_amd import ingest


def _patristic(tree):
    parent, dist, names = tree['parentIndex'], tree['distanceToParent'], tree['node_names']
    depth = np.zeros(len(parent))
    anc = [[0] for _ in parent]
    for n in range(1, len(parent)):
        depth[n] = depth[parent[n]] + dist[n]
        anc[n] = anc[parent[n]] + [n]
    leaves = {names[n]: n for n in range(len(parent)) if n not in set(parent[1:])}
    out = {}
    for a, b in itertools.combinations(sorted(leaves), 2):
        x, y = leaves[a], leaves[b]
        lca = max(set(anc[x]) & set(anc[y]), key=lambda n: depth[n])
        out[(a, b)] = depth[x] + depth[y] - 2 * depth[lca]
    return out


def test_newick_known_answers():
    r = ingest.parse_newick("((A:0.1,B:0.2):0.3,C:0.4);")
    assert list(r['parentIndex']) == [-1, 0, 1, 1, 0] and r['leaf_names'] == ['A', 'B', 'C']
    np.testing.assert_allclose(r['distanceToParent'], [0, 0.3, 0.1, 0.2, 0.4])
    r = ingest.parse_newick("((A:0.1,B:0.2)X:0.3,C:0.4)root;")
    assert r['node_names'][0] == 'root' and r['node_names'][1] == 'X'
    assert np.all(ingest.parse_newick("((A,B),C);")['distanceToParent'] == 0)
    r = ingest.parse_newick("A:0.5;")
    assert list(r['parentIndex']) == [-1] and r['leaf_names'] == ['A'] and r['distanceToParent'][0] == 0
    assert ingest.parse_newick("('leaf one':0.1,'it''s':0.2);")['leaf_names'] == ['leaf one', "it's"]
    r = ingest.parse_newick("(A:1.5e-3,B:2.0E+1);")
    np.testing.assert_allclose(r['distanceToParent'][1:], [1.5e-3, 20.0])
    assert ingest.parse_newick("((A[comment]:0.1,B:0.2):0.3,C:0.4);")['leaf_names'] == ['A', 'B', 'C']
    assert len(ingest.parse_newick("(A:0.1,B:0.2,C:0.3);")['parentIndex']) == 4
    assert ingest.parse_newick("(A:0.1,B:0.2)")['leaf_names'] == ['A', 'B']
    assert ingest.parse_newick(" ( (A : 0.1 , B:0.2) : 0.3 ,\n C:0.4 ) ;")['leaf_names'] == ['A', 'B', 'C']
    with pytest.raises(ValueError):
        ingest.parse_newick("")
    with pytest.raises(ValueError):
        ingest.parse_newick("((A,B),C")


def _random_newick(rng, names):
    nodes = [f'{n}:{rng.uniform(0.01, 0.5):.6f}' for n in names]
    while len(nodes) > 1:
        i, j = sorted(rng.choice(len(nodes), 2, replace=False))
        b = nodes.pop(j)
        a = nodes.pop(i)
        nodes.append(f'({a},{b}):{rng.uniform(0.01, 0.5):.6f}')
    return nodes[0] + ';'


@pytest.mark.parametrize('seed', range(5))
def test_prune_tree_keeps_leaf_distances(seed):
    rng = np.random.default_rng(seed)
    names = [f'sp{i}' for i in range(15)]
    tree = ingest.parse_newick(_random_newick(rng, names))
    full = _patristic(tree)
    for k in (2, 3, 7, 15):
        keep = set(rng.choice(names, k, replace=False))
        p = ingest.prune_tree(tree, keep)
        assert sorted(p['leaf_names']) == sorted(keep)
        par = p['parentIndex']
        assert par[0] == -1 and all(0 <= par[n] < n for n in range(1, len(par)))
        cnt = np.bincount(par[1:], minlength=len(par))
        assert set(cnt[cnt > 0]) == {2}                     # still binary
        sub = _patristic(p)
        for pair, d in sub.items():
            assert abs(d - full[pair]) < 1e-9
    assert ingest.prune_tree(tree, {'sp0'}) is None and ingest.prune_tree(tree, set()) is None


MAF = """##maf version=1
a score=10
s human.chr1 100 6 + 1000 ACG-TN
s mouse.chr2 200 5 + 900  AC.GTT

a score=3
s human.chr1 106 3 + 1000 GTa
s dog.chrX   10  3 - 500  GTC
"""


def test_maf_blocks_tokens_and_features(tmp_path):
    blocks = list(ingest.iter_maf_blocks(io.StringIO(MAF)))
    assert [b['species'] for b in blocks] == [['human', 'mouse'], ['human', 'dog']]
    assert blocks[1]['strands'] == ['+', '-'] and blocks[0]['starts'] == [100, 200]
    p = tmp_path / 'x.maf.gz'
    with gzip.open(p, 'wt') as f:
        f.write(MAF)
    assert [b['seqs'] for b in ingest.iter_maf_blocks(str(p))] == [b['seqs'] for b in blocks]
    aln = ingest.seqs_to_alignment(blocks[0]['seqs'])
    assert aln.tolist() == [[0, 1, 2, 5, 3, 4], [0, 1, 5, 2, 3, 3]]
    tree = ingest.parse_newick('((human:0.1,dog:0.2):0.05,(mouse:0.3,rat:0.3):0.1);')
    pr = ingest.prune_tree(tree, {'human', 'mouse'})
    ex = ingest.expand_alignment_to_tree(aln, ['human', 'mouse'], pr)
    assert ex.shape == (3, 6) and np.all(ex[0] == 5)
    assert ex[pr['node_names'].index('human')].tolist() == aln[0].tolist()
    c = np.arange(4 * 4 * 3, dtype=np.float32).reshape(4, 4, 3)
    f = ingest.counts_to_features(c)
    assert f.shape == (3, 20)
    assert f[2].tolist() == c[:, :, 2].reshape(16).tolist() + [c[i, i, 2] for i in range(4)]
    assert torch.equal(ingest.counts_to_features(torch.from_numpy(c)), torch.from_numpy(f))
