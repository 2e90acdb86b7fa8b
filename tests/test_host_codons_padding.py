"""Host-side helpers: the genetic code, GY94 codon model construction and the
padding helpers (no GPU).

The genetic code is checked against the NCBI translation table 1 string,
written in TCAG order (an independent encoding of the same table); GY94 against
the structure that defines it (one-nucleotide neighbours only, detailed
balance, mean rate 1, and for omega = kappa = 1 with uniform pi a constant rate
on the codon graph)."""
import itertools

import numpy as np
import pytest
import torch

import subby_amd as sb
from subby_amd.padding import strip_dummy_leaves

NCBI1 = 'FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG'


def test_genetic_code_matches_ncbi_table_1():
    gc = sb.genetic_code()
    table = {a + b + c: NCBI1[i] for i, (a, b, c) in enumerate(itertools.product('TCAG', repeat=3))}
    assert [table[c] for c in gc['codons']] == gc['amino_acids']
    assert list(np.nonzero(~gc['sense_mask'])[0]) == [48, 50, 56]   # TAA, TAG, TGA
    assert len(gc['sense_codons']) == 61
    aln = np.array([[0, 48, 63, 64, -1, 65]])
    out, names = sb.codon_to_sense(aln)
    assert list(out[0]) == [0, -1, 60, 61, -1, -1] and names[60] == 'TTT'


def test_gy94_structure():
    rng = np.random.default_rng(0)
    pi = rng.dirichlet(np.ones(61) * 5)
    Q = sb.gy94_rate_matrix(0.3, 2.5, pi)
    sc = sb.genetic_code()['sense_codons']
    ham = np.array([[sum(x != y for x, y in zip(a, b)) for b in sc] for a in sc])
    off = ~np.eye(61, dtype=bool)
    assert np.all(Q[off & (ham != 1)] == 0) and np.all(Q[off & (ham == 1)] > 0)
    assert int((Q[off] > 0).sum()) == 526
    np.testing.assert_allclose(Q.sum(1), 0, atol=1e-13)
    np.testing.assert_allclose(-(pi @ np.diag(Q)), 1.0, rtol=1e-13)
    F = pi[:, None] * Q
    np.testing.assert_allclose(F, F.T, atol=1e-14)
    assert sb.check_detailed_balance(Q, pi)
    # omega = kappa = 1, uniform pi: one rate on every edge of the codon graph
    Q1 = sb.gy94_rate_matrix(1.0, 1.0)
    vals = Q1[off & (ham == 1)]
    np.testing.assert_allclose(vals, vals[0], rtol=1e-14)
    m = sb.gy94_model(0.3, 2.5, pi)
    V, mu = m.eigenvectors, m.eigenvalues
    s = np.sqrt(pi)
    np.testing.assert_allclose((V * mu) @ V.T / s[:, None] * s[None, :], Q, atol=1e-12)


@pytest.mark.parametrize('C,bin_size,expected', [(1, 8, 8), (8, 8, 8), (9, 8, 16), (15, 8, 16),
                                                 (1, 128, 128), (127, 128, 128), (129, 128, 256)])
def test_padded_shape(C, bin_size, expected):
    aln = np.zeros((5, C), dtype=np.int32)
    p, C0 = sb.pad_alignment(aln, bin_size=bin_size)
    assert C0 == C and p.shape == (5, expected) and np.all(p[:, C:] == -1)
    pt, c0 = sb.pad_alignment(torch.zeros((5, C), dtype=torch.int8), bin_size=bin_size)
    assert pt.shape == (5, expected) and bool((pt[:, C:] == -1).all())
    assert sb.unpad_columns(p, C0).shape == (5, C)


def test_geom_bins_and_tree_padding():
    bins = [sb.pad_to_geom_bin(n) for n in range(1, 3000)]
    assert all(b >= n for n, b in zip(range(1, 3000), bins)) and bins == sorted(bins)
    assert len(set(bins)) < 60
    t = sb.synthetic.random_tree(5, np.random.default_rng(1), 0.1)
    aln = np.zeros((9, 7), dtype=np.int8)
    pt, pa, R0, C0 = sb.pad_tree_and_alignment(t, aln)
    R, C = pa.shape
    assert (R0, C0) == (9, 7) and R >= 9 and C >= 7 and len(pt.parentIndex) == R
    assert np.all(np.asarray(pt.parentIndex)[9:] == 0) and np.all(np.asarray(pt.distanceToParent)[9:] == 0)
    # the API drops exactly the dummy leaves
    keep = strip_dummy_leaves(np.asarray(pt.parentIndex), np.asarray(pt.distanceToParent), torch.as_tensor(pa))
    assert keep == 9 if R > 9 else keep == R
    # a real leaf under the root with data is kept
    pa2 = pa.copy()
    if R > 9:
        pa2[R - 1, 0] = 2
        assert strip_dummy_leaves(np.asarray(pt.parentIndex), np.asarray(pt.distanceToParent),
                                  torch.as_tensor(pa2)) == R
