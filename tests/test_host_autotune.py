"""Autotuner search spaces (host logic, no GPU)."""
import numpy as np

import subby_amd as sb
from subby_amd.autotune import candidates, family
from subby_amd.batch import Block


def _blk(A, f81=False, irr=False):
    pi = np.full(A, 1.0 / A)
    if irr:
        Q = np.ones((A, A)) + np.eye(A)
        Q[0, 1] += 1.0
        np.fill_diagonal(Q, 0.0)
        np.fill_diagonal(Q, -Q.sum(1))
        m = sb.diagonalize_irreversible(Q, pi)
    else:
        m = sb.jukes_cantor_model(A)
    t = sb.Tree(np.array([-1, 0, 0]), np.array([0.0, 0.1, 0.2]))
    return Block(np.zeros((3, 5), np.int8), t, m, f81_fast=f81)


def test_families_and_spaces():
    assert family([_blk(4)]) == 'small' and family([_blk(8)]) == 'small'
    assert family([_blk(20)]) == 'gen2' and family([_blk(64)]) == 'gen2'
    assert family([_blk(20)], 'fp64') == 'large'
    assert family([_blk(20, f81=True)]) == 'large'
    assert family([_blk(12, irr=True)]) == 'large'
    small = candidates('small')
    assert len(small) == 3 * 6 and all(set(c) == {'waves_per_group', 'waves_per_cu'} for c in small)
    assert [c['gen2_round'] for c in candidates('gen2')][0] == 0
    assert all(set(c) == {'waves_per_cu'} for c in candidates('large'))
