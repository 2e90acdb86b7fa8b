"""Host-side branch-mask shapes for stacked models (no GPU needed): an explicit
mask is (R, C) (shared by every model) or (*H, R, C) (slice h for model h,
subby/jax/__init__.py:355-356); anything else raises before a launch."""
import numpy as np
import pytest
import torch

from subby_amd.api import _stack_masks


def test_stack_masks_shapes():
    R, C = 5, 7
    rng = np.random.default_rng(0)
    per = rng.random((2, 3, R, C)) < 0.5
    out = _stack_masks(per, (2, 3), R, C)
    assert len(out) == 6
    for h in range(6):
        assert torch.equal(out[h], torch.as_tensor(per.reshape(6, R, C)[h]))
    shared = _stack_masks(per[0, 0], (2, 3), R, C)
    assert all(torch.equal(x, torch.as_tensor(per[0, 0])) for x in shared)
    row = _stack_masks(per[:, :1], (2, 3), R, C)        # (2, 1, R, C) broadcasts over the second axis
    assert torch.equal(row[4], torch.as_tensor(per[1, 0]))
    assert _stack_masks('auto', (4,), R, C) == ['auto'] * 4
    assert _stack_masks(None, (2,), R, C) == [None, None]


@pytest.mark.parametrize('shape', [(4, 5, 7), (2, 3, 5, 6), (5, 7, 2), (1, 2, 3, 5, 7), (7,)])
def test_stack_masks_reject(shape):
    with pytest.raises(ValueError, match='branch_mask'):
        _stack_masks(np.ones(shape, dtype=bool), (2, 3), 5, 7)
