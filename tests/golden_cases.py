"""Loader for the reference's golden fixtures (tests/golden/*.npz)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
CASES = ['5node_jc4', '7node_hky85_4', '7node_jc64', '7node_jc4_mixed_gaps', '7node_jc4_all_gaps']


def load(case):
    g = np.load(os.path.join(GOLDEN, case + '.npz'))
    return {k: g[k] for k in g.files}
