"""Convert the reference's golden fixtures (JSON) into compact .npz files.

Reads the six data files of the reference's own test fixtures,
``tests/golden/*.json`` (written there by ``scripts/generate_golden_tests.py``),
and stores every array they hold, unchanged, as float64 / int arrays in
``tests/golden/<case>.npz`` here.  Only data is carried over: inputs,
intermediates and outputs.  Run once in a container that has the reference
checkout; the .npz files are committed.

    python tests/golden/convert_reference_goldens.py [/root/reference]
"""
import json
import os
import sys

import numpy as np

CASES = ['5node_jc4', '7node_hky85_4', '7node_jc64', '7node_jc4_mixed_gaps',
         '7node_jc4_all_gaps', '7node_mixture_jc4']


def flatten(d, prefix=''):
    out = {}
    for k, v in d.items():
        key = f'{prefix}{k}'
        if isinstance(v, dict):
            out.update(flatten(v, key + '.'))
        else:
            out[key] = np.asarray(v)
    return out


def main(ref_root='/root/reference'):
    here = os.path.dirname(os.path.abspath(__file__))
    for case in CASES:
        with open(os.path.join(ref_root, 'tests', 'golden', case + '.json')) as f:
            d = json.load(f)
        arrs = flatten(d)
        np.savez_compressed(os.path.join(here, case + '.npz'), **arrs)
        print(case, {k: v.shape for k, v in arrs.items()})


if __name__ == '__main__':
    main(*sys.argv[1:])
