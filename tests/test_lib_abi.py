"""The C-ABI library loads and exports every entry point include/subby_amd.h declares.

No compute call is made (no GPU needed)."""
import os
import re
import ctypes

import pytest

from subby_amd import _lib

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'include', 'subby_amd.h')


def declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r'^\s*(?:int|const char\*)\s+(sb_\w+)\s*\(', txt, re.M)))


def test_header_declares_the_binding_list():
    assert declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_symbol():
    lib = _lib.load()
    for name in declared():
        assert hasattr(lib, name), name
    assert lib.sb_version() == 1


def test_errors_without_gpu_are_clean():
    """A bad context pointer fails with an error code, not a crash."""
    lib = _lib.load()
    assert lib.sb_model_put(None, 4, None, None, None, None) == -1
    assert b'null' in lib.sb_last_error()
    assert lib.sb_loglike(None, 0, None, None, None, 3, 1, 0, None, None) != 0


def test_sbblock_layout_matches_header():
    # sb_block: 3 ints, 3 pointers, 3 ints, 4 pointers (LP64)
    assert ctypes.sizeof(_lib.SbBlock) == 3 * 4 + 4 + 3 * 8 + 3 * 4 + 4 + 4 * 8
