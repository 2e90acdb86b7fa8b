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


def _c_layout():
    """sizeof / offsetof of the ABI structs as the C compiler lays them out."""
    import subprocess
    import tempfile
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "subby_amd.h"
#define F(T, m) printf("%s.%s %zu\n", #T, #m, offsetof(T, m));
int main(void) {
  printf("sb_block %zu\n", sizeof(sb_block));
  printf("sb_request %zu\n", sizeof(sb_request));
  F(sb_block, d_alignment) F(sb_block, want) F(sb_block, d_mask) F(sb_block, d_loglike)
  F(sb_request, eigenvalues) F(sb_request, R) F(sb_request, alignment) F(sb_request, mask_mode)
  F(sb_request, counts) F(sb_request, loglike) F(sb_request, tag)
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c, exe = os.path.join(d, 'l.c'), os.path.join(d, 'l')
        open(c, 'w').write(src)
        subprocess.run(['gcc', '-I', os.path.dirname(HDR), '-o', exe, c], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return dict(line.rsplit(' ', 1) for line in out.strip().splitlines())


def test_struct_layouts_match_the_c_compiler():
    lay = {k: int(v) for k, v in _c_layout().items()}
    assert lay['sb_block'] == ctypes.sizeof(_lib.SbBlock)
    assert lay['sb_request'] == ctypes.sizeof(_lib.SbRequest)
    for name, cls in (('sb_block', _lib.SbBlock), ('sb_request', _lib.SbRequest)):
        for key, off in lay.items():
            if key.startswith(name + '.'):
                assert getattr(cls, key.split('.', 1)[1]).offset == off, key


def test_server_entry_points_reject_bad_arguments():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.sb_server_create(None, 3, 1 << 30, 16, 1 << 20, ctypes.byref(h)) == -1   # SB_ERR_ARG
    assert lib.sb_server_submit(None, 1, None) == -1
    assert lib.sb_server_drain(None) == -1
    assert lib.sb_server_destroy(None) == 0
