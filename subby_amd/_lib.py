"""ctypes binding of libsubby_amd.so (the C ABI in include/subby_amd.h).

The library is built in-tree by ``subby_amd.build``.  Nothing here falls back
to a CPU path: a missing library or a missing GPU raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libsubby_amd.so')

SB_OK = 0
SB_F32, SB_F64 = 0, 1
SB_MASK_NONE, SB_MASK_AUTO, SB_MASK_EXPLICIT = 0, 1, 2
SB_WANT_LOGLIKE, SB_WANT_ROOT, SB_WANT_COUNTS = 1, 2, 4
SB_TUNE_WAVES_PER_GROUP, SB_TUNE_WAVES_PER_CU, SB_TUNE_GEN2_ROUND, SB_TUNE_SPLIT = 0, 1, 3, 4
SB_TUNE_LAST_SPLIT, SB_TUNE_SPLIT_MIN = 5, 6

# Every exported symbol of include/subby_amd.h.
EXPORTS = [
    'sb_last_error', 'sb_version', 'sb_create', 'sb_destroy', 'sb_model_put',
    'sb_model_drop', 'sb_loglike', 'sb_root_prob', 'sb_counts', 'sb_loglike_grad',
    'sb_batch_create', 'sb_batch_destroy', 'sb_batch_submit', 'sb_batch_set_tiling',
    'sb_batch_sync', 'sb_batch_set_timing', 'sb_batch_kernel_ms', 'sb_counts_plain',
    'sb_server_create', 'sb_server_destroy', 'sb_server_submit', 'sb_server_poll',
    'sb_server_drain', 'sb_server_replay', 'sb_inside_outside', 'sb_model_put_irrev',
    'sb_counts_multi', 'sb_branch_counts_multi', 'sb_inside_outside_multi', 'sb_mixture_posterior',
    'sb_arena_stats', 'sb_batch_set_tune', 'sb_batch_get_tune', 'sb_counts_noalloc',
    'sb_server_replay_timed', 'sb_context_set_tune',
]


class SbBlock(ctypes.Structure):
    """sb_block of include/subby_amd.h."""
    _fields_ = [
        ('model_id', ctypes.c_int),
        ('R', ctypes.c_int),
        ('C', ctypes.c_int),
        ('d_alignment', ctypes.c_void_p),
        ('parent_index', ctypes.c_void_p),
        ('distances', ctypes.c_void_p),
        ('want', ctypes.c_int),
        ('f81_fast', ctypes.c_int),
        ('mask_mode', ctypes.c_int),
        ('d_mask', ctypes.c_void_p),
        ('d_counts', ctypes.c_void_p),
        ('d_root', ctypes.c_void_p),
        ('d_loglike', ctypes.c_void_p),
    ]


class SbRequest(ctypes.Structure):
    """sb_request of include/subby_amd.h."""
    _fields_ = [
        ('model_id', ctypes.c_int),
        ('A', ctypes.c_int),
        ('eigenvalues', ctypes.c_void_p),
        ('eigenvectors', ctypes.c_void_p),
        ('pi', ctypes.c_void_p),
        ('R', ctypes.c_int),
        ('C', ctypes.c_int),
        ('alignment', ctypes.c_void_p),
        ('parent_index', ctypes.c_void_p),
        ('distances', ctypes.c_void_p),
        ('mask_mode', ctypes.c_int),
        ('want', ctypes.c_int),
        ('counts', ctypes.c_void_p),
        ('root', ctypes.c_void_p),
        ('loglike', ctypes.c_void_p),
        ('tag', ctypes.c_longlong),
    ]


class SbRequestTiming(ctypes.Structure):
    """sb_request_timing of include/subby_amd.h."""
    _fields_ = [(f, ctypes.c_double) for f in (
        'admitted_s', 'enqueued_s', 'gpu_done_s', 'done_s', 'upload_ms', 'kernel_ms', 'download_ms',
        'plan_ms', 'models_ms')] + [
        ('batch_columns', ctypes.c_longlong),
        ('batch_requests', ctypes.c_int),
        ('lane', ctypes.c_int),
    ]


TIMING_DTYPE = np.dtype({
    'names': ['admitted_s', 'enqueued_s', 'gpu_done_s', 'done_s', 'upload_ms', 'kernel_ms',
              'download_ms', 'plan_ms', 'models_ms', 'batch_columns', 'batch_requests', 'lane'],
    'formats': ['<f8'] * 9 + ['<i8', '<i4', '<i4'],
    'offsets': [0, 8, 16, 24, 32, 40, 48, 56, 64, 72, 80, 84],
    'itemsize': 88})
assert TIMING_DTYPE.itemsize == ctypes.sizeof(SbRequestTiming)


class SubbyError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f'subby_amd error {code}: {msg}')
        self.code = code


_lib = None


def load(path=LIB_PATH):
    """Load the shared library and declare signatures (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f'{path} is missing: build it with `python -m subby_amd.build` '
            '(hipcc --offload-arch=gfx950); there is no CPU fallback')
    lib = ctypes.CDLL(path)
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER
    lib.sb_last_error.restype = ctypes.c_char_p
    lib.sb_last_error.argtypes = []
    lib.sb_version.restype = i
    lib.sb_create.argtypes = [P(vp), i]
    lib.sb_destroy.argtypes = [vp]
    lib.sb_model_put.argtypes = [vp, i, vp, vp, vp, P(i)]
    lib.sb_model_put_irrev.argtypes = [vp, i, vp, vp, vp, vp, i, P(i)]
    lib.sb_model_drop.argtypes = [vp, i]
    lib.sb_arena_stats.argtypes = [vp, vp, vp, vp]
    lib.sb_loglike.argtypes = [vp, i, vp, vp, vp, i, i, i, vp, vp]
    lib.sb_root_prob.argtypes = [vp, i, vp, vp, vp, i, i, i, vp, vp, vp]
    lib.sb_counts.argtypes = [vp, i, vp, vp, vp, i, i, i, i, i, vp, vp, vp, vp, vp]
    lib.sb_counts_noalloc.argtypes = [vp, i, vp, vp, vp, i, i, i, i, i, vp, vp, vp, vp, vp]
    lib.sb_loglike_grad.argtypes = [vp, i, vp, vp, vp, i, i, i, vp, vp, i, vp, vp, vp, vp]
    lib.sb_batch_create.argtypes = [vp, vp, P(vp)]
    lib.sb_batch_destroy.argtypes = [vp]
    lib.sb_batch_submit.argtypes = [vp, i, vp, i, i]
    lib.sb_batch_set_tiling.argtypes = [vp, i, i]
    lib.sb_batch_set_tune.argtypes = [vp, i, i]
    lib.sb_batch_get_tune.argtypes = [vp, i, P(i)]
    lib.sb_context_set_tune.argtypes = [vp, i, i]
    lib.sb_batch_sync.argtypes = [vp]
    lib.sb_batch_set_timing.argtypes = [vp, i]
    lib.sb_batch_kernel_ms.argtypes = [vp, P(d), P(i)]
    lib.sb_counts_plain.argtypes = [vp, i, vp, vp, vp, i, i, i, vp, vp, vp, vp]
    lib.sb_inside_outside.argtypes = [vp, i, vp, vp, vp, i, i, i, i, i, vp, vp, vp, vp, vp, vp]
    lib.sb_counts_multi.argtypes = [vp, i, i, vp, vp, vp, vp, i, i, i, i, i, vp, vp, vp, vp, vp]
    lib.sb_branch_counts_multi.argtypes = [vp, i, i, vp, vp, vp, vp, i, i, i, i, i, vp, vp, vp, vp]
    lib.sb_inside_outside_multi.argtypes = [vp, i, vp, vp, vp, vp, i, i, i, vp, vp, vp, vp]
    lib.sb_mixture_posterior.argtypes = [vp, i, vp, vp, vp, vp, i, i, i, vp, vp, vp, vp]
    ll = ctypes.c_longlong
    lib.sb_server_create.argtypes = [vp, i, ll, i, ll, P(vp)]
    lib.sb_server_destroy.argtypes = [vp]
    lib.sb_server_submit.argtypes = [vp, i, vp]
    lib.sb_server_poll.argtypes = [vp, i, vp, vp, P(i)]
    lib.sb_server_drain.argtypes = [vp]
    lib.sb_server_replay.argtypes = [vp, i, vp, vp, vp]
    lib.sb_server_replay_timed.argtypes = [vp, i, vp, vp, vp]
    for name in EXPORTS:
        if name not in ('sb_last_error',):
            getattr(lib, name).restype = i
    _lib = lib
    return lib


def check(rc):
    if rc != SB_OK:
        msg = _lib.sb_last_error().decode(errors='replace') if _lib else '?'
        raise SubbyError(rc, msg)
    return rc
