"""Build the gfx950 HIP library in-tree: subby_amd/libsubby_amd.so.

    python -m subby_amd.build        (or subby_amd.build.build())

Each source is compiled to its own object in subby_amd/_build/ (in parallel), then the
objects are linked into the shared library.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'csrc')
OBJ = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libsubby_amd.so')
SOURCES = ['sb_host.cpp', 'sb_serve.cpp', 'sb_fused.hip', 'sb_large.hip', 'sb_f81w.hip', 'sb_gen2.hip', 'sb_post.hip', 'sb_plain.hip', 'sb_multi.hip', 'sb_noalloc.hip',
           'sb_stage.hip']

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = [
    '-O3', '-std=c++17', '-fPIC', '--offload-arch=gfx950',
    # no packed-FP32 instructions in device code
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
]


def _jobs():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16, len(SOURCES)))


def _compile(src, verbose):
    obj = os.path.join(OBJ, os.path.basename(src) + '.o')
    cmd = [HIPCC] + FLAGS + ['-I', os.path.join(HERE, '..', 'include'), '-I', SRC, '-c', src, '-o', obj]
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    return obj, r


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [os.path.join(SRC, s) for s in SOURCES if os.path.exists(os.path.join(SRC, s))]
    with ThreadPoolExecutor(_jobs()) as ex:
        results = list(ex.map(lambda s: _compile(s, verbose), srcs))
    failed = [(o, r) for o, r in results if r.returncode != 0]
    for o, r in failed:
        sys.stderr.write(r.stdout + r.stderr)
    if failed:
        raise RuntimeError('hipcc failed on ' + ', '.join(os.path.basename(o) for o, _ in failed))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '-shared', '-fPIC', '--offload-arch=gfx950', '-o', tmp] + [o for o, _ in results]
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('hipcc failed linking libsubby_amd.so')
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
