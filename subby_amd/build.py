"""Build the gfx950 HIP library in-tree: subby_amd/libsubby_amd.so.

    python -m subby_amd.build        (or subby_amd.build.build())
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libsubby_amd.so')
SOURCES = ['sb_host.cpp', 'sb_serve.cpp', 'sb_fused.hip', 'sb_large.hip', 'sb_f81w.hip', 'sb_gen2.hip', 'sb_post.hip', 'sb_plain.hip', 'sb_multi.hip']

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = [
    '-O3', '-std=c++17', '-fPIC', '-shared', '--offload-arch=gfx950',
    # no packed-FP32 instructions in device code
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
]


def build(verbose=False):
    srcs = [os.path.join(SRC, s) for s in SOURCES if os.path.exists(os.path.join(SRC, s))]
    tmp = LIB + '.tmp'
    cmd = [HIPCC] + FLAGS + ['-I', os.path.join(HERE, '..', 'include'), '-o', tmp] + srcs
    if verbose:
        print(' '.join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('hipcc failed building libsubby_amd.so')
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
