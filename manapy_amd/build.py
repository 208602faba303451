"""Build libmfv.so (hipcc, gfx950) in-tree.

The library is compiled with ``-ffp-contract=off`` so that no multiply-add is fused
(every kernel's arithmetic is compared bit for bit with the CPU restatement) and with
packed-FP32 instruction selection disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libmfv.so')
SOURCES = ['mfv_core.hip', 'mfv_sw.hip', 'mfv_sw_mixed.hip', 'mfv_srnh.hip', 'mfv_euler.hip', 'mfv_diamond.hip',
           'mfv_krylov.hip', 'mfv_tiles.hip']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('MFV_ARCH', 'gfx950')

FLAGS = ['-O3', '-std=c++17', '-fPIC', '-shared', '--offload-arch=' + ARCH, '-ffp-contract=off',
         '-fno-fast-math', '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-I' + os.path.join(ROOT, 'include'), '-Wall', '-Wno-unused-function']


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES]


def build_library(verbose=False, extra=()):
    cmd = [HIPCC] + FLAGS + list(extra) + sources() + ['-o', LIB + '.tmp']
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + '.tmp', LIB)
    return LIB


if __name__ == '__main__':
    build_library(verbose=True, extra=sys.argv[1:])
