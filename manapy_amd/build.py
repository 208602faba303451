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
           'mfv_krylov.hip', 'mfv_tiles.hip', 'mfv_domain.hip']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('MFV_ARCH', 'gfx950')

FLAGS = ['-O3', '-std=c++17', '-fPIC', '-shared', '--offload-arch=' + ARCH, '-ffp-contract=off',
         '-fno-fast-math', '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-I' + os.path.join(ROOT, 'include'), '-Wall', '-Wno-unused-function']


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES]


def build_library(verbose=False, extra=(), jobs=None):
    """Compile every source to an object in parallel (one hipcc per file), then link."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in FLAGS if f != '-shared'] + list(extra)
    objs = [os.path.join(objdir, os.path.basename(s) + '.o') for s in sources()]

    def compile_one(pair):
        src, obj = pair
        cmd = [HIPCC] + cflags + ['-c', src, '-o', obj]
        if verbose:
            print(' '.join(cmd), flush=True)
        return subprocess.run(cmd, capture_output=True, text=True)

    jobs = jobs or min(len(objs), os.cpu_count() or 4, 8)
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(compile_one, zip(sources(), objs)))
    for src, r in zip(sources(), results):
        err = '\n'.join(l for l in r.stderr.splitlines() if 'packed-fp32-ops' not in l)
        if err.strip():
            sys.stderr.write(err + '\n')
        if r.returncode != 0:
            raise subprocess.CalledProcessError(r.returncode, 'hipcc ' + src)
    cmd = [HIPCC, '-shared', '-fPIC', '--offload-arch=' + ARCH] + objs + ['-o', LIB + '.tmp']
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + '.tmp', LIB)
    return LIB


if __name__ == '__main__':
    build_library(verbose=True, extra=sys.argv[1:])
