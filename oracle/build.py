"""ORACLE build recipe (test infrastructure): compiles the C restatements (mfv_oracle.c, mixed32.c, srnh.c, euler32.c)
to oracle/liboracle.so with strict IEEE semantics (no contraction, no fast-math), and the
OpenMP CPU baseline (omp_baseline.c) to oracle/liboracle_omp.so."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'liboracle.so')
LIB_OMP = os.path.join(HERE, 'liboracle_omp.so')


FLAGS = ['-O3', '-march=x86-64-v3', '-fPIC', '-shared', '-std=c99', '-ffp-contract=off', '-fno-fast-math', '-Wall']


def build_oracle(verbose=False):
    """liboracle.so (the serial restatements) and liboracle_omp.so (the same code, with the
    OpenMP all-cores baseline of the shallow-water step). -O3 without fast-math or contraction
    keeps every result bit-identical to -O0."""
    cmds = [(['gcc'] + FLAGS + [os.path.join(HERE, 'mfv_oracle.c'), os.path.join(HERE, 'mixed32.c'),
                                os.path.join(HERE, 'srnh.c'), os.path.join(HERE, 'euler32.c'), '-lm'], LIB),
            (['gcc'] + FLAGS + ['-fopenmp', os.path.join(HERE, 'omp_baseline.c'), '-lm'], LIB_OMP)]
    for cmd, out in cmds:
        cmd = cmd + ['-o', out + '.tmp']
        if verbose:
            print(' '.join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(out + '.tmp', out)
    return LIB


if __name__ == '__main__':
    build_oracle(verbose=True)
