"""ORACLE build recipe (test infrastructure): compiles the C restatements (mfv_oracle.c, mixed32.c, srnh.c)
to oracle/liboracle.so with strict IEEE semantics (no contraction, no fast-math)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'liboracle.so')


def build_oracle(verbose=False):
    cmd = ['gcc', '-O2', '-fPIC', '-shared', '-std=c99', '-ffp-contract=off', '-fno-fast-math',
           '-Wall', os.path.join(HERE, 'mfv_oracle.c'), os.path.join(HERE, 'mixed32.c'), os.path.join(HERE, 'srnh.c'), '-lm', '-o', LIB + '.tmp']
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + '.tmp', LIB)
    return LIB


if __name__ == '__main__':
    build_oracle(verbose=True)
