"""Build recipe for the CPU oracle (test infrastructure only).

Compiles oracle/wt_oracle.cpp with g++ into oracle/liboracle.so. Strict IEEE
arithmetic: -ffp-contract=off and no fast-math, so each floating-point
operation of the restatement is rounded exactly once, in source order.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'wt_oracle.cpp')
OUT = os.path.join(HERE, 'liboracle.so')


def build(verbose=False):
    cmd = ['g++', '-O2', '-std=c++17', '-fPIC', '-shared', '-fopenmp',
           '-ffp-contract=off', '-fno-fast-math', '-fno-finite-math-only',
           '-Wall', '-Wextra', '-Wno-unused-parameter', SRC, '-o', OUT + '.tmp']
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + '.tmp', OUT)
    return OUT


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
