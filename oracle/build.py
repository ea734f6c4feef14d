"""Build recipe of the oracle's C port (test infrastructure only).

    python -m oracle.build

Writes ``oracle/liboracle_c.so`` (gcc, OpenMP).  The reference itself is pure
Python/JAX with no C sources, so there is no ``oracle/_ref`` build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'rqx_oracle.c')
OUT = os.path.join(HERE, 'liboracle_c.so')


def build(verbose=False):
    cmd = ['gcc', '-O3', '-march=x86-64-v2', '-fopenmp', '-fPIC', '-shared', SRC, '-o', OUT + '.tmp']
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f'gcc failed with status {res.returncode}')
    os.replace(OUT + '.tmp', OUT)
    return OUT


if __name__ == '__main__':
    print(build(verbose=True))
