"""Build librqx.so (gfx950) in-tree with hipcc.

    python -m rqutils_amd._build

The library is written next to this file as ``librqx.so``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, 'csrc', 'rqx.hip')
OUT = os.path.join(HERE, 'librqx.so')

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

FLAGS = [
    '--offload-arch=gfx950',
    '-O3', '-std=c++17', '-fPIC', '-shared',
    # packed FP32 instructions are kept out of every kernel
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
]


def build(verbose=False):
    cmd = [HIPCC, *FLAGS, '-I', os.path.join(ROOT, 'include'), SRC, '-o', OUT + '.tmp']
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f'hipcc failed with status {res.returncode}')
    os.replace(OUT + '.tmp', OUT)
    return OUT


if __name__ == '__main__':
    print(build(verbose=True))
