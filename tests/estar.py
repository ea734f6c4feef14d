"""Reference ground energies computed by the oracle (``tests/golden/estar_*.json``,
written by ``tests/golden/make_estar.py``): load one and check that the workload
regenerated here is the instance it was computed for."""
import json
import os

from rqutils_amd import workloads

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def load(name, w):
    with open(os.path.join(HERE, f'estar_{name}.json')) as f:
        ref = json.load(f)
    got = workloads.fingerprint(w)
    want = ref['workload']
    for k, v in want.items():
        if k == 'coef_abs_sum':
            assert abs(got[k] - v) <= 1e-9 * abs(v), (name, k, got[k], v)
        else:
            assert got[k] == v, (name, k, got[k], v)
    return ref['e_star']
