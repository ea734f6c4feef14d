"""Write the golden fixtures used by tests (inputs + expected outputs from the oracle).

    python tests/golden/make_golden.py

The reference ships no golden vectors for this path and cannot be imported
here (it needs Python >= 3.12 and JAX), so the expected outputs come from the
oracle (``oracle/``), whose own pins are in ``tests/test_oracle.py``: the
textbook per-qubit Pauli action, ``numpy.searchsorted`` and dense ``eigh``.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import dense, sqdref, terms  # noqa: E402


def make_case(name, n, num_terms, num_states, seed, probs=(0.55, 0.15, 0.15, 0.15), states_size=None):
    rng = np.random.default_rng(seed)
    codes = rng.choice(4, size=(num_terms, n), p=probs)
    paulis = [''.join('IXYZ'[c] for c in row) for row in codes]
    coeffs = rng.standard_normal(num_terms)
    keys = np.sort(rng.choice(1 << n, size=num_states, replace=False).astype(np.uint64))
    states = ((keys[:, None] >> np.arange(n - 1, -1, -1, dtype=np.uint64)) & np.uint64(1)).astype(np.uint8)
    states_size = states_size or num_states
    x, z, c, _ = terms.from_paulisum_tuple(paulis, coeffs, add_padding=True)
    states_u = sqdref.uniquify_states(sqdref.pack_states(states), states_size)
    xsrc = sqdref.xsources_all(x, states_u)
    v = rng.standard_normal(states_size) + 1j * rng.standard_normal(states_size)
    v[num_states:] = 0
    y = sqdref.apply_h_x_cached(v, xsrc, z, c, states_u)
    hs = dense.projected_matrix(paulis, coeffs, keys)
    e0 = np.linalg.eigvalsh(hs)[0]
    np.savez_compressed(os.path.join(HERE, f'{name}.npz'),
                        paulis=np.array(paulis), coeffs=coeffs, states=states, keys=keys,
                        states_size=states_size, xsource=xsrc[:, :num_states].astype(np.int32),
                        v=v, y=y, e0=e0, x=x, z=z, c=c)


if __name__ == '__main__':
    make_case('small_10q', n=10, num_terms=60, num_states=300, seed=11)
    make_case('filler_12q', n=12, num_terms=100, num_states=1000, seed=12, states_size=1100)
    print('written')
