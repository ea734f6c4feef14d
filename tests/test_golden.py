"""The oracle reproduces the committed golden fixtures (no GPU)."""
import os

import numpy as np
import pytest

from oracle import dense, sqdref, terms

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


@pytest.mark.parametrize('name', ['small_10q', 'filler_12q'])
def test_oracle_reproduces_golden(name):
    f = np.load(os.path.join(HERE, f'{name}.npz'))
    paulis = [str(p) for p in f['paulis']]
    x, z, c, _ = terms.from_paulisum_tuple(paulis, f['coeffs'], add_padding=True)
    assert np.array_equal(x, f['x']) and np.array_equal(z, f['z']) and np.allclose(c, f['c'])
    n_states = f['keys'].shape[0]
    states_u = sqdref.uniquify_states(sqdref.pack_states(f['states']), int(f['states_size']))
    xs = sqdref.xsources_all(x, states_u)
    assert np.array_equal(xs[:, :n_states], f['xsource'])
    y = sqdref.apply_h_x_cached(f['v'], xs, z, c, states_u)
    assert np.allclose(y, f['y'], rtol=0, atol=1e-13 * np.abs(f['y']).max())
    hs = dense.projected_matrix(paulis, f['coeffs'], f['keys'])
    assert np.allclose(hs @ f['v'][:n_states], f['y'][:n_states], rtol=0, atol=1e-12 * np.abs(f['y']).max())
    assert abs(np.linalg.eigvalsh(hs)[0] - f['e0']) < 1e-10
