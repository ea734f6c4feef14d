"""The torch LOCG drivers on CPU tensors (dense matrices), pinned against the oracle and eigh."""
import numpy as np
import pytest
import torch

from oracle import locgref
from rqutils_amd.ground_locg import ground_locg
from rqutils_amd.locg_mixed import ground_locg_mixed


def _rand_herm(n, seed, complex_=False, gap_shift=0.0):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    if complex_:
        a = a + 1j * rng.standard_normal((n, n))
    a = (a + a.conj().T) / 2
    w, v = np.linalg.eigh(a)
    w[0] -= gap_shift
    return (v * w) @ v.conj().T


@pytest.mark.parametrize('complex_', [False, True])
def test_plain_locg_follows_oracle_trajectory(complex_):
    a = _rand_herm(120, 1, complex_, gap_shift=1.0)
    rng = np.random.default_rng(2)
    x0 = rng.standard_normal(120) + (1j * rng.standard_normal(120) if complex_ else 0)
    tr = []
    th_o, v_o, it_o = locgref.ground_locg(lambda v: a @ v, x0, maxiter=25, tol=1e-300, trace=tr)
    at = torch.from_numpy(a)
    th, v, it = ground_locg(lambda v: at @ v, torch.from_numpy(x0), maxiter=25, tol=1e-300)
    assert it == it_o == 25
    assert abs(float(th) - th_o) < 1e-11 * abs(th_o)
    w = np.linalg.eigvalsh(a)[0]
    th, v, it = ground_locg(lambda v: at @ v, torch.from_numpy(x0), maxiter=500)
    assert abs(float(th) - w) < 1e-10 * abs(w)


@pytest.mark.parametrize('complex_', [False, True])
def test_mixed_locg_reaches_fp64_accuracy(complex_):
    n = 400
    a = _rand_herm(n, 3, complex_, gap_shift=2.0)
    w = np.linalg.eigvalsh(a)[0]
    at = torch.from_numpy(a)
    a32 = at.to(torch.complex64 if complex_ else torch.float32)
    rng = np.random.default_rng(4)
    x0 = torch.from_numpy(rng.standard_normal(n) + (1j * rng.standard_normal(n) if complex_ else 0))
    res = ground_locg_mixed(lambda v: at @ v, lambda v: a32 @ v, x0, maxiter=500, tol=1e-9, record=True)
    assert abs(float(res.theta) - w) < 1e-10 * abs(w)
    assert res.matvecs_lo >= res.matvecs_hi > 0
    phases = [h[3] for h in res.history]
    assert phases[0] == 'lo' and phases[-1] == 'hi'
    # Rayleigh-Ritz values from FP64 images never increase in the FP64 phase
    hi_theta = [h[1] for h in res.history if h[3] == 'hi']
    assert all(b <= a_ + 1e-12 * abs(w) for a_, b in zip(hi_theta, hi_theta[1:]))


def test_mixed_locg_exact_start_vector():
    a = np.diag(np.arange(1., 30.))
    at = torch.from_numpy(a)
    x0 = torch.zeros(29, dtype=torch.float64)
    x0[0] = 1.
    res = ground_locg_mixed(lambda v: at @ v, None, x0, maxiter=10, tol=1e-12)
    assert float(res.theta) == 1.0 and res.niter == 0


@pytest.mark.parametrize('complex_', [False, True])
def test_mixed_locg_orth_deflation(complex_):
    """orth deflation: with the two lowest eigenvectors projected out, the mixed solver
    returns the third eigenvalue (as the reference's orth argument, ground_locg.py:221-222)."""
    n = 300
    a = _rand_herm(n, 5, complex_, gap_shift=1.0)
    w, v = np.linalg.eigh(a)
    at = torch.from_numpy(a)
    a32 = at.to(torch.complex64 if complex_ else torch.float32)
    rng = np.random.default_rng(6)
    x0 = torch.from_numpy(rng.standard_normal(n) + (1j * rng.standard_normal(n) if complex_ else 0))
    orth = tuple(torch.from_numpy(v[:, k].copy()) for k in (0, 1))
    res = ground_locg_mixed(lambda u: at @ u, lambda u: a32 @ u, x0, maxiter=2000, tol=1e-10, orth=orth)
    assert abs(float(res.theta) - w[2]) < 1e-9 * abs(w[2])
    for o in orth:
        assert abs(complex(torch.vdot(o.to(res.x.dtype), res.x))) < 1e-7
