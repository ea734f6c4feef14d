"""Backward of y = H(alpha) x and Hellmann-Feynman gradients (GPU) against FD and dense autograd."""
import sys

import numpy as np
import pytest
import torch

from oracle import dense
from rqutils_amd import workloads
from rqutils_amd.autograd import hellmann_feynman, pauli_matvec
from rqutils_amd.ground_locg import ground_locg
from rqutils_amd.paulis.symplectic import TermTables
from rqutils_amd.plan import Plan

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


def _say(msg):
    print(f'[case] {msg}', flush=True)
    sys.stdout.flush()


def _strings(t: TermTables):
    n = t.num_qubits
    out = []
    for xm, zm in zip(t.term_xmask(), t.zmask):
        out.append(''.join('IZXY'[((int(xm) >> (n - 1 - c)) & 1) * 2 + ((int(zm) >> (n - 1 - c)) & 1)]
                           for c in range(n)))
    return out


def _problem(seed, n=9, t=50, ns=250, probs=(0.5, 0.2, 0.15, 0.15)):
    rng = np.random.default_rng(seed)
    codes = workloads.random_pauli_codes(n, t, probs, rng)
    xm, zm = workloads.codes_to_masks(codes)
    tables = TermTables.from_masks(xm, zm, rng.standard_normal(t), n)
    keys = np.sort(rng.choice(1 << n, size=ns, replace=False).astype(np.uint64))
    return tables, keys, rng


def _dense_terms(tables, keys):
    """Projected dense matrix of every term's Pauli string (unit coefficient)."""
    return np.stack([dense.projected_matrix([s], [1.], keys) for s in _strings(tables)])


def test_grad_coef_matches_dense_and_fd():
    _say('grad coef dense + fd')
    tables, keys, rng = _problem(1)
    plan = Plan(tables, keys)
    alpha = torch.from_numpy(tables.alphas().real.copy()).to(DEV)
    x = torch.from_numpy(rng.standard_normal(250) + 1j * rng.standard_normal(250)).to(DEV)
    gy = torch.from_numpy(rng.standard_normal(250) + 1j * rng.standard_normal(250)).to(DEV)
    g = plan.grad_coef(x, gy).cpu().numpy()
    P = _dense_terms(tables, keys)
    want = np.array([np.vdot(gy.cpu().numpy(), P[k] @ x.cpu().numpy()).real for k in range(P.shape[0])])
    assert np.allclose(g, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    # central finite differences of L(alpha) = Re <gy, H(alpha) x> (step 1e-6, FP64)
    h = 1e-6
    for k in range(0, tables.num_terms, 7):
        a_p, a_m = alpha.clone(), alpha.clone()
        a_p[k] += h
        a_m[k] -= h
        plan.set_alphas(a_p)
        lp = torch.vdot(gy, plan.matvec(x)).real.item()
        plan.set_alphas(a_m)
        lm = torch.vdot(gy, plan.matvec(x)).real.item()
        assert abs((lp - lm) / (2 * h) - g[k]) < 1e-6 * max(1., abs(g[k]))
    plan.set_alphas(alpha)


def test_autograd_matches_dense_torch_autograd():
    _say('autograd vs dense')
    tables, keys, rng = _problem(2)
    plan = Plan(tables, keys)
    a0 = tables.alphas().real.copy()
    x0 = rng.standard_normal(250) + 1j * rng.standard_normal(250)
    # ours
    alpha = torch.from_numpy(a0).to(DEV).requires_grad_(True)
    x = torch.from_numpy(x0).to(DEV).requires_grad_(True)
    y = pauli_matvec(plan, alpha, x)
    loss = (y.abs() ** 2).sum() + torch.vdot(x, y).real
    loss.backward()
    # dense torch autograd on the CPU (float64 / complex128)
    P = torch.from_numpy(_dense_terms(tables, keys))
    ad = torch.from_numpy(a0).requires_grad_(True)
    xd = torch.from_numpy(x0).requires_grad_(True)
    H = (ad.to(torch.complex128)[:, None, None] * P).sum(0)
    yd = H @ xd
    loss_d = (yd.abs() ** 2).sum() + torch.vdot(xd, yd).real
    loss_d.backward()
    assert abs(loss.item() - loss_d.item()) < 1e-10 * abs(loss_d.item())
    assert torch.allclose(alpha.grad.cpu(), ad.grad, rtol=1e-10, atol=1e-10 * ad.grad.abs().max().item())
    assert torch.allclose(x.grad.cpu(), xd.grad, rtol=1e-10, atol=1e-10 * xd.grad.abs().max().item())


def test_hellmann_feynman_matches_energy_fd():
    _say('hellmann feynman')
    tables, keys, rng = _problem(3, n=8, t=40, ns=200)
    plan = Plan(tables, keys)
    a0 = tables.alphas().real.copy()
    P = _dense_terms(tables, keys)

    def e0(a):
        return np.linalg.eigvalsh(np.tensordot(a, P, axes=1))[0]

    w, v = np.linalg.eigh(np.tensordot(a0, P, axes=1))
    psi = torch.from_numpy(v[:, 0] * 3.0).to(DEV)         # unnormalised on purpose
    hf = hellmann_feynman(plan, psi).cpu().numpy()
    h = 1e-6
    for k in range(0, tables.num_terms, 5):
        ap, am = a0.copy(), a0.copy()
        ap[k] += h
        am[k] -= h
        fd = (e0(ap) - e0(am)) / (2 * h)
        assert abs(fd - hf[k]) < 1e-6 * max(1., abs(fd))
    # sum_k alpha_k <P_k> = E
    assert abs(np.dot(a0, hf) - w[0]) < 1e-10 * abs(w[0])


def test_grad_real_dtype_jw():
    _say('grad real dtype jw')
    w = workloads.jw32(n_alpha=30, n_beta=30)
    plan = Plan(w.tables, w.states)
    rng = np.random.default_rng(4)
    x = torch.from_numpy(rng.standard_normal(plan.n_total)).to(DEV)
    gy = torch.from_numpy(rng.standard_normal(plan.n_total)).to(DEV)
    g_real = plan.grad_coef(x, gy)
    g_cplx = plan.grad_coef(x.to(torch.complex128), gy.to(torch.complex128))
    assert torch.allclose(g_real, g_cplx, rtol=1e-10, atol=1e-10 * g_real.abs().max().item())
    g32 = plan.grad_coef(x.float(), gy.float())
    assert torch.allclose(g32, g_real, rtol=0, atol=2e-4 * g_real.abs().max().item())


def test_grad_identities_at_36q_10kt_1m_shape():
    """36 qubits, 10k terms, 1M states: sum_k alpha_k g_k = Re <gy, H x> (size-independent)."""
    _say('grad identities rand36')
    w = workloads.random36_10kt()
    plan = Plan(w.tables, w.states)
    rng = np.random.default_rng(3003)
    n = plan.n_total
    x = torch.from_numpy(rng.standard_normal(n) + 1j * rng.standard_normal(n)).to(DEV)
    gy = torch.from_numpy(rng.standard_normal(n) + 1j * rng.standard_normal(n)).to(DEV)
    alpha = torch.from_numpy(w.tables.alphas().real.copy()).to(DEV)
    g = plan.grad_coef(x, gy)
    lhs = torch.dot(alpha, g).item()
    rhs = torch.vdot(gy, plan.matvec(x)).real.item()
    assert abs(lhs - rhs) < 1e-9 * (abs(rhs) + torch.linalg.vector_norm(gy).item())
    # grad_x of Re<gy, H x> is H gy: <H gy, x> = <gy, H x>
    assert abs(torch.vdot(plan.matvec(gy), x).real.item() - rhs) < 1e-9 * (abs(rhs) + 1)
    # central finite differences (step 1e-6, FP64) of L(alpha) = Re <gy, H(alpha) x> at full
    # size, for the 6 largest gradient entries and 6 drawn at random
    gh = g.cpu().numpy()
    ks = list(np.argsort(-np.abs(gh))[:6]) + list(rng.choice(len(gh), size=6, replace=False))
    step = 1e-6
    worst = 0.0
    for k in ks:
        val = []
        for sgn in (1.0, -1.0):
            a = alpha.clone()
            a[k] += sgn * step
            plan.set_alphas(a)
            val.append(torch.vdot(gy, plan.matvec(x)).real.item())
        fd = (val[0] - val[1]) / (2 * step)
        err = abs(fd - gh[k]) / max(abs(gh[k]), 1e-3 * np.abs(gh).max())
        worst = max(worst, err)
    plan.set_alphas(alpha)
    print(f'[case] fd worst relative difference {worst:.2e} over {len(ks)} terms', flush=True)
    assert worst < 1e-6, worst


def test_hellmann_feynman_jw_ground_state():
    _say('hf jw ground state')
    w = workloads.jw32(n_alpha=40, n_beta=40)
    plan = Plan(w.tables, w.states)
    diag = plan.diagonal().real
    theta, psi, _ = ground_locg(plan.matvec, torch.argmin(diag), maxiter=1500,
                                vspace=(plan.n_total, torch.float64, plan.device))
    hf = hellmann_feynman(plan, psi)
    alpha = torch.from_numpy(w.tables.alphas().real.copy()).to(DEV)
    assert abs(torch.dot(alpha, hf).item() - float(theta)) < 1e-9 * abs(float(theta))
