"""GPU LOCG and the drop-in sqd / run_sqd / hproj / matmul entry points against the oracle."""
import sys

import numpy as np
import pytest
import scipy.linalg
import torch

from oracle import dense, locgref, sqdref
from rqutils_amd import workloads
from rqutils_amd.ground_locg import ground_locg
from rqutils_amd.paulis.symplectic import PauliSumXZ
from rqutils_amd.plan import Plan
from rqutils_amd.sqd import SQDOptions, hproj, run_sqd, sqd

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


def _say(msg):
    print(f'[case] {msg}', flush=True)
    sys.stdout.flush()


def _tfim(nq, hx):
    paulis, coeffs = [], []
    for q in range(nq):
        s = ['I'] * nq
        s[q] = 'Z'
        s[(q + 1) % nq] = 'Z'
        paulis.append(''.join(s))
        coeffs.append(-1.)
    for q in range(nq):
        s = ['I'] * nq
        s[q] = 'X'
        paulis.append(''.join(s))
        coeffs.append(hx)
    return paulis, coeffs


def _tfim_e0(nq, hx):
    k = np.pi * (2 * np.arange(nq) + 1) / nq
    return -np.sum(np.sqrt(1. + hx * hx - 2. * hx * np.cos(k)))


def _random_problem(seed, n=10, t=80, ns=400):
    rng = np.random.default_rng(seed)
    codes = rng.choice(4, size=(t, n), p=(0.5, 0.2, 0.15, 0.15))
    paulis = [''.join('IXYZ'[c] for c in row) for row in codes]
    coeffs = rng.standard_normal(t)
    keys = np.sort(rng.choice(1 << n, size=ns, replace=False).astype(np.uint64))
    return paulis, coeffs, keys


def test_locg_tfim_closed_form():
    _say('locg tfim')
    nq, hx = 12, 0.8
    h = PauliSumXZ.from_paulisum(_tfim(nq, hx))
    plan = Plan.full_space(h.tables)
    x0 = torch.from_numpy(np.random.default_rng(0).standard_normal(1 << nq)).to(DEV)
    theta, vec, niter = ground_locg(plan.matvec, x0, maxiter=1000)
    assert abs(float(theta) - _tfim_e0(nq, hx)) < 1e-9 * nq
    assert abs(torch.linalg.vector_norm(vec).item() - 1) < 1e-10


def test_locg_matches_oracle_trajectory():
    _say('locg vs oracle')
    paulis, coeffs, keys = _random_problem(1)
    h = PauliSumXZ.from_paulisum((paulis, coeffs))
    plan = Plan(h.tables, keys)
    hs = dense.projected_matrix(paulis, coeffs, keys)
    x0 = np.random.default_rng(2).standard_normal(400) + 1j * np.random.default_rng(3).standard_normal(400)
    theta_o, vec_o, it_o = locgref.ground_locg(lambda v: hs @ v, x0, maxiter=40, tol=1e-300)
    theta_g, vec_g, it_g = ground_locg(plan.matvec, torch.from_numpy(x0).to(DEV), maxiter=40, tol=1e-300)
    assert it_o == it_g == 40
    w = scipy.linalg.eigvalsh(hs)[0]
    # same iteration sequence: the 40-iteration Rayleigh values agree far below the remaining error
    assert abs(float(theta_g) - theta_o) < 1e-9 * abs(w)
    theta, vec, _ = ground_locg(plan.matvec, torch.from_numpy(x0).to(DEV), maxiter=3000)
    assert abs(float(theta) - w) < 1e-9 * abs(w)


def test_locg_orth_and_dense_matrix():
    _say('locg orth/dense')
    nq, hx = 8, 1.3
    hmat = dense.full_matrix(*_tfim(nq, hx), nq).real
    w, v = np.linalg.eigh(hmat)
    x0 = np.random.default_rng(4).standard_normal(1 << nq)
    theta, vec, _ = ground_locg(torch.from_numpy(hmat).to(DEV), torch.from_numpy(x0).to(DEV),
                                orth=torch.from_numpy(v[:, 0]).to(DEV), maxiter=2000)
    assert abs(float(theta) - w[1]) < 1e-8 * abs(w[1])
    # integer start index with a dense matrix
    theta, vec, _ = ground_locg(torch.from_numpy(hmat).to(DEV), 3, maxiter=2000)
    assert abs(float(theta) - w[0]) < 1e-8 * abs(w[0])


def test_sqd_end_to_end_matches_dense_eigh():
    """The method of tests/sqd.ipynb: dense eigh of the projected Hamiltonian."""
    _say('sqd end to end')
    rng = np.random.default_rng(5)
    nq, nsample, nterm = 10, 3000, 100
    states = rng.choice(2, size=(nsample, nq)).astype(np.uint8)
    paulis = [''.join('IXYZ'[ip] for ip in row) for row in rng.choice(4, size=(nterm, nq))]
    coeffs = rng.uniform(-1., 1., size=nterm)
    eigval, eigvec, basis = sqd((paulis, coeffs), states, options={'locg_maxiter': 3000})
    hmat = dense.full_matrix(paulis, coeffs, nq)
    states_u = np.unique(states, axis=0)
    idx = np.sum(states_u * (1 << np.arange(nq)[::-1])[None, :], axis=1)
    hp = hmat[idx][:, idx]
    w, vv = np.linalg.eigh(hp)
    assert np.array_equal(basis, states_u)
    assert abs(eigval - w[0]) < 1e-8 * abs(w[0])
    assert abs(abs(np.vdot(vv[:, 0], eigvec)) - 1) < 1e-6
    # energy only
    e = sqd((paulis, coeffs), states, options=SQDOptions(return_eigvec=False, locg_maxiter=3000))
    assert abs(e - w[0]) < 1e-8 * abs(w[0])


def test_run_sqd_packed_with_fillers():
    _say('run_sqd packed + fillers')
    paulis, coeffs, keys = _random_problem(6, n=11, t=60, ns=500)
    states = ((keys[:, None] >> np.arange(10, -1, -1, dtype=np.uint64)) & np.uint64(1)).astype(np.uint8)
    states = np.concatenate([states, states[:50]])       # duplicates
    states_p = sqdref.pack_states(states)
    h = PauliSumXZ.from_paulisum((paulis, coeffs), force_real=True, add_padding=True)
    e, vec, states_u, dim = run_sqd(h, states_p, 600, options=SQDOptions(locg_maxiter=3000))
    assert dim == 500
    want_u = sqdref.uniquify_states(states_p, 600)
    assert np.array_equal(states_u, want_u)
    assert vec.shape == (600,) and torch.count_nonzero(vec[500:]) == 0
    hs = dense.projected_matrix(paulis, coeffs, keys)
    w = scipy.linalg.eigvalsh(hs)[0]
    assert abs(float(e) - w) < 1e-8 * abs(w)
    # vinit rule: one-hot at argmin of the diagonal
    x, z, c, _ = (h.x, h.z, h.c, h.num_qubits)
    assert sqdref.vinit(x, z, c, want_u) == int(np.argmin(np.real(np.diag(hs))))


def test_hproj_matches_dense():
    _say('hproj')
    paulis, coeffs, keys = _random_problem(7, n=9, t=50, ns=200)
    states = ((keys[:, None] >> np.arange(8, -1, -1, dtype=np.uint64)) & np.uint64(1)).astype(np.uint8)
    m = hproj((paulis, coeffs), states[::-1])
    hs = dense.projected_matrix(paulis, coeffs, keys)
    assert np.allclose(m.toarray(), hs, rtol=0, atol=1e-12)


def test_matmul_full_space_matches_dense():
    _say('matmul')
    rng = np.random.default_rng(8)
    codes = rng.choice(4, size=(40, 8))
    paulis = [''.join('IXYZ'[c] for c in row) for row in codes]
    coeffs = rng.standard_normal(40)
    h = PauliSumXZ.from_paulisum((paulis, coeffs))
    v = rng.standard_normal((256, 3)) + 1j * rng.standard_normal((256, 3))
    got = h.matmul(v)
    assert np.allclose(got, dense.full_matrix(paulis, coeffs, 8) @ v, atol=1e-12)


def test_locg_jw_matches_eigsh():
    """Plain LOCG on a JW subspace reaches the eigsh energy of the oracle's CSR matrix."""
    _say('locg jw small vs eigsh')
    import scipy.sparse.linalg
    w = workloads.jw32(n_alpha=40, n_beta=40)
    plan = Plan(w.tables, w.states)
    t = w.tables
    hs = sqdref.csr_u64(t.xmask, t.grp_off, t.zmask, t.coef, w.states)
    e_ref = scipy.sparse.linalg.eigsh(hs, k=1, which='SA', tol=1e-13)[0][0]
    n = plan.n_states
    diag = plan.diagonal().real
    theta, vec, niter = ground_locg(plan.matvec, torch.argmin(diag), maxiter=2000,
                                   vspace=(n, torch.float64, plan.device))
    assert abs(float(theta) - e_ref) < 1e-9 * abs(e_ref)


def test_hproj_nonzero_triplets_match_lookup_and_scale():
    """hproj from per-hit triplets: equal to the J x N lookup construction; beyond J x N = 2^32 it matches the matvec."""
    _say('hproj triplets')
    import scipy.sparse
    from rqutils_amd import workloads
    from rqutils_amd.paulis.symplectic import PauliSumXZ
    from rqutils_amd.plan import Plan
    from rqutils_amd.sqd import u64_to_states
    w = workloads.jw32(n_alpha=40, n_beta=30)
    h = PauliSumXZ.from_tables(w.tables)
    m = hproj(h, u64_to_states(w.states, 32), unique_states=True)
    plan = Plan(w.tables, w.states)
    t = w.tables
    ref = sqdref.csr_u64(t.xmask, t.grp_off, t.zmask, t.coef, w.states)
    assert m.nnz == ref.nnz == plan.count()['hits']
    # same sparsity; elements agree to rounding (a long group's sum may be split across lanes)
    assert (abs(m - ref)).max() <= 1e-13 * abs(ref).max()
    assert np.array_equal(m.indices, ref.indices) and np.array_equal(m.indptr, ref.indptr)
    # a table J x N beyond 2^32 entries would not fit the lookup construction at scale;
    # the triplets give H v = plan.matvec(v)
    w = workloads.jw32(n_alpha=600, n_beta=600)
    plan = Plan(w.tables, w.states)
    assert plan.info['n_groups'] * plan.n_total > 2 ** 32
    m = hproj(PauliSumXZ.from_tables(w.tables), u64_to_states(w.states, 32), unique_states=True)
    rng = np.random.default_rng(31)
    v = rng.standard_normal(plan.n_total)
    got = m @ v
    want = plan.matvec(torch.from_numpy(v).to(DEV)).cpu().numpy()
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    rows = np.sort(rng.choice(plan.n_states, size=300, replace=False))
    t = w.tables
    ora = sqdref.matvec_rows_u64(t.xmask, t.grp_off, t.zmask, t.coef, w.states, v, rows)
    assert np.max(np.abs(got[rows] - ora)) <= 1e-12 * np.max(np.abs(ora))


def test_mixed_and_batched_locg_deflate_to_the_first_excited_state():
    """orth deflation in the mixed and graph-captured batched solvers: with the ground state
    projected out, both converge to the first excited energy of the periodic TFIM (eigh)."""
    _say('orth in mixed / batched locg')
    from rqutils_amd.locg_batch import batch_ground_locg
    from rqutils_amd.locg_mixed import ground_locg_mixed
    nq, hx = 10, 1.3
    paulis, coeffs = _tfim(nq, hx)
    w, v = np.linalg.eigh(dense.full_matrix(paulis, coeffs, nq).real)
    assert w[1] - w[0] > 1e-3
    h = PauliSumXZ.from_paulisum((paulis, coeffs))
    plan = Plan.full_space(h.tables)
    ground = torch.from_numpy(v[:, 0].copy()).to(DEV)
    rng = np.random.default_rng(41)
    x0 = torch.from_numpy(rng.standard_normal(1 << nq)).to(DEV)
    res = ground_locg_mixed(plan.matvec, plan.matvec, x0, maxiter=3000, tol=1e-10, orth=(ground,))
    assert abs(float(res.theta) - w[1]) < 1e-8 * abs(w[1]), (float(res.theta), w[1])
    assert abs(float(torch.dot(res.x, ground))) < 1e-6
    for graph in (False, True):
        theta, x, it, info = batch_ground_locg(plan, x0, tol=1e-10, maxiter=3000, check_every=5, use_graph=graph,
                                               orth=(ground,))
        assert abs(float(theta[0]) - w[1]) < 1e-8 * abs(w[1]), (graph, float(theta[0]), w[1])
        assert abs(float(torch.dot(x, ground))) < 1e-6
    # without deflation the same solver finds the ground state
    theta, _, _, _ = batch_ground_locg(plan, x0, tol=1e-10, maxiter=3000, check_every=5)
    assert abs(float(theta[0]) - w[0]) < 1e-8 * abs(w[0])


@pytest.mark.parametrize('dtype', [torch.complex128, torch.float64, torch.complex64])
def test_full_apply_matches_dense_and_subspace_kernel(dtype):
    """The direct full-space kernel (no index): dense textbook matrix for several columns,
    and the subspace kernel over the complete basis at 20 qubits."""
    _say(f'full apply {dtype}')
    rng = np.random.default_rng(9)
    n = 9
    codes = rng.choice(4, size=(50, n))
    if not dtype.is_complex:                      # a real Hamiltonian: even numbers of Y
        odd = np.nonzero(np.sum(codes == 2, axis=1) % 2 == 1)[0]
        codes[odd, np.argmax(codes[odd] == 2, axis=1)] = 3
    paulis = [''.join('IXYZ'[c] for c in row) for row in codes]
    coeffs = rng.standard_normal(50)
    h = PauliSumXZ.from_paulisum((paulis, coeffs))
    assert h.tables.is_real == (not dtype.is_complex)
    v = rng.standard_normal((1 << n, 4))
    if dtype.is_complex:
        v = v + 1j * rng.standard_normal((1 << n, 4))
    plan = Plan(h.tables, np.zeros(0, dtype=np.uint64))
    got = plan.full_apply(torch.from_numpy(v).to(DEV, dtype)).cpu().numpy()
    want = dense.full_matrix(paulis, coeffs, n) @ v
    tol = 1e-12 if dtype != torch.complex64 else 2e-5
    assert np.max(np.abs(got - want)) <= tol * np.max(np.abs(want))
    # 20 qubits: the direct kernel against the subspace kernel over all 2^20 states
    w = workloads.random20_300t()
    full = Plan.full_space(w.tables)
    x = torch.from_numpy(rng.standard_normal(1 << 20) + 1j * rng.standard_normal(1 << 20)).to(DEV)
    direct = Plan(w.tables, np.zeros(0, dtype=np.uint64)).full_apply(x)
    ref = full.matvec(x)
    assert float((direct - ref).abs().max()) <= 1e-12 * float(ref.abs().max())


def test_full_apply_beyond_thirty_qubits():
    """31 qubits (past the subspace index's 2^31-row limit): sampled rows against the
    CPU evaluation of sum_j C_j(b) x[b ^ x_j] with x given by a closed form."""
    _say('full apply 31 qubits')
    n = 31
    rng = np.random.default_rng(10)
    codes = workloads.random_pauli_codes(n, 12, (0.7, 0.1, 0.1, 0.1), rng)
    xm, zm = workloads.codes_to_masks(codes)
    from rqutils_amd.paulis.symplectic import TermTables
    t = TermTables.from_masks(xm, zm, rng.standard_normal(12), n)

    def phase(b):                                   # x[b] = exp(i phi(b)), phi from b alone
        return ((b % 1009) * 0.37 + (b % 997) * 0.11)

    b = torch.arange(1 << n, dtype=torch.int64, device=DEV)
    phi = (b % 1009).to(torch.float64) * 0.37 + (b % 997).to(torch.float64) * 0.11   # phase(b) in FP64
    phi = torch.remainder(phi, 2 * np.pi).to(torch.float32)                            # reduced, then narrowed
    del b
    x = torch.polar(torch.ones_like(phi), phi)
    del phi
    plan = Plan(t, np.zeros(0, dtype=np.uint64))
    y = plan.full_apply(x)
    rows = rng.integers(0, 1 << n, size=64, dtype=np.uint64)
    got = y[torch.from_numpy(rows.astype(np.int64)).to(DEV)].cpu().numpy()
    want = np.zeros(64, dtype=np.complex128)
    for xj, k0, k1 in zip(t.xmask, t.grp_off[:-1], t.grp_off[1:]):
        c = np.zeros(64, dtype=np.complex128)
        for k in range(k0, k1):
            c += t.coef[k] * (1. - 2. * (np.bitwise_count(rows & t.zmask[k]) & 1))
        src = (rows ^ np.uint64(xj)).astype(np.int64)
        want += c * np.exp(1j * phase(src))
    assert np.max(np.abs(got - want)) <= 1e-5 * np.max(np.abs(want))
    del x, y
    torch.cuda.empty_cache()
