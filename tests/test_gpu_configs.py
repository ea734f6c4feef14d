"""The named workloads at full size on the GPU (25M-determinant JW model, 64-problem
batch, 1M-determinant JW model), checked against the oracle on sampled rows and
against the oracle's eigsh ground energies (``tests/golden/estar_*.json``)."""
import sys

import numpy as np
import pytest
import scipy.sparse.linalg
import torch

import estar
from oracle import sqdref
from rqutils_amd import workloads
from rqutils_amd.ground_locg import ground_locg
from rqutils_amd.locg_batch import batch_ground_locg
from rqutils_amd.plan import BatchPlan, Plan, sort_unique_u64

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


def _say(msg):
    print(f'[case] {msg}', flush=True)
    sys.stdout.flush()


def _cvec(rng, n, dtype):
    v = torch.from_numpy(rng.standard_normal(n) + 1j * rng.standard_normal(n)).to(DEV)
    return v.to(dtype)


def _rel(a, b):
    return float((a - b).abs().max()) / max(float(b.abs().max()), 1e-300)


# ---------------------------------------------------------------- 25M determinants, 40 spin-orbitals

@pytest.fixture(scope='module')
def jw25m():
    _say('generate jw-40so-5000x5000')
    w = workloads.jw40()
    plan = Plan(w.tables, w.states)
    yield w, plan
    plan.close()
    torch.cuda.empty_cache()


def test_jw25m_rows_and_diagonal_match_oracle(jw25m):
    """Sampled rows of H x (complex128, matrix-free) and of the diagonal against the oracle."""
    w, plan = jw25m
    _say('jw25m rows vs oracle')
    assert plan.n_states == 25_000_000 and plan.info['n_groups'] == w.tables.num_groups
    rng = np.random.default_rng(401)
    x = _cvec(rng, plan.n_total, torch.complex128)
    y = plan.matvec(x)
    rows = np.sort(rng.choice(plan.n_states, size=400, replace=False))
    t = w.tables
    xh = x.cpu().numpy()
    want = sqdref.matvec_rows_u64(t.xmask, t.grp_off, t.zmask, t.coef, w.states, xh, rows)
    got = y[torch.from_numpy(rows).to(DEV)].cpu().numpy()
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    d = plan.diagonal()[torch.from_numpy(rows).to(DEV)].cpu().numpy()
    dz = sqdref.diagonal_u64(t.zmask[t.grp_off[0]:t.grp_off[1]], t.coef[t.grp_off[0]:t.grp_off[1]], w.states[rows])
    assert t.has_diagonal and np.max(np.abs(d - dz)) <= 1e-13 * np.max(np.abs(dz))


@pytest.mark.parametrize('dtype', [torch.complex64, torch.complex128])
def test_jw25m_hermitian_and_linear(jw25m, dtype):
    w, plan = jw25m
    _say(f'jw25m hermitian/linear {dtype}')
    rng = np.random.default_rng(402)
    u, v = _cvec(rng, plan.n_total, dtype), _cvec(rng, plan.n_total, dtype)
    hu, hv = plan.matvec(u), plan.matvec(v)
    tol = 1e-12 if dtype == torch.complex128 else 2e-5
    lhs, rhs = torch.vdot(u.to(torch.complex128), hv.to(torch.complex128)), \
        torch.vdot(hu.to(torch.complex128), v.to(torch.complex128))
    assert abs(complex(lhs - rhs)) <= tol * abs(complex(lhs))
    a, b = 0.75 - 0.5j, -1.25 + 0.25j
    lin = plan.matvec(a * u + b * v)
    assert _rel(lin, a * hu + b * hv) <= (1e-13 if dtype == torch.complex128 else 1e-5)


def test_jw25m_cache_identical_to_matrix_free(jw25m):
    """Both cache precisions built in one pass: every cached matvec equals the matrix-free one (the
    product kernel; complex128 blocks of 5000 rows take more than 64 KiB of LDS) bit for bit, and
    the default mode takes whichever the build timed faster."""
    w, plan = jw25m
    _say('jw25m cache == matrix-free')
    rng = np.random.default_rng(403)
    xs = {dt: _cvec(rng, plan.n_total, dt) for dt in (torch.complex64, torch.complex128)}
    assert all(plan.matvec_path(dt) == 'product' for dt in xs)
    free = {dt: plan.matvec(x) for dt, x in xs.items()}
    info = plan.build_cache(0)
    try:
        assert info['built'] and info['nnz'] == plan.count()['hits'] - plan.n_states
        plan.use_cache(True)
        for dt, x in xs.items():
            assert plan.matvec_path(dt) == 'cache'
            assert torch.equal(plan.matvec(x), free[dt]), dt
        plan.use_cache('auto')
        for dt, x in xs.items():
            assert plan.matvec_path(dt) in ('cache', 'product')
            assert torch.equal(plan.matvec(x), free[dt]), dt
    finally:
        plan.use_cache('auto')
        plan.drop_cache()


def test_jw25m_device_build_identical_to_host_build(jw25m):
    """The plan built from device keys (the device sort/unique output, no host copy)
    holds the same index as the host build: equal info and array checksums."""
    w, plan = jw25m
    _say('jw25m device build == host build')
    keys = sort_unique_u64(torch.from_numpy(w.states.view(np.int64)).to(DEV))
    dev = Plan(w.tables, keys)
    try:
        assert dev.info == plan.info and dev.problems == plan.problems
        assert dev.digest() == plan.digest()
    finally:
        dev.close()
        del keys
        torch.cuda.empty_cache()


# ---------------------------------------------------------------- 64 problems of 40 qubits

def _lowest_eigenvalue(h):
    """Lowest eigenvalue of a sparse Hermitian matrix as the minimum over the connected
    components of its graph (dense eigvalsh each): random 40-qubit subspaces are nearly
    diagonal, where Lanczos breaks down."""
    import scipy.sparse.csgraph
    ncomp, label = scipy.sparse.csgraph.connected_components(abs(h), directed=False)
    order = np.argsort(label, kind='stable')
    bounds = np.searchsorted(label[order], np.arange(ncomp + 1))
    best = np.inf
    for c in range(ncomp):
        idx = order[bounds[c]:bounds[c + 1]]
        if idx.shape[0] > 3000:
            raise AssertionError('component too large for a dense solve')
        best = min(best, float(np.linalg.eigvalsh(h[idx][:, idx].toarray())[0]))
    return best


@pytest.fixture(scope='module')
def batch64():
    _say('generate 64 x 40-qubit problems')
    probs = [workloads.batch40_problem(i) for i in range(64)]
    bp = BatchPlan([(p.tables, p.states) for p in probs])
    yield probs, bp
    bp.close()
    torch.cuda.empty_cache()


@pytest.mark.parametrize('dtype', [torch.complex64, torch.complex128])
def test_batch64_matvec_equals_single_plans(batch64, dtype):
    probs, bp = batch64
    _say(f'batch64 == single plans {dtype}')
    rng = np.random.default_rng(501)
    x = _cvec(rng, bp.n_total, dtype)
    y = bp.matvec(x)
    for k, p in enumerate(probs):
        a, n = bp.problems[k]['row_offset'], bp.problems[k]['n_states']
        plan = Plan(p.tables, p.states)
        assert torch.equal(y[a:a + n], plan.matvec(x[a:a + n].contiguous())), k
        plan.close()


def test_batch64_rows_match_oracle(batch64):
    probs, bp = batch64
    _say('batch64 rows vs oracle')
    rng = np.random.default_rng(502)
    x = _cvec(rng, bp.n_total, torch.complex128)
    y = bp.matvec(x).cpu().numpy()
    xh = x.cpu().numpy()
    for k, p in enumerate(probs):
        a, n = bp.problems[k]['row_offset'], bp.problems[k]['n_states']
        rows = np.sort(rng.choice(n, size=min(n, 150), replace=False))
        t = p.tables
        want = sqdref.matvec_rows_u64(t.xmask, t.grp_off, t.zmask, t.coef, p.states, xh[a:a + n], rows)
        got = y[a + rows]
        assert np.max(np.abs(got - want)) <= 1e-12 * max(np.max(np.abs(want)), 1e-300), k


def test_batch64_graph_locg_converges_and_matches_eigsh(batch64):
    """One graph-captured mixed-precision solve of all 64 problems to relative residual 1e-7;
    energies against the oracle's CSR matrices (exact per connected component)."""
    probs, bp = batch64
    _say('batch64 graph locg')
    rng = np.random.default_rng(503)
    x0 = _cvec(rng, bp.n_total, torch.complex128)
    theta, x, it, info = batch_ground_locg(bp, x0, tol=1e-7, maxiter=1000, check_every=10, use_graph=True)
    rel = info['rel_residual'].cpu()
    assert bool(info['converged'].all()) and float(rel.max()) <= 1e-7, float(rel.max())
    # problems with a diagonal and some 2 x 2 blocks, small enough for the oracle's CSR
    # (random 40-qubit subspaces of <= 1e6 states are nearly diagonal: the smallest
    # problems have no nonzero at all)
    for k in (8, 54, 26, 4):
        t, s = probs[k].tables, probs[k].states
        e = _lowest_eigenvalue(sqdref.csr_u64(t.xmask, t.grp_off, t.zmask, t.coef, s))
        assert abs(float(theta[k]) - e) <= 1e-8 * max(abs(e), 1.0), (k, float(theta[k]), e)


# ---------------------------------------------------------------- ground energies against the oracle

def _x0(plan):
    diag = plan.diagonal().real[:plan.n_states]
    x0 = torch.zeros(plan.n_total, dtype=torch.complex128, device=DEV)
    x0[int(torch.argmin(diag))] = 1
    return x0


@pytest.mark.parametrize('name', ['jw-32so-1000x1000', 'jw-40so-1000x1000'])
def test_jw_ground_energy_matches_oracle(name):
    """Plain FP64 LOCG (reference sequence, FP64 hit cache) and the graph-captured
    mixed-precision LOCG (both cache precisions) reach the oracle's eigsh energy to 1e-8."""
    _say(f'{name} ground energy vs oracle')
    w = workloads.jw32() if name.startswith('jw-32') else workloads.jw40(1000, 1000)
    e_star = estar.load(name, w)
    plan = Plan(w.tables, w.states)
    x0 = _x0(plan)
    plan.build_cache(64)
    theta, _, niter = ground_locg(plan.matvec, x0.clone(), maxiter=1000)
    assert abs(float(theta) - e_star) <= 1e-8 * abs(e_star), (float(theta), e_star)
    plan.drop_cache()
    plan.build_cache(0)
    th, _, it, info = batch_ground_locg(plan, x0, tol=1e-12, maxiter=1000, check_every=5, energy_tol=1e-9)
    assert abs(float(th[0]) - e_star) <= 1e-8 * abs(e_star), (float(th[0]), e_star)
    plan.close()
