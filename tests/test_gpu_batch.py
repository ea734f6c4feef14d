"""Batched plans: one launch for many problems equals the per-problem kernels (GPU)."""
import sys

import numpy as np
import pytest
import scipy.sparse
import scipy.sparse.linalg
import torch

from oracle import sqdref
from rqutils_amd import workloads
from rqutils_amd.autotune import autotune
from rqutils_amd.locg_batch import batch_ground_locg
from rqutils_amd.paulis.symplectic import TermTables
from rqutils_amd.plan import BatchPlan, Plan, eigenpair_batch

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


def _say(msg):
    print(f'[case] {msg}', flush=True)
    sys.stdout.flush()


def _random_problem(seed, n, t, ns, probs=(0.55, 0.15, 0.15, 0.15)):
    rng = np.random.default_rng(seed)
    codes = workloads.random_pauli_codes(n, t, probs, rng)
    xm, zm = workloads.codes_to_masks(codes)
    tables = TermTables.from_masks(xm, zm, rng.standard_normal(t), n)
    states = workloads.random_subspace(n, ns, rng)
    return tables, states


@pytest.fixture(scope='module')
def problems():
    out = [_random_problem(1, 10, 60, 500), _random_problem(2, 12, 120, 2000),
           _random_problem(3, 16, 300, 3000, probs=(0.7, 0.1, 0.1, 0.1)), _random_problem(4, 9, 20, 50)]
    w = workloads.jw32(n_alpha=30, n_beta=25)
    out.append((w.tables, w.states))
    return out


_CSR = {}


def _csr(t, s):
    """The oracle's projected Hamiltonian (CPU: searchsorted sources, term-by-term sums)."""
    key = (id(t), id(s))
    if key not in _CSR:
        _CSR[key] = sqdref.csr_u64(t.xmask, t.grp_off, t.zmask, t.coef, s)
    return _CSR[key]


def _e0(h):
    if h.shape[0] <= 400:
        return np.linalg.eigvalsh(h.toarray())[0]
    return scipy.sparse.linalg.eigsh(h, k=1, which='SA', tol=1e-13)[0][0]


def test_batch_matvec_equals_single_plans(problems):
    _say('batch matvec == single')
    bp = BatchPlan(problems)
    assert bp.info['n_problems'] == len(problems) and bp.info['persistent'] == 1
    rng = np.random.default_rng(7)
    x = torch.from_numpy(rng.standard_normal(bp.n_total) + 1j * rng.standard_normal(bp.n_total)).to(DEV)
    for dt in (torch.complex128, torch.complex64):
        y = bp.matvec(x.to(dt))
        for k, (t, s) in enumerate(problems):
            a, n = bp.problems[k]['row_offset'], bp.problems[k]['n_states']
            yk = Plan(t, s).matvec(x[a:a + n].to(dt).contiguous())
            assert torch.equal(y[a:a + n], yk), k
    # the same plan launched one workgroup per tile gives identical numbers
    y_q = bp.matvec(x)
    bp.set_launch(False)
    assert torch.equal(bp.matvec(x), y_q)


def test_batch_of_listed_products_equals_single_plans_and_oracle():
    """Two product subspaces (fully listed at split 10) between random problems in one batched plan:
    every problem's rows equal its own plan bit for bit and the oracle; the cache gives the same bits."""
    _say('batch of listed products')
    w1, w2 = workloads.jw32(n_alpha=40, n_beta=30), workloads.jw32(n_alpha=25, n_beta=45)
    probs = [_random_problem(11, 12, 100, 900), (w1.tables, w1.states), _random_problem(12, 10, 40, 300),
             (w2.tables, w2.states)]
    bp = BatchPlan(probs, split_bits=10)
    assert [p['list_mode'] for p in bp.problems][1::2] == [2, 2]
    assert [p['alpha_rows'] for p in bp.problems][1::2] == [40, 25]
    rng = np.random.default_rng(17)
    x = torch.from_numpy(rng.standard_normal(bp.n_total) + 1j * rng.standard_normal(bp.n_total)).to(DEV)
    for dt in (torch.complex128, torch.complex64):
        y = bp.matvec(x.to(dt))
        for k, (t, s) in enumerate(probs):
            a, n = bp.problems[k]['row_offset'], bp.problems[k]['n_states']
            yk = Plan(t, s, split_bits=10).matvec(x[a:a + n].to(dt).contiguous())
            assert torch.equal(y[a:a + n], yk), k
            if dt == torch.complex128:
                want = _csr(t, s) @ x[a:a + n].cpu().numpy()
                got = y[a:a + n].cpu().numpy()
                assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), k
    y_free = bp.matvec(x)
    bp.build_cache(64)
    assert torch.equal(bp.matvec(x), y_free)


def test_persistent_workgroups_take_several_tiles(problems):
    """More tiles than persistent workgroups: every mode equals the one-tile-per-workgroup launch."""
    _say('persistent multi-tile')
    w = workloads.jw32(n_alpha=300, n_beta=300)
    bp = BatchPlan(list(problems) + [(w.tables, w.states)])
    bp.set_launch(True, 1)                       # grid = number of CUs
    assert bp.info['n_tiles'] > 2 * bp.info['n_cu']
    rng = np.random.default_rng(12)
    x = torch.from_numpy(rng.standard_normal(bp.n_total) + 1j * rng.standard_normal(bp.n_total)).to(DEV)
    gy = torch.from_numpy(rng.standard_normal(bp.n_total) + 1j * rng.standard_normal(bp.n_total)).to(DEV)
    c_q, y_q, g_q = bp.count(), bp.matvec(x), bp.grad_coef(x, gy)
    bp.set_launch(False)
    c_t, y_t, g_t = bp.count(), bp.matvec(x), bp.grad_coef(x, gy)
    assert c_q == c_t
    assert torch.equal(y_q, y_t)
    assert torch.allclose(g_q, g_t, rtol=1e-12, atol=1e-12 * g_t.abs().max().item())


def test_batch_grad_equals_single(problems):
    _say('batch grad == single')
    bp = BatchPlan(problems)
    rng = np.random.default_rng(8)
    x = torch.from_numpy(rng.standard_normal(bp.n_total) + 1j * rng.standard_normal(bp.n_total)).to(DEV)
    gy = torch.from_numpy(rng.standard_normal(bp.n_total) + 1j * rng.standard_normal(bp.n_total)).to(DEV)
    g = bp.grad_coef(x, gy)
    for k, (t, s) in enumerate(problems):
        pk = bp.problems[k]
        a, n, ta, tn = pk['row_offset'], pk['n_states'], pk['term_offset'], pk['n_terms']
        gk = Plan(t, s).grad_coef(x[a:a + n].contiguous(), gy[a:a + n].contiguous())
        assert torch.allclose(g[ta:ta + tn], gk, rtol=1e-12, atol=1e-12 * gk.abs().max().item())


def test_batch_vector_kernels(problems):
    _say('batch dots/combine/eigen')
    bp = BatchPlan(problems)
    rng = np.random.default_rng(9)
    vs = [torch.from_numpy(rng.standard_normal(bp.n_total) + 1j * rng.standard_normal(bp.n_total)).to(DEV)
          for _ in range(3)]
    d = bp.dots([(vs[0], vs[1]), (vs[2], vs[2])])
    coef = torch.from_numpy(rng.standard_normal((bp.n_problems, 3)) + 1j * rng.standard_normal((bp.n_problems, 3))).to(DEV)
    out = torch.empty_like(vs[0])
    bp.combine(vs, coef, out)
    for k in range(bp.n_problems):
        a, n = bp.problems[k]['row_offset'], bp.problems[k]['n_states']
        sl = slice(a, a + n)
        assert abs(d[k, 0] - torch.vdot(vs[0][sl], vs[1][sl])) < 1e-10 * n
        assert abs(d[k, 1] - torch.vdot(vs[2][sl], vs[2][sl])) < 1e-10 * n
        want = coef[k, 0] * vs[0][sl] + coef[k, 1] * vs[1][sl] + coef[k, 2] * vs[2][sl]
        assert torch.allclose(out[sl], want, rtol=1e-13, atol=1e-12)
    # in-place combine
    v0 = vs[0].clone()
    bp.combine([v0, vs[1]], coef[:, :2], v0)
    for k in range(bp.n_problems):
        a, n = bp.problems[k]['row_offset'], bp.problems[k]['n_states']
        sl = slice(a, a + n)
        assert torch.allclose(v0[sl], coef[k, 0] * vs[0][sl] + coef[k, 1] * vs[1][sl], rtol=1e-13, atol=1e-12)
    m = torch.from_numpy(rng.standard_normal((50, 3, 3)) + 1j * rng.standard_normal((50, 3, 3)))
    m = (m + m.conj().transpose(1, 2)) / 2
    th, c = eigenpair_batch(m.to(DEV))
    w = np.linalg.eigvalsh(m.numpy())[:, 0]
    assert np.allclose(th.cpu().numpy(), w, atol=1e-10)
    res = torch.linalg.vector_norm(m.to(DEV) @ c[:, :, None] - th[:, None, None] * c[:, :, None], dim=(1, 2))
    assert float(res.max()) < 1e-8


@pytest.mark.parametrize('use_graph', [False, True])
def test_batch_locg_matches_eigsh(problems, use_graph):
    _say(f'batch locg graph={use_graph}')
    bp = BatchPlan(problems)
    rng = np.random.default_rng(10)
    x0 = torch.from_numpy(rng.standard_normal(bp.n_total) + 1j * rng.standard_normal(bp.n_total)).to(DEV)
    theta, x, it, info = batch_ground_locg(bp, x0, tol=1e-7, maxiter=3000, check_every=10, use_graph=use_graph)
    for k, (t, s) in enumerate(problems):
        e = _e0(_csr(t, s))
        assert abs(float(theta[k]) - e) < 1e-8 * max(1., abs(e)), (k, float(theta[k]), e)
    assert info['switch_iteration'] is not None


@pytest.mark.parametrize('cached', [False, True])
def test_single_plan_graph_locg_energy_stop(problems, cached):
    """A single Plan drives the graph-captured step; the energy-tail stop reaches 1e-8 of eigsh."""
    _say(f'single-plan graph locg cached={cached}')
    t, s = problems[4]                       # JW model, real
    plan = Plan(t, s, n_total=len(s) + 11)   # filler rows stay 0
    if cached:
        plan.build_cache(32)
        plan.build_cache(64)
    diag = plan.diagonal().real[:plan.n_states]
    x0 = torch.zeros(plan.n_total, dtype=torch.complex128, device=DEV)
    x0[int(torch.argmin(diag))] = 1
    theta, x, it, info = batch_ground_locg(plan, x0, tol=1e-12, maxiter=3000, check_every=5,
                                           energy_tol=3e-10)
    e = _e0(_csr(t, s))
    assert abs(float(theta[0]) - e) < 1e-8 * abs(e), (float(theta[0]), e)
    assert it < 3000 and info['switch_iteration'] is not None
    assert torch.count_nonzero(x[plan.n_states:]) == 0


@pytest.mark.parametrize('mixed', [False, True])
def test_fused_step_follows_reference_step(problems, mixed):
    """rqx_locg_step against the same iteration built from per-operation kernels."""
    _say(f'fused step vs reference mixed={mixed}')
    from rqutils_amd.locg_batch import BatchLOCG
    bp = BatchPlan(problems[:3] + problems[4:])
    rng = np.random.default_rng(12)
    x0 = torch.from_numpy(rng.standard_normal(bp.n_total) + 1j * rng.standard_normal(bp.n_total)).to(DEV)
    a = BatchLOCG(bp, x0, tol=0.0, mixed=mixed, lo_storage=False)
    b = BatchLOCG(bp, x0, tol=0.0, mixed=mixed, lo_storage=False)
    for it in range(25):
        a.step()
        b.step_reference()
        torch.cuda.synchronize()
        assert torch.allclose(a.theta, b.theta, rtol=1e-9, atol=0), (it, a.theta, b.theta)
    for u, v in ((a.x, b.x), (a.ax, b.ax), (a.y, b.y)):
        assert float((u - v).abs().max()) < 1e-6 * float(v.abs().max())


def test_fused_step_real_vectors(problems):
    """float64 vectors (real Hamiltonian): the fused step keeps real coefficients."""
    _say('fused step real')
    from rqutils_amd.locg_batch import BatchLOCG
    t, s = problems[4]
    plan = Plan(t, s)
    rng = np.random.default_rng(13)
    x0 = torch.from_numpy(rng.standard_normal(plan.n_total)).to(DEV)
    a = BatchLOCG(plan, x0, tol=0.0, mixed=True, lo_storage=False)
    b = BatchLOCG(plan, x0, tol=0.0, mixed=True, lo_storage=False)
    for _ in range(20):
        a.step()
        b.step_reference()
    torch.cuda.synchronize()
    assert torch.allclose(a.theta, b.theta, rtol=1e-9, atol=0)
    assert a.x.dtype == torch.float64 and bool(torch.isfinite(a.x).all())
    # float32-stored bulk phase on real vectors converges to the same ground energy
    e = _e0(_csr(t, s))
    theta, x, it, info = batch_ground_locg(plan, x0, tol=1e-9, maxiter=3000, check_every=10, lo_storage=True)
    assert x.dtype == torch.float64 and info['switch_iteration'] is not None
    assert abs(float(theta[0]) - e) < 1e-9 * abs(e)


@pytest.mark.parametrize('lo_storage', [False, True])
def test_bulk_phase_storage_reaches_the_same_ground_state(problems, lo_storage):
    """Bulk phase with FP32-stored vectors or FP64 vectors + low copies: both converge to eigsh."""
    _say(f'bulk storage lo={lo_storage}')
    bp = BatchPlan(problems[:2] + problems[4:])
    rng = np.random.default_rng(14)
    x0 = torch.from_numpy(rng.standard_normal(bp.n_total) + 1j * rng.standard_normal(bp.n_total)).to(DEV)
    theta, x, it, info = batch_ground_locg(bp, x0, tol=1e-9, maxiter=3000, check_every=10, lo_storage=lo_storage)
    assert x.dtype == torch.complex128 and info['switch_iteration'] is not None
    for k, (t, s) in enumerate(problems[:2] + problems[4:]):
        e = _e0(_csr(t, s))
        assert abs(float(theta[k]) - e) < 1e-9 * max(1., abs(e)), (k, float(theta[k]), e)


def test_batch_graph_replays_match_eager(problems):
    """The captured step replayed N times gives the eager trajectory."""
    _say('graph vs eager')
    bp = BatchPlan(problems[:3])
    rng = np.random.default_rng(11)
    x0 = torch.from_numpy(rng.standard_normal(bp.n_total) + 1j * rng.standard_normal(bp.n_total)).to(DEV)
    th_e, _, _, _ = batch_ground_locg(bp, x0, tol=0.0, maxiter=41, check_every=20, use_graph=False, mixed=False)
    th_g, _, _, _ = batch_ground_locg(bp, x0, tol=0.0, maxiter=41, check_every=20, use_graph=True, mixed=False)
    assert torch.allclose(th_e, th_g, rtol=1e-10, atol=1e-12)


def test_autotune_picks_a_valid_plan():
    _say('autotune')
    w = workloads.jw32(n_alpha=200, n_beta=200)
    plan, results = autotune(w.tables, w.states, reps=2)
    assert results and results[0]['ms'] <= results[-1]['ms']
    assert plan.info['split_bits'] == results[0]['split_bits']
    ref = Plan(w.tables, w.states)
    x = torch.randn(plan.n_total, dtype=torch.float64, device=DEV)
    assert torch.allclose(plan.matvec(x), ref.matvec(x), rtol=0, atol=1e-12)


def test_autotune_batch_picks_a_launch_shape():
    """The batched autotuner leaves the fastest measured launch shape set; the matvec is unchanged."""
    _say('autotune batch')
    from rqutils_amd.autotune import autotune_batch
    probs = [_random_problem(70 + k, 12, 60, 400 + 300 * k) for k in range(4)]
    ref = BatchPlan(probs)
    x = torch.randn(ref.n_total, dtype=torch.complex128, device=DEV)
    want = ref.matvec(x)
    plan, results = autotune_batch(probs, reps=2)
    assert len(results) == 6 and results[0]['ms'] <= results[-1]['ms']
    assert plan.info['persistent'] == int(results[0]['persistent'])
    assert torch.equal(plan.matvec(x), want)


def test_batch_locg_is_deterministic(problems):
    """Per-problem dots are summed in a fixed order: two solves give the same bits and iteration count."""
    _say('batch locg deterministic')
    bp = BatchPlan(problems)
    rng = np.random.default_rng(19)
    x0 = torch.from_numpy(rng.standard_normal(bp.n_total) + 1j * rng.standard_normal(bp.n_total)).to(DEV)
    runs = [batch_ground_locg(bp, x0, tol=1e-8, maxiter=3000, check_every=5, use_graph=True) for _ in range(2)]
    (th_a, x_a, it_a, _), (th_b, x_b, it_b, _) = runs
    assert it_a == it_b
    assert torch.equal(th_a, th_b) and torch.equal(x_a, x_b)
    d1 = bp.dots([(x_a, x0), (x0, x0)])
    d2 = bp.dots([(x_a, x0), (x0, x0)])
    assert torch.equal(d1, d2)


def test_zero_ground_energy_converges():
    """A Hamiltonian shifted so that its ground energy is 0 still meets the relative residual."""
    _say('zero ground energy')
    t, s = _random_problem(5, 10, 60, 300)
    e0 = scipy.sparse.linalg.eigsh(_csr(t, s), k=1, which='SA', tol=1e-13)[0][0] if len(s) > 400 else \
        np.linalg.eigvalsh(_csr(t, s).toarray())[0]
    shifted = TermTables.from_masks(np.concatenate([t.term_xmask(), [0]]).astype(np.uint64),
                                    np.concatenate([t.zmask, [0]]).astype(np.uint64),
                                    np.concatenate([t.alphas().real, [-e0]]), t.num_qubits)
    plan = Plan(shifted, s)
    rng = np.random.default_rng(20)
    x0 = torch.from_numpy(rng.standard_normal(plan.n_total) + 1j * rng.standard_normal(plan.n_total)).to(DEV)
    theta, x, it, info = batch_ground_locg(plan, x0, tol=1e-8, maxiter=3000, check_every=5)
    assert bool(info['converged'][0]) and it < 3000
    assert abs(float(theta[0])) <= 1e-7 * abs(e0)
