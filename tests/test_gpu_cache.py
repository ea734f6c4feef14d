"""Hit cache (sliced ELL of the hits): cached matvec equals the matrix-free kernel bit for bit (GPU)."""
import sys

import numpy as np
import pytest
import torch

from oracle import sqdref
import estar
from rqutils_amd import _capi, workloads
from rqutils_amd.paulis.symplectic import TermTables
from rqutils_amd.plan import BatchPlan, Plan

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'
RTOL64 = 1e-12


def _say(msg):
    print(f'[case] {msg}', flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _cache_always(monkeypatch):
    """These tests exercise the cached kernel itself: after a build, matvec reads the cache
    always (the default mode may take the product kernel where the build timed it faster)."""
    build = Plan.build_cache

    def build_and_use(self, *args, **kwargs):
        info = build(self, *args, **kwargs)
        self.use_cache(True)
        return info

    monkeypatch.setattr(Plan, 'build_cache', build_and_use)


def _random_problem(seed, n, t, ns, probs=(0.55, 0.15, 0.15, 0.15), drop_diagonal=False):
    rng = np.random.default_rng(seed)
    codes = workloads.random_pauli_codes(n, t, probs, rng)
    xm, zm = workloads.codes_to_masks(codes)
    if drop_diagonal:
        keep = xm != 0
        xm, zm = xm[keep], zm[keep]
    tables = TermTables.from_masks(xm, zm, rng.standard_normal(xm.shape[0]), n)
    states = workloads.random_subspace(n, ns, rng)
    return tables, states


@pytest.fixture(scope='module')
def cases():
    w = workloads.jw32(n_alpha=40, n_beta=30)
    return {
        'complex12': _random_problem(2, 12, 120, 2000),
        'real_jw': (w.tables, w.states),
        'no_diagonal': _random_problem(5, 11, 80, 1500, drop_diagonal=True),
        'wide16': _random_problem(3, 16, 300, 3000, probs=(0.7, 0.1, 0.1, 0.1)),
    }


def _dtypes(plan):
    return (torch.complex128, torch.complex64) if not plan.is_real else \
        (torch.float64, torch.complex128, torch.float32, torch.complex64)


def _vec(n, dtype, seed):
    rng = np.random.default_rng(seed)
    v = torch.from_numpy(rng.standard_normal(n) + 1j * rng.standard_normal(n)).to(DEV)
    return (v.real if not dtype.is_complex else v).to(dtype).contiguous()


def _check_identical(plan, seed=11):
    for dt in _dtypes(plan):
        x = _vec(plan.n_total, dt, seed)
        plan.use_cache(False)
        y_free = plan.matvec(x)
        plan.use_cache(True)
        y_cache = plan.matvec(x)
        torch.cuda.synchronize()
        assert torch.equal(y_cache, y_free), dt


@pytest.mark.parametrize('name', ['complex12', 'real_jw', 'no_diagonal', 'wide16'])
def test_cached_matvec_identical(cases, name):
    _say(f'cache {name}')
    tables, states = cases[name]
    n_total = len(states) + 37          # filler rows past the states
    plan = Plan(tables, states, n_total=n_total)
    hits = plan.count()['hits']
    for precision in (64, 32):
        info = plan.build_cache(precision)
        assert info['built'] and info['complex'] == (not plan.is_real)
        assert info['nnz'] == hits - (plan.n_states if plan.info['has_diagonal'] else 0)
        # section B: packets of 4 entries for the 8 rows of a chunk; section A: levels of 64 lanes
        assert info['b_entries'] % 32 == 0 and info['a_entries'] % 64 == 0
        assert (info['col_entries'] - info['b_entries']) % 64 == 0
        assert info['entries'] == info['b_entries'] + info['a_entries'] + info['a_levels']
    _check_identical(plan)
    # the cached values are the oracle's matrix elements
    x = _vec(plan.n_total, torch.complex128, 3)
    y = plan.matvec(x).cpu().numpy()
    groups = [(tables.zmask[tables.grp_off[j]:tables.grp_off[j + 1]],
               tables.coef[tables.grp_off[j]:tables.grp_off[j + 1]]) for j in range(tables.num_groups)]
    want = sqdref.matvec_u64(tables.xmask, groups, np.asarray(states, dtype=np.uint64), x.cpu().numpy())
    assert np.max(np.abs(y[:plan.n_states] - want)) <= RTOL64 * max(np.max(np.abs(want)), 1e-300)
    assert np.all(y[plan.n_states:] == 0)
    plan.drop_cache(64)
    assert not plan.cache_info(64)['built'] and plan.cache_info(32)['built']
    plan.drop_cache()
    assert not plan.cache_info(32)['built']


def test_product_subspace_levels_store_offsets():
    """Product subspaces: most first-section levels are a row offset (no stored columns), same matvec."""
    _say('cache product-subspace levels')
    w = workloads.jw32(n_alpha=300, n_beta=200)
    plan = Plan(w.tables, w.states)
    info = plan.build_cache(32)
    stored_a = (info['col_entries'] - info['b_entries']) // 64       # levels that store their columns
    valued_a = info['a_entries'] // 64                                # levels that store 64 values
    print(f"first-section levels {info['a_levels']}, stored columns {stored_a}, stored values {valued_a}", flush=True)
    assert info['a_levels'] > 0 and stored_a < info['a_levels'] and valued_a < info['a_levels']
    _check_identical(plan, seed=19)
    # new coefficients rewrite the single and stored values in place (a full value pass)
    rng = np.random.default_rng(20)
    plan.set_alphas(torch.from_numpy(rng.standard_normal(plan.info['n_terms'])).to(DEV))
    _check_identical(plan, seed=21)
    plan.build_cache(0)                      # both precisions: section B in the column pass, A after
    _check_identical(plan, seed=22)
    plan.set_alphas(torch.from_numpy(rng.standard_normal(plan.info['n_terms'])).to(DEV))
    _check_identical(plan, seed=23)


def test_cached_batch_matvec_identical(cases):
    _say('cache batch')
    bp = BatchPlan([cases[k] for k in ('complex12', 'real_jw', 'no_diagonal', 'wide16')])
    assert bp.info['persistent'] == 1
    bp.build_cache(64)
    bp.build_cache(32)
    _check_identical(bp)
    bp.set_launch(False)            # fill pass by one workgroup per tile gives the same cache
    bp.build_cache(64)
    _check_identical(bp, seed=12)


def test_set_alphas_rewrites_cached_values(cases):
    _say('cache set_alphas')
    tables, states = cases['complex12']
    plan = Plan(tables, states)
    plan.build_cache(64)
    plan.build_cache(32)
    rng = np.random.default_rng(9)
    plan.set_alphas(torch.from_numpy(rng.standard_normal(plan.info['n_terms'])).to(DEV))
    assert plan.cache_info(64)['built']
    _check_identical(plan, seed=13)


def test_real_to_complex_coefficients_release_the_cache(cases):
    _say('cache real->complex')
    tables, states = cases['real_jw']
    plan = Plan(tables, states)
    assert plan.is_real
    plan.build_cache(64)
    rng = np.random.default_rng(4)
    coef = torch.from_numpy(rng.standard_normal(plan.info['n_terms']) +
                            1j * rng.standard_normal(plan.info['n_terms'])).to(DEV).contiguous()
    stream = _capi.current_stream(plan.device)
    _capi.check(_capi.lib().rqx_plan_update_coef(plan._handle, torch.view_as_real(coef).data_ptr(), 0, stream))
    plan.is_real = False
    assert not plan.cache_info(64)['built']
    plan.build_cache(64)
    assert plan.cache_info(64)['complex']
    _check_identical(plan, seed=14)


def test_cache_limit_is_an_error(cases):
    _say('cache limit')
    tables, states = cases['wide16']
    plan = Plan(tables, states)
    with pytest.raises(_capi.RqxError):
        plan.build_cache(64, max_bytes=4096)
    assert not plan.cache_info(64)['built']
    x = _vec(plan.n_total, torch.complex128, 5)
    assert torch.isfinite(plan.matvec(x)).all()


def test_run_sqd_cache_strategies_identical(cases):
    """Every cache_strategy of the reference (``rqutils/sqd.py:363-401``: source cache or not,
    times no / sign-bit / diagonal cache) gives the same eigenpair bit for bit: (1, *) runs on the
    hit cache, (0, *) matrix-free.  The energy is the oracle's dense eigenvalue."""
    _say('cache run_sqd strategies')
    from oracle import dense
    from rqutils_amd.paulis.symplectic import PauliSumXZ, masks_to_packed
    from rqutils_amd.sqd import SQDOptions, run_sqd
    rng = np.random.default_rng(21)
    codes = workloads.random_pauli_codes(11, 70, (0.55, 0.15, 0.15, 0.15), rng)
    paulis = workloads.codes_to_strings(codes)
    coeffs = rng.standard_normal(len(paulis))
    keys = workloads.random_subspace(11, 700, rng)
    h = PauliSumXZ.from_paulisum((paulis, coeffs), force_real=True, add_padding=True)
    states_p = masks_to_packed(keys, 11, padded=True)
    strategies = ((0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (1, 2))
    out = {}
    for cs in strategies:
        out[cs] = run_sqd(h, states_p, 720, options=SQDOptions(cache_strategy=cs, locg_maxiter=2000))
    for cs in strategies[1:]:
        assert float(out[cs][0]) == float(out[(0, 0)][0]), cs
        assert torch.equal(out[cs][1], out[(0, 0)][1]), cs
        assert out[cs][3] == 700
    e_ref = np.linalg.eigvalsh(dense.projected_matrix(paulis, coeffs, keys))[0]
    assert abs(float(out[(0, 0)][0]) - e_ref) <= 1e-8 * abs(e_ref), (float(out[(0, 0)][0]), e_ref)


@pytest.mark.parametrize('name', ['complex12', 'real_jw', 'no_diagonal'])
def test_both_precisions_in_one_pass(cases, name):
    """build_cache(0): one pass fills FP64 and FP32 values over shared columns; both bit-identical."""
    _say(f'cache both {name}')
    tables, states = cases[name]
    plan = Plan(tables, states, n_total=len(states) + 5)
    plan.build_cache(0)
    i64, i32 = plan.cache_info(64), plan.cache_info(32)
    assert i64['built'] and i32['built'] and i64['entries'] == i32['entries'] and i64['nnz'] == i32['nnz']
    _check_identical(plan, seed=15)
    rng = np.random.default_rng(16)
    plan.set_alphas(torch.from_numpy(rng.standard_normal(plan.info['n_terms'])).to(DEV))
    _check_identical(plan, seed=17)
    plan.drop_cache(32)                      # built together, released together
    assert not plan.cache_info(64)['built'] and not plan.cache_info(32)['built']
    plan.build_cache(0)
    plan.build_cache(64)                     # rebuilding one precision replaces the pair
    assert plan.cache_info(64)['built'] and not plan.cache_info(32)['built']
    _check_identical(plan, seed=18)


@pytest.mark.parametrize('cs', [(0, 0), (1, 0)])
def test_run_sqd_mixed_precision_matches_fp64(cs):
    """SQDOptions(locg_precision='mixed'): same ground state as the reference LOCG."""
    _say(f'run_sqd mixed cs={cs}')
    from rqutils_amd.paulis.symplectic import PauliSumXZ, masks_to_packed
    from rqutils_amd.sqd import SQDOptions, run_sqd
    w = workloads.jw32(n_alpha=60, n_beta=50)
    h = PauliSumXZ.from_tables(w.tables, add_padding=True)
    states_p = masks_to_packed(w.states, w.tables.num_qubits, padded=True)
    size = len(w.states) + 9
    e64, v64, _, dim = run_sqd(h, states_p, size, options=SQDOptions(cache_strategy=cs, locg_maxiter=3000))
    emx, vmx, _, dim2 = run_sqd(h, states_p, size,
                                options=SQDOptions(cache_strategy=cs, locg_maxiter=3000, locg_precision='mixed'))
    assert dim == dim2 == len(w.states)
    e_star = estar.load('jw-32so-60x50', w)      # eigsh of the oracle's CSR matrix
    assert abs(float(e64) - e_star) <= 1e-9 * abs(e_star)
    assert abs(float(emx) - e_star) <= 1e-9 * abs(e_star)
    assert abs(float(emx) - float(e64)) <= 1e-9 * abs(float(e64))
    assert abs(abs(torch.vdot(v64, vmx.to(v64.dtype)).item()) - 1.0) < 1e-6
    assert torch.count_nonzero(vmx[dim:]) == 0


def test_section_split_follows_the_product_structure(cases):
    """Product subspaces (alpha strings x beta strings) get the alpha/beta boundary as the cache's
    section split, also where the search split is smaller; other subspaces keep the search split.
    Cached and matrix-free matvecs stay identical and match the oracle."""
    _say('section split')
    w = workloads.jw32(n_alpha=128, n_beta=90)    # beta blocks of two whole 64-row slices
    for split in (-1, 12, 16):
        plan = Plan(w.tables, w.states, split_bits=split)
        pi = plan.problems[0]
        # never below the search split (the first section is a subset of its classes' first groups)
        assert pi['section_bits'] == max(16, pi['split_bits']), pi
        plan.build_cache(0)
        _check_identical(plan, seed=21)
        info = plan.cache_info(64)
        if pi['split_bits'] <= 16:
            # beta flips only, slices inside one beta block: every first-section level is a row
            # offset (no stored first-section column)
            assert info['col_entries'] == info['b_entries'] and info['a_pad_columns'] == 0, info
        rng = np.random.default_rng(22)
        x = rng.standard_normal(plan.n_total) + 1j * rng.standard_normal(plan.n_total)
        y = plan.matvec(torch.from_numpy(x).to(DEV)).cpu().numpy()
        rows = np.sort(rng.choice(plan.n_states, size=200, replace=False))
        t = w.tables
        want = sqdref.matvec_rows_u64(t.xmask, t.grp_off, t.zmask, t.coef, w.states, x, rows)
        assert np.max(np.abs(y[rows] - want)) <= RTOL64 * np.max(np.abs(want))
        plan.close()
    tables, states = cases['complex12']
    plan = Plan(tables, states)
    assert plan.problems[0]['section_bits'] == plan.problems[0]['split_bits']


def test_cached_matvec_graph_replays_match_eager():
    """The cached matvec (both cache precisions built in one pass) captured in a HIP
    graph: every replay, with eager work in between, equals the eager result and the
    matrix-free kernel bit for bit."""
    _say('cached matvec graph replays')
    w = workloads.jw32(n_alpha=200, n_beta=150)
    plan = Plan(w.tables, w.states)
    plan.build_cache(0)
    assert plan.cache_info(64)['a_levels'] > 0 and plan.cache_info(32)['a_levels'] > 0
    rng = np.random.default_rng(41)
    for dt in (torch.complex64, torch.complex128):
        x = torch.from_numpy(rng.standard_normal(plan.n_total) + 1j * rng.standard_normal(plan.n_total)).to(DEV).to(dt)
        want = plan.matvec(x)
        plan.use_cache(False)
        assert torch.equal(plan.matvec(x), want)      # the matrix-free kernel, bit for bit
        plan.use_cache(True)
        y = torch.empty_like(x)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            plan.matvec(x, out=y)
        torch.cuda.current_stream().wait_stream(side)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            plan.matvec(x, out=y)
        for r in range(4):
            y.zero_()
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(y, want), (dt, r)
            other = plan.matvec(2 * x)             # eager work between the replays
            torch.cuda.synchronize()
            assert torch.equal(other, plan.matvec(2 * x))
    plan.close()
