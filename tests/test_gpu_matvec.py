"""Parity of the gfx950 lookup / matvec kernels with the oracle (GPU)."""
import os
import sys

import numpy as np
import pytest
import torch

from oracle import dense, sqdref
from rqutils_amd import workloads
from rqutils_amd.paulis.symplectic import PauliSumXZ, TermTables
from rqutils_amd.plan import Plan, eigenpair_small, sort_unique_u64

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
DEV = 'cuda:0'

# tolerances written per dtype: FP64 paths must agree to 1e-12 relative (max norm)
RTOL64 = 1e-12
RTOL32 = 2e-5


def _say(msg):
    print(f'[case] {msg}', flush=True)
    sys.stdout.flush()


def _groups(t: TermTables):
    return [(t.zmask[t.grp_off[j]:t.grp_off[j + 1]], t.coef[t.grp_off[j]:t.grp_off[j + 1]])
            for j in range(t.num_groups)]


def _oracle_lookup(t, keys, n_total):
    out = np.full((t.num_groups, n_total), -1, dtype=np.int64)
    for j, xm in enumerate(t.xmask):
        out[j, :keys.shape[0]] = sqdref.xsource_searchsorted(keys, xm)
    return out


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope='module')
def rand20():
    return workloads.random20_300t()


def test_lookup_bit_exact_rand20(rand20):
    _say('lookup rand20')
    plan = Plan(rand20.tables, rand20.states)
    got = plan.lookup().cpu().numpy()
    want = _oracle_lookup(rand20.tables, rand20.states, plan.n_total)
    assert np.array_equal(got, want)


@pytest.mark.parametrize('split', [-1, 0, 7, 13])
def test_matvec_c128_rand20_all_splits(rand20, split):
    _say(f'matvec c128 rand20 split={split}')
    plan = Plan(rand20.tables, rand20.states, split_bits=split)
    rng = np.random.default_rng(2)
    v = rng.standard_normal(plan.n_total) + 1j * rng.standard_normal(plan.n_total)
    y = plan.matvec(torch.from_numpy(v).to(DEV)).cpu().numpy()
    want = sqdref.matvec_u64(rand20.tables.xmask, _groups(rand20.tables), rand20.states, v)
    assert _rel(y, want) < RTOL64


def test_matvec_split_invariance(rand20):
    """The split only moves work between the two accumulation sections: same values to rounding."""
    _say('matvec split invariance')
    rng = np.random.default_rng(3)
    v = torch.from_numpy(rng.standard_normal(8192) + 1j * rng.standard_normal(8192)).to(DEV)
    ys = [Plan(rand20.tables, rand20.states, split_bits=s).matvec(v) for s in (0, 5, 11, 13)]
    scale = float(ys[0].abs().max())
    for y in ys[1:]:
        assert float((y - ys[0]).abs().max()) <= 1e-14 * scale


def test_matvec_c64_rand20(rand20):
    _say('matvec c64 rand20')
    plan = Plan(rand20.tables, rand20.states)
    rng = np.random.default_rng(4)
    v = rng.standard_normal(plan.n_total) + 1j * rng.standard_normal(plan.n_total)
    y = plan.matvec(torch.from_numpy(v.astype(np.complex64)).to(DEV)).cpu().numpy()
    want = sqdref.matvec_u64(rand20.tables.xmask, _groups(rand20.tables), rand20.states, v.astype(np.complex64))
    assert _rel(y, want) < RTOL32


def test_real_dtype_rejected_for_complex_hamiltonian(rand20):
    plan = Plan(rand20.tables, rand20.states)
    with pytest.raises(ValueError, match='complex'):
        plan.matvec(torch.zeros(plan.n_total, dtype=torch.float64, device=DEV))


@pytest.fixture(scope='module')
def jw_small():
    return workloads.jw32(n_alpha=60, n_beta=50)


def test_lookup_and_matvec_jw_small(jw_small):
    _say('jw small lookup/matvec')
    t, keys = jw_small.tables, jw_small.states
    plan = Plan(t, keys, n_total=keys.shape[0] + 17)
    assert plan.is_real
    got = plan.lookup().cpu().numpy()
    assert np.array_equal(got, _oracle_lookup(t, keys, plan.n_total))
    rng = np.random.default_rng(5)
    v = rng.standard_normal(plan.n_total)
    v[keys.shape[0]:] = rng.standard_normal(17)      # filler inputs must not leak
    want = np.zeros(plan.n_total, dtype=np.complex128)
    want[:keys.shape[0]] = sqdref.matvec_u64(t.xmask, _groups(t), keys, v[:keys.shape[0]])
    y64 = plan.matvec(torch.from_numpy(v).to(DEV)).cpu().numpy()
    assert _rel(y64, want.real) < RTOL64
    assert np.all(y64[keys.shape[0]:] == 0)
    yc = plan.matvec(torch.from_numpy(v.astype(np.complex128)).to(DEV)).cpu().numpy()
    assert _rel(yc, want) < RTOL64
    y32 = plan.matvec(torch.from_numpy(v.astype(np.float32)).to(DEV)).cpu().numpy()
    assert _rel(y32, want.real) < RTOL32
    d = plan.diagonal().cpu().numpy()
    z0, c0 = _groups(t)[0]
    assert np.allclose(d[:keys.shape[0]], sqdref.diagonal_u64(z0, c0, keys), rtol=0, atol=1e-12)
    assert np.all(d[keys.shape[0]:] == 0)


def test_alpha_lists_product_subspace_match_oracle():
    """A product subspace (60 alpha x 50 beta strings): the matvec takes every hit from lists -- per alpha
    string for the groups that flip only alpha bits (split values for those whose terms hold beta Z
    bits), per beta string for the groups that flip only beta bits, and alpha x beta partners through a
    pattern map for the rest.  Every dtype equals the oracle, also after complex coefficients replace
    the real ones (values and G tables rewritten); one state less is no product and has no lists."""
    _say('alpha lists')
    from rqutils_amd import _capi
    w = workloads.jw32(n_alpha=60, n_beta=50)
    t, keys = w.tables, w.states
    plan = Plan(t, keys, split_bits=10)          # segments below the alpha bits (a tiny problem's default is wider)
    assert plan.problems[0]['section_bits'] == 16
    assert plan.problems[0]['alpha_rows'] == 60
    assert plan.problems[0]['list_mode'] == 2          # every group from the lists: no search
    rng = np.random.default_rng(51)
    v = rng.standard_normal(plan.n_total) + 1j * rng.standard_normal(plan.n_total)
    want = sqdref.matvec_u64(t.xmask, _groups(t), keys, v)
    assert _rel(plan.matvec(torch.from_numpy(v).to(DEV)).cpu().numpy(), want) < RTOL64
    assert _rel(plan.matvec(torch.from_numpy(v.astype(np.complex64)).to(DEV)).cpu().numpy(), want) < RTOL32
    want_re = sqdref.matvec_u64(t.xmask, _groups(t), keys, v.real)
    assert _rel(plan.matvec(torch.from_numpy(v.real.copy()).to(DEV)).cpu().numpy(), want_re.real) < RTOL64
    assert _rel(plan.matvec(torch.from_numpy(v.real.astype(np.float32)).to(DEV)).cpu().numpy(), want_re.real) < RTOL32
    coef = rng.standard_normal(plan.info['n_terms']) + 1j * rng.standard_normal(plan.info['n_terms'])
    ct = torch.from_numpy(coef).to(DEV).contiguous()
    _capi.check(_capi.lib().rqx_plan_update_coef(plan._handle, torch.view_as_real(ct).data_ptr(), 0,
                                                 _capi.current_stream(plan.device)))
    plan.is_real = False
    t2 = TermTables(t.xmask, t.grp_off, t.zmask, coef, t.num_qubits)
    want2 = sqdref.matvec_u64(t2.xmask, _groups(t2), keys, v)
    assert _rel(plan.matvec(torch.from_numpy(v).to(DEV)).cpu().numpy(), want2) < RTOL64
    assert _rel(plan.matvec(torch.from_numpy(v.astype(np.complex64)).to(DEV)).cpu().numpy(), want2) < RTOL32
    # the hit cache sums the same values in the same order: bit for bit the matrix-free result;
    # the default path ('auto') is whichever build_cache timed faster, the same bits again
    plan.build_cache(0)
    for dt in (torch.complex128, torch.complex64):
        x = torch.from_numpy(v).to(DEV).to(dt)
        plan.use_cache(False)
        assert plan.matvec_path(dt) == 'product'
        y_free = plan.matvec(x)
        plan.use_cache(True)
        assert plan.matvec_path(dt) == 'cache'
        assert torch.equal(plan.matvec(x), y_free), dt
        plan.use_cache('auto')
        assert plan.matvec_path(dt) in ('cache', 'product')
        assert torch.equal(plan.matvec(x), y_free), dt
    plan.drop_cache()
    assert plan.matvec_path(torch.complex64) == 'product'
    no_product = Plan(t, keys[1:], split_bits=10).problems[0]
    assert no_product['alpha_rows'] == 0 and no_product['list_mode'] == 0
    # the plan built from device-resident states finds the same product and the same lists
    host = Plan(t, keys, split_bits=10)
    dev = Plan(t, torch.from_numpy(keys.view(np.int64)).to(DEV), split_bits=10)
    assert dev.problems[0]['list_mode'] == 2 and dev.problems[0]['alpha_rows'] == 60
    xv = torch.from_numpy(v).to(DEV)
    assert torch.equal(dev.matvec(xv), host.matvec(xv))


def test_product_partners_mixed_evaluation_match_oracle():
    """Partner hits of a product subspace whose map holds groups of three kinds: sign-table groups whose
    first term matches the alpha pattern's reference Z part but for the pattern's X bits (values staged
    per beta partner), table groups with an extra alpha Z bit outside the pattern (evaluated per row),
    and groups of one or two terms without a table (evaluated per row).  Every dtype equals the oracle,
    and the hit cache equals the matrix-free result bit for bit."""
    _say('product partners, mixed evaluation')
    w = workloads.jw32(n_alpha=60, n_beta=50)
    t, keys = w.tables, w.states
    amask = np.uint64((1 << 16) - 1)
    rng = np.random.default_rng(77)
    xm, zm, cf, off = [], [], [], [0]
    n_thin = n_moved = 0
    for j in range(t.num_groups):
        x = np.uint64(t.xmask[j])
        z = t.zmask[t.grp_off[j]:t.grp_off[j + 1]].astype(np.uint64).copy()
        c = t.coef[t.grp_off[j]:t.grp_off[j + 1]].copy()
        mixed = (x & amask) != 0 and (x & ~amask) != 0
        if mixed and z.size >= 3:
            r = rng.random()
            if r < 0.3:                    # one or two terms: no sign table
                k = int(rng.integers(1, 3))
                z, c = z[:k], c[:k]
                n_thin += 1
            elif r < 0.5:                  # an alpha Z bit outside the pattern on every term
                free = [b for b in range(16) if not (int(x) >> b) & 1]
                z = z ^ np.uint64(1 << int(rng.choice(free)))
                n_moved += 1
        xm.append(x)
        zm.extend(z.tolist())
        cf.extend(c.tolist())
        off.append(len(zm))
    assert n_thin > 0 and n_moved > 0
    t3 = TermTables(np.array(xm, dtype=np.uint64), np.array(off, dtype=np.int64), np.array(zm, dtype=np.uint64),
                    np.array(cf), t.num_qubits)
    plan = Plan(t3, keys, split_bits=10)
    assert plan.problems[0]['list_mode'] == 2
    rng = np.random.default_rng(78)
    v = rng.standard_normal(plan.n_total) + 1j * rng.standard_normal(plan.n_total)
    want = sqdref.matvec_u64(t3.xmask, _groups(t3), keys, v)
    assert _rel(plan.matvec(torch.from_numpy(v).to(DEV)).cpu().numpy(), want) < RTOL64
    assert _rel(plan.matvec(torch.from_numpy(v.astype(np.complex64)).to(DEV)).cpu().numpy(), want) < RTOL32
    want_re = sqdref.matvec_u64(t3.xmask, _groups(t3), keys, v.real)
    assert _rel(plan.matvec(torch.from_numpy(v.real.copy()).to(DEV)).cpu().numpy(), want_re.real) < RTOL64
    assert _rel(plan.matvec(torch.from_numpy(v.real.astype(np.float32)).to(DEV)).cpu().numpy(), want_re.real) < RTOL32
    plan.build_cache(0)
    for dt in (torch.complex128, torch.complex64, torch.float64, torch.float32):
        x = torch.from_numpy(v.real.copy() if not dt.is_complex else v).to(DEV).to(dt)
        plan.use_cache(False)
        y_free = plan.matvec(x)
        plan.use_cache(True)
        assert torch.equal(plan.matvec(x), y_free), dt


def test_product_kernel_real_coefficient_update_matches_oracle():
    """A real coefficient update on a fully listed product plan rewrites the alpha-list entries,
    their one-word form for real coefficients (what the product kernel reads), the G tables and
    the beta lists: the matrix-free matvec equals the oracle with the new coefficients, and the
    cache built afterwards equals it bit for bit."""
    _say('product kernel, real coefficient update')
    from rqutils_amd import _capi
    w = workloads.jw32(n_alpha=60, n_beta=50)
    t, keys = w.tables, w.states
    plan = Plan(t, keys, split_bits=10)
    assert plan.problems[0]['list_mode'] == 2 and plan.is_real
    rng = np.random.default_rng(95)
    coef = rng.standard_normal(plan.info['n_terms']).astype(np.complex128)
    ct = torch.from_numpy(coef).to(DEV).contiguous()
    _capi.check(_capi.lib().rqx_plan_update_coef(plan._handle, torch.view_as_real(ct).data_ptr(), 1,
                                                 _capi.current_stream(plan.device)))
    t2 = TermTables(t.xmask, t.grp_off, t.zmask, coef, t.num_qubits)
    v = rng.standard_normal(plan.n_total) + 1j * rng.standard_normal(plan.n_total)
    want = sqdref.matvec_u64(t2.xmask, _groups(t2), keys, v)
    free = {}
    for dt, tol in ((torch.complex128, RTOL64), (torch.complex64, RTOL32)):
        assert plan.matvec_path(dt) == 'product'
        free[dt] = plan.matvec(torch.from_numpy(v).to(DEV).to(dt))
        assert _rel(free[dt].cpu().numpy(), want) < tol, dt
    want_re = sqdref.matvec_u64(t2.xmask, _groups(t2), keys, v.real)
    assert _rel(plan.matvec(torch.from_numpy(v.real.astype(np.float32)).to(DEV)).cpu().numpy(), want_re.real) < RTOL32
    plan.build_cache(0)
    plan.use_cache(True)
    for dt, y in free.items():
        assert torch.equal(plan.matvec(torch.from_numpy(v).to(DEV).to(dt)), y), dt
    plan.use_cache('auto')


def test_product_kernel_blocks_beyond_64k_lds():
    """4200 alpha x 3 beta strings: a complex128 beta block (4200 x 16 bytes of x plus its staged
    values) needs more than 64 KiB of LDS, so the product kernel runs one workgroup per CU with the
    device's LDS; complex64 blocks fit 64 KiB.  Both equal the oracle and the cache bit for bit."""
    _say('product kernel, large blocks')
    w = workloads.jw32(n_alpha=4200, n_beta=3)
    t, keys = w.tables, w.states
    plan = Plan(t, keys, split_bits=10)          # segments below the alpha bits
    assert plan.problems[0]['list_mode'] == 2 and plan.problems[0]['alpha_rows'] == 4200
    rng = np.random.default_rng(91)
    v = rng.standard_normal(plan.n_total) + 1j * rng.standard_normal(plan.n_total)
    want = sqdref.matvec_u64(t.xmask, _groups(t), keys, v)
    free = {}
    for dt, tol in ((torch.complex128, RTOL64), (torch.complex64, RTOL32)):
        assert plan.matvec_path(dt) == 'product'
        free[dt] = plan.matvec(torch.from_numpy(v).to(DEV).to(dt))
        assert _rel(free[dt].cpu().numpy(), want) < tol, dt
    plan.build_cache(0)
    plan.use_cache(True)
    for dt, y in free.items():
        assert torch.equal(plan.matvec(torch.from_numpy(v).to(DEV).to(dt)), y), dt
    plan.use_cache('auto')


@pytest.mark.parametrize('name', ['small_10q', 'filler_12q'])
def test_golden_fixtures(name):
    _say(f'golden {name}')
    f = np.load(os.path.join(HERE, f'{name}.npz'))
    h = PauliSumXZ.from_paulisum(([str(p) for p in f['paulis']], f['coeffs']), add_padding=True)
    plan = Plan(h.tables, f['keys'], n_total=int(f['states_size']))
    n = f['keys'].shape[0]
    assert np.array_equal(plan.lookup().cpu().numpy()[:, :n], f['xsource'])
    y = plan.matvec(torch.from_numpy(f['v']).to(DEV)).cpu().numpy()
    assert _rel(y, f['y']) < RTOL64
    # the hit cache (both precisions) reproduces the golden vector as well
    plan.build_cache(0)
    yc = plan.matvec(torch.from_numpy(f['v']).to(DEV)).cpu().numpy()
    assert _rel(yc, f['y']) < RTOL64 and np.array_equal(yc, y)
    # and the nonzero triplets are the fixture's (source, element) pairs
    rows, cols, vals = plan.hits()
    xs = f['xsource']
    assert rows.shape[0] == int((xs >= 0).sum())


def test_dense_projection_small():
    _say('dense projection')
    rng = np.random.default_rng(8)
    n = 10
    codes = rng.choice(4, size=(80, n), p=(0.5, 0.2, 0.15, 0.15))
    paulis = [''.join('IXYZ'[c] for c in row) for row in codes]
    coeffs = rng.standard_normal(80)
    keys = np.sort(rng.choice(1 << n, size=400, replace=False).astype(np.uint64))
    h = PauliSumXZ.from_paulisum((paulis, coeffs))
    plan = Plan(h.tables, keys)
    hs = dense.projected_matrix(paulis, coeffs, keys)
    eye = torch.eye(400, dtype=torch.complex128, device=DEV)
    cols = torch.stack([plan.matvec(eye[:, k].contiguous()) for k in range(400)], dim=1).cpu().numpy()
    assert np.allclose(cols, hs, rtol=0, atol=1e-12 * np.abs(hs).max())


def test_edge_cases():
    _say('edge cases')
    # x=3, z=3 is a YY-type term: phase (-i)^2 = -1 keeps the Hamiltonian real
    t = TermTables.from_masks(np.array([0, 3], np.uint64), np.array([1, 3], np.uint64), np.array([0.5, -1.]), 4)
    assert t.is_real
    # one state: only the diagonal survives
    plan = Plan(t, np.array([5], np.uint64), n_total=3)
    y = plan.matvec(torch.tensor([2., 7., 7.], dtype=torch.float64, device=DEV)).cpu().numpy()
    want_d = 0.5 * (1 - 2 * (bin(5 & 1).count('1') & 1))
    assert np.allclose(y, [want_d * 2., 0, 0])
    # empty subspace
    plan = Plan(t, np.zeros(0, np.uint64), n_total=4)
    y = plan.matvec(torch.ones(4, dtype=torch.complex128, device=DEV))
    assert torch.count_nonzero(y) == 0
    # no diagonal group
    t2 = TermTables.from_masks(np.array([1, 6], np.uint64), np.array([0, 6], np.uint64), np.array([1., 2.]), 3)
    assert t2.is_real
    keys = np.arange(8, dtype=np.uint64)
    plan = Plan(t2, keys)
    assert plan.info['has_diagonal'] == 0
    v = np.arange(8, dtype=np.float64) + 1
    want = sqdref.matvec_u64(t2.xmask, _groups(t2), keys, v).real
    assert np.allclose(plan.matvec(torch.from_numpy(v).to(DEV)).cpu().numpy(), want)
    # no terms at all
    t3 = TermTables.from_masks(np.zeros(0, np.uint64), np.zeros(0, np.uint64), np.zeros(0), 5)
    plan = Plan(t3, keys)
    assert torch.count_nonzero(plan.matvec(torch.ones(8, dtype=torch.float64, device=DEV))) == 0


def test_plan_state_errors_large_subspace():
    """Subspaces of 2^18 states or more are checked on host threads: the first bad state in row
    order still decides the error."""
    t = TermTables.from_masks(np.array([0], np.uint64), np.array([1], np.uint64), np.array([1.]), 20)
    big = np.arange(1 << 19, dtype=np.uint64)
    bad = big.copy()
    bad[400000] = bad[399999]                 # first: a repeated state
    bad[450000] = np.uint64(1) << np.uint64(25)   # later: bits beyond 20 qubits
    with pytest.raises(ValueError, match='sorted'):
        Plan(t, bad)
    bad = big.copy()
    bad[100] = np.uint64(1) << np.uint64(21)  # first: bits beyond (row 101 is then out of order too)
    bad[300000] = 5
    with pytest.raises(ValueError, match='beyond'):
        Plan(t, bad)
    plan = Plan(t, big)
    assert plan.n_states == big.shape[0]


def test_plan_argument_errors():
    t = TermTables.from_masks(np.array([0], np.uint64), np.array([1], np.uint64), np.array([1.]), 4)
    with pytest.raises(ValueError, match='sorted'):
        Plan(t, np.array([3, 1], np.uint64))
    with pytest.raises(ValueError, match='beyond'):
        Plan(t, np.array([1, 99], np.uint64))
    plan = Plan(t, np.array([1, 2], np.uint64))
    with pytest.raises(ValueError, match='shape'):
        plan.matvec(torch.zeros(3, dtype=torch.float64, device=DEV))


def test_sort_unique():
    _say('sort unique')
    rng = np.random.default_rng(9)
    keys = rng.integers(0, 1 << 40, size=100_000, dtype=np.uint64)
    keys[::7] = keys[0]
    got = sort_unique_u64(torch.from_numpy(keys.view(np.int64)).to(DEV)).cpu().numpy().view(np.uint64)
    assert np.array_equal(got, np.unique(keys))
    big = rng.integers(0, 1 << 63, size=1000, dtype=np.uint64) | np.uint64(1 << 63)
    got = sort_unique_u64(torch.from_numpy(big.view(np.int64)).to(DEV)).cpu().numpy().view(np.uint64)
    assert np.array_equal(got, np.unique(big))


@pytest.mark.parametrize('n', [2, 3])
def test_eigenpair_small_matches_eigh(n):
    rng = np.random.default_rng(20 + n)
    for k in range(40):
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        if k % 4 == 1:
            a = a.real.astype(complex)
        a = (a + a.conj().T) / 2
        if k % 4 == 2:
            a = np.diag(np.diag(a).real).astype(complex)
        theta, kappa = eigenpair_small(torch.from_numpy(a).to(DEV))
        theta, kappa = float(theta), kappa.cpu().numpy()
        w = np.linalg.eigvalsh(a)
        assert abs(theta - w[0]) < 1e-10 * max(1, abs(w).max())
        assert abs(np.linalg.norm(kappa) - 1) < 1e-10
        assert np.linalg.norm(a @ kappa - theta * kappa) < 1e-8
    # degenerate lowest eigenvalue and a multiple of the identity
    for a in (np.diag([1., 1., 3.])[:n, :n], 2. * np.eye(n)):
        theta, kappa = eigenpair_small(torch.from_numpy(a.astype(complex)).to(DEV))
        kappa = kappa.cpu().numpy()
        assert abs(float(theta) - np.linalg.eigvalsh(a)[0]) < 1e-12
        assert abs(np.linalg.norm(kappa) - 1) < 1e-10
        assert np.linalg.norm(a @ kappa - float(theta) * kappa) < 1e-10


def test_hermitian_and_linear_jw32_full():
    """Size-independent properties at the full 1M-determinant shape."""
    _say('hermiticity jw32 full')
    w = workloads.jw32()
    plan = Plan(w.tables, w.states)
    rng = np.random.default_rng(6)
    u = torch.from_numpy(rng.standard_normal(plan.n_total)).to(DEV)
    v = torch.from_numpy(rng.standard_normal(plan.n_total)).to(DEV)
    hu, hv = plan.matvec(u), plan.matvec(v)
    lhs, rhs = torch.dot(u, hv).item(), torch.dot(hu, v).item()
    assert abs(lhs - rhs) < 1e-10 * (abs(lhs) + 1)
    hw = plan.matvec(2.5 * u - 0.5 * v)
    assert torch.allclose(hw, 2.5 * hu - 0.5 * hv, rtol=0, atol=1e-11 * hu.abs().max().item())
    # diagonal vs oracle on a sample of rows, and the counts line up
    d = plan.diagonal().real.cpu().numpy()
    z0, c0 = _groups(w.tables)[0]
    sample = np.linspace(0, plan.n_states - 1, 2000).astype(np.int64)
    assert np.allclose(d[sample], sqdref.diagonal_u64(z0, c0, w.states[sample]).real, atol=1e-12)
    cnt = plan.count()
    assert cnt['hits'] > plan.n_states
    # row sample of the matvec against the oracle (only the sampled rows' sources are needed)
    e = torch.zeros(plan.n_total, dtype=torch.float64, device=DEV)
    e[12345] = 1.
    col = plan.matvec(e).cpu().numpy()      # column 12345 of H_S (H is symmetric)
    t = w.tables
    keys = w.states
    s = keys[12345]
    want = np.zeros(plan.n_total)
    for j in range(t.num_groups):
        tgt = s ^ t.xmask[j]
        pos = np.searchsorted(keys, tgt)
        if pos < keys.shape[0] and keys[pos] == tgt:
            zs, cs = t.zmask[t.grp_off[j]:t.grp_off[j + 1]], t.coef[t.grp_off[j]:t.grp_off[j + 1]]
            # H[pos, 12345] = C_j(s_pos) for the group mapping s_pos -> s
            val = np.sum(cs * (1. - 2. * (np.bitwise_count(keys[pos] & zs) & 1)))
            want[pos] += val.real
    assert _rel(col, want) < RTOL64


def test_lookup_table_beyond_2_32_entries():
    """J x N above 2^32 entries (a 26 GB table): every miss, up to the last entry, reads -1."""
    _say('lookup J*N > 2^32')
    w = workloads.jw32(n_alpha=600, n_beta=600)
    plan = Plan(w.tables, w.states)
    total = plan.info['n_groups'] * plan.n_total
    assert total > 2 ** 32
    idx = plan.lookup()
    hits = plan.count()['hits']
    assert int((idx >= 0).sum()) == hits
    assert int(idx[-1].min()) >= -1 and int(idx.max()) < plan.n_states
    del idx
    torch.cuda.empty_cache()


def _table_problem(seed, n=14, n_groups=60, n_states=3000):
    """Groups whose terms differ only on their X bits (sign tables), mixed with ordinary groups."""
    rng = np.random.default_rng(seed)
    xs, zs = [], []
    for g in range(n_groups):
        bits = rng.choice(n, size=int(rng.integers(1, 6)), replace=False)
        x = int(sum(1 << int(b) for b in bits))
        z0 = int(rng.integers(0, 1 << n))
        k = int(rng.integers(1, 9))
        for _ in range(k):
            if g % 3 == 2:                        # ordinary: z anywhere
                z = int(rng.integers(0, 1 << n))
            else:                                 # z_0 ^ (a subset of x)
                z = z0 ^ (x & int(rng.integers(0, 1 << n)))
            xs.append(x)
            zs.append(z)
    xs.append(0)                                  # a diagonal term
    zs.append(5)
    xm = np.array(xs, dtype=np.uint64)
    zm = np.array(zs, dtype=np.uint64)
    tables = TermTables.from_masks(xm, zm, rng.standard_normal(xm.shape[0]), n)
    states = workloads.random_subspace(n, n_states, rng)
    return tables, states, xm, zm


def test_sign_table_groups_match_oracle_and_follow_updates():
    """Groups evaluated by sign table (incl. complex phases) equal the oracle; set_alphas rebuilds the tables."""
    _say('sign tables')
    tables, states, xm, zm = _table_problem(31)
    sizes = np.diff(tables.grp_off)
    bits = np.array([bin(int(x)).count('1') for x in tables.xmask])
    assert np.any((sizes >= 3) & (bits <= 4) & (tables.xmask != 0))   # some groups take the table path
    assert np.any(tables.coef.imag != 0)                              # Y phases: complex tables
    plan = Plan(tables, states)
    rng = np.random.default_rng(32)
    v = rng.standard_normal(plan.n_total) + 1j * rng.standard_normal(plan.n_total)
    want = sqdref.matvec_u64(tables.xmask, _groups(tables), states, v)
    y = plan.matvec(torch.from_numpy(v).to(DEV)).cpu().numpy()
    assert _rel(y, want) < RTOL64
    y32 = plan.matvec(torch.from_numpy(v.astype(np.complex64)).to(DEV)).cpu().numpy()
    assert _rel(y32, want) < RTOL32
    # new coefficients: the tables follow (oracle rebuilt from the same masks)
    alpha = rng.standard_normal(plan.info['n_terms'])
    plan.set_alphas(torch.from_numpy(alpha).to(DEV))
    phases = plan.term_phases().cpu().numpy()
    coef = alpha * phases
    t2 = TermTables(tables.xmask, tables.grp_off, tables.zmask, coef, tables.num_qubits)
    want2 = sqdref.matvec_u64(t2.xmask, _groups(t2), states, v)
    y2 = plan.matvec(torch.from_numpy(v).to(DEV)).cpu().numpy()
    assert _rel(y2, want2) < RTOL64


@pytest.mark.parametrize('split', [12, 18, 22, 26])
def test_hashed_and_binary_search_targets(split):
    """Splits up to 18 bits find targets by hash table, wider ones by binary search: both exact."""
    _say(f'target search split={split}')
    rng = np.random.default_rng(41)
    codes = workloads.random_pauli_codes(30, 150, (0.6, 0.14, 0.13, 0.13), rng)
    xm, zm = workloads.codes_to_masks(codes)
    tables = TermTables.from_masks(xm, zm, rng.standard_normal(xm.shape[0]), 30)
    states = workloads.random_subspace(30, 4000, rng)
    # states sharing their high bits, so that targets exist for many groups
    states = np.unique(np.concatenate([states, states ^ np.uint64(1 << 3), states ^ np.uint64(1 << 20)]))
    plan = Plan(tables, states, split_bits=split)
    assert plan.info['split_bits'] == split
    got = plan.lookup().cpu().numpy()
    assert np.array_equal(got, _oracle_lookup(tables, states, plan.n_total))
    v = rng.standard_normal(plan.n_total) + 1j * rng.standard_normal(plan.n_total)
    y = plan.matvec(torch.from_numpy(v).to(DEV)).cpu().numpy()
    want = sqdref.matvec_u64(tables.xmask, _groups(tables), states, v)
    assert _rel(y, want) < RTOL64
