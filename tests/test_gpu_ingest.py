"""Device ingest: packed rows -> uint64 keys on the GPU, and plans built from device
keys (``rqx_plan_create_device``) identical to the host build of the same states."""
import ctypes
import sys

import numpy as np
import pytest
import torch

from rqutils_amd import _capi, workloads
from rqutils_amd.paulis.symplectic import masks_to_packed
from rqutils_amd.plan import Plan, pack_to_u64, sort_unique_u64

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


def _say(msg):
    print(f'[case] {msg}', flush=True)
    sys.stdout.flush()


def _dev_keys(states):
    return torch.from_numpy(np.ascontiguousarray(states, dtype=np.uint64).view(np.int64)).to(DEV)


@pytest.mark.parametrize('n', [1, 5, 7, 8, 17, 40, 63, 64])
def test_pack_to_u64_matches_host_unpacking(n):
    """Real rows in input order, filler rows (pad bit set, 0xFF) dropped."""
    _say(f'pack_to_u64 n={n}')
    rng = np.random.default_rng(900 + n)
    keys = rng.integers(0, 2 ** min(n, 63), size=3000, dtype=np.uint64)
    if n == 64:
        keys |= rng.integers(0, 2, size=3000, dtype=np.uint64) << np.uint64(63)
    packed = masks_to_packed(keys, n, padded=True)
    fill = rng.random(3000) < 0.2
    rows = packed.copy()
    rows[fill] = 255
    got = pack_to_u64(torch.from_numpy(rows).to(DEV), n).cpu().numpy().view(np.uint64)
    assert np.array_equal(got, keys[~fill])
    empty = pack_to_u64(torch.zeros((0, rows.shape[1]), dtype=torch.uint8, device=DEV), n)
    assert empty.numel() == 0


def test_pack_to_u64_rejects_wrong_row_width():
    rows = torch.zeros((4, 3), dtype=torch.uint8, device=DEV)
    with pytest.raises(ValueError, match='row_bytes'):
        pack_to_u64(rows, 40)


def _cases():
    w32 = workloads.jw32()
    yield 'jw-32so-1000x1000', w32.tables, w32.states, -1
    yield 'jw-32so-1000x1000 split 9', w32.tables, w32.states, 9
    b8 = workloads.batch40_problem(8)
    yield 'batch40 p8', b8.tables, b8.states, -1
    r36 = workloads.random36_10kt(num_states=40_000, num_terms=300)
    yield 'random36 40k', r36.tables, r36.states, -1
    yield 'random36 40k split 14', r36.tables, r36.states, 14
    yield 'one state', w32.tables, w32.states[:1], -1
    yield 'two states', w32.tables, w32.states[[0, 5000]], -1
    yield 'full 12 qubits', workloads.jw32().tables, None, -1


@pytest.mark.parametrize('case', list(range(8)))
def test_device_build_is_identical_to_host_build(case):
    name, tables, states, split = list(_cases())[case]
    if states is None:   # the whole 2^12 basis of a 12-qubit model
        from rqutils_amd.paulis.symplectic import PauliSumXZ
        rng = np.random.default_rng(77)
        paulis = [''.join(rng.choice(list('IXYZ'), 12)) for _ in range(200)]
        tables = PauliSumXZ.from_paulisum((paulis, rng.standard_normal(200)), force_real=True).tables
        states = np.arange(1 << 12, dtype=np.uint64)
    _say(f'device build == host build: {name}')
    host = Plan(tables, states, split_bits=split)
    dev = Plan(tables, _dev_keys(states), split_bits=split)
    try:
        assert dev.info == host.info
        assert dev.problems == host.problems
        assert dev.digest() == host.digest()
        rng = np.random.default_rng(31 + case)
        n = host.n_total
        x = torch.from_numpy(rng.standard_normal(n) + 1j * rng.standard_normal(n)).to(DEV)
        assert torch.equal(dev.matvec(x), host.matvec(x))
        if host.n_total * host.info['n_groups'] <= 50_000_000:   # (J, N) tables of a small case
            assert torch.equal(dev.lookup(values=True)[1], host.lookup(values=True)[1])
        for p in (host, dev):
            p.build_cache(0)
        assert dev.cache_info(64) == host.cache_info(64) and dev.cache_info(32) == host.cache_info(32)
        assert torch.equal(dev.matvec(x), host.matvec(x))
        assert np.array_equal(dev.states, host.states)
    finally:
        host.close()
        dev.close()


def test_device_build_rejects_bad_states_like_the_host():
    w = workloads.jw32()
    t = w.tables
    cases = {
        'unsorted': w.states[[3, 1, 2]],
        'duplicate': w.states[[1, 2, 2, 3]],
        'beyond n_qubits': np.array([1, 2, 1 << 40], dtype=np.uint64),
    }
    for name, states in cases.items():
        _say(f'device build rejects {name}')
        with pytest.raises(ValueError) as eh:
            Plan(t, states)
        with pytest.raises(ValueError) as ed:
            Plan(t, _dev_keys(states))
        assert str(ed.value) == str(eh.value), name


def test_device_create_rejects_host_memory():
    """A host array passed as device states is refused before any kernel reads it."""
    w = workloads.jw32()
    t = w.tables
    states = np.ascontiguousarray(w.states[:100])
    xmask = np.ascontiguousarray(t.xmask, dtype=np.uint64)
    grp_off = np.ascontiguousarray(t.grp_off, dtype=np.int64)
    zmask = np.ascontiguousarray(t.zmask, dtype=np.uint64)
    coef = np.ascontiguousarray(t.coef, dtype=np.complex128).view(np.float64)
    handle = ctypes.c_void_p()
    rc = _capi.lib().rqx_plan_create_device(ctypes.byref(handle), t.num_qubits, _capi.ptr(xmask),
                                            _capi.ptr(grp_off), t.num_groups, _capi.ptr(zmask), _capi.ptr(coef),
                                            t.num_terms, _capi.ptr(states), 100, 100, -1,
                                            _capi.current_stream(torch.device(DEV)))
    assert rc == _capi.RQX_ERR_ARG and not handle.value
    assert 'device memory' in _capi.lib().rqx_last_error().decode()


def test_make_plan_uniquifies_on_the_device():
    """make_plan: keys with duplicates, unsorted, from numpy or a device tensor."""
    from rqutils_amd.sqd import make_plan
    _say('make_plan device ingest')
    w = workloads.jw32()
    rng = np.random.default_rng(5)
    keys = rng.choice(w.states, size=30_000, replace=True)
    want = np.unique(keys)

    class _H:   # make_plan only reads .tables
        tables = w.tables

    for src in (keys, _dev_keys(keys)):
        plan = make_plan(_H, src, states_size=40_000)
        assert plan.n_states == want.shape[0] and plan.n_total == 40_000
        assert np.array_equal(plan.states, want)
        ref = Plan(w.tables, want, n_total=40_000)
        assert plan.digest() == ref.digest()
        plan.close()
        ref.close()
    assert sort_unique_u64(_dev_keys(keys)).shape[0] == want.shape[0]
