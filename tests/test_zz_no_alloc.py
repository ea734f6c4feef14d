"""rqx_locg_step_static (work space in the library's device array) against rqx_locg_step
(work space from the caller), through the C ABI: identical results, call after call,
also when the second plan is smaller than the first."""
import ctypes
import sys

import numpy as np
import pytest
import torch

from rqutils_amd import _capi, workloads
from rqutils_amd.locg_batch import BatchLOCG
from rqutils_amd.paulis.symplectic import TermTables
from rqutils_amd.plan import BatchPlan, Plan

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'
_CODES = {torch.complex128: _capi.RQX_C128, torch.float64: _capi.RQX_F64}


def _say(msg):
    print(f'[case] {msg}', flush=True)
    sys.stdout.flush()


def _problem(seed, n, t, ns):
    rng = np.random.default_rng(seed)
    codes = workloads.random_pauli_codes(n, t, (0.55, 0.15, 0.15, 0.15), rng)
    xm, zm = workloads.codes_to_masks(codes)
    return TermTables.from_masks(xm, zm, rng.standard_normal(t), n), workloads.random_subspace(n, ns, rng)


def _static_step(solver, low):
    """One step of `solver` through rqx_locg_step_static (same vectors and outputs as BatchLOCG.step)."""
    vecs = (solver.x, solver.ax, solver.y, solver.ay, solver.p, solver.ap, solver.p_lo, solver.ap_lo)
    ptrs = (ctypes.c_void_p * 8)(*[v.data_ptr() for v in vecs])
    stream = _capi.current_stream(solver.plan.device)
    rc = _capi.lib().rqx_locg_step_static(solver.plan._handle, _CODES[solver.hi], int(low), ptrs,
                                          solver.theta.data_ptr(), solver.rel.data_ptr(), solver.active.data_ptr(),
                                          solver.scale.data_ptr(), float(solver.tol), stream)
    _capi.check(rc)


def _compare(plan, x0, low, steps):
    mixed = bool(low)
    a = BatchLOCG(plan, x0, tol=1e-10, mixed=mixed, lo_storage=False)
    b = BatchLOCG(plan, x0, tol=1e-10, mixed=mixed, lo_storage=False)
    for it in range(steps):
        a.step()
        _static_step(b, low)
        torch.cuda.synchronize()
        assert torch.equal(a.theta, b.theta), it
        assert torch.equal(a.rel, b.rel) and torch.equal(a.active, b.active), it
    for u, v in ((a.x, b.x), (a.ax, b.ax), (a.y, b.y), (a.ay, b.ay)):
        assert torch.equal(u, v)
    return a


def test_static_step_equals_caller_work_space():
    _say('locg step with the static work space, two plans in a row')
    big = BatchPlan([_problem(1, 12, 120, 2000), _problem(2, 10, 60, 500), _problem(3, 11, 80, 900)])
    rng = np.random.default_rng(4)
    x0 = torch.from_numpy(rng.standard_normal(big.n_total) + 1j * rng.standard_normal(big.n_total)).to(DEV)
    a = _compare(big, x0, low=1, steps=30)
    assert bool((a.theta < 0).all())
    # second call sequence: a smaller single plan, real vectors; the array keeps the
    # first plan's contents, which the step must not read
    w = workloads.jw32(n_alpha=30, n_beta=20)
    small = Plan(w.tables, w.states)
    x1 = torch.from_numpy(np.random.default_rng(5).standard_normal(small.n_total)).to(DEV)
    _compare(small, x1, low=0, steps=25)


def test_static_step_rejects_plans_beyond_its_size():
    _say('static step size limit')
    assert _capi.RQX_LOCG_STATIC_PROBLEMS == 4096
    t, s = _problem(6, 6, 10, 3)
    plan = BatchPlan([(t, s)] * (_capi.RQX_LOCG_STATIC_PROBLEMS + 1))
    x0 = torch.ones(plan.n_total, dtype=torch.complex128, device=DEV)
    solver = BatchLOCG(plan, x0, tol=1e-10, mixed=False, lo_storage=False)
    with pytest.raises(_capi.RqxError):
        _static_step(solver, 0)
