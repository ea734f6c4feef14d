"""ctypes binding of the oracle's C port (``oracle/rqx_oracle.c``); test infrastructure only."""
import ctypes
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'liboracle_c.so')
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f'{LIB} missing: run `python -m oracle.build`')
        _lib = ctypes.CDLL(LIB)
        vp, i64, u64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64
        _lib.oracle_xsource.argtypes = [vp, i64, u64, vp]
        _lib.oracle_xsource.restype = ctypes.c_int
        _lib.oracle_apply_group.argtypes = [vp, i64, vp, vp, vp, i64, vp, vp]
        _lib.oracle_apply_group.restype = None
        _lib.oracle_num_threads.restype = ctypes.c_int
        _lib.oracle_xsource_bsearch.argtypes = [vp, i64, u64, vp]
        _lib.oracle_xsource_bsearch.restype = None
        _lib.oracle_apply_group_c64.argtypes = [vp, i64, vp, vp, vp, i64, vp, vp]
        _lib.oracle_apply_group_c64.restype = None
    return _lib


def xsource(keys, xmask):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.empty(keys.shape[0], dtype=np.int32)
    if lib().oracle_xsource(keys.ctypes.data, keys.shape[0], int(xmask), out.ctypes.data) != 0:
        raise MemoryError('oracle_xsource')
    return out


def xsource_bsearch(keys, xmask):
    """The same indices as ``xsource`` by OpenMP binary search (untimed baseline setup)."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.empty(keys.shape[0], dtype=np.int32)
    lib().oracle_xsource_bsearch(keys.ctypes.data, keys.shape[0], int(xmask), out.ctypes.data)
    return out


def apply_group(keys, src, zmask, coef, v, y):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    z = np.ascontiguousarray(zmask, dtype=np.uint64)
    c = np.ascontiguousarray(coef, dtype=np.complex128)
    fn = lib().oracle_apply_group_c64 if v.dtype == np.complex64 else lib().oracle_apply_group
    if v.dtype not in (np.complex64, np.complex128) or y.dtype != v.dtype:
        raise TypeError('v and y must both be complex64 or both complex128')
    fn(keys.ctypes.data, keys.shape[0], src.ctypes.data, z.ctypes.data, c.ctypes.data, z.shape[0],
       v.ctypes.data, y.ctypes.data)


def matvec(tables, keys, v, bsearch=False):
    """Full strategy-(1,0) matvec (small inputs): cache all sources, then apply every group
    (complex128, or complex64 for a complex64 ``v``)."""
    v = np.ascontiguousarray(v, dtype=np.complex64 if np.asarray(v).dtype == np.complex64 else np.complex128)
    y = np.zeros_like(v)
    for j in range(tables.num_groups):
        src = (xsource_bsearch if bsearch else xsource)(keys, tables.xmask[j])
        a, b = tables.grp_off[j], tables.grp_off[j + 1]
        apply_group(keys, src, tables.zmask[a:b], tables.coef[a:b], v, y)
    return y


def timed_sample(tables, keys, budget_s=15.0, max_groups=None, seed=0, dtype=np.complex128):
    """Time the per-matvec work of strategy (1, 0) on a sample of groups.

    Source indices are computed first (the reference caches them once,
    ``sqd.py:347-351``; here by binary search) and not timed; the timed part is
    the per-matvec diagonal recomposition + gather of each sampled group, in
    ``dtype`` (complex128, or complex64 with single-precision sums).  Groups are
    sampled uniformly until ``budget_s`` of wall time (sources included) has
    passed.  Returns ``(seconds_per_matvec_extrapolated, groups_timed, threads,
    sampled_fraction)`` where the fraction is the sampled share of the
    ``sum_j (K_j + 1)`` work units by which the time is extrapolated.
    """
    rng = np.random.default_rng(seed)
    J = tables.num_groups
    order = rng.permutation(J)
    n = keys.shape[0]
    v = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dtype)
    y = np.zeros_like(v)
    timed = 0.0
    terms = 0
    done = 0
    wall0 = time.perf_counter()
    for j in order:
        src = xsource_bsearch(keys, tables.xmask[j])
        a, b = tables.grp_off[j], tables.grp_off[j + 1]
        t0 = time.perf_counter()
        apply_group(keys, src, tables.zmask[a:b], tables.coef[a:b], v, y)
        timed += time.perf_counter() - t0
        terms += b - a
        done += 1
        if time.perf_counter() - wall0 >= budget_s or (max_groups and done >= max_groups):
            break
    # extrapolate in proportion to the term count (each group costs ~ (K_j + 1) N)
    frac = (terms + done) / (tables.num_terms + J)
    return timed / frac, done, lib().oracle_num_threads(), frac
