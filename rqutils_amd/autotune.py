"""Autotuner for the tile kernels: segment split and launch shape, picked by measurement.

The plan's host cost model picks a split L (rows are segmented by ``s >> L``);
this module times the matvec with HIP events for splits around that choice
and for both launch shapes (one workgroup per tile, or persistent workgroups
on a tile queue at the occupancy limit or twice fewer), and returns the
fastest plan.  Timings are kept so callers can report them.  ``autotune_batch``
does the same for a batched plan over the persistent grid size (the batch's
splits are the cost model's, one per problem).
"""
import torch

from rqutils_amd import _capi
from rqutils_amd.plan import BatchPlan, Plan


def time_matvec(plan, dtype=torch.complex64, reps=3, warmup=1):
    """Average matvec time (ms) on the plan's device, measured with HIP events."""
    gen = torch.Generator(device='cpu').manual_seed(1234)
    x = torch.randn(plan.n_total, generator=gen, dtype=torch.float64)
    x = x.to(plan.device, dtype if not dtype.is_complex else torch.complex128).to(dtype)
    y = torch.empty_like(x)
    for _ in range(warmup):
        plan.matvec(x, out=y)
    stream = torch.cuda.current_stream(plan.device)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        plan.matvec(x, out=y)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def autotune(tables, states, n_total=None, dtype=torch.complex64, deltas=(-3, -2, -1, 0, 1, 2), reps=3,
             device=None, launch_shapes=((False, 0), (True, 0))):
    """Best ``Plan`` for (tables, states) by measured matvec time.

    Returns ``(plan, results)`` with ``results`` a list of
    ``{'split_bits', 'persistent', 'blocks_per_cu', 'ms'}`` sorted fastest first.
    """
    base = Plan(tables, states, n_total=n_total, device=device)
    l0 = base.info['split_bits']
    lmax = min(32, tables.num_qubits)
    results = []
    best = None
    for d in deltas:
        L = l0 + d
        if L < 0 or L > lmax:
            continue
        if d == 0:
            plan = base
        else:
            try:
                plan = Plan(tables, states, n_total=n_total, device=device, split_bits=L)
            except (ValueError, _capi.RqxError):
                continue            # segments beyond the LDS capacity
        for persistent, bpc in launch_shapes:
            plan.set_launch(persistent, bpc)
            ms = time_matvec(plan, dtype, reps)
            results.append({'split_bits': L, 'persistent': persistent, 'blocks_per_cu': bpc, 'ms': ms})
            if best is None or ms < best[0]:
                best = (ms, plan, persistent, bpc)
    results.sort(key=lambda r: r['ms'])
    _, plan, persistent, bpc = best
    plan.set_launch(persistent, bpc)
    return plan, results


def autotune_batch(problems, dtype=torch.complex64, reps=3, device=None,
                   launch_shapes=((True, 0), (True, 2), (True, 4), (True, 6), (True, 8), (False, 0))):
    """Best launch shape of a ``BatchPlan`` over ``problems`` by measured matvec time.

    ``launch_shapes`` are ``(persistent, blocks_per_cu)`` pairs (0 = the
    occupancy limit).  Returns ``(plan, results)`` with ``results`` sorted
    fastest first, each ``{'persistent', 'blocks_per_cu', 'ms'}``.
    """
    plan = problems if isinstance(problems, BatchPlan) else BatchPlan(problems, device=device)
    results = []
    for persistent, bpc in launch_shapes:
        plan.set_launch(persistent, bpc)
        results.append({'persistent': persistent, 'blocks_per_cu': bpc, 'ms': time_matvec(plan, dtype, reps)})
    results.sort(key=lambda r: r['ms'])
    plan.set_launch(results[0]['persistent'], results[0]['blocks_per_cu'])
    return plan, results
