"""Benchmark: SQD subspace matvec throughput on one MI355X (+ CPU baseline).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload jw1m|jw25m|rand8k]
                    [--dtype c64|f32|f64|c128] [--no-cpu] [--no-locg] [--no-extras]

A step is one matvec ``y = H_S x`` over the whole subspace with the plan and
``x`` resident in HBM.  The headline path is the one ``run_sqd`` takes by
default (``cache_strategy`` (1, *)): the hits, built once by the matrix-free
tile kernel, streamed from the device hit cache.  ``value`` is that matvec's
byte rate (cache + diagonal + x once + y, divided by the kernel time measured
with HIP events on the launch stream).  The matrix-free kernel is timed over
the same steps and reported under ``matrix_free``.  One JSON line is printed
on stdout.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'matvec GB/s (% HBM peak); time to 1e-8 rel. ground-energy err vs FP64 plain'
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def load_workload(name):
    from rqutils_amd import workloads
    if name == 'jw1m':
        return workloads.jw32()
    if name == 'jw25m':
        return workloads.jw40()
    if name == 'rand8k':
        return workloads.random20_300t()
    raise ValueError(f'unknown workload {name}')


def traffic_from_profiles(workload, dtype):
    """HBM bytes per matvec measured with rocprofv3 --pmc (committed summary), if present."""
    path = os.path.join(ROOT, 'profiles', 'matvec_traffic.json')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    entry = data.get(f'{workload}:{dtype}')
    return None if entry is None else entry.get('hbm_bytes_per_launch')


def cpu_baseline(w, alg_bytes, budget_s):
    """Oracle C port of the reference's strategy (1, 0), timed on a sample of groups.

    ``value`` puts the CPU matvec on the headline's scale: the headline's bytes
    per matvec divided by the CPU seconds per matvec.
    """
    from oracle import build as obuild, cport
    if not os.path.exists(cport.LIB):
        obuild.build()
    t_mv, groups, threads = cport.timed_sample(w.tables, w.states, budget_s=budget_s)
    return {
        'value': alg_bytes / t_mv / 1e9,
        'unit': 'GB/s',
        'cores': threads,
        'kind': 'port',
        'sample': (f'strategy (1,0) per-matvec work (diagonal recomposition + gather) on {groups} of '
                   f'{w.tables.num_groups} X-groups, extrapolated by term count to one matvec '
                   f'({t_mv:.3f} s/matvec), complex128'),
        'seconds_per_matvec': t_mv,
    }


def locg_time_to_solution(plan, complex_, maxiter, mixed_tol, energy_tol, target=1e-8):
    """Time to ``target`` relative ground-energy error: plain FP64 LOCG vs mixed precision.

    Both start from the reference's one-hot vector at the argmin of the x = 0
    diagonal (``sqd.py:406-426``).  The plain run follows the reference
    sequence (4 matvecs per iteration) to its own stopping rule; its final
    energy is E*, and its time-to-target is the first iteration within
    ``target`` of E*.  The mixed run stops by its own rule (no knowledge of
    E*); its time is the whole call, its error is measured against E*.
    """
    import torch
    from rqutils_amd.ground_locg import ground_locg
    from rqutils_amd.locg_mixed import ground_locg_mixed

    hi = torch.complex128 if complex_ else torch.float64
    n = plan.n_total
    diag = plan.diagonal().real[:plan.n_states]
    v0 = int(torch.argmin(diag))
    x0 = torch.zeros(n, dtype=hi, device=plan.device)
    x0[v0] = 1
    trace = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    theta, _, niter = ground_locg(plan.matvec, x0.clone(), maxiter=maxiter,
                                  callback=lambda k, th: trace.append((time.perf_counter() - t0, th)))
    torch.cuda.synchronize()
    t_plain_total = time.perf_counter() - t0
    e_star = float(theta)
    t_plain = next((t for t, th in trace if abs(th - e_star) <= target * abs(e_star)), t_plain_total)
    log(f'plain LOCG: E*={e_star:.12f} in {niter} iterations, {t_plain_total:.2f} s '
        f'(within {target:g} after {t_plain:.2f} s)')
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = ground_locg_mixed(plan.matvec, plan.matvec, x0.clone(), maxiter=maxiter, tol=mixed_tol,
                            energy_tol=energy_tol, record=True)
    torch.cuda.synchronize()
    t_mixed = time.perf_counter() - t0
    err = abs(float(res.theta) - e_star) / abs(e_star)
    log(f'mixed LOCG: E={float(res.theta):.12f} rel err {err:.2e} in {t_mixed:.2f} s '
        f'({res.niter} iterations, {res.matvecs_lo} low + {res.matvecs_hi} FP64 matvecs)')
    lo_name = 'c64' if complex_ else 'f32'
    hi_name = 'c128' if complex_ else 'f64'
    g = graph_mixed(plan, x0, maxiter, energy_tol, e_star)
    g['speedup'] = round(t_plain / g['time_s'], 3)
    return {
        'e_star': e_star,
        'target_rel_err': target,
        'plain': {'dtype': hi_name, 'iterations': niter, 'matvecs': 4 * niter + 3,
                  'time_to_target_s': round(t_plain, 4), 'total_s': round(t_plain_total, 4)},
        'mixed': {'dtype': f'{lo_name} matvec, {hi_name} vectors/RR', 'iterations': res.niter,
                  'matvecs_low': res.matvecs_lo, 'matvecs_fp64': res.matvecs_hi,
                  'time_s': round(t_mixed, 4), 'rel_err_vs_plain': err, 'stop_rel_residual': mixed_tol,
                  'stop_energy_tail': energy_tol},
        'speedup': round(t_plain / t_mixed, 3),
        'mixed_graph': g,
    }


def graph_mixed(plan, x0, maxiter, energy_tol, e_star, check_every=5):
    """Mixed LOCG with the fused per-problem kernels, one HIP graph per step (single problem)."""
    import torch
    from rqutils_amd.locg_batch import batch_ground_locg
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    theta, _, it, info = batch_ground_locg(plan, x0.clone(), tol=1e-12, maxiter=maxiter, check_every=check_every,
                                           energy_tol=energy_tol, use_graph=True)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    err = abs(float(theta[0]) - e_star) / abs(e_star)
    log(f'graph mixed LOCG: rel err {err:.2e} in {t:.3f} s ({it} iterations, switch at {info["switch_iteration"]})')
    return {'time_s': round(t, 4), 'iterations': it, 'switch_iteration': info['switch_iteration'],
            'rel_err_vs_plain': err, 'stop_energy_tail': energy_tol, 'check_every': check_every,
            'dtype': 'c64 matvec, c128 vectors/RR', 'hip_graph': True}


def _event_ms(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(torch.cuda.current_stream())
    for _ in range(reps):
        fn()
    e1.record(torch.cuda.current_stream())
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def cache_leg(plan, dtype, reps=10):
    """Build the hit cache for ``dtype``'s precision and time the cached matvec against its bytes."""
    import numpy as np
    import torch
    precision = 64 if dtype in (torch.float64, torch.complex128) else 32
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info = plan.build_cache(precision)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    rng = np.random.default_rng(5005)
    x = torch.from_numpy(rng.standard_normal(plan.n_total) + 1j * rng.standard_normal(plan.n_total)).to(plan.device)
    x = (x if dtype.is_complex else x.real).to(dtype).contiguous()
    y = torch.empty_like(x)
    ms = _event_ms(lambda: plan.matvec(x, out=y), reps)
    plan.use_cache(False)
    y_free = plan.matvec(x)
    plan.use_cache(True)
    identical = bool(torch.equal(plan.matvec(x), y_free))
    nbytes = plan.cached_bytes(dtype)
    gbs = nbytes / ms / 1e6
    name = {torch.complex64: 'c64', torch.complex128: 'c128', torch.float32: 'f32', torch.float64: 'f64'}[dtype]
    log(f'cached {name} matvec: {ms:.3f} ms, cache {info["bytes"] / 1e9:.2f} GB built in {t_build:.2f} s, '
        f'{gbs:.0f} GB/s of cache+vector bytes, identical to matrix-free: {identical}')
    return {'dtype': name, 'ms_per_matvec': round(ms, 4), 'cache_bytes': info['bytes'], 'stored_hits': info['nnz'],
            'entries_incl_padding': info['entries'], 'build_s': round(t_build, 3), 'bytes_per_matvec': nbytes,
            'achieved_GBs': round(gbs, 1), 'frac_of_8TBs': round(gbs / HBM_PEAK_GBS, 4),
            'identical_to_matrix_free': identical}


def locg_cached(plan, complex_, maxiter, mixed_tol, energy_tol, e_star, target=1e-8):
    """Plain and mixed LOCG with their matvecs read from hit caches (cache builds included).

    The cached matvec returns the matrix-free kernel's values, so the plain run
    repeats the matrix-free plain run iteration for iteration (same E*).
    """
    import torch
    from rqutils_amd.ground_locg import ground_locg
    from rqutils_amd.locg_mixed import ground_locg_mixed
    hi = torch.complex128 if complex_ else torch.float64
    diag = plan.diagonal().real[:plan.n_states]
    x0 = torch.zeros(plan.n_total, dtype=hi, device=plan.device)
    x0[int(torch.argmin(diag))] = 1
    plan.drop_cache()
    torch.cuda.synchronize()
    trace = []
    t0 = time.perf_counter()
    plan.build_cache(64)
    torch.cuda.synchronize()
    t_build64 = time.perf_counter() - t0
    theta, _, niter = ground_locg(plan.matvec, x0.clone(), maxiter=maxiter,
                                  callback=lambda k, th: trace.append((time.perf_counter() - t0, th)))
    torch.cuda.synchronize()
    t_plain_total = time.perf_counter() - t0
    t_plain = next((t for t, th in trace if abs(th - e_star) <= target * abs(e_star)), t_plain_total)
    same = float(theta) == e_star
    log(f'cached plain LOCG: {niter} iterations, within {target:g} after {t_plain:.3f} s (total {t_plain_total:.3f} s, '
        f'same E* as matrix-free: {same})')
    plan.drop_cache()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.build_cache(0)                      # FP32 and FP64 values in one pass
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    res = ground_locg_mixed(plan.matvec, plan.matvec, x0.clone(), maxiter=maxiter, tol=mixed_tol,
                            energy_tol=energy_tol)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    err = abs(float(res.theta) - e_star) / abs(e_star)
    log(f'cached mixed LOCG: rel err {err:.2e} in {t:.3f} s incl. {t_build:.3f} s cache build ({res.niter} iterations)')
    g = graph_mixed(plan, x0, maxiter, energy_tol, e_star)
    g['time_s'] = round(g['time_s'] + t_build, 4)
    g['cache_build_s'] = round(t_build, 4)
    g['speedup'] = round(t_plain / g['time_s'], 3)
    return {'plain': {'time_to_target_s': round(t_plain, 4), 'total_s': round(t_plain_total, 4), 'iterations': niter,
                      'cache_build_s': round(t_build64, 4), 'same_energy_as_matrix_free': same},
            'mixed': {'time_s': round(t, 4), 'cache_build_s': round(t_build, 4), 'iterations': res.niter,
                      'matvecs_low': res.matvecs_lo, 'matvecs_fp64': res.matvecs_hi, 'rel_err_vs_plain': err},
            'speedup': round(t_plain / t, 3), 'mixed_graph': g}


def locg_large(plan, maxiter=1000, energy_tol=1e-9, target=1e-8):
    """Ground energy of the 25M-determinant model to 1e-8: plain FP64 LOCG vs mixed, both on hit caches.

    E* is the plain run's own converged energy (the cached matvec equals the
    matrix-free one bit for bit, so this is the matrix-free plain run too).
    """
    import torch
    from rqutils_amd.ground_locg import ground_locg
    from rqutils_amd.locg_batch import batch_ground_locg
    diag = plan.diagonal().real[:plan.n_states]
    x0 = torch.zeros(plan.n_total, dtype=torch.complex128, device=plan.device)
    x0[int(torch.argmin(diag))] = 1
    del diag
    plan.drop_cache()
    torch.cuda.synchronize()
    trace = []
    t0 = time.perf_counter()
    plan.build_cache(64)
    theta, _, niter = ground_locg(plan.matvec, x0.clone(), maxiter=maxiter,
                                  callback=lambda k, th: trace.append((time.perf_counter() - t0, th)))
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t0
    e_star = float(theta)
    t_plain = next((t for t, th in trace if abs(th - e_star) <= target * abs(e_star)), t_total)
    log(f'25M plain LOCG (cached): E*={e_star:.10f}, {niter} iterations, within {target:g} after {t_plain:.2f} s')
    plan.drop_cache()
    torch.cuda.empty_cache()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.build_cache(0)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    th, _, it, info = batch_ground_locg(plan, x0, tol=1e-12, maxiter=maxiter, check_every=5, energy_tol=energy_tol)
    torch.cuda.synchronize()
    t_mixed = time.perf_counter() - t0
    err = abs(float(th[0]) - e_star) / abs(e_star)
    log(f'25M mixed LOCG (cached, graph): rel err {err:.2e} in {t_mixed:.2f} s incl. {t_build:.2f} s build '
        f'({it} iterations)')
    plan.drop_cache()
    return {'e_star': e_star, 'plain_c128_cached': {'time_to_target_s': round(t_plain, 3), 'total_s': round(t_total, 3),
                                                     'iterations': niter},
            'mixed_graph_cached': {'time_s': round(t_mixed, 3), 'cache_build_s': round(t_build, 3), 'iterations': it,
                                   'switch_iteration': info['switch_iteration'], 'rel_err_vs_plain': err},
            'speedup': round(t_plain / t_mixed, 3)}


def extra_large_matvec(reps=3):
    """25M-determinant JW model (40 spin-orbitals, 5000 x 5000): complex64 matvec and LOCG to 1e-8."""
    import numpy as np
    import torch
    from rqutils_amd import workloads
    from rqutils_amd.plan import Plan
    w = workloads.jw40()
    t0 = time.perf_counter()
    plan = Plan(w.tables, w.states)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    hits = plan.count()['hits']
    alg = plan.algorithmic_bytes(torch.complex64, hits=hits)
    rng = np.random.default_rng(4002)
    x = torch.from_numpy(rng.standard_normal(plan.n_total) + 1j * rng.standard_normal(plan.n_total)).to(
        plan.device, torch.complex64)
    y = torch.empty_like(x)
    ms = _event_ms(lambda: plan.matvec(x, out=y), reps)
    log(f'25M matvec: {ms:.1f} ms, {alg / ms / 1e6:.0f} GB/s algorithmic')
    del x, y
    try:
        cached = cache_leg(plan, torch.complex64, reps=5)
    except Exception as exc:  # the cache may not fit next to other allocations
        cached = {'error': str(exc)}
    try:
        locg = locg_large(plan)
    except Exception as exc:
        locg = {'error': str(exc)}
    out = {'workload': w.name, 'n_states': plan.n_states, 'n_groups': plan.info['n_groups'],
           'n_terms': plan.info['n_terms'], 'hits_per_matvec': hits, 'split_bits': plan.info['split_bits'],
           'dtype': 'c64', 'ms_per_matvec': round(ms, 3), 'algorithmic_bytes_per_matvec': alg,
           'achieved_GBs': round(alg / ms / 1e6, 1), 'frac_of_8TBs': round(alg / ms / 1e6 / HBM_PEAK_GBS, 4),
           'setup_s': round(setup, 2), 'cached': cached, 'locg': locg}
    del plan
    torch.cuda.empty_cache()
    return out


def extra_batch(maxiter=1000):
    """64 independent 40-qubit random Pauli sums in one batched plan: matvec + batched LOCG."""
    import numpy as np
    import torch
    from rqutils_amd import workloads
    from rqutils_amd.autotune import autotune_batch
    from rqutils_amd.plan import BatchPlan
    from rqutils_amd.locg_batch import batch_ground_locg
    probs = [workloads.batch40_problem(i) for i in range(64)]
    t0 = time.perf_counter()
    bp = BatchPlan([(p.tables, p.states) for p in probs])
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    rng = np.random.default_rng(2002)
    x = torch.from_numpy(rng.standard_normal(bp.n_total) + 1j * rng.standard_normal(bp.n_total)).to(bp.device)
    x64 = x.to(torch.complex64)
    y = torch.empty_like(x64)
    bp.set_launch(False)
    ms_t = _event_ms(lambda: bp.matvec(x64, out=y), 5)
    # the persistent grid picked by measurement (the autotuner's batch mode)
    bp, tuned = autotune_batch(bp, reps=3)
    ms_q = _event_ms(lambda: bp.matvec(x64, out=y), 5)
    theta, _, it, info = batch_ground_locg(bp, x, tol=1e-7, maxiter=maxiter, check_every=10, use_graph=True)
    rel = info['rel_residual']
    log(f'batch: matvec {ms_q:.1f} ms (queue) / {ms_t:.1f} ms (per tile); LOCG {info["seconds"]:.2f} s, '
        f'{int((rel <= 1e-7).sum())}/64 converged')
    # the same solves on hit caches of all 64 problems (both precisions, one pass)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cinfo = bp.build_cache(0)
    torch.cuda.synchronize()
    t_cache = time.perf_counter() - t0
    ms_c = _event_ms(lambda: bp.matvec(x64, out=y), 5)
    theta_c, _, it_c, info_c = batch_ground_locg(bp, x, tol=1e-7, maxiter=maxiter, check_every=10, use_graph=True)
    rel_c = info_c['rel_residual']
    log(f'batch cached: matvec {ms_c:.2f} ms, cache {cinfo["bytes"] / 1e9:.2f} GB in {t_cache:.2f} s; '
        f'LOCG {info_c["seconds"]:.2f} s, {int((rel_c <= 1e-7).sum())}/64 converged')
    out = {'problems': 64, 'sum_states': bp.n_states, 'sum_terms': bp.info['n_terms'], 'tiles': bp.info['n_tiles'],
           'matvec_ms_work_queue': round(ms_q, 3), 'matvec_ms_one_block_per_tile': round(ms_t, 3),
           'autotuned_launch': tuned[0], 'autotune_results': tuned,
           'locg_seconds': round(info['seconds'], 3), 'locg_iterations': it,
           'converged_rel_residual_1e-7': int((rel <= 1e-7).sum()), 'worst_rel_residual': float(rel.max()),
           'precision_switch_iteration': info['switch_iteration'], 'hip_graph': True,
           'setup_s': round(setup, 2), 'dtype': 'c64 matvec, c128 vectors/RR',
           'cached': {'cache_bytes': cinfo['bytes'], 'build_s': round(t_cache, 3), 'matvec_ms': round(ms_c, 3),
                      'locg_seconds': round(info_c['seconds'], 3), 'locg_iterations': it_c,
                      'converged_rel_residual_1e-7': int((rel_c <= 1e-7).sum()),
                      'max_theta_diff_vs_matrix_free': float((theta_c - theta).abs().max())}}
    del bp
    torch.cuda.empty_cache()
    return out


def extra_grad(reps=5):
    """36-qubit, 10k-term random Pauli sum on 1M states: forward, backward and Hellmann-Feynman."""
    import numpy as np
    import torch
    from rqutils_amd import workloads
    from rqutils_amd.plan import Plan
    w = workloads.random36_10kt()
    plan = Plan(w.tables, w.states)
    rng = np.random.default_rng(3003)
    n = plan.n_total
    x = torch.from_numpy(rng.standard_normal(n) + 1j * rng.standard_normal(n)).to(plan.device)
    gy = torch.from_numpy(rng.standard_normal(n) + 1j * rng.standard_normal(n)).to(plan.device)
    out = {'workload': w.name, 'n_terms': plan.info['n_terms'], 'dtype': 'c128'}
    out['matvec_ms'] = round(_event_ms(lambda: plan.matvec(x), reps), 3)
    out['grad_x_ms'] = out['matvec_ms']
    out['grad_coef_ms'] = round(_event_ms(lambda: plan.grad_coef(x, gy), reps), 3)
    out['hellmann_feynman_ms'] = round(_event_ms(lambda: plan.expect_terms(x), reps), 3)
    # identity check at full size: sum_k alpha_k g_k = Re <gy, H x>
    alpha = torch.from_numpy(w.tables.alphas().real.copy()).to(plan.device)
    g = plan.grad_coef(x, gy)
    lhs, rhs = torch.dot(alpha, g).item(), torch.vdot(gy, plan.matvec(x)).real.item()
    out['identity_rel_err'] = abs(lhs - rhs) / max(abs(rhs), 1e-300)
    log(f'grad: matvec {out["matvec_ms"]} ms, grad_coef {out["grad_coef_ms"]} ms')
    del plan
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--workload', default='jw1m')
    ap.add_argument('--dtype', default='c64', choices=['c64', 'f32', 'f64', 'c128'])
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-budget', type=float, default=20.0)
    ap.add_argument('--split-bits', type=int, default=-1)
    ap.add_argument('--no-locg', action='store_true')
    ap.add_argument('--locg-real', action='store_true', help='float64/float32 LOCG instead of complex')
    ap.add_argument('--locg-maxiter', type=int, default=1000)
    ap.add_argument('--mixed-tol', type=float, default=1e-6)
    ap.add_argument('--energy-tol', type=float, default=1e-9)
    ap.add_argument('--no-extras', action='store_true',
                    help='skip the 25M matvec, the 64-problem batch and the gradient timings')
    ap.add_argument('--calibrate-copy', action='store_true',
                    help='only copy 256 MiB three times (a kernel of known traffic for the PMC byte counters)')
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        src = torch.randn(64 * 1024 * 1024, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({'copy_bytes_per_dispatch': src.numel() * 4, 'dispatches': 3}))
        return

    if args.gpus != 1:
        print(json.dumps({'error': 'this build runs on one GPU; --gpus must be 1 '
                                   '(multi-GPU sharding is design only)'}))
        sys.exit(2)

    import numpy as np
    import torch
    from rqutils_amd.plan import Plan

    dev = torch.device('cuda', 0)
    torch.cuda.set_device(dev)
    dt = {'c64': torch.complex64, 'f32': torch.float32, 'f64': torch.float64, 'c128': torch.complex128}[args.dtype]

    t0 = time.perf_counter()
    w = load_workload(args.workload)
    t_gen = time.perf_counter() - t0
    log(f'workload {w.name}: N={w.states.shape[0]} J={w.tables.num_groups} T={w.tables.num_terms} '
        f'(generated in {t_gen:.1f} s)')
    t0 = time.perf_counter()
    plan = Plan(w.tables, w.states, split_bits=args.split_bits)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    log(f'plan: {plan.info} (setup {t_setup:.2f} s)')
    counts = plan.count()
    alg = plan.algorithmic_bytes(dt, hits=counts['hits'])
    log(f'counts: {counts}; algorithmic bytes/matvec {alg / 1e9:.3f} GB')

    rng = np.random.default_rng(1002)
    xr = rng.standard_normal(plan.n_total)
    if dt.is_complex:
        x = torch.from_numpy(xr + 1j * rng.standard_normal(plan.n_total)).to(dev, dt)
    else:
        x = torch.from_numpy(xr).to(dev, dt)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream(dev)

    def timed(label):
        for _ in range(args.warmup):
            plan.matvec(x, out=y)
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        wall0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            plan.matvec(x, out=y)
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - wall0
        ms = ev0.elapsed_time(ev1) / args.steps
        log(f'{label}: {args.steps} matvecs after {args.warmup} warmup, {ms:.4f} ms/matvec (events), '
            f'wall {wall * 1e3 / args.steps:.4f} ms')
        return ms

    # matrix-free kernel (the tile kernel searching every (row, group) pair)
    ms_mf = timed('matrix-free')
    ach_mf = alg / (ms_mf * 1e-3) / 1e9
    matrix_free = {
        'ms_per_matvec': round(ms_mf, 4),
        'model_bytes_per_matvec': alg,
        'roofline': {'bound': 'lds', 'achieved': round(ach_mf, 2), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(ach_mf / HBM_PEAK_GBS, 4), 'traffic': traffic_from_profiles(w.name, args.dtype),
                     'note': 'model bytes count index probes that LDS/L2 serve; measured traffic is HBM bytes'},
    }
    # hit cache (run_sqd's default cache_strategy (1, *)): built once, then every matvec streams it
    precision = 64 if dt in (torch.float64, torch.complex128) else 32
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cinfo = plan.build_cache(precision)
    torch.cuda.synchronize()
    t_cache = time.perf_counter() - t0
    log(f'hit cache: {cinfo} built in {t_cache:.3f} s')
    ms = timed('cached')
    y_cached = y.clone()
    plan.use_cache(False)
    plan.matvec(x, out=y)
    plan.use_cache(True)
    identical = bool(torch.equal(y, y_cached))
    del y_cached
    nbytes = plan.cached_bytes(dt)
    achieved = nbytes / (ms * 1e-3) / 1e9
    traffic = traffic_from_profiles(w.name, args.dtype + ':cached')
    result = {
        'metric': METRIC,
        'value': round(achieved, 2),
        'unit': 'GB/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(ms, 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': args.dtype,
        'data': 'synthetic (seeded generators, rqutils_amd/workloads.py)',
        'config': {
            'workload': w.name,
            'path': 'hit cache (cache_strategy (1, *), the run_sqd default); matrix-free kernel in matrix_free',
            'n_states': plan.n_states,
            'n_groups': plan.info['n_groups'],
            'n_terms': plan.info['n_terms'],
            'hits_per_matvec': counts['hits'],
            'segment_probes_per_matvec': counts['probes'],
            'split_bits': plan.info['split_bits'],
            'bytes_per_matvec': nbytes,
            'cache_bytes': cinfo['bytes'],
            'cache_entries_incl_padding': cinfo['entries'],
            'cache_build_s': round(t_cache, 4),
            'identical_to_matrix_free': identical,
            'setup_s': round(t_setup, 3),
        },
        'roofline': {
            'bound': 'hbm',
            'achieved': round(achieved, 2),
            'peak': HBM_PEAK_GBS,
            'unit': 'GB/s',
            'frac': round(achieved / HBM_PEAK_GBS, 4),
            'traffic': traffic,
        },
        'matrix_free': matrix_free,
    }
    plan.drop_cache()
    if not args.no_locg:
        result['locg'] = locg_time_to_solution(plan, not args.locg_real, args.locg_maxiter, args.mixed_tol,
                                               args.energy_tol)
    if not args.no_locg:
        loc = locg_cached(plan, not args.locg_real, args.locg_maxiter, args.mixed_tol, args.energy_tol,
                          result['locg']['e_star'])
        loc['speedup_vs_matrix_free_plain'] = round(result['locg']['plain']['time_to_target_s'] /
                                                    loc['mixed_graph']['time_s'], 3)
        result['locg_cached'] = loc
        plan.drop_cache()
    if not args.no_extras:
        del y, x
        torch.cuda.empty_cache()
        result['extras'] = {'jw25m_matvec': extra_large_matvec(), 'batch64_locg': extra_batch(),
                            'grad36q': extra_grad()}
    if not args.no_cpu:
        log(f'cpu baseline (budget {args.cpu_budget} s) ...')
        result['cpu_baseline'] = cpu_baseline(w, nbytes, args.cpu_budget)
        log(f"cpu baseline {result['cpu_baseline']['value']:.3f} GB/s")
    print(json.dumps(result), flush=True)


if __name__ == '__main__':
    main()
