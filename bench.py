"""Benchmark: SQD subspace matvec throughput on one MI355X (+ CPU baseline).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload jw25m|jw1m|jw40-1m|rand8k]
                    [--dtype c64|f32|f64|c128] [--traffic live|none] [--no-cpu] [--no-locg] [--no-extras]

A step is one matvec ``y = H_S x`` over the whole subspace with the plan and
``x`` resident in HBM.  The default workload is the 25M-determinant JW model
(40 spin-orbitals, 5000 x 5000 sampled alpha/beta strings) in complex64.  The
headline path is the one ``run_sqd`` takes by default (``cache_strategy``
(1, *)): the hits, built once by the matrix-free tile kernel, streamed from
the device hit cache.  ``value`` is that matvec's rate over the non-padding
bytes of the cache plus the diagonal, ``x`` once and ``y``, divided by the
kernel time measured with HIP events on the launch stream.  The matrix-free
kernel is timed over the same steps and reported under ``matrix_free``.
With ``--traffic live`` (default) the measured HBM bytes per launch come from
two ``rocprofv3 --pmc`` passes (FETCH_SIZE, then WRITE_SIZE) over a child run
of the same matvecs, scaled by the same counters on a 256 MiB device copy.
One JSON line is printed on stdout.
"""
import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'matvec GB/s (% HBM peak); time to 1e-8 rel. ground-energy err vs FP64 plain'
HBM_PEAK_GBS = 8000.0
# reference ground energies computed by the oracle (tests/golden/make_estar.py)
ESTAR = {'jw1m': 'jw-32so-1000x1000', 'jw40-1m': 'jw-40so-1000x1000'}
DTYPES = ('c64', 'f32', 'f64', 'c128')


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def load_workload(name):
    from rqutils_amd import workloads
    if name == 'jw1m':
        return workloads.jw32()
    if name == 'jw25m':
        return workloads.jw40()
    if name == 'jw40-1m':
        return workloads.jw40(1000, 1000)
    if name == 'rand8k':
        return workloads.random20_300t()
    raise ValueError(f'unknown workload {name}')


def torch_dtype(name):
    import torch
    return {'c64': torch.complex64, 'f32': torch.float32, 'f64': torch.float64, 'c128': torch.complex128}[name]


def oracle_estar(name, w):
    """The oracle's eigsh ground energy of workload ``name`` (a committed fixture), or None."""
    fixture = ESTAR.get(name)
    if fixture is None:
        return None
    path = os.path.join(ROOT, 'tests', 'golden', f'estar_{fixture}.json')
    if not os.path.exists(path):
        return None
    from rqutils_amd import workloads
    with open(path) as f:
        ref = json.load(f)
    got = workloads.fingerprint(w)
    for k, v in ref['workload'].items():
        same = abs(got[k] - v) <= 1e-9 * abs(v) if k == 'coef_abs_sum' else got[k] == v
        if not same:
            log(f'oracle fixture {fixture} does not match the generated workload ({k})')
            return None
    return ref['e_star']


def random_vector(n, dtype, seed, device):
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    x = torch.from_numpy(rng.standard_normal(n) + 1j * rng.standard_normal(n)).to(device)
    return (x if dtype.is_complex else x.real).to(dtype).contiguous()


def event_ms(fn, steps, warmup):
    """Average milliseconds per call of ``fn`` from HIP events on the current stream."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


# ---------------------------------------------------------------------------- HBM traffic (rocprofv3 --pmc)

# the matrix-free matvec runs k_product on fully listed product subspaces, k_tiles otherwise
KERNELS = {'cached': ('k_spmv_sell<',), 'matrix_free': ('k_tiles<0,', 'k_product<')}


def pmc_child(args):
    """The matvecs whose counters the parent collects: matrix-free, then cached (3 launches each)."""
    import torch
    from rqutils_amd.plan import Plan
    w = load_workload(args.workload)
    dt = torch_dtype(args.dtype)
    plan = Plan(w.tables, w.states)
    x = random_vector(plan.n_total, dt, 1002, plan.device)
    y = torch.empty_like(x)
    for _ in range(3):
        plan.matvec(x, out=y)
    plan.build_cache(64 if dt in (torch.float64, torch.complex128) else 32)
    plan.use_cache(True)                    # the cached kernel (the build timed one launch of each path too)
    for _ in range(3):
        plan.matvec(x, out=y)
    torch.cuda.synchronize()
    print(json.dumps({'done': True}), flush=True)


def _rocprof_counter(counter, child_args, tmp):
    """Run ``rocprofv3 --pmc counter`` over ``python3 bench.py child_args``; per-kernel average KiB per dispatch."""
    out_dir = os.path.join(tmp, f'pmc_{counter}_{abs(hash(tuple(child_args))) % 10000}')
    cmd = ['rocprofv3', '--pmc', counter, '--output-format', 'csv', '-d', out_dir, '-o', 'run', '--',
           sys.executable, os.path.join(ROOT, 'bench.py'), *child_args]
    env = dict(os.environ, TMPDIR='/tmp')
    try:
        res = subprocess.run(cmd, cwd=tmp, env=env, capture_output=True, text=True, timeout=150)
    except subprocess.TimeoutExpired:
        raise RuntimeError(f'rocprofv3 --pmc {counter} timed out')
    files = glob.glob(os.path.join(out_dir, '**', '*counter_collection.csv'), recursive=True)
    if not files:
        raise RuntimeError(f'rocprofv3 --pmc {counter}: no counter file (exit {res.returncode}): '
                           f'{res.stderr.strip()[-300:]}')
    sums, counts = {}, {}
    for path in files:
        with open(path, newline='') as f:
            for row in csv.DictReader(f):
                if row.get('Counter_Name') != counter:
                    continue
                name = row.get('Kernel_Name', '')
                sums[name] = sums.get(name, 0.0) + float(row['Counter_Value'])
                counts[name] = counts.get(name, 0) + 1
    return {k: (sums[k] / counts[k], counts[k]) for k in sums}


def _kernel_avg(table, prefix):
    names = (prefix,) if isinstance(prefix, str) else prefix
    hits = [(v, n) for k, (v, n) in table.items() if any(q in k for q in names)]
    if not hits:
        raise RuntimeError(f'no dispatch of {" or ".join(names)} in the counter file')
    total = sum(v * n for v, n in hits)
    n = sum(n for _, n in hits)
    return total / n, n


def live_traffic(workload, dtype):
    """Measured HBM bytes per launch of the cached and matrix-free matvecs (scaled PMC counters)."""
    if shutil.which('rocprofv3') is None:
        raise RuntimeError('rocprofv3 not found')
    tmp = tempfile.mkdtemp(prefix='rqx_pmc_', dir='/tmp')
    try:
        child = ['--pmc-child', '--workload', workload, '--dtype', dtype]
        copy = ['--calibrate-copy']
        log('traffic: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE over the copy calibration ...')
        cf, _ = _kernel_avg(_rocprof_counter('FETCH_SIZE', copy, tmp), 'copyBuffer')
        cw, _ = _kernel_avg(_rocprof_counter('WRITE_SIZE', copy, tmp), 'copyBuffer')
        copy_kib = 64 * 1024 * 1024 * 4 / 1024.0
        f_scale, w_scale = copy_kib / cf, copy_kib / cw
        log(f'traffic: copy scales fetch x{f_scale:.4f}, write x{w_scale:.4f}; matvec counters ...')
        fetch = _rocprof_counter('FETCH_SIZE', child, tmp)
        write = _rocprof_counter('WRITE_SIZE', child, tmp)
        out = {'fetch_scale': f_scale, 'write_scale': w_scale, 'copy_fetch_kib': cf, 'copy_write_kib': cw}
        for key, prefix in KERNELS.items():
            fk, nf = _kernel_avg(fetch, prefix)
            wk, nw = _kernel_avg(write, prefix)
            out[key] = {'hbm_bytes_per_launch': int(round((fk * f_scale + wk * w_scale) * 1024)),
                        'fetch_kib_raw': round(fk, 1), 'write_kib_raw': round(wk, 1), 'dispatches': [nf, nw]}
        return out
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


# ---------------------------------------------------------------------------- LOCG legs

def x0_onehot(plan, dtype):
    """The reference's start vector: one-hot at the argmin of the x = 0 diagonal (``sqd.py:406-426``)."""
    import torch
    diag = plan.diagonal().real[:plan.n_states]
    x0 = torch.zeros(plan.n_total, dtype=dtype, device=plan.device)
    x0[int(torch.argmin(diag))] = 1
    return x0


def graph_mixed(plan, x0, maxiter, energy_tol, e_ref, check_every=5, probe_below=1e-4, target=1e-8):
    """Mixed LOCG with the fused per-problem kernels, one HIP graph per step (single problem).

    Below relative residual ``probe_below`` every poll also takes the FP64 Rayleigh quotient
    of the current vector (one FP64 matvec, timed with the run): the first probe within
    ``target`` of ``e_ref`` is the time to target, measured as for the plain run.
    """
    import torch
    from rqutils_amd.locg_batch import batch_ground_locg
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    theta, _, it, info = batch_ground_locg(plan, x0.clone(), tol=1e-12, maxiter=maxiter, check_every=check_every,
                                           energy_tol=energy_tol, use_graph=True, energy_probe=probe_below)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    err = abs(float(theta[0]) - e_ref) / abs(e_ref)
    hit = next(((ts, k, float(th[0])) for ts, k, th, _ in info['probes']
                if abs(float(th[0]) - e_ref) <= target * abs(e_ref)), None)
    t_target = hit[0] if hit is not None else t
    log(f'graph mixed LOCG: within {target:g} after {t_target:.3f} s (iteration {hit[1] if hit else it}); '
        f'own stop: rel err {err:.2e} in {t:.3f} s ({it} iterations, switch at {info["switch_iteration"]})')
    return {'time_to_target_s': round(t_target, 4), 'target_iteration': hit[1] if hit else it,
            'target_energy': hit[2] if hit else float(theta[0]), 'probes': len(info['probes']),
            'time_s': round(t, 4), 'iterations': it, 'switch_iteration': info['switch_iteration'],
            'energy': float(theta[0]), 'rel_err': err, 'stop_energy_tail': energy_tol, 'check_every': check_every,
            'dtype': 'c64 matvec, c128 vectors/RR', 'hip_graph': True}


_WARM = []


def locg_warmup():
    """Both solvers once on a small JW problem (seconds: 0.1), so that no timed run pays the
    process's first use of a torch operator or the capture of its first graph (torch
    loads its GPU code objects lazily; a first call of one operator costs up to 0.1 s)."""
    if _WARM:
        return
    import torch
    from rqutils_amd import workloads
    from rqutils_amd.ground_locg import ground_locg
    from rqutils_amd.locg_batch import batch_ground_locg
    from rqutils_amd.locg_mixed import ground_locg_mixed
    from rqutils_amd.plan import Plan
    w = workloads.jw32(60, 50)
    plan = Plan(w.tables, w.states)
    x0 = x0_onehot(plan, torch.complex128)
    for build in (64, 0):
        plan.build_cache(build)
        ground_locg(plan.matvec, x0.clone(), maxiter=20)
        batch_ground_locg(plan, x0.clone(), tol=1e-12, maxiter=60, check_every=5, energy_tol=1e-9,
                          use_graph=True, energy_probe=1.0)
        plan.drop_cache()
    ground_locg_mixed(plan.matvec, plan.matvec, x0.clone(), maxiter=40, tol=1e-9, energy_tol=1e-9)
    torch.cuda.synchronize()
    plan.close()
    _WARM.append(True)


def locg_legs(plan, e_oracle, maxiter=1000, energy_tol=1e-9, target=1e-8, matrix_free=True):
    """Time to ``target`` relative ground-energy error: plain FP64 LOCG (the reference sequence,
    complex128) against mixed-precision LOCG, on the hit caches (builds included) and, with
    ``matrix_free``, without them.

    The reference energy ``e_ref`` is the oracle's eigsh energy when a fixture exists for the
    workload, else the plain run's own converged energy.  Both times to target are measured
    the same way: the first moment the run's energy estimate is within ``target`` of
    ``e_ref`` -- for the plain run its theta (an FP64 Rayleigh quotient every iteration),
    for the mixed run the FP64 Rayleigh quotient of its vector taken at its polls (the
    bulk phase's own theta carries the FP32 image's error, so only these probes count;
    their matvecs are inside the timed run).  The mixed run's own stop (its settle rule,
    no knowledge of ``e_ref``) is reported beside it.
    """
    import torch
    from rqutils_amd.ground_locg import ground_locg
    from rqutils_amd.locg_mixed import ground_locg_mixed
    out = {}
    locg_warmup()
    x0 = x0_onehot(plan, torch.complex128)

    def plain(label):
        trace = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if label == 'cached':
            plan.build_cache(64)
        theta, _, niter = ground_locg(plan.matvec, x0.clone(), maxiter=maxiter,
                                      callback=lambda k, th: trace.append((time.perf_counter() - t0, th)))
        torch.cuda.synchronize()
        total = time.perf_counter() - t0
        return float(theta), niter, total, trace

    def to_target(trace, e_ref, total):
        return next((t for t, th in trace if abs(th - e_ref) <= target * abs(e_ref)), total)

    plan.drop_cache()
    e_plain, niter, total, trace = plain('cached')
    plain_path = plan.matvec_path(torch.complex128)
    e_ref = e_oracle if e_oracle is not None else e_plain
    t_plain = to_target(trace, e_ref, total)
    plan.drop_cache()
    log(f'plain LOCG (FP64 cache): E={e_plain:.12f}, {niter} iterations, within {target:g} of '
        f'{"the oracle" if e_oracle is not None else "its own energy"} after {t_plain:.3f} s (total {total:.3f} s)')
    out['reference_energy'] = {'value': e_ref, 'source': 'oracle eigsh fixture' if e_oracle is not None
                               else 'plain FP64 LOCG on the GPU (no oracle at this size)'}
    out['plain_cached'] = {'dtype': 'c128', 'energy': e_plain, 'rel_err': abs(e_plain - e_ref) / abs(e_ref),
                           'iterations': niter, 'matvecs': 4 * niter + 3, 'time_to_target_s': round(t_plain, 4),
                           'total_s': round(total, 4), 'matvec_path': plain_path}
    torch.cuda.empty_cache()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.build_cache(0)                     # FP32 and FP64 values in one pass
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    g = graph_mixed(plan, x0, maxiter, energy_tol, e_ref, target=target)
    g['matvec_paths'] = {'c64': plan.matvec_path(torch.complex64), 'c128': plan.matvec_path(torch.complex128)}
    g['time_to_target_s'] = round(g['time_to_target_s'] + t_build, 4)
    g['time_s'] = round(g['time_s'] + t_build, 4)
    g['cache_build_s'] = round(t_build, 4)
    g['speedup_vs_plain'] = round(t_plain / g['time_to_target_s'], 3)
    g['speedup_own_stop_vs_plain_target'] = round(t_plain / g['time_s'], 3)
    out['mixed_graph_cached'] = g
    plan.drop_cache()
    if matrix_free:
        e_mf, niter, total, trace = plain('free')
        t_mf = to_target(trace, e_ref, total)
        log(f'plain LOCG (matrix-free): {niter} iterations, within {target:g} after {t_mf:.3f} s')
        out['plain_matrix_free'] = {'dtype': 'c128', 'energy': e_mf, 'rel_err': abs(e_mf - e_ref) / abs(e_ref),
                                    'iterations': niter, 'time_to_target_s': round(t_mf, 4), 'total_s': round(total, 4)}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = ground_locg_mixed(plan.matvec, plan.matvec, x0.clone(), maxiter=maxiter, tol=1e-9,
                                energy_tol=energy_tol)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        err = abs(float(res.theta) - e_ref) / abs(e_ref)
        log(f'mixed LOCG (matrix-free): rel err {err:.2e} in {t:.3f} s')
        out['mixed_matrix_free'] = {'time_s': round(t, 4), 'iterations': res.niter, 'matvecs_low': res.matvecs_lo,
                                    'matvecs_fp64': res.matvecs_hi, 'energy': float(res.theta), 'rel_err': err,
                                    'speedup_vs_plain': round(t_mf / t, 3)}
        g = graph_mixed(plan, x0, maxiter, energy_tol, e_ref, target=target)
        g['speedup_vs_plain'] = round(t_mf / g['time_to_target_s'], 3)
        out['mixed_graph_matrix_free'] = g
    return out


# ---------------------------------------------------------------------------- matvec legs

def matvec_leg(plan, dt, steps, warmup, seed=1002):
    """Matrix-free and cached matvec times, bytes and the bitwise check between them."""
    import torch
    counts = plan.count()
    alg = plan.algorithmic_bytes(dt, hits=counts['hits'])
    x = random_vector(plan.n_total, dt, seed, plan.device)
    y = torch.empty_like(x)
    ms_mf = event_ms(lambda: plan.matvec(x, out=y), steps, warmup)
    y_free = y.clone()
    log(f'matrix-free matvec: {ms_mf:.4f} ms')
    precision = 64 if dt in (torch.float64, torch.complex128) else 32
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cinfo = plan.build_cache(precision)
    torch.cuda.synchronize()
    t_cache = time.perf_counter() - t0
    auto_path = plan.matvec_path(dt)          # the default mode's pick (build_cache timed both)
    plan.use_cache(True)                      # the cached kernel itself
    ms = event_ms(lambda: plan.matvec(x, out=y), steps, warmup)
    plan.use_cache('auto')
    identical = bool(torch.equal(y, y_free))
    del y_free
    stored = plan.cached_bytes(dt)
    padding = plan.cached_padding_bytes(dt)
    useful = stored - padding
    log(f'cached matvec: {ms:.4f} ms, cache {cinfo["bytes"] / 1e9:.2f} GB built in {t_cache:.3f} s; '
        f'{useful / 1e9:.3f} GB non-padding of {stored / 1e9:.3f} GB stored -> {useful / ms / 1e6:.0f} GB/s; '
        f'identical to matrix-free: {identical}; default path with the cache built: {auto_path}')
    plan.drop_cache()
    return {'counts': counts, 'alg_bytes': alg, 'ms_free': ms_mf, 'ms': ms, 'cache': cinfo, 'cache_build_s': t_cache,
            'stored_bytes': stored, 'padding_bytes': padding, 'useful_bytes': useful, 'identical': identical,
            'auto_path': auto_path, 'free_path': plan.matvec_path(dt)}


def line_for(w, plan, leg, dt_name, steps, warmup, traffic):
    """The JSON line's fields for one workload's matvec leg."""
    useful, ms = leg['useful_bytes'], leg['ms']
    achieved = useful / (ms * 1e-3) / 1e9
    listed = all(pr['list_mode'] == 2 for pr in plan.problems)
    ach_mf = leg['alg_bytes'] / (leg['ms_free'] * 1e-3) / 1e9
    cinfo = leg['cache']
    t_cached = traffic.get('cached') if isinstance(traffic, dict) else None
    t_free = traffic.get('matrix_free') if isinstance(traffic, dict) else None
    hbm_mf = t_free['hbm_bytes_per_launch'] / (leg['ms_free'] * 1e-3) / 1e9 if t_free else None
    return {
        'value': round(achieved, 2),
        'ms_per_step': round(ms, 4),
        'config': {
            'workload': w.name,
            'path': ('hit cache kernel (cache_strategy (1, *), the run_sqd default; timed with the cache forced: '
                     'the default mode runs the kernel named by default_path_with_cache, the faster one the '
                     'cache build timed, same bits); matrix-free kernel in matrix_free'),
            'n_states': plan.n_states, 'n_groups': plan.info['n_groups'], 'n_terms': plan.info['n_terms'],
            'hits_per_matvec': leg['counts']['hits'], 'segment_probes_per_matvec': leg['counts']['probes'],
            'split_bits': plan.info['split_bits'],
            'bytes_per_matvec': useful, 'stored_bytes_per_matvec': leg['stored_bytes'],
            'padding_bytes_per_matvec': leg['padding_bytes'],
            'stored_GBs': round(leg['stored_bytes'] / (ms * 1e-3) / 1e9, 1),
            'cache_bytes': cinfo['bytes'], 'cache_entries_incl_padding': cinfo['entries'],
            'cache_info': {k: v for k, v in cinfo.items() if k != 'built'},
            'cache_build_s': round(leg['cache_build_s'], 4), 'identical_to_matrix_free': leg['identical'],
            'default_path_with_cache': leg['auto_path'],
        },
        'roofline': {
            'bound': 'hbm', 'achieved': round(achieved, 2), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
            'frac': round(achieved / HBM_PEAK_GBS, 4),
            'traffic': t_cached['hbm_bytes_per_launch'] if t_cached else None,
        },
        'matrix_free': {
            'kernel': leg['free_path'],
            'ms_per_matvec': round(leg['ms_free'], 4),
            'model_bytes_per_matvec': leg['alg_bytes'],
            'list_mode': [pr['list_mode'] for pr in plan.problems],
            'roofline': {'bound': 'hbm' if listed else 'lds',
                         'achieved': round(hbm_mf, 2) if hbm_mf is not None else None, 'peak': HBM_PEAK_GBS,
                         'unit': 'GB/s', 'frac': round(hbm_mf / HBM_PEAK_GBS, 4) if hbm_mf is not None else None,
                         'traffic': t_free['hbm_bytes_per_launch'] if t_free else None,
                         'model_GBs': round(ach_mf, 2), 'model_frac': round(ach_mf / HBM_PEAK_GBS, 4),
                         'note': ('achieved = measured HBM bytes per launch / kernel time.  Model bytes (every hit '
                                  'from the product-subspace lists, no index probe: keys, one gathered x entry per '
                                  'hit, y, term tables) count the x gathers that the kernel serves from LDS, '
                                  'so model_frac is no HBM bound')
                         if listed else
                         'achieved = measured HBM bytes per launch / kernel time; the model bytes count index '
                         'probes that LDS/L2 serve'},
        },
    }


def cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_baseline(w, dt_name, useful_bytes, gpu_ms, budget_s):
    """Oracle C port of the reference's strategy (1, 0), timed on a sample of groups of the same workload."""
    import numpy as np
    from oracle import build as obuild, cport
    if not os.path.exists(cport.LIB):
        obuild.build()
    np_dt = np.complex64 if dt_name in ('c64', 'f32') else np.complex128
    t_mv, groups, threads, frac = cport.timed_sample(w.tables, w.states, budget_s=budget_s, dtype=np_dt)
    return {
        'value': round(t_mv, 3),
        'unit': 's/matvec',
        'cores': threads,
        'kind': 'port',
        'sample': (f'{w.name}: strategy (1,0) per-matvec work (group diagonal recomposition + gather, '
                   f'{"complex64" if np_dt == np.complex64 else "complex128"}) on {groups} of {w.tables.num_groups} '
                   f'X-groups ({100 * frac:.2f}% of the sum_j (K_j + 1) N work units), extrapolated by term count; '
                   f'source indices (cached once by the reference) precomputed, not timed'),
        'cpu_model': cpu_model(),
        'sampled_fraction': frac,
        'equivalent_GBs': round(useful_bytes / t_mv / 1e9, 3),
        'gpu_speedup': round(t_mv / (gpu_ms * 1e-3), 1),
    }


# ---------------------------------------------------------------------------- extras

def setup_leg(w, dup_frac=0.1, seed=1003, host_path=True):
    """run_sqd's setup from the reference's packed rows (``sqd.py:255`` layout: pad
    bit first, big endian) to a ready plan, on the device: upload the rows,
    unpack (``pack_to_u64``), sort/unique, build the index from the device keys.
    The rows are the workload's states shuffled, with ``dup_frac`` of them
    repeated (samples repeat); one untimed pass over 100k of them first loads the
    kernels.  Returns the plan and the stage times; with
    ``host_path`` also the time of the host route (unpack with numpy, sort/unique
    on the device, index built from host keys) for comparison."""
    import numpy as np
    import torch
    from rqutils_amd.paulis.symplectic import masks_to_packed, packed_to_masks
    from rqutils_amd.plan import Plan, pack_to_u64, sort_unique_u64
    from rqutils_amd.sqd import uniquify_states
    rng = np.random.default_rng(seed)
    n = w.states.shape[0]
    rows_keys = np.concatenate([w.states, rng.choice(w.states, size=int(dup_frac * n))])
    rng.shuffle(rows_keys)
    t0 = time.perf_counter()
    rows = masks_to_packed(rows_keys, w.num_qubits, padded=True)
    log(f'setup: {rows.shape[0]} packed rows of {rows.shape[1]} bytes (packed in {time.perf_counter() - t0:.1f} s)')
    del rows_keys
    dev = torch.device('cuda', torch.cuda.current_device())
    # the same route once on the first 100k rows: the process's first launch of each
    # kernel (code object loading) and the allocator's first blocks are not timed
    warm = Plan(w.tables, sort_unique_u64(pack_to_u64(torch.from_numpy(rows[:100_000]).to(dev), w.num_qubits)),
                device=dev)
    warm.close()
    torch.cuda.synchronize()
    times = {}
    t = time.perf_counter()
    t0 = t

    def mark(name):
        nonlocal t
        torch.cuda.synchronize()
        now = time.perf_counter()
        times[name] = round(now - t, 4)
        t = now

    rows_d = torch.from_numpy(rows).to(dev)
    mark('upload_s')
    keys = pack_to_u64(rows_d, w.num_qubits)
    mark('unpack_s')
    keys_u = sort_unique_u64(keys)
    mark('sort_unique_s')
    plan = Plan(w.tables, keys_u, device=dev)
    mark('index_s')
    times['total_s'] = round(time.perf_counter() - t0, 4)
    assert plan.n_states == n
    del rows_d, keys, keys_u
    out = {'rows': int(rows.shape[0]), 'row_bytes': int(rows.shape[1]), 'n_unique': n, 'device_path': times}
    log(f'setup (device path): {times}')
    if host_path:
        t0 = time.perf_counter()
        hk = packed_to_masks(rows, w.num_qubits, padded=True)
        t_unpack = time.perf_counter() - t0
        hu = uniquify_states(hk)
        t_uniq = time.perf_counter() - t0 - t_unpack
        hp = Plan(w.tables, hu, device=dev)
        torch.cuda.synchronize()
        total = time.perf_counter() - t0
        same = hp.digest() == plan.digest()
        hp.close()
        del hk, hu
        out['host_path'] = {'unpack_s': round(t_unpack, 4), 'sort_unique_s': round(t_uniq, 4),
                            'index_s': round(total - t_unpack - t_uniq, 4), 'total_s': round(total, 4),
                            'same_index': same}
        log(f"setup (host path): {out['host_path']}")
    torch.cuda.empty_cache()
    return plan, out


def extra_workload(name, dt, steps, warmup, locg=True, matrix_free_locg=True):
    """Another workload's matvec legs and LOCG times (ground energies against the oracle where it has one)."""
    import torch
    from rqutils_amd.plan import Plan
    w = load_workload(name)
    t0 = time.perf_counter()
    plan = Plan(w.tables, w.states)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    leg = matvec_leg(plan, dt, steps, warmup)
    line = line_for(w, plan, leg, None, steps, warmup, None)
    out = {'workload': w.name, 'setup_s': round(setup, 3), 'ms_per_matvec': line['ms_per_step'],
           'achieved_GBs': line['value'], 'frac_of_8TBs': line['roofline']['frac'], 'config': line['config'],
           'matrix_free': line['matrix_free']}
    if locg:
        out['locg'] = locg_legs(plan, oracle_estar(name, w), matrix_free=matrix_free_locg)
    plan.close()
    torch.cuda.empty_cache()
    return out


def extra_batch(maxiter=1000):
    """64 independent 40-qubit random Pauli sums in one batched plan: matvec + batched LOCG."""
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
    x = random_vector(bp.n_total, torch.complex128, 2002, bp.device)
    x64 = x.to(torch.complex64)
    y = torch.empty_like(x64)
    bp.set_launch(False)
    ms_t = event_ms(lambda: bp.matvec(x64, out=y), 5, 1)
    bp, tuned = autotune_batch(bp, reps=3)   # the persistent grid picked by measurement
    ms_q = event_ms(lambda: bp.matvec(x64, out=y), 5, 1)
    theta, _, it, info = batch_ground_locg(bp, x, tol=1e-7, maxiter=maxiter, check_every=10, use_graph=True)
    rel = info['rel_residual']
    log(f'batch: matvec {ms_q:.1f} ms (queue) / {ms_t:.1f} ms (per tile); LOCG {info["seconds"]:.2f} s, '
        f'{int((rel <= 1e-7).sum())}/64 converged')
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cinfo = bp.build_cache(0)
    torch.cuda.synchronize()
    t_cache = time.perf_counter() - t0
    ms_c = event_ms(lambda: bp.matvec(x64, out=y), 5, 1)
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
    bp.close()
    torch.cuda.empty_cache()
    return out


def extra_grad(reps=5):
    """36-qubit, 10k-term random Pauli sum on 1M states: forward, backward and Hellmann-Feynman."""
    import torch
    from rqutils_amd import workloads
    from rqutils_amd.plan import Plan
    w = workloads.random36_10kt()
    plan = Plan(w.tables, w.states)
    n = plan.n_total
    x = random_vector(n, torch.complex128, 3003, plan.device)
    gy = random_vector(n, torch.complex128, 3004, plan.device)
    out = {'workload': w.name, 'n_terms': plan.info['n_terms'], 'dtype': 'c128'}
    out['matvec_ms'] = round(event_ms(lambda: plan.matvec(x), reps, 1), 3)
    out['grad_x_ms'] = out['matvec_ms']
    out['grad_coef_ms'] = round(event_ms(lambda: plan.grad_coef(x, gy), reps, 1), 3)
    out['hellmann_feynman_ms'] = round(event_ms(lambda: plan.expect_terms(x), reps, 1), 3)
    # identity at full size: sum_k alpha_k g_k = Re <gy, H x>
    alpha = torch.from_numpy(w.tables.alphas().real.copy()).to(plan.device)
    g = plan.grad_coef(x, gy)
    lhs, rhs = torch.dot(alpha, g).item(), torch.vdot(gy, plan.matvec(x)).real.item()
    out['identity_rel_err'] = abs(lhs - rhs) / max(abs(rhs), 1e-300)
    log(f'grad: matvec {out["matvec_ms"]} ms, grad_coef {out["grad_coef_ms"]} ms')
    plan.close()
    torch.cuda.empty_cache()
    return out


# ---------------------------------------------------------------------------- main

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--workload', default='jw25m')
    ap.add_argument('--dtype', default='c64', choices=DTYPES)
    ap.add_argument('--traffic', default='live', choices=['live', 'none'])
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-budget', type=float, default=20.0)
    ap.add_argument('--split-bits', type=int, default=-1)
    ap.add_argument('--no-setup', action='store_true',
                    help='build the plan from host states instead of timing the setup from packed rows')
    ap.add_argument('--no-host-setup', action='store_true', help='skip timing the host setup route')
    ap.add_argument('--no-locg', action='store_true')
    ap.add_argument('--locg-maxiter', type=int, default=1000)
    ap.add_argument('--energy-tol', type=float, default=1e-9)
    ap.add_argument('--no-extras', action='store_true',
                    help='skip the 1M-determinant models, the 64-problem batch and the gradient timings')
    ap.add_argument('--pmc-child', action='store_true', help=argparse.SUPPRESS)
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
    if args.pmc_child:
        pmc_child(args)
        return

    if args.gpus != 1:
        print(json.dumps({'error': 'this build runs on one GPU; --gpus must be 1 '
                                   '(multi-GPU sharding is design only)'}))
        sys.exit(2)

    import torch
    from rqutils_amd.plan import Plan

    dev = torch.device('cuda', 0)
    torch.cuda.set_device(dev)
    dt = torch_dtype(args.dtype)

    t0 = time.perf_counter()
    w = load_workload(args.workload)
    log(f'workload {w.name}: N={w.states.shape[0]} J={w.tables.num_groups} T={w.tables.num_terms} '
        f'(generated in {time.perf_counter() - t0:.1f} s)')
    setup = None
    if args.split_bits < 0 and not args.no_setup:
        plan, setup = setup_leg(w, host_path=not args.no_host_setup)
        t_setup = setup['device_path']['total_s']
    else:
        t0 = time.perf_counter()
        plan = Plan(w.tables, w.states, split_bits=args.split_bits)
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t0
    log(f'plan: {plan.info} (setup {t_setup:.2f} s)')
    leg = matvec_leg(plan, dt, args.steps, args.warmup)

    traffic = None
    if args.traffic == 'live':
        try:
            traffic = live_traffic(args.workload, args.dtype)
            log(f'traffic: cached {traffic["cached"]["hbm_bytes_per_launch"] / 1e9:.3f} GB/launch, '
                f'matrix-free {traffic["matrix_free"]["hbm_bytes_per_launch"] / 1e9:.3f} GB/launch')
        except Exception as exc:   # the profiler is optional: report why the figure is missing
            log(f'traffic: not measured ({exc})')
            traffic = {'error': str(exc)}

    line = line_for(w, plan, leg, args.dtype, args.steps, args.warmup, traffic)
    result = {
        'metric': METRIC,
        'value': line['value'],
        'unit': 'GB/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': line['ms_per_step'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': args.dtype,
        'data': 'synthetic (seeded generators, rqutils_amd/workloads.py)',
        'config': dict(line['config'], setup_s=round(t_setup, 3)),
        'roofline': line['roofline'],
        'matrix_free': line['matrix_free'],
    }
    if setup is not None:
        result['setup'] = setup
    if isinstance(traffic, dict):
        result['traffic'] = traffic
    if not args.no_locg:
        result['locg'] = locg_legs(plan, oracle_estar(args.workload, w), maxiter=args.locg_maxiter,
                                   energy_tol=args.energy_tol, matrix_free=plan.n_states <= 4_000_000)
    plan.close()
    del plan
    torch.cuda.empty_cache()
    if not args.no_extras:
        extras = {}
        for name in ('jw1m', 'jw40-1m'):
            if name != args.workload:
                extras[name] = extra_workload(name, dt, args.steps, args.warmup, locg=not args.no_locg)
        extras['batch64_locg'] = extra_batch()
        extras['grad36q'] = extra_grad()
        result['extras'] = extras
    if not args.no_cpu:
        log(f'cpu baseline (budget {args.cpu_budget} s) ...')
        result['cpu_baseline'] = cpu_baseline(w, args.dtype, leg['useful_bytes'], leg['ms'], args.cpu_budget)
        log(f"cpu baseline {result['cpu_baseline']['value']} s/matvec")
    print(json.dumps(result), flush=True)


if __name__ == '__main__':
    main()
