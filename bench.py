"""Benchmark: SQD subspace matvec throughput on one MI355X (+ CPU baseline).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload jw1m|jw25m|rand8k]
                    [--dtype c64|f32|f64|c128] [--no-cpu] [--locg]

A step is one matvec ``y = H_S x`` over the whole subspace with the term
tables, the index and ``x`` resident in HBM.  ``value`` is the algorithmic
byte rate of the matvec (bytes per matvec from the HBM model divided by the
kernel time measured with HIP events on the launch stream).  One JSON line is
printed on stdout.
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
        return workloads.config1()
    if name == 'jw25m':
        return workloads.config4()
    if name == 'rand8k':
        return workloads.config0()
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
    """Oracle C port of the reference's strategy (1, 0), timed on a sample of groups."""
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
    args = ap.parse_args()

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

    for _ in range(args.warmup):
        plan.matvec(x, out=y)
    torch.cuda.synchronize()
    log(f'warmup done ({args.warmup} matvecs)')
    stream = torch.cuda.current_stream(dev)
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
    log(f'timed {args.steps} matvecs: {ms:.3f} ms/matvec (events), wall {wall * 1e3 / args.steps:.3f} ms')

    achieved = alg / (ms * 1e-3) / 1e9
    traffic = traffic_from_profiles(w.name, args.dtype)
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
            'n_states': plan.n_states,
            'n_groups': plan.info['n_groups'],
            'n_terms': plan.info['n_terms'],
            'hits_per_matvec': counts['hits'],
            'segment_probes_per_matvec': counts['probes'],
            'split_bits': plan.info['split_bits'],
            'algorithmic_bytes_per_matvec': alg,
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
    }
    if not args.no_cpu:
        log(f'cpu baseline (budget {args.cpu_budget} s) ...')
        result['cpu_baseline'] = cpu_baseline(w, alg, args.cpu_budget)
        log(f"cpu baseline {result['cpu_baseline']['value']:.3f} GB/s")
    print(json.dumps(result), flush=True)


if __name__ == '__main__':
    main()
