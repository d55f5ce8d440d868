"""Benchmark of the daily transmission step on one MI355X.

A step is one simulated day of the plain hybrid population (1M agents; layers h/s/w/c with betas
3.0/0.6/0.6/0.3, community resampled daily, 180 days). The whole run is captured as one graph; after
``--warmup`` untimed replays (at least one), one replay is timed, and inside it ``--steps K`` days
spread evenly over the run are timed in place by the device wall clock that every day's first pass
stamps (a day: from its first pass to the next day's). The whole replay's host-clock time is
reported beside them (``full_run``).

Prints one JSON line: agent-days/s over the K timed days, the edge-pass roofline (algorithmic bytes
per launch over the launch's duration, every day of the timed replay, against the 8 TB/s HBM3E peak;
the duration is the launch's span from its first workgroup's start to its last one's end by the
device wall clock; an eager run of the same days gives the HIP-event figure beside it),
the whole-day roofline (algorithmic bytes of the days over their time) and the CPU baselines (the
oracle's C restatement, one thread and OpenMP, timed on the same sampled days of the same population
on this host).
"""

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'agent-days/sec at 1M hybrid; batched vs plain sims/sec; HBM GB/s vs peak'
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    'hybrid1m': dict(pop_size=1_000_000, pop_type='hybrid', n_days=180, households='categorical',
                     pars=dict(beta=0.016, use_waning=False, pop_infected=20, dynam_layer=dict(c=1)),
                     desc='hybrid 1M agents, 180 days, households 1..7 (mean 2.6), s Poisson(20) ages 6-22, '
                          'w Poisson(16) ages 22-65, c Poisson(20) resampled daily; layer betas h3.0 s0.6 w0.6 '
                          'c0.3; beta 0.016; float32 state'),
    'hybrid10m_variants': dict(pop_size=10_000_000, pop_type='hybrid', n_days=120, households='categorical',
                               pars=dict(beta=0.016, use_waning=True, pop_infected=20, dynam_layer=dict(c=1)),
                               variants=[dict(rel_beta=1.7, days=[30], n_imports=100)],
                               desc='hybrid 10M agents, 120 days, waning NAbs, 2 variants (second rel_beta 1.7, '
                                    '100 imports on day 30); layers as hybrid1m'),
    'batch256': dict(n_sims=256, pop_lo=10_000, pop_hi=500_000, beta_lo=0.010, beta_hi=0.025, n_days=120,
                     pop_type='hybrid', households='categorical',
                     pars=dict(use_waning=False, pop_infected=20, dynam_layer=dict(c=1)),
                     desc='256 independent hybrid sims, pop sizes log-uniform over 10k-500k, beta uniform over '
                          '0.010-0.025, per-sim seeds, 120 days, layers as hybrid1m, one launch per phase per day'),
    'relax1m': dict(pop_size=1_000_000, pop_type='hybrid', n_days=60, households='categorical', relax=True,
                    pars=dict(beta=0.016, use_waning=False, pop_infected=20, dynam_layer=dict(c=1)),
                    desc='differentiable relaxation on the hybrid 1M population (h/s/w held, c resampled daily '
                         'and taken in expectation), 60 days: '
                         'float32 forward + hand-written adjoint with checkpoints every 10 days, gradient wrt '
                         '4 layer betas, beta and 10 age-bin rel_sus; calibration to a stochastic incidence curve'),
    'interv1m': dict(pop_size=1_000_000, pop_type='hybrid', n_days=180, households='categorical', interventions=True,
                     pars=dict(beta=0.016, use_waning=False, pop_infected=20, dynam_layer=dict(c=1)),
                     desc='hybrid1m with interventions: change_beta(30, 0.6) and clip_edges([40, 90], [0.5, 1.0], '
                          'layers s, w); clip days run through the individual passes, the days between them as '
                          'captured graphs; the change_beta-only run is one graph; a third run tests '
                          '(test_prob symp 0.1, asymp 0.001, delay 1) and traces contacts (h 1.0, s/w 0.5 after '
                          '1 day, c 0.1 after 2 days) every day, one graph'),
    'popgen10m': dict(pop_size=10_000_000, pop_type='hybrid', households='categorical', popgen=True,
                      pars=dict(),
                      desc='population build of the hybrid 10M population on the device: ages, age-correlated '
                           'households 1..7, s/w/c random layers, prognosis arrays (cvk_pop_*)'),
    'random20k': dict(pop_size=20_000, pop_type='random', n_days=60, households='poisson',
                      pars=dict(beta=0.016, use_waning=True, pop_infected=20),
                      desc='random 20k agents, one layer a with Poisson(20) contacts, 60 days, beta 0.016'),
}


def measured_traffic(workload, tuning):
    """HBM bytes per edge-pass launch from the committed counter summary (profiles/), or None.

    The summary holds FETCH_SIZE and WRITE_SIZE means over the edge-pass dispatches of this workload,
    already scaled by what the same counters read for a device copy of known size; it counts only
    if it was taken with the same tuning and on a library built from the same sources (cvk_source_hash).
    """
    path = os.path.join(ROOT, 'profiles', f'edge_pass_traffic_{workload}.txt')
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ':' in line and not line.startswith('#'):
            k, v = line.split(':', 1)
            vals[k.strip()] = v.strip()
    if vals.get('tuning', 'default') != (json.dumps(tuning, sort_keys=True) if tuning else 'default'):
        return None
    from covasim_amd import _lib
    if vals.get('source_hash') != f'{_lib.lib().cvk_source_hash():016x}':  # counted on another kernel build
        return None
    try:
        return float(vals['fetch_bytes_per_launch']) + float(vals['write_bytes_per_launch'])
    except (KeyError, ValueError):
        return None


def build_population(w, seed):
    from covasim_amd import parameters as cvpar
    from covasim_amd import population as cvpop
    pars = cvpar.make_pars(pop_size=w['pop_size'], pop_type=w['pop_type'])
    pars['prognoses'] = cvpar.get_prognoses()
    rng = np.random.default_rng(seed)
    pop = cvpop.make_randpop(pars, rng, households=w['households'])
    pop.update(cvpop.make_people_arrays(pop['age'], pars, rng))
    for lay in pop['contacts'].values():
        lay['beta'] = np.ones(len(lay['p1']), np.float32)
    return pop


def cpu_baseline_days(w, pop, seed, days, budget_s=90.0):
    """The oracle's compiled C restatement on the same population, timed on the same sampled days
    as the GPU.

    The run advances through the unsampled days untimed with the OpenMP build; each sampled day is
    run once by the single-threaded build (numba's default mode, numba_parallel='none') and, from
    the same state, once by the OpenMP build over the host's cores (numba's parallel modes). Both
    builds give identical results, so the state stays that of one run. Stops early, on a sampled-day
    boundary, once budget_s has elapsed."""
    from oracle import covasim_oracle as co
    from oracle.cport.port import PortSim, load
    opars = co.make_pars(pop_size=w['pop_size'], pop_type=w['pop_type'], n_days=w['n_days'], rand_seed=seed,
                         variants=list(w.get('variants', [])), **w['pars'])
    keys = ('symp_prob', 'severe_prob', 'crit_prob', 'death_prob', 'rel_sus', 'rel_trans')
    init = dict(age=pop['age'], **{k: pop[k] for k in keys})
    layers = {k: (pop['contacts'][k]['p1'], pop['contacts'][k]['p2'], pop['contacts'][k]['beta'])
              for k in opars['beta_layer']}
    sim = PortSim(opars, init, layers, seed, omp=True)
    one, omp = load(False), load(True)
    threads = omp.port_threads()
    want = set(days)
    t_one, t_omp = {}, {}
    t0 = time.perf_counter()
    while sim.t < sim.npts and time.perf_counter() - t0 < budget_s:
        t = sim.t
        if t not in want:
            sim.lib = omp
            sim.step()
            continue
        state = {k: v.copy() for k, v in sim.a.items()}
        lay = [[c.copy() for c in l] for l in sim.layers_np]
        sim.lib = one
        c0 = time.perf_counter()
        sim.step()
        t_one[t] = time.perf_counter() - c0
        for k, v in state.items():
            np.copyto(sim.a[k], v)
        for l, saved in zip(sim.layers_np, lay):
            for c, v in zip(l, saved):
                np.copyto(c, v)
        sim.t = t
        sim.lib = omp
        c0 = time.perf_counter()
        sim.step()
        t_omp[t] = time.perf_counter() - c0
    done = [d for d in days if d in t_one]
    n = w['pop_size']
    host = host_cpu()
    def line(times, cores, label):
        el = sum(times[d] for d in done)
        return dict(value=n * len(done) / el if el > 0 else None, unit='agent-days/s', cores=cores, kind='port',
                    sample=f'oracle C restatement (gcc -O3{label}, Philox draws) on the same population, the same '
                           f'{len(done)} sampled days as the GPU ({done[0] if done else "-"}..'
                           f'{done[-1] if done else "-"}), {el:.2f} s on {cores} host thread(s)', host=host)
    return line(t_one, 1, ''), line(t_omp, threads, ', OpenMP')


def host_cpu():
    """The host CPU's model and logical core count (as lscpu / nproc report them)."""
    model = ''
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    return {'model': model, 'logical_cores': os.cpu_count(),
            'usable_cores': len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else None}


def run_batch(w, args):
    """Batched many-sim workload: all sims per launch vs the same sims run one at a time."""
    import torch
    import covasim_amd as cv
    from covasim_amd import _lib
    rng = np.random.default_rng(args.seed)
    S = w['n_sims']
    sizes = np.exp(rng.uniform(np.log(w['pop_lo']), np.log(w['pop_hi']), S)).astype(np.int64)
    betas = rng.uniform(w['beta_lo'], w['beta_hi'], S)
    t0 = time.perf_counter()
    sims = []
    for i in range(S):
        pop = build_population(dict(w, pop_size=int(sizes[i])), args.seed * 100003 + i)
        sim = cv.Sim(pop_size=int(sizes[i]), pop_type=w['pop_type'], n_days=w['n_days'], rand_seed=args.seed + i,
                     beta=float(betas[i]), popdict=pop, households=w['households'], log_capacity=1024, **w['pars'])
        sim.initialize()
        sims.append(sim)
        if (i + 1) % 32 == 0:
            print(f'built {i + 1}/{S} sims ({time.perf_counter() - t0:.1f} s)', file=sys.stderr, flush=True)
    batch = cv.BatchSim(sims)
    setup_s = time.perf_counter() - t0
    n_agents = int(sizes.sum())
    n_edges = int(batch.edge_counts.sum())
    steps = w['n_days'] + 1
    tuned = batch.autotune(days=8)
    print(f'autotune: {batch.tuning}', file=sys.stderr, flush=True)
    stream = torch.cuda.Stream()
    L = _lib.lib()
    snap = batch.snapshot_state()
    with torch.cuda.stream(stream):
        gexec = batch.graph_create(stream)
        batch.restore_state(snap)  # warmup replay
        _lib.check(L.cvk_graph_launch(gexec, C.c_void_p(stream.cuda_stream)), 'batch graph launch')
        torch.cuda.synchronize()
        batch.restore_state(snap)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        _lib.check(L.cvk_graph_launch(gexec, C.c_void_p(stream.cuda_stream)), 'batch graph launch')
        torch.cuda.synchronize()
        wall_batch = time.perf_counter() - t1
        _lib.check(L.cvk_graph_destroy(gexec), 'graph destroy')
        res, _ = batch.raw_results()
        batch_inf = res[:, :, _lib.R_COL['new_infections']].sum(axis=1).astype(np.int64)
        # the same sims one at a time: one captured graph per sim, launched back to back
        execs = []
        for sim in sims:
            e = C.c_void_p()
            _lib.check(L.cvk_graph_create(sim.people.struct(), sim._layers, sim._params, sim._work, 0, sim.npts,
                                          int(sim.pars['pop_infected']), sim._imports,
                                          C.c_void_p(stream.cuda_stream), C.byref(e)), 'graph create')
            execs.append(e)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for e in execs:
            _lib.check(L.cvk_graph_launch(e, C.c_void_p(stream.cuda_stream)), 'graph launch')
        torch.cuda.synchronize()
        wall_single = time.perf_counter() - t2
        for e in execs:
            L.cvk_graph_destroy(e)
    single_inf = np.array([int(sim._w['results'][:, _lib.R_COL['new_infections']].sum().item()) for sim in sims])
    # edge-pass roofline: the same run eagerly, HIP events around each day's batched edge launch
    batch.restore_state(snap)
    _lib.check(L.cvk_batch_time_edges(batch._handle, 1), 'time edges')
    with torch.cuda.stream(stream):
        _lib.check(L.cvk_batch_run_days(batch._handle, 0, steps, batch.pop_infected.ctypes.data,
                                        C.byref(batch._imports), 0, C.c_void_p(stream.cuda_stream)), 'batch eager run')
    torch.cuda.synchronize()
    edge_ms = (C.c_float * steps)()
    got = L.cvk_batch_edge_times(batch._handle, edge_ms, steps)
    _lib.check(0 if got == steps else got, 'edge times')
    _lib.check(L.cvk_batch_time_edges(batch._handle, 0), 'time edges')
    edge_ms = np.array(edge_ms[:steps], dtype=np.float64)
    eager_inf = batch.raw_results()[0][:, :, _lib.R_COL['new_infections']].sum(axis=1).astype(np.int64)
    dyn_keys = [k for k, v in w['pars'].get('dynam_layer', {}).items() if v]
    dyn_total = int(sum(batch.edge_counts[batch.layer_keys.index(k)].sum() for k in dyn_keys))
    mode = 2 if args.generate_dynamic is None else args.generate_dynamic  # CVK_TUNE_GENERATE_DYNAMIC's rule
    gen = dyn_total > 0 and (mode == 1 or (mode == 2 and dyn_total >= 2 ** 25))
    edge_bytes = 8 * (n_edges - (dyn_total if gen else 0)) + n_agents / 8
    edge_achieved = edge_bytes / (float(edge_ms.mean()) / 1e3) / 1e9
    # results of all sims: device finalisation (sim.py:764-1001) and reduction over sims (run.py:220-303)
    batch.device_results()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    out_rows, out_vrows = batch.device_results()
    torch.cuda.synchronize()
    fin_ms = (time.perf_counter() - t3) * 1e3
    from covasim_amd import results as cvres
    t4 = time.perf_counter()
    red = [cvres.reduce_tables(out_rows), cvres.reduce_tables(out_vrows)]
    torch.cuda.synchronize()
    red_ms = (time.perf_counter() - t4) * 1e3
    del red
    results_line = {'finalize_ms': fin_ms, 'reduce_ms': red_ms, 'rows': int(out_rows.shape[1]),
                    'variant_rows': int(out_vrows.shape[1])}
    if not args.no_cpu_baseline:
        from oracle import results_ref as rr
        t5 = time.perf_counter()
        res_h, resf_h = batch.raw_results()
        dates = {k: batch._d[k].cpu().numpy() for k in ('date_infectious', 'date_recovered', 'date_dead')}
        host_rows = []
        for si, sim in enumerate(sims):
            a, b = int(batch.agent_offsets[si]), int(batch.agent_offsets[si + 1])
            dur = sim.pars['dur']
            R = rr.finalize(res_h[si], resf_h[si], _lib.R_COL, _lib.R_BYVAR, _lib.MAX_VARIANTS, batch.n_variants,
                            float(b - a), sim.pars['pop_infected'], sim.pars['use_waning'], sim.import_schedule(),
                            dates['date_infectious'][a:b], dates['date_recovered'][a:b], dates['date_dead'][a:b],
                            dur['exp2inf']['par1'] + dur['asym2rec']['par1'])
            host_rows.append(np.stack([np.asarray(R[k], dtype=float) for k in cvres.F_ROW_NAMES]))
        rr.reduce(np.stack(host_rows))
        results_line['cpu_finalize_reduce_ms'] = (time.perf_counter() - t5) * 1e3
        results_line['cpu_sample'] = ('numpy restatement on 1 host thread, including the copy of the result tables '
                                      'and the three date arrays to the host')
    value = n_agents * steps / wall_batch
    single = n_agents * steps / wall_single
    b_agent = 56
    day_bytes_total = steps * (n_agents * b_agent + edge_bytes + (0 if gen else 8 * dyn_total)) + 64 * int(batch_inf.sum())
    out = {
        'metric': METRIC, 'value': value, 'unit': 'agent-days/s', 'n_gpus': 1, 'steps': steps, 'warmup': steps,
        'ms_per_step': wall_batch / steps * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': w['desc'], 'sims': S, 'agents': n_agents, 'edges': n_edges, 'days': steps,
                   'setup_s': round(setup_s, 1)},
        'batched': {'sims_per_s': S / wall_batch, 'agent_days_per_s': value, 'wall_s': wall_batch,
                    'one_at_a_time_sims_per_s': S / wall_single, 'one_at_a_time_agent_days_per_s': single,
                    'one_at_a_time_wall_s': wall_single, 'speedup': wall_single / wall_batch,
                    'identical_infection_totals': bool(np.array_equal(batch_inf, single_inf) and
                                                       np.array_equal(batch_inf, eager_inf)),
                    'dynamic_layers': 'generated from the resample stream' if gen else 'resampled daily',
                    'autotuned_shape': list(tuned), 'autotune_trials_ms': batch.tuning['trials_ms']},
        'roofline': {'bound': 'hbm', 'kernel': 'batch_edge_kernel', 'achieved': edge_achieved, 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': edge_achieved / HBM_PEAK_GBS, 'traffic': None,
                     'avg_launch_us': float(edge_ms.mean()) * 1e3, 'launches': int(steps),
                     'algorithmic_bytes_per_launch': edge_bytes,
                     'launch_us_p10_p50_p90_max': [round(float(x) * 1e3, 1) for x in
                                                   np.percentile(edge_ms, [10, 50, 90, 100])],
                     'whole_day': {'bytes_per_agent': b_agent, 'achieved': day_bytes_total / wall_batch / 1e9,
                                   'frac': day_bytes_total / wall_batch / 1e9 / HBM_PEAK_GBS,
                                   'formula': 'per day N*56 + edge-pass bytes + 8*E_dyn written (0 generated) + '
                                              '64*new infections, over the batch graph wall time'}},
        'results': results_line,
    }
    if not args.no_cpu_baseline:
        out['cpu_baseline'], out['cpu_baseline_omp'] = cpu_baseline_sims(w, sims, sizes, betas, args, 30.0)
    print(json.dumps(out), flush=True)


def cpu_baseline_sims(w, sims, sizes, betas, args, budget_s):
    """The batch's sims one after another through the oracle's C restatement (as the reference's
    MultiSim runs independent sims), one thread and OpenMP, until budget_s each; agent-days/s."""
    from oracle.cport.port import PortSim, load
    from oracle import covasim_oracle as co
    out = []
    for omp in (False, True):
        t0 = time.perf_counter()
        agent_days, done = 0, 0
        for i, sim in enumerate(sims):
            if time.perf_counter() - t0 > budget_s:
                break
            people = sim.people
            opars = co.make_pars(pop_size=int(sizes[i]), pop_type=w['pop_type'], n_days=w['n_days'],
                                 rand_seed=args.seed + i, beta=float(betas[i]), **w['pars'])
            init = dict(age=people.age, **{k: people.get(k) for k in ('symp_prob', 'severe_prob', 'crit_prob',
                                                                       'death_prob', 'rel_sus', 'rel_trans')})
            layers = {k: tuple(people.contacts[k][c] for c in ('p1', 'p2', 'beta')) for k in opars['beta_layer']}
            PortSim(opars, init, layers, args.seed + i, omp=omp).run()
            agent_days += int(sizes[i]) * (w['n_days'] + 1)
            done += 1
        el = time.perf_counter() - t0
        cores = load(omp).port_threads()
        out.append(dict(value=agent_days / el, unit='agent-days/s', cores=cores, kind='port',
                        sample=f'oracle C restatement (gcc -O3{", OpenMP" if omp else ""}) running the first {done} '
                               f'of the batch\'s sims one after another, all {w["n_days"] + 1} days each, '
                               f'{el:.1f} s on {cores} host thread(s), set-up included',
                        host=host_cpu()))
    return out[0], out[1]


def run_relax(w, args):
    """Relaxation workload: forward + adjoint timing, float64 finite-difference check, calibration."""
    import torch
    import covasim_amd as cv
    from covasim_amd.relax import Relaxation, calibrate, log_incidence_loss
    t0 = time.perf_counter()
    pop = build_population(w, args.seed)

    def make_sim(seed):
        return cv.Sim(pop_size=w['pop_size'], pop_type=w['pop_type'], n_days=w['n_days'], rand_seed=seed,
                      popdict=pop, households=w['households'], **w['pars']).initialize()

    truth_sim = make_sim(args.seed)
    truth_sim.run()
    observed = truth_sim.results['new_infections'].values.copy()
    r32 = Relaxation(make_sim(args.seed), precision=32, checkpoint_every=10)
    r64 = Relaxation(r32.sim, precision=64, checkpoint_every=10)
    setup_s = time.perf_counter() - t0
    theta = r32.theta0()
    c = np.random.default_rng(args.seed).uniform(0.5, 1.5, r32.n_days) / 1e4
    for _ in range(2):  # warmup
        r32.forward_np(theta)
        r32.backward_np(theta, c)
    reps = 5
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        r32.forward_np(theta)
    fwd_s = (time.perf_counter() - t1) / reps
    stream = torch.cuda.current_stream()  # the stream the relaxation's kernels are launched on
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t2 = time.perf_counter()
    ev0.record(stream)
    for _ in range(reps):
        r32.forward_np(theta)
        g = r32.backward_np(theta, c)
    ev1.record(stream)
    torch.cuda.synchronize()
    fb_s = (time.perf_counter() - t2) / reps
    fb_ev_s = ev0.elapsed_time(ev1) / 1e3 / reps
    h = 1e-4
    fd = np.zeros_like(theta)
    for k in range(len(theta)):
        tp, tm = theta.copy(), theta.copy()
        tp[k] *= 1 + h
        tm[k] *= 1 - h
        fd[k] = (c @ r64.forward_np(tp) - c @ r64.forward_np(tm)) / (2 * h * theta[k])
    grad_err = float(np.linalg.norm(g - fd) / np.linalg.norm(fd))
    start = theta * np.random.default_rng(args.seed + 1).uniform(0.6, 1.5, len(theta))
    t3 = time.perf_counter()
    fitted, hist = calibrate(r32, observed, start, iters=100, lr=0.05)
    cal_s = time.perf_counter() - t3
    obs_t = torch.as_tensor(observed[:r32.n_days])
    loss_truth = float(log_incidence_loss(torch.as_tensor(r32.forward_np(theta)), obs_t))
    n, T = w['pop_size'], r32.n_days
    # algorithmic HBM bytes of one forward + adjoint (every kernel; the neighbour gathers of tr, lbar and
    # age bin are served from the caches and not counted): per day t a forward (relax_trans_kernel: rel_trans,
    # bin, tr and min(t, kmax) past deltas; relax_forward_kernel: indptr (int64) per layer, bin, s read and
    # written, delta, 4 per neighbour slot), the adjoint's recomputed forward (plus L and s recorded), and
    # the adjoint (relax_adj_agent_kernel: bin, L, s, sbar r/w, lbar, min(kmax, T-1-t) ibar rows;
    # relax_pack_kernel: tr, lbar, bin, the 16-byte neighbour record written; relax_adj_edge_kernel:
    # indptr per layer, bin, tr, lbar, rel_trans, ibar written, 4 per slot)
    slots = int(sum(lay[1].numel() for lay in r32._csr if lay is not None))
    n_dyn = sum(1 for c_l in r32.dyn_slots if c_l)
    L_ = r32.n_layers
    kmax = min(r32.w.shape[1] - 1, T)
    L_ = L_ - n_dyn  # the resampled layers have no lists: their expected term reads tr once (4 B per agent)
    fwd = sum(n * (9 + 4 * min(t, kmax)) + n * (8 * L_ + 13) + 4 * n * (n_dyn > 0) + 4 * slots for t in range(T))
    rec = 8 * n * T
    adj = sum(n * (21 + 4 * min(kmax, T - 1 - t)) + n * 25 + n * (8 * L_ + 17) + 9 * n * (n_dyn > 0) + 4 * slots
              for t in range(T))
    fb_bytes = 2 * fwd + rec + adj
    fb_gbs = fb_bytes / fb_ev_s / 1e9
    out = {
        'metric': METRIC, 'value': n * T / fb_s, 'unit': 'agent-days/s', 'n_gpus': 1, 'steps': T, 'warmup': 2 * T,
        'ms_per_step': fb_s / T * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': w['desc'], 'agents': n, 'days': T, 'params': int(len(theta)),
                   'setup_s': round(setup_s, 1)},
        'roofline': {'bound': 'hbm', 'kernel': 'relaxation forward + adjoint (every kernel, 60 days)',
                     'achieved': fb_gbs, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': fb_gbs / HBM_PEAK_GBS,
                     'traffic': None, 'algorithmic_bytes_per_step': fb_bytes, 'step_ms_hip_events': fb_ev_s * 1e3,
                     'neighbour_slots': slots, 'note': 'the neighbour gathers (one 16-byte record per slot in the '
                     'adjoint, one tr value in the forward) are served from L2 / MALL and bound the two edge '
                     'kernels; they are not in the bytes'},
        'relaxation': {'forward_ms': fwd_s * 1e3, 'forward_adjoint_ms': fb_s * 1e3,
                       'grad_rel_err_vs_fp64_central_fd': grad_err, 'fd_step_relative': h,
                       'calibration_iters': len(hist), 'calibration_s': cal_s,
                       'loss_start': hist[0], 'loss_end': hist[-1], 'loss_at_true_params': loss_truth,
                       'fitted_over_true': [round(float(x), 4) for x in fitted / theta]},
    }
    print(json.dumps(out), flush=True)


def run_popgen(w, args):
    """One step = one build of the whole population on the device (ages, households, layers,
    per-agent arrays), host synchronisations for the layer sizes included."""
    import torch
    from covasim_amd import parameters as cvpar
    from covasim_amd import popgen as pg
    n = w['pop_size']
    pars = cvpar.make_pars(pop_size=n, pop_type=w['pop_type'], rand_seed=args.seed)
    warm_pars = cvpar.make_pars(pop_size=10_000, pop_type=w['pop_type'], rand_seed=args.seed)
    pg.make_population(warm_pars, households=w['households'])
    steps = args.steps or 3
    warmup = 1 if args.warmup is None else args.warmup
    for _ in range(warmup):
        pop = pg.make_population(pars, households=w['households'])
        del pop
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pop = pg.make_population(pars, households=w['households'])
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    n_edges = sum(int(v['p1'].numel()) for v in pop['contacts'].values())
    out_bytes = 12 * n_edges + 29 * n  # p1, p2, beta per edge; age, sex and six float32 arrays per agent
    achieved = out_bytes / wall / 1e9
    line = {
        'metric': 'population build throughput (agents/s)', 'value': n / wall, 'unit': 'agents/s', 'n_gpus': 1,
        'steps': steps, 'warmup': warmup, 'ms_per_step': wall * 1e3, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64 draws, int32 edges', 'data': 'synthetic',
        'config': {'workload': w['desc'], 'agents': n, 'edges': n_edges,
                   'layers': {k: int(v['p1'].numel()) for k, v in pop['contacts'].items()}},
        'roofline': {'bound': 'hbm', 'kernel': 'whole build (all cvk_pop_* launches and host synchronisations)',
                     'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                     'traffic': None, 'output_bytes': out_bytes},
    }
    del pop
    if not args.no_cpu_baseline:
        from oracle.cport import popgen as op
        sample = max(10_000, int(n // 10))
        spars = cvpar.make_pars(pop_size=sample, pop_type=w['pop_type'], rand_seed=args.seed)
        t1 = time.perf_counter()
        op.make_population(pg.pop_tables(spars, households=w['households']), pg.layer_plan(spars), sample)
        el = time.perf_counter() - t1
        line['cpu_baseline'] = dict(value=sample / el, unit='agents/s', cores=1, kind='port',
                                    sample=f'oracle C restatement (gcc -O3) building the same kind of population '
                                           f'with {sample} agents, {el:.1f} s on 1 host thread')
    print(json.dumps(line))


def device_copy_gbs(mib=1024, reps=5):
    """Measured device-to-device copy bandwidth (bytes read + written per second, best of reps), for
    context beside the 8 TB/s peak."""
    import torch
    src = torch.ones((mib << 20) // 4, dtype=torch.float32, device='cuda')
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = None
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    del src, dst
    return round(2 * (mib << 20) / (best / 1e3) / 1e9, 1)


def run_interventions(w, args):
    """hybrid1m with interventions: a change_beta-only run (one captured graph, per-day rates), a
    run that also clips edges (clip days through the individual passes, graphs between them) and a
    testing + contact-tracing run (one captured graph). One
    step = one day; value = the clip run's agent-days/s, wall time of Sim.run without finalisation."""
    import torch
    import covasim_amd as cv
    pop = build_population(w, args.seed)
    n = w['pop_size']

    def copy():
        out = {k: v for k, v in pop.items() if k != 'contacts'}
        out['contacts'] = {k: {c: a.copy() for c, a in lay.items()} for k, lay in pop['contacts'].items()}
        return out

    def one(ivs):
        sim = cv.Sim(pop_size=n, pop_type=w['pop_type'], n_days=w['n_days'], rand_seed=args.seed, popdict=copy(),
                     households=w['households'], interventions=ivs, **w['pars'])
        sim.initialize()
        torch.cuda.synchronize()
        sim.run()
        return sim
    one([cv.change_beta(30, 0.6)])  # warm-up
    beta_only = one([cv.change_beta(30, 0.6)])
    clipped = one([cv.change_beta(30, 0.6), cv.clip_edges([40, 90], [0.5, 1.0], layers=['s', 'w'])])
    days = clipped.npts
    # testing and contact tracing every day, one captured graph
    traced = one([cv.test_prob(symp_prob=0.1, asymp_prob=0.001, test_delay=1),
                  cv.contact_tracing(trace_probs=dict(h=1.0, s=0.5, w=0.5, c=0.1), trace_time=dict(h=0, s=1, w=1, c=2))])
    # device-state snapshot mid-run (checkpoint.py): save at day 89, load, finish
    import tempfile
    sim = cv.Sim(pop_size=n, pop_type=w['pop_type'], n_days=w['n_days'], rand_seed=args.seed, popdict=copy(),
                 households=w['households'],
                 interventions=[cv.change_beta(30, 0.6), cv.clip_edges([40, 90], [0.5, 1.0], layers=['s', 'w'])],
                 **w['pars'])
    sim.run(until=89)
    with tempfile.TemporaryDirectory() as tmp:
        fn = os.path.join(tmp, 'mid.sim')
        t0 = time.perf_counter()
        sim.save(fn)
        save_s = time.perf_counter() - t0
        size = os.path.getsize(fn)
        t0 = time.perf_counter()
        resumed = cv.load(fn)
        load_s = time.perf_counter() - t0
    resumed.run()
    snapshot = {'save_s': save_s, 'load_s': load_s, 'bytes': size,
                'resumed_equals_uninterrupted': bool(np.array_equal(resumed.results['new_infections'].values,
                                                                    clipped.results['new_infections'].values))}
    line = {
        'metric': 'agent-days/sec at 1M hybrid with interventions', 'value': n * days / clipped.elapsed,
        'unit': 'agent-days/s', 'n_gpus': 1, 'steps': days, 'warmup': days, 'ms_per_step': clipped.elapsed / days * 1e3,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': w['desc'], 'agents': n, 'days': days},
        'change_beta_only': {'ms_per_day': beta_only.elapsed / days * 1e3,
                             'agent_days_per_s': n * days / beta_only.elapsed,
                             'cum_infections': float(beta_only.results['cum_infections'].values[-1])},
        'clip_run': {'cum_infections': float(clipped.results['cum_infections'].values[-1])},
        'testing_tracing': {'ms_per_day': traced.elapsed / days * 1e3, 'agent_days_per_s': n * days / traced.elapsed,
                            **{k: float(traced.results[k].values[-1]) for k in
                               ('cum_infections', 'cum_tests', 'cum_diagnoses', 'cum_quarantined')}},
        'snapshot_day_89': snapshot,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=None,
                    help='timed days, spread evenly over the run (default 20)')
    ap.add_argument('--warmup', type=int, default=None)
    ap.add_argument('--workload', default='hybrid1m', choices=sorted(WORKLOADS))
    ap.add_argument('--seed', type=int, default=1)
    ap.add_argument('--cpu-budget', type=float, default=90.0,
                    help='seconds of CPU restatement work at most (stops on a sampled-day boundary)')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--edge-chunks', type=int, default=None, help='edge-pass wave tile depth (2 or 4 chunks of 4 edges per lane)')
    ap.add_argument('--edge-lds', type=int, default=None, help='1: stage the infectious bitmap in LDS, 0: read it from L2')
    ap.add_argument('--edge-prefetch', type=int, default=None,
                    help='1: issue the next wave tile\'s endpoint loads before this tile\'s candidates')
    ap.add_argument('--edge-coarse', type=int, default=None,
                    help='1: with the bitmap in L2, filter its reads by a coarse map in LDS; 0: no filter')
    ap.add_argument('--overlap-resample', type=int, default=None,
                    help='1: the next day\'s resample runs on a side branch beside the infect pass; 0: in the agent pass')
    ap.add_argument('--generate-dynamic', type=int, default=None,
                    help='1: the day loop computes dynamic-layer endpoints in place; 0: writes and reads them daily; '
                         '2 (default): 1 from 2^25 dynamic edges up')
    ap.add_argument('--source-lists', type=int, default=None,
                    help='static layers walked through source lists: 0 never, 1 always, 2 (default) on quiet days')
    ap.add_argument('--source-permille', type=int, default=None,
                    help='with --source-lists 2: most infectious agents per thousand for which lists are walked')
    ap.add_argument('--target-bits', type=int, default=None,
                    help='1 (default): with the bitmap out of LDS, mixed edges read target bits instead of target '
                         'codes; 0: codes')
    ap.add_argument('--eager-only', action='store_true',
                    help='counter-collection mode: the whole run once through the uncaptured day loop (the graph\'s '
                         'kernels, launch by launch; no CPU baseline)')
    ap.add_argument('--copy-probe', type=int, default=0,
                    help='MiB of one device-to-device copy after the run (a known-size kernel for counter scaling)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({'error': 'this build runs on one GPU; --gpus above 1 is not supported'}))
        sys.exit(2)

    import torch
    import covasim_amd as cv
    from covasim_amd import _lib

    w = WORKLOADS[args.workload]
    tuning = {}
    if args.overlap_resample is not None:
        _lib.check(_lib.lib().cvk_set_tuning(_lib.TUNE_OVERLAP_RESAMPLE, args.overlap_resample), 'tuning')
        tuning['overlap_resample'] = args.overlap_resample
    if args.generate_dynamic is not None:
        _lib.check(_lib.lib().cvk_set_tuning(_lib.TUNE_GENERATE_DYNAMIC, args.generate_dynamic), 'tuning')
        tuning['generate_dynamic'] = args.generate_dynamic
    if args.edge_chunks is not None:
        _lib.check(_lib.lib().cvk_set_tuning(_lib.TUNE_EDGE_CHUNKS, args.edge_chunks), 'tuning')
        tuning['edge_chunks'] = args.edge_chunks
    if args.edge_prefetch is not None:
        _lib.check(_lib.lib().cvk_set_tuning(_lib.TUNE_EDGE_PREFETCH, args.edge_prefetch), 'tuning')
        tuning['edge_prefetch'] = args.edge_prefetch
    if args.edge_coarse is not None:
        _lib.check(_lib.lib().cvk_set_tuning(_lib.TUNE_EDGE_COARSE, args.edge_coarse), 'tuning')
        tuning['edge_coarse'] = args.edge_coarse
    if args.source_lists is not None:
        _lib.check(_lib.lib().cvk_set_tuning(_lib.TUNE_SOURCE_LISTS, args.source_lists), 'tuning')
        tuning['source_lists'] = args.source_lists
    if args.source_permille is not None:
        _lib.check(_lib.lib().cvk_set_tuning(_lib.TUNE_SOURCE_PERMILLE, args.source_permille), 'tuning')
        tuning['source_permille'] = args.source_permille
    if args.edge_lds is not None:
        _lib.check(_lib.lib().cvk_set_tuning(_lib.TUNE_EDGE_LDS, args.edge_lds), 'tuning')
        tuning['edge_lds'] = args.edge_lds
    if args.target_bits is not None:
        _lib.check(_lib.lib().cvk_set_tuning(_lib.TUNE_TARGET_BITS, args.target_bits), 'tuning')
        tuning['target_bits'] = args.target_bits
    if 'n_sims' in w:
        return run_batch(w, args)
    if w.get('popgen'):
        return run_popgen(w, args)
    if w.get('interventions'):
        return run_interventions(w, args)
    if w.get('relax'):
        return run_relax(w, args)
    return run_single(w, args, tuning)


def sample_days(npts, k):
    """k days spread evenly over a run of npts days (the middle of k equal slices)."""
    return [min(npts - 1, (2 * i + 1) * npts // (2 * k)) for i in range(k)]


def day_bytes(n, b_agent, edge_read, resample_write, new_inf):
    """Algorithmic HBM bytes of one simulated day: the agent pass (b_agent per agent), the edge pass
    (edge_read: 8 B per stored edge read, 12 B where the layer carries weights other than 1, none for
    a generated dynamic layer; plus the N/8-byte bitmap), the dynamic layer's resample writes (8 B
    per written edge) and 64 B per new infection."""
    return n * b_agent + edge_read + n / 8 + resample_write + 64 * new_inf


def run_single(w, args, tuning):
    """One sim: K days spread evenly over the whole run, timed in place inside one replay of the run's
    captured graph by the device wall clock, beside that replay's host-clock time."""
    import torch
    import covasim_amd as cv
    from covasim_amd import _lib
    n_days = w['n_days']
    npts = n_days + 1
    K = args.steps if args.steps else 20
    W = 5 if args.warmup is None else args.warmup
    t_build = time.perf_counter()
    print(f'building population ({args.workload})', file=sys.stderr, flush=True)
    pop = build_population(w, args.seed)
    n_edges = sum(len(l['p1']) for l in pop['contacts'].values())
    dyn_keys = [k for k, v in w['pars'].get('dynam_layer', {}).items() if v]
    dyn_edges = sum(len(pop['contacts'][k]['p1']) for k in dyn_keys)
    variants = [cv.variant(dict(rel_beta=v['rel_beta']), days=v['days'], n_imports=v['n_imports'])
                for v in w.get('variants', [])]
    sim = cv.Sim(pop_size=w['pop_size'], pop_type=w['pop_type'], n_days=n_days, rand_seed=args.seed,
                 popdict=pop, households=w['households'], variants=variants, **w['pars'])
    sim.initialize()
    build_s = time.perf_counter() - t_build
    L = _lib.lib()
    stream = torch.cuda.Stream()
    st = C.c_void_p(stream.cuda_stream)
    people = sim.people
    P = people.struct()
    pr, wk = sim._params, sim._work
    pop_inf = int(w['pars']['pop_infected'])
    init_state = {k: v.clone() for k, v in people._d.items()}
    init_layers = {lk: {c: lay.tensor(c).clone() for c in ('p1', 'p2', 'beta')} for lk, lay in people.contacts.items()}
    plan = (C.c_int32 * 2)()
    _lib.check(L.cvk_day_plan(sim._layers, pr, sim._imports, npts, plan), 'day plan')
    gen = bool(plan[0])

    def reset():
        for k, v in init_state.items():
            people._d[k].copy_(v)
        for lk, lay in people.contacts.items():
            for c in ('p1', 'p2', 'beta'):
                lay.tensor(c).copy_(init_layers[lk][c])
        _lib.reset_workspace(sim._w)

    def counts():
        res = sim._w['results'].cpu().numpy().view(np.uint32)
        return res.copy()

    def resample(t):
        for li in range(pr.n_layers):
            if sim._layers[li].dynamic:
                _lib.check(L.cvk_resample_layer(C.byref(sim._layers[li]), li, pr, t, st), 'resample')

    def eager_day(t, events=None):
        """Day t through the individual entry points; HIP events around the edge pass when events is a
        list. Day t+1's contact resample runs right after day t's infect pass, where the captured day loop
        runs it (beside that pass), so each edge pass meets the caches as it does in a graph run."""
        if t == 0:
            _lib.check(L.cvk_choose(P, pr, wk, 0, _lib.STREAM_CHOOSE, pop_inf, 0, _lib.LAYER_SEED, 0, st), 'seed')
            _lib.check(L.cvk_infect_pass(P, sim._layers, pr, wk, 0, 0, st), 'seed')
            resample(0)
        _lib.check(L.cvk_agent_pass(P, pr, wk, t, st), 'agent')
        days_i, vars_i, nimp_i = sim._imp_arrays
        for j in range(len(days_i)):
            if days_i[j] == t:
                _lib.check(L.cvk_choose(P, pr, wk, t, _lib.STREAM_IMPORT | int(vars_i[j]), int(nimp_i[j]), 1,
                                        _lib.LAYER_IMPORT, int(vars_i[j]), st), 'imports')
        if events is not None:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
        _lib.check(L.cvk_edge_pass(P, sim._layers, pr, wk, t, None, st), 'edge')
        if events is not None:
            e1.record(stream)
            events.append((e0, e1))
        _lib.check(L.cvk_infect_pass(P, sim._layers, pr, wk, t, 1, st), 'infect')
        if t + 1 < npts:
            resample(t + 1)
        _lib.check(L.cvk_count_pass(P, pr, wk, t, st), 'count')

    def graph(t0, t1):
        g = C.c_void_p()
        _lib.check(L.cvk_graph_create(P, sim._layers, pr, wk, t0, t1, pop_inf, sim._imports, st, C.byref(g)),
                   'graph create')
        return g

    days = sample_days(npts, K)
    clock = sim._w['day_clock']
    hz = C.c_int64(0)
    _lib.check(L.cvk_wall_clock_rate(C.byref(hz)), 'wall clock rate')
    tick_s = 1.0 / hz.value
    if args.eager_only:  # counter collection: the whole run once through the day loop, launch by launch
        # (cvk_run_days without capture: the same kernels and tunables as the captured graph, every launch in
        # the trace)
        with torch.cuda.stream(stream):
            reset()
            _lib.check(L.cvk_run_days(P, sim._layers, pr, wk, 0, npts, pop_inf, sim._imports, 0, st), 'run days')
            torch.cuda.synchronize()
        line = {'mode': 'the day loop, uncaptured (counter collection)', 'days': npts,
                'source_hash': f'{L.cvk_source_hash():016x}',
                'cum_infections': int(counts()[:, _lib.R_COL['new_infections']].sum()) + pop_inf}
        if args.copy_probe:
            nbytes = args.copy_probe << 20
            src = torch.ones(nbytes // 4, dtype=torch.float32, device='cuda')
            dst = torch.empty_like(src)
            torch.cuda.synchronize()
            dst.copy_(src)
            torch.cuda.synchronize()
            line['copy_probe_bytes'] = nbytes
        print(json.dumps(line), flush=True)
        return
    with torch.cuda.stream(stream):
        # (1) the whole run as one captured graph: warmup replays, then one timed replay; every day's
        # first pass stamps the device wall clock, so each day's duration inside the run is known
        full = graph(0, npts)
        for _ in range(max(1, -(-W // npts))):
            reset()
            _lib.check(L.cvk_graph_launch(full, st), 'graph launch')
        reset()
        clock.zero_()
        clock[npts + 1::2] = -1  # edge-pass spans: earliest workgroup start ~0, latest end 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _lib.check(L.cvk_graph_launch(full, st), 'graph launch')
        torch.cuda.synchronize()
        full_wall = time.perf_counter() - t0
        _lib.check(L.cvk_graph_destroy(full), 'graph destroy')
        res_full = counts()
        ticks = clock.cpu().numpy().astype(np.int64)
        day_s = np.diff(ticks[:npts + 1]) * tick_s  # day t: from its first pass to the next day's (or the closing pass)
        # each day's edge pass inside the same replay: its first workgroup's start to its last one's end
        edge_ms = (ticks[npts + 2::2] - ticks[npts + 1::2]) * tick_s * 1e3
        print(f'full run: {npts} days in {full_wall * 1e3:.1f} ms', file=sys.stderr, flush=True)
        # (2) every day eagerly: HIP events around each edge pass (the roofline kernel), and the
        # source-list counters read after each day
        reset()
        clock[npts + 1::2] = -1
        clock[npts + 2::2] = 0
        _lib.check(L.cvk_set_tuning(_lib.TUNE_COUNT_SOURCES, 1), 'tuning')
        edge_ev = []
        ctl = sim._w['ctl']
        src_counts = np.zeros((npts, 3), dtype=np.int64)
        for t in range(npts):
            eager_day(t, edge_ev)
            src_counts[t] = ctl[_lib.CTL_SRC_SLOTS:_lib.CTL_SRC_SLOTS + 6].cpu().numpy().view(np.int64)
        _lib.check(L.cvk_set_tuning(_lib.TUNE_COUNT_SOURCES, 0), 'tuning')
        torch.cuda.synchronize()
        edge_ev_ms = np.array([a.elapsed_time(b) for a, b in edge_ev])
        ticks_eager = clock.cpu().numpy().astype(np.int64)
        edge_eager_ms = (ticks_eager[npts + 2::2] - ticks_eager[npts + 1::2]) * tick_s * 1e3
        res_eager = counts()
    # per day: source-list slots walked, lists walked, 1 if the listed layers were streamed instead
    src_day = np.diff(np.vstack([np.zeros((1, 3), np.int64), src_counts]), axis=0)
    step_s = np.array([day_s[d] for d in days])
    n = w['pop_size']
    new_inf = res_full[:, _lib.R_COL['new_infections']].astype(np.int64)
    cum_inf = int(new_inf.sum()) + pop_inf
    timed_s = float(step_s.sum())
    value = n * K / timed_s
    # roofline of the dominant kernel: the edge pass over every day of the run. Algorithmic bytes per
    # launch: 8 per streamed edge (12 where the layer carries weights), none for a generated layer, 8 per
    # source-list slot walked (target + edge) and per list (two offsets), and the N/8-byte bitmap. The
    # layers with source lists count as streamed on the days the walk decision streams them.
    listed = set(sim._sources) if tuning.get('source_lists', 2) else set()

    def layer_bytes(keys):
        return sum((12 if not np.all(pop['contacts'][k]['beta'] == 1) else 8) * len(pop['contacts'][k]['p1'])
                   for k in keys)
    streamed = [k for k in pop['contacts'] if k not in listed]
    stream_bytes = layer_bytes(streamed) - (8 * dyn_edges if gen else 0)
    edge_bytes_day = (stream_bytes + n / 8 + 8 * src_day[:, 0] + 8 * src_day[:, 1]
                      + layer_bytes(listed) * (src_day[:, 2] > 0))
    achieved = float(edge_bytes_day.sum()) / (float(edge_ms.sum()) / 1e3) / 1e9
    nv = 1 + len(w.get('variants', []))
    b_agent = 56 + 8 * (nv - 1) + ((20 + 12 * nv) if w['pars'].get('use_waning') else 0)
    bday = np.array([n * b_agent + edge_bytes_day[t] + (0 if gen else 8 * dyn_edges) + 64 * int(new_inf[t])
                     for t in range(npts)])
    whole_sampled = float(sum(bday[d] for d in days)) / timed_s / 1e9
    whole_full = float(bday.sum()) / full_wall / 1e9
    out = {
        'metric': METRIC, 'value': value, 'unit': 'agent-days/s', 'n_gpus': 1, 'steps': K, 'warmup': W,
        'ms_per_step': timed_s / K * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': w['desc'], 'agents': n, 'days': npts, 'edges': int(n_edges),
                   'dynamic_edges': int(dyn_edges), 'source_listed_layers': sorted(listed),
                   'dynamic_layers': 'generated from the resample stream' if gen else 'resampled into the layer arrays daily',
                   'step': 'one simulated day; the K timed days are spread evenly over the whole run and timed in '
                           'place inside one replay of the whole run as a captured graph (host synchronisation '
                           'before and after it), each from its first pass to the next day\'s, by the device '
                           'wall clock that the day loop stamps',
                   'timed_days': days, 'cum_infections': cum_inf,
                   'runs_identical': bool(np.array_equal(res_full, res_eager)),
                   'setup_s': round(build_s, 2), **({'tuning': tuning} if tuning else {})},
        'full_run': {'days': npts, 'wall_ms': full_wall * 1e3, 'ms_per_day': full_wall / npts * 1e3,
                     'agent_days_per_s': n * npts / full_wall, 'device_clock_ms': float(day_s.sum()) * 1e3,
                     'note': 'all days as one captured graph, one launch, host clock with synchronisation'},
        'roofline': {'bound': 'hbm', 'kernel': 'edge_pass_kernel', 'achieved': achieved, 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                     'traffic': measured_traffic(args.workload, tuning),
                     'avg_launch_us': float(edge_ms.mean()) * 1e3, 'launches': int(len(edge_ms)),
                     'algorithmic_bytes_per_launch': float(edge_bytes_day.mean()),
                     'launch_us_p10_p50_p90_max': [round(float(x) * 1e3, 1) for x in
                                                   np.percentile(edge_ms, [10, 50, 90, 100])],
                     'timed_days_avg_launch_us': float(np.mean([edge_ms[d] for d in days])) * 1e3,
                     'launch_timing': 'each day\'s launch inside the timed replay of the whole run, from its '
                                      'first workgroup start to its last workgroup end by the device wall clock',
                     'eager_run_avg_launch_us': float(edge_eager_ms.mean()) * 1e3,
                     'eager_run_avg_launch_us_hip_events': float(edge_ev_ms.mean()) * 1e3,
                     'listed_layers_streamed_days': int((src_day[:, 2] > 0).sum()),
                     'source_slots_per_day_mean': float(src_day[:, 0].mean()),
                     'whole_day': {'bytes_per_agent': b_agent, 'achieved_timed_days': whole_sampled,
                                   'frac_timed_days': whole_sampled / HBM_PEAK_GBS,
                                   'achieved_full_run': whole_full, 'frac_full_run': whole_full / HBM_PEAK_GBS,
                                   'bytes_full_run': float(bday.sum()),
                                   'formula': 'B_day = N*b_agent + edge-pass bytes + 8*E_dyn written + '
                                              '64*new infections'}},
    }
    print('edge pass per-day us: ' + ' '.join(f'{x * 1e3:.0f}' for x in edge_ms), file=sys.stderr, flush=True)
    print('infectious per thousand: ' + ' '.join(f'{x * 1e3 / n:.1f}' for x in
                                                  res_full[:, _lib.R_COL['n_infectious']]), file=sys.stderr, flush=True)
    print('source slots per day: ' + ' '.join(str(int(x)) for x in src_day[:, 0]), file=sys.stderr, flush=True)
    print('day us (device clock): ' + ' '.join(f'{x * 1e6:.0f}' for x in day_s), file=sys.stderr, flush=True)
    print('timed day us: ' + ' '.join(f'{d}:{x * 1e6:.0f}' for d, x in zip(days, step_s)), file=sys.stderr,
          flush=True)
    if args.copy_probe:
        nbytes = args.copy_probe << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        out['copy_probe_bytes'] = nbytes
    out['roofline']['device_copy_gbs'] = device_copy_gbs()
    if not args.no_cpu_baseline:
        one, omp = cpu_baseline_days(w, pop, args.seed, days, args.cpu_budget)
        out['cpu_baseline'] = one
        out['cpu_baseline_omp'] = omp
    print(json.dumps(out), flush=True)

if __name__ == '__main__':
    main()
