"""Whole runs at full workload sizes, GPU against the oracle's compiled C restatement.

The small-size parity tests (test_gpu_parity.py) cover every code path with forced tunables. These
tests run the production defaults at the sizes the throughput figures are quoted on, where the
launch shapes, bitmap placements (LDS bitmap up to ~1.06M agents, coarse map + L2 bitmap above)
and dynamic-layer handling (written below 2^25 dynamic edges, generated from there up) are the
ones a user gets, and where ~10^6 infections exercise the device's double-precision exp / log1p /
cos against the host C library's in every duration and NAb draw:

- hybrid 1M agents, 180 days, households 1..7, community layer resampled daily (sim.py:558-685);
- hybrid 2M agents, 40 days, beta 0.05, waning, two variants: busy days with the coarse map;
- hybrid 10M agents, two variants (rel_beta 1.7, 100 imports on day 30), waning NAbs, 60 days:
  coarse map, L2 bitmap and generated dynamic contacts (immunity.py:117-350);
- 256 hybrid sims of 10k-500k agents (log-uniform), beta 0.010-0.025, 120 days, one batch: every
  sim equals the same sim run alone, and the smallest, median and largest equal the restatement.

Every per-day count, every agent array (byte for byte, NaN included) and the contact layers after
the run must be identical. The restatement runs with OpenMP over the host's cores.
"""

import ctypes as C
import time

import numpy as np
import pytest

from tests.helpers import (COMPARE_KEYS, assert_matches_port, copy_pop, make_population, port_sim, port_view)

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


def log(msg):
    print(f'[full-size] {msg}', flush=True)


def sim_results(sim):
    from oracle.cport.port import ROW_KEYS
    out = {k: sim.results[k].values for k in ROW_KEYS if k != 'n_susceptible'}
    out['n_susceptible'] = sim.raw['n_susceptible_agents']
    return out


def sim_layers(sim):
    return [tuple(sim.people.contacts[lk].tensor(c).cpu().numpy() for c in ('p1', 'p2', 'beta'))
            for lk in sim.layer_keys()]


def run_gpu_and_port(n, n_days, seed, pars, variants=(), pop_seed=1):
    import covasim_amd as cv
    t0 = time.perf_counter()
    pop = make_population(n=n, seed=pop_seed, pop_type='hybrid', households='categorical')
    log(f'n={n} population built in {time.perf_counter() - t0:.1f} s '
        f'({sum(len(l["p1"]) for l in pop["contacts"].values())} edges)')
    gvars = [cv.variant(dict(rel_beta=v['rel_beta']), days=v['days'], n_imports=v['n_imports']) for v in variants]
    sim = cv.Sim(pop_size=n, pop_type='hybrid', n_days=n_days, rand_seed=seed, popdict=copy_pop(pop),
                 households='categorical', variants=gvars, **pars)
    t0 = time.perf_counter()
    sim.run()
    log(f'n={n} GPU run of {n_days} days: {time.perf_counter() - t0:.1f} s, '
        f'cum_infections {sim.results["cum_infections"].values[-1]:.0f}')
    port = port_sim(pop, seed=seed, n_days=n_days, pop_type='hybrid', variants=list(variants), **pars)
    del pop
    t0 = time.perf_counter()
    port.run()
    log(f'n={n} C restatement: {time.perf_counter() - t0:.1f} s')
    return sim, port


def test_hybrid_1m_180_days_equals_restatement(gpu):
    pars = dict(beta=0.016, use_waning=False, pop_infected=20, dynam_layer=dict(c=1))
    sim, port = run_gpu_and_port(1_000_000, 180, 1, pars)
    assert sim.results['cum_infections'].values[-1] > 500_000  # the whole epidemic, peak included
    assert_matches_port(sim_results(sim), port_view(sim.people.get, 1), sim_layers(sim), port, 'hybrid1m')


def test_hybrid_2m_busy_days_coarse_map_equals_restatement(gpu):
    pars = dict(beta=0.05, use_waning=True, pop_infected=200, dynam_layer=dict(c=1))
    variants = [dict(rel_beta=1.7, days=[8], n_imports=100)]
    sim, port = run_gpu_and_port(2_000_000, 40, 5, pars, variants, pop_seed=2)
    assert sim.results['new_infections'].values.max() > 50_000  # busy days: most coarse groups hold a case
    assert_matches_port(sim_results(sim), port_view(sim.people.get, 2), sim_layers(sim), port, 'hybrid2m')


def test_hybrid_10m_two_variants_waning_equals_restatement(gpu):
    from covasim_amd import _lib
    pars = dict(beta=0.016, use_waning=True, pop_infected=20, dynam_layer=dict(c=1))
    variants = [dict(rel_beta=1.7, days=[30], n_imports=100)]
    sim, port = run_gpu_and_port(10_000_000, 60, 1, pars, variants)
    plan = (C.c_int32 * 2)()
    _lib.check(_lib.lib().cvk_day_plan(sim._layers, sim._params, None, 60, plan), 'plan')
    assert plan[0] == 1  # the dynamic layer was generated, not written, each day
    assert sim.results['variant']['new_infections_by_variant'].values[1].sum() > 100  # the imports spread
    assert_matches_port(sim_results(sim), port_view(sim.people.get, 2), sim_layers(sim), port, 'hybrid10m')


def test_batch_256_sims_equal_single_runs_and_restatement(gpu):
    import torch

    import covasim_amd as cv
    S, n_days = 256, 120
    rng = np.random.default_rng(1)
    sizes = np.exp(rng.uniform(np.log(10_000), np.log(500_000), S)).astype(np.int64)
    betas = rng.uniform(0.010, 0.025, S)
    common = dict(use_waning=False, pop_infected=20, dynam_layer=dict(c=1))
    t0 = time.perf_counter()
    pops = [make_population(n=int(sizes[i]), seed=1000 + i, pop_type='hybrid', households='categorical')
            for i in range(S)]
    log(f'{S} populations ({sizes.sum()} agents) built in {time.perf_counter() - t0:.1f} s')

    def make(i):
        return cv.Sim(pop_size=int(sizes[i]), pop_type='hybrid', n_days=n_days, rand_seed=1 + i, beta=float(betas[i]),
                      popdict=copy_pop(pops[i]), households='categorical', log_capacity=1024, **common)
    batch = cv.BatchSim([make(i) for i in range(S)])
    t0 = time.perf_counter()
    batch.run()
    log(f'batch run: {time.perf_counter() - t0:.1f} s')
    keep = {int(np.argmin(sizes)), int(np.argsort(sizes)[S // 2]), int(np.argmax(sizes))}
    kept = {}
    t0 = time.perf_counter()
    for si in range(S):
        alone = make(si)
        alone.run()
        bs = batch.sims[si]
        for key in COMPARE_KEYS + ['n_exposed', 'n_infectious', 'n_recovered', 'n_dead', 'n_susceptible']:
            assert np.array_equal(bs.results[key].values, alone.results[key].values), (si, key)
        view = port_view(lambda k: batch.get(k, si), 1)
        mine = port_view(alone.people.get, 1)
        for k in view:
            assert np.array_equal(view[k], mine[k]), (si, k)
        if si in keep:
            lay = []
            for li, lk in enumerate(batch.layer_keys):
                a, c = int(batch.edge_starts[li, si]), int(batch.edge_counts[li, si])
                lay.append(tuple(batch._layers[lk][col][a:a + c].cpu().numpy() for col in ('p1', 'p2', 'beta')))
            kept[si] = (sim_results(bs), view, lay)
        del alone
    torch.cuda.synchronize()
    log(f'{S} single runs compared in {time.perf_counter() - t0:.1f} s')
    assert sum(float(s.results['cum_infections'].values[-1]) for s in batch.sims) > 1e6
    for si in sorted(keep):
        port = port_sim(pops[si], seed=1 + si, n_days=n_days, pop_type='hybrid', beta=float(betas[si]), **common)
        t0 = time.perf_counter()
        port.run()
        log(f'sim {si} (n={sizes[si]}): C restatement {time.perf_counter() - t0:.1f} s')
        assert_matches_port(*kept[si], port, f'batch sim {si}')
