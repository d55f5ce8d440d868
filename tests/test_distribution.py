"""Whole-run trajectories agree in distribution with the reference-order oracle (200 seeds).

CPU: the compiled Philox restatement (the GPU path's draws, equal to the GPU bit for bit
on the same seed) against the reference-order fixture, plus a planted mistake (beta
+15%) that the same checks must reject. GPU: the HIP day loop on 200 seeds.
"""

import numpy as np
import pytest

from tests.distcheck import SERIES, compare, load_reference
from tests.helpers import make_population

CAND_SEEDS = range(1000, 1200)


def _port_runs(config, beta_scale=1.0):
    from oracle import covasim_oracle as co
    from oracle.cport.port import PortSim
    out = {k: [] for k in SERIES}
    for seed in CAND_SEEDS:
        pop = make_population(config['n'], seed=seed, pop_type=config['pop_type'])
        pars = co.make_pars(pop_size=config['n'], pop_type=config['pop_type'], n_days=config['n_days'],
                            rand_seed=seed, beta=config['beta'] * beta_scale, pop_infected=config['pop_infected'],
                            use_waning=config['use_waning'])
        init = dict(age=pop['age'], **{k: pop[k] for k in ('symp_prob', 'severe_prob', 'crit_prob', 'death_prob',
                                                             'rel_sus', 'rel_trans')})
        layers = {k: (pop['contacts'][k]['p1'], pop['contacts'][k]['p2'], pop['contacts'][k]['beta'])
                  for k in pars['beta_layer']}
        sim = PortSim(pars, init, layers, seed, omp=True).run()
        for k in SERIES:
            out[k].append(sim.result(k).copy())
    return {k: np.asarray(v, np.float64) for k, v in out.items()}


def test_philox_restatement_matches_reference_distribution():
    ref, config = load_reference()
    failures, report = compare(ref, _port_runs(config))
    assert not failures, (failures, report)
    # power: a 15% higher beta is detected by the same checks
    failures, _ = compare(ref, _port_runs(config, beta_scale=1.15))
    assert failures


@pytest.mark.gpu
def test_gpu_matches_reference_distribution(gpu):
    import covasim_amd as cv
    ref, config = load_reference()
    out = {k: [] for k in SERIES}
    for seed in CAND_SEEDS:
        pop = make_population(config['n'], seed=seed, pop_type=config['pop_type'])
        sim = cv.Sim(pop_size=config['n'], pop_type=config['pop_type'], n_days=config['n_days'], rand_seed=seed,
                     beta=config['beta'], pop_infected=config['pop_infected'], use_waning=config['use_waning'],
                     popdict=pop)
        sim.run()
        for k in SERIES:
            out[k].append(np.asarray(sim.results[k].values, np.float64))
    failures, report = compare(ref, {k: np.asarray(v) for k, v in out.items()})
    print(report, flush=True)
    assert not failures, (failures, report)
