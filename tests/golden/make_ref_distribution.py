"""Writes ref_dist_random20k.npz: the reference-order distribution of whole-run trajectories.

200 seeds (0..199) of the default parity sim: 20 000 agents, pop_type 'random' (one
layer 'a', Poisson(20) contacts), beta 0.016, pop_infected 20, 60 days, waning on
(parameters.py:70 default). Each seed draws its own population (as Sim.initialize
does per seed) and runs the numpy oracle with RefRNG: two MT19937 streams consumed
in the reference's call order (numba's stream for transmission and seed choice,
numpy's for outcomes; utils.py:271-298). Stored: daily new_infections, new_severe,
new_deaths as int32 [200, 61]. Run in the build container (about 20 s on 8 cores):
    python tests/golden/make_ref_distribution.py
"""
import json
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests.helpers import make_population, oracle_sim  # noqa: E402

CONFIG = dict(n=20_000, pop_type='random', beta=0.016, pop_infected=20, n_days=60, use_waning=True)
SERIES = ('new_infections', 'new_severe', 'new_deaths')
SEEDS = list(range(200))


def run_one(seed):
    c = CONFIG
    pop = make_population(c['n'], seed=seed, pop_type=c['pop_type'])
    sim = oracle_sim(pop, rng_kind='ref', seed=seed, pop_type=c['pop_type'], beta=c['beta'],
                     pop_infected=c['pop_infected'], n_days=c['n_days'], use_waning=c['use_waning']).run()
    res = sim.finalize()
    return np.stack([np.asarray(res[k], np.int32) for k in SERIES])


def main():
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        runs = pool.map(run_one, SEEDS)
    arr = np.stack(runs)  # [seed, series, day]
    out = {k: arr[:, i, :] for i, k in enumerate(SERIES)}
    np.savez_compressed(os.path.join(HERE, 'ref_dist_random20k.npz'), seeds=np.asarray(SEEDS, np.int32),
                        config=np.asarray(json.dumps(CONFIG)), **out)
    print({k: float(v.sum(1).mean()) for k, v in out.items()})


if __name__ == '__main__':
    main()
