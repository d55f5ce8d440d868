"""Shared builders: one synthetic population fed to both the GPU path and the oracle."""

import numpy as np

from covasim_amd import parameters as cvpar
from covasim_amd import population as cvpop
from oracle import covasim_oracle as co

PERSON_KEYS = ('symp_prob', 'severe_prob', 'crit_prob', 'death_prob', 'rel_sus', 'rel_trans')


def make_population(n=2000, seed=1, pop_type='random', households='poisson', contacts=None, weighted=False):
    pars = cvpar.make_pars(pop_size=n, pop_type=pop_type, contacts=contacts or {})
    pars['prognoses'] = cvpar.get_prognoses()
    rng = np.random.default_rng(seed + 12345)
    pop = cvpop.make_randpop(pars, rng, households=households)
    pop.update(cvpop.make_people_arrays(pop['age'], pars, rng))
    for lay in pop['contacts'].values():
        if weighted:  # per-edge weights other than 1 (base.py:1671-1674 allows any float32)
            lay['beta'] = rng.choice(np.array([0.5, 1.0, 2.0, 0.0], np.float32), len(lay['p1']))
        else:
            lay['beta'] = np.ones(len(lay['p1']), np.float32)
    return pop


def oracle_sim(pop, rng_kind='philox', seed=1, **pars):
    """An oracle Sim over the given population with the same parameters as a covasim_amd Sim."""
    variants = pars.pop('variants', [])
    opars = co.make_pars(pop_size=len(pop['age']), rand_seed=seed, variants=variants, **pars)
    init = dict(age=pop['age'], **{k: pop[k] for k in PERSON_KEYS})
    keys = list(opars['beta_layer'].keys())
    layers = {k: (pop['contacts'][k]['p1'], pop['contacts'][k]['p2'], pop['contacts'][k]['beta']) for k in keys}
    rng = co.PhiloxRNG(seed) if rng_kind == 'philox' else co.RefRNG(seed)
    return co.Sim(opars, init, layers, rng)


def copy_pop(pop):
    out = {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v) for k, v in pop.items() if k != 'contacts'}
    out['contacts'] = {k: {c: np.array(a, copy=True) for c, a in v.items()} for k, v in pop['contacts'].items()}
    return out


COMPARE_KEYS = ['new_infections', 'new_reinfections', 'new_infectious', 'new_symptomatic', 'new_severe',
                'new_critical', 'new_recoveries', 'new_deaths', 'new_known_deaths']
STOCK_KEYS = ['susceptible', 'exposed', 'infectious', 'symptomatic', 'severe', 'critical', 'recovered', 'dead']
