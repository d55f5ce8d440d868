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


def port_sim(pop, seed=1, n_days=40, omp=True, **pars):
    """The oracle's compiled C restatement (oracle/cport) over the given population, with the same
    parameters as a covasim_amd Sim (variants as dicts of rel_beta, days, n_imports)."""
    from oracle.cport.port import PortSim
    variants = list(pars.pop('variants', []))
    opars = co.make_pars(pop_size=len(pop['age']), rand_seed=seed, n_days=n_days, variants=variants, **pars)
    init = dict(age=pop['age'], **{k: pop[k] for k in PERSON_KEYS})
    layers = {k: (pop['contacts'][k]['p1'], pop['contacts'][k]['p2'], pop['contacts'][k]['beta'])
              for k in opars['beta_layer']}
    return PortSim(opars, init, layers, seed, omp=omp)


PORT_FLAG_KEYS = ['susceptible', 'naive', 'exposed', 'infectious', 'symptomatic', 'severe', 'critical', 'diagnosed',
                  'recovered', 'known_dead', 'dead']
PORT_BIT_KEYS = ['exposed_variant', 'infectious_variant', 'recovered_variant', 'date_exposed', 'date_infectious',
                 'date_symptomatic', 'date_severe', 'date_critical', 'date_recovered', 'date_dead', 'dur_exp2inf',
                 'dur_inf2sym', 'dur_sym2sev', 'dur_sev2crit', 'dur_disease', 'rel_trans', 'rel_sus', 'nab',
                 'peak_nab', 'n_infections', 'n_breakthroughs', 't_nab_event', 'sus_imm', 'symp_imm', 'sev_imm']


def port_view(get, n_variants):
    """Agent arrays of a finished run as the port holds them, from a key -> host array accessor."""
    out = {k: np.asarray(get(k)).astype(np.uint8) for k in PORT_FLAG_KEYS}
    for k in PORT_BIT_KEYS:
        out[k] = np.ascontiguousarray(get(k)).view(np.uint8)
    for k in ['exposed_by_variant', 'infectious_by_variant']:
        out[k] = np.asarray(get(k))[:n_variants].astype(np.uint8)
    return out


def assert_matches_port(results, state, layers, port, label=''):
    """A finished covasim_amd run equals the C restatement's: every per-day count of the port's rows,
    every agent array byte for byte (NaN included), and the contact layers as they stand after the
    run. results: key -> per-day values; state: port_view(...) of the device arrays; layers:
    key -> (p1, p2, beta) host arrays in the port's layer order."""
    from oracle.cport.port import ROW_KEYS
    for k in ROW_KEYS:
        got = np.asarray(results[k], dtype=np.float64)
        want = port.result(k)
        if not np.array_equal(got, want):
            bad = np.flatnonzero(got != want)
            raise AssertionError(f'{label} {k}: first difference on day {bad[0]}: {got[bad[0]]} vs {want[bad[0]]}')
    pa = port.a
    for k in PORT_FLAG_KEYS + ['exposed_by_variant', 'infectious_by_variant']:
        assert np.array_equal(state[k], pa[k].astype(np.uint8)), (label, k, int(np.count_nonzero(state[k] != pa[k])))
    for k in PORT_BIT_KEYS:
        want = np.ascontiguousarray(pa[k]).view(np.uint8)
        assert np.array_equal(state[k], want), (label, k)
    for li, (p1, p2, beta) in enumerate(layers):
        q1, q2, qb = port.layers_np[li]
        assert np.array_equal(p1, q1) and np.array_equal(p2, q2), (label, 'layer', li)
        assert np.array_equal(np.asarray(beta, np.float32).view(np.uint32), qb.view(np.uint32)), (label, 'beta', li)
