"""Dynamic-layer extension points (covasim/base.py:1849-1876, examples/t11_dynamic_layers.py:9-26).

CPU: the oracle's partial resample (Layer.update with frac < 1) is pinned by what the reference's
steps fix: exactly round(n * frac) edges (numpy's half-to-even rounding) get new endpoints and weight
1 and the rest are untouched; frac = 1 is the full resample; over many days every edge is chosen
equally often (chi-squared).
GPU: cvk_resample_layer_frac equals the oracle; whole runs with a Layer subclass whose update()
replaces the layer by one of a different length every day (t11's CustomLayer), and with one that
resamples a fraction of its edges, equal the oracle with the same per-day layers.
"""

import numpy as np
import pytest
from scipy import stats

from oracle import philox as px
from tests.helpers import COMPARE_KEYS, copy_pop, make_population, oracle_sim


def fresh(n):
    return np.full(n, -1, np.int32), np.full(n, -1, np.int32), np.full(n, 0.5, np.float32)


@pytest.mark.parametrize('n, frac, k', [(10_001, 0.3, 3000), (5, 0.5, 2), (7, 0.5, 4), (1000, 0.0, 0),
                                        (999, 1.0, 999)])
def test_partial_resample_changes_round_n_frac_edges(n, frac, k):
    p1, p2, beta = px.partial_resample(3, 7, 2, *fresh(n), n_agents=500, frac=frac)
    changed = p1 >= 0
    assert changed.sum() == k  # np.round(n * frac): 2.5 -> 2, 3.5 -> 4 (base.py:1868)
    assert np.array_equal(changed, p2 >= 0)
    assert np.all(beta[changed] == 1.0) and np.all(beta[~changed] == 0.5)
    assert p1.max() < 500 and p2.max() < 500


def test_partial_resample_full_is_the_resample():
    q1, q2 = px.resample_edges(9, 4, 3, 4001, 1234)
    p1, p2, beta = px.partial_resample(9, 4, 3, *fresh(4001), n_agents=1234, frac=1.0)
    assert np.array_equal(p1, q1) and np.array_equal(p2, q2) and np.all(beta == 1.0)
    # a chosen edge gets exactly its full-resample endpoints
    r1, r2, _ = px.partial_resample(9, 4, 3, *fresh(4001), n_agents=1234, frac=0.4)
    ch = r1 >= 0
    assert np.array_equal(r1[ch], q1[ch]) and np.array_equal(r2[ch], q2[ch])


def test_partial_resample_choice_is_uniform():
    n, frac, days = 200, 0.25, 800
    hits = np.zeros(n)
    for t in range(days):
        p1, _, _ = px.partial_resample(5, t, 1, *fresh(n), n_agents=50, frac=frac)
        hits += p1 >= 0
    assert hits.sum() == days * 50
    exp = days * frac
    chi2 = ((hits - exp) ** 2 / (exp * (1 - frac))).sum()  # hypergeometric variance per edge
    assert stats.chi2.sf(chi2, n - 1) > 1e-3
    assert len(set(map(tuple, [np.flatnonzero(px.partial_resample(5, t, 1, *fresh(n), 50, frac)[0] >= 0)
                                for t in range(3)]))) == 3  # a new subset every day


# ---------------------------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize('frac', [0.0, 0.3, 0.5, 1.0])
def test_resample_frac_matches_oracle(gpu, frac):
    import ctypes as C
    import torch
    from covasim_amd import _lib
    n_edges, n_agents, t, layer = 100_001, 4000, 6, 3
    p1 = torch.full((n_edges,), -1, dtype=torch.int32, device='cuda')
    p2 = torch.full((n_edges,), -1, dtype=torch.int32, device='cuda')
    beta = torch.full((n_edges,), 0.5, dtype=torch.float32, device='cuda')
    lay = _lib.LayerS(p1.data_ptr(), p2.data_ptr(), beta.data_ptr(), n_edges, 1, 0)
    prm = _lib.Params()
    prm.n_agents = n_agents
    prm.seed = 77
    L = _lib.lib()
    k = int(np.round(n_edges * frac))
    nb = int(L.cvk_resample_frac_work_bytes())
    work = torch.empty(nb // 4, dtype=torch.int32, device='cuda')
    _lib.check(L.cvk_resample_layer_frac(C.byref(lay), layer, C.byref(prm), t, k, work.data_ptr(), nb,
                                         _lib.stream_ptr()), 'partial resample')
    torch.cuda.synchronize()
    w1, w2, wb = px.partial_resample(77, t, layer, *fresh(n_edges)[:2], np.full(n_edges, 0.5, np.float32),
                                     n_agents, frac)
    assert np.array_equal(p1.cpu().numpy(), w1) and np.array_equal(p2.cpu().numpy(), w2)
    assert np.array_equal(beta.cpu().numpy(), wb)
    assert L.cvk_resample_layer_frac(C.byref(lay), layer, C.byref(prm), t, n_edges + 1, work.data_ptr(), nb,
                                     None) == _lib.E_INVALID


def t11_layer_class():
    """examples/t11_dynamic_layers.py:9-26, with the day's contacts drawn from a generator keyed by the day."""
    import covasim_amd as cv

    class CustomLayer(cv.Layer):
        def __init__(self, layer, contact_data):
            for k, v in layer.items():  # no Layer.__init__, as in the example
                self[k] = v
            self.contact_data = contact_data

        def update(self, people):
            n_new = int(self.contact_data[people.t])
            self['p1'], self['p2'], self['beta'] = custom_edges(people.t, len(people), n_new)
    return CustomLayer


def custom_edges(t, n_agents, n_new):
    rng = np.random.default_rng([41, t])
    return (rng.integers(0, n_agents, n_new).astype(np.int32), rng.integers(0, n_agents, n_new).astype(np.int32),
            np.ones(n_new, np.float32))


def compare_runs(sim, osim):
    ores = osim.finalize()
    for key in COMPARE_KEYS:
        assert np.array_equal(sim.results[key].values, ores[key]), key
    gp, op = sim.people, osim.people
    for key in ['susceptible', 'exposed', 'infectious', 'recovered', 'dead']:
        assert np.array_equal(gp.get(key), getattr(op, key)), key
    for key in ['date_exposed', 'date_infectious', 'date_recovered', 'date_dead', 'rel_trans', 'nab']:
        assert np.array_equal(gp.get(key).view(np.uint32), getattr(op, key).view(np.uint32)), key
    for lk, (p1, p2, beta) in osim.layers.items():
        lay = gp.contacts[lk]
        assert np.array_equal(lay['p1'], p1) and np.array_equal(lay['p2'], p2), lk
        assert np.array_equal(lay['beta'].view(np.uint32), np.asarray(beta, np.float32).view(np.uint32)), lk


@pytest.mark.gpu
def test_custom_variable_length_layer_matches_oracle(gpu):
    import covasim_amd as cv
    n, days = 4000, 30
    pop = make_population(n=n, pop_type='hybrid')
    pars = dict(pop_type='hybrid', use_waning=True, beta=0.03, dynam_layer=dict(c=1))
    sim = cv.Sim(pop_size=n, n_days=days, rand_seed=4, popdict=copy_pop(pop), **pars)
    sim.initialize()
    contact_data = np.random.default_rng(2).integers(n * 2, n * 6, days + 1)
    CustomLayer = t11_layer_class()
    sim.people.contacts['c'] = CustomLayer(sim.people.contacts['c'], contact_data)
    assert sim._custom_update == ['c']
    sim.run()
    osim = oracle_sim(pop, seed=4, n_days=days, **pars)
    osim.layer_updates['c'] = lambda t, li, lay, n_agents: custom_edges(t, n_agents, int(contact_data[t]))
    osim.run()
    assert len(sim.people.contacts['c']) == contact_data[days]  # the last day's layer
    assert sim.results['cum_infections'].values[-1] > 100
    compare_runs(sim, osim)


@pytest.mark.gpu
def test_partial_resample_layer_subclass_matches_oracle(gpu):
    import covasim_amd as cv
    n, days, frac = 4000, 30, 0.3

    class PartialLayer(cv.Layer):
        def update(self, people):
            super().update(people, frac=frac)

    pop = make_population(n=n, pop_type='hybrid')
    pars = dict(pop_type='hybrid', use_waning=False, beta=0.03, dynam_layer=dict(c=1, w=1))
    sim = cv.Sim(pop_size=n, n_days=days, rand_seed=6, popdict=copy_pop(pop), **pars)
    sim.initialize()
    for lk in ('c', 'w'):
        old = sim.people.contacts[lk]
        sim.people.contacts[lk] = PartialLayer({k: old[k] for k in ('p1', 'p2', 'beta')}, label=lk)
    sim.run()
    osim = oracle_sim(pop, seed=6, n_days=days, **pars)
    for lk in ('c', 'w'):
        osim.layer_updates[lk] = lambda t, li, lay, n_agents: px.partial_resample(6, t, li, *lay, n_agents, frac)
    osim.run()
    assert sim.results['cum_infections'].values[-1] > 100
    compare_runs(sim, osim)
