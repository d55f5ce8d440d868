"""change_beta (covasim/interventions.py:530-586) on the transmission step.

CPU: the oracle's restatement stops transmission at change 0, equals a sim with the scaled beta
when applied from day 0, and the host builds the per-day rate table the day loop reads.
GPU: runs with change_beta interventions (overall and per layer, several days) equal the oracle
with the same interventions bit for bit, in a captured graph, eagerly and stepped day by day.
"""

import numpy as np
import pytest

from oracle import interventions_ref as oiv
from tests.helpers import COMPARE_KEYS, STOCK_KEYS, copy_pop, make_population, oracle_sim


def run_oracle(pop, ivs, **pars):
    osim = oracle_sim(pop, **pars)
    osim.interventions = [iv.initialize(osim) for iv in ivs]
    osim.run()
    return osim, osim.finalize()


def test_oracle_change_beta_zero_stops_transmission():
    pop = make_population(n=3000)
    _, res = run_oracle(pop, [oiv.change_beta(0, 0.0)], seed=2, n_days=30, beta=0.05)
    assert res['new_infections'].sum() == 0
    _, res = run_oracle(pop, [oiv.change_beta(12, 0.0, layers='a')], seed=2, n_days=30, beta=0.05)
    assert res['new_infections'][:12].sum() > 0 and res['new_infections'][12:].sum() == 0


def test_oracle_change_from_day_zero_equals_scaled_beta():
    pop = make_population(n=3000)
    _, a = run_oracle(pop, [oiv.change_beta(0, 0.5)], seed=3, n_days=30, beta=0.05)
    _, b = run_oracle(pop, [], seed=3, n_days=30, beta=0.025)
    for k in COMPARE_KEYS:
        assert np.array_equal(a[k], b[k]), k


def test_host_rate_table():
    import covasim_amd as cv
    from covasim_amd import interventions as cviv

    class FakeSim:
        npts = 8

        def __init__(self):
            self.pars = dict(beta=0.02, beta_layer=dict(h=3.0, c=0.3), interventions=[
                cv.change_beta(2, 0.5), cv.change_beta([3, 6], [0.0, 2.0], layers='c')])

    sim = FakeSim()
    for iv in sim.pars['interventions']:
        iv.initialize(sim)

    def rates(pars):
        return [np.float32(pars['beta'])], [np.float32(pars['beta_layer'][k]) for k in ('h', 'c')]
    bv, bl = cviv.beta_schedule(sim, rates)
    assert np.array_equal(bv[:, 0], np.float32([0.02, 0.02, 0.01, 0.01, 0.01, 0.01, 0.01, 0.01]))
    assert np.array_equal(bl[:, 1], np.float32([0.3, 0.3, 0.3, 0.0, 0.0, 0.0, 0.6, 0.6]))
    assert np.all(bl[:, 0] == np.float32(3.0)) and np.all(bl[:, 2:] == 0)
    assert sim.pars['beta'] == 0.02 and sim.pars['beta_layer']['c'] == 0.3  # the replay leaves pars alone


@pytest.mark.gpu
@pytest.mark.parametrize('mode', ['graph', 'eager', 'step'])
def test_change_beta_matches_oracle(mode):
    import covasim_amd as cv
    print(f'change_beta mode={mode}', flush=True)
    pop = make_population(n=20000, pop_type='hybrid')
    kw = dict(n_days=60, pop_type='hybrid', beta=0.03, dynam_layer=dict(c=1))
    ivs = [cv.change_beta(10, 0.4), cv.change_beta([5, 25], [0.2, 1.0], layers='c'),
           cv.change_beta(40, 0.0, layers=['s', 'w'])]
    sim = cv.Sim(pop_size=20000, rand_seed=6, popdict=copy_pop(pop), interventions=ivs, **kw)
    if mode == 'step':
        sim.initialize()
        while not sim.complete:
            sim.step()
        sim.finalize()
    else:
        sim.run(use_graph=(mode == 'graph'))
    oivs = [oiv.change_beta(10, 0.4), oiv.change_beta([5, 25], [0.2, 1.0], layers='c'),
            oiv.change_beta(40, 0.0, layers=['s', 'w'])]
    osim, ores = run_oracle(pop, oivs, seed=6, **kw)
    for key in COMPARE_KEYS:
        assert np.array_equal(sim.results[key].values, ores[key]), key
    for key in STOCK_KEYS[1:]:
        assert np.array_equal(sim.results[f'n_{key}'].values, ores[f'n_{key}']), key
    assert sim.pars['beta'] == 0.03  # parameters are left as given


# ---------------------------------------------------------------------------------------------
# clip_edges (interventions.py:589-673)
# ---------------------------------------------------------------------------------------------
def test_edge_choice_is_exact_and_uniform():
    n, k = 200, 60
    assert oiv.chosen_edges(1, 0, 0, n, 0).sum() == 0 and oiv.chosen_edges(1, 0, 0, n, n).all()
    counts = np.zeros(n)
    for t in range(400):
        m = oiv.chosen_edges(3, t, 2, n, k)
        assert m.sum() == k
        counts += m
    from scipy import stats
    assert stats.chisquare(counts).pvalue > 1e-3  # each edge chosen with probability k / n


def test_oracle_clip_edges_moves_and_restores():
    pop = make_population(n=3000)
    osim = oracle_sim(pop, seed=2, n_days=20, beta=0.05)
    iv = oiv.clip_edges([3, 8, 12], [0.4, 0.0, 1.0]).initialize(osim)
    osim.interventions = [iv]
    n0 = len(osim.layers['a'][0])
    before = sorted(zip(*[a.tolist() for a in osim.layers['a']]))
    sizes = []
    for t in range(osim.npts):
        osim.step()
        sizes.append((len(osim.layers['a'][0]), len(iv.contacts['a'][0])))
    assert sizes[2] == (n0, 0) and sizes[3][0] + sizes[3][1] == n0
    assert sizes[3][0] == n0 - int((1.0 - 0.4) * n0)
    assert sizes[8] == (0, n0) and sizes[12] == (n0, 0)
    assert sorted(zip(*[a.tolist() for a in osim.layers['a']])) == before  # every edge came back
    res = osim.finalize()
    assert res['new_infections'][8:12].sum() == 0  # no contacts while the layer is empty


@pytest.mark.gpu
@pytest.mark.parametrize('k', [0, 1, 37, 50_000, 99_999, 100_000])
def test_move_edges_matches_oracle(k):
    import torch
    from covasim_amd import interventions as cviv
    rng = np.random.default_rng(k)
    n = 100_000
    cols = [rng.integers(0, 10**6, n).astype(np.int32), rng.integers(0, 10**6, n).astype(np.int32),
            rng.random(n).astype(np.float32)]
    dest = [rng.integers(0, 10**6, 777).astype(np.int32), rng.integers(0, 10**6, 777).astype(np.int32),
            rng.random(777).astype(np.float32)]
    tc = [torch.as_tensor(c, device='cuda') for c in cols]
    td = [torch.as_tensor(c, device='cuda') for c in dest]
    keep, out = cviv.move_edges(tc, n, k, td, seed=99, t=7, slot=3)
    mask = oiv.chosen_edges(99, 7, 3, n, k)
    for i in range(3):
        assert np.array_equal(keep[i].cpu().numpy(), cols[i][~mask])
        assert np.array_equal(out[i].cpu().numpy(), np.concatenate([dest[i], cols[i][mask]]))


@pytest.mark.gpu
@pytest.mark.parametrize('mode', ['run', 'step'])
def test_clip_edges_matches_oracle(mode):
    import covasim_amd as cv
    print(f'clip_edges mode={mode}', flush=True)
    pop = make_population(n=20000, pop_type='hybrid')
    kw = dict(n_days=45, pop_type='hybrid', beta=0.03, dynam_layer=dict(c=1))
    ivs = [cv.clip_edges([5, 20, 30], [0.3, 0.0, 1.0], layers=['s', 'c']), cv.clip_edges(12, 0.5),
           cv.change_beta(25, 0.5, layers='w')]
    sim = cv.Sim(pop_size=20000, rand_seed=8, popdict=copy_pop(pop), interventions=ivs, **kw)
    if mode == 'step':
        sim.initialize()
        while not sim.complete:
            sim.step()
        sim.finalize()
    else:
        sim.run()
    oivs = [oiv.clip_edges([5, 20, 30], [0.3, 0.0, 1.0], layers=['s', 'c']), oiv.clip_edges(12, 0.5),
            oiv.change_beta(25, 0.5, layers='w')]
    osim = oracle_sim(pop, seed=8, **kw)
    osim.interventions = [iv.initialize(osim) for iv in oivs]
    osim.run()
    ores = osim.finalize()
    for key in COMPARE_KEYS:
        assert np.array_equal(sim.results[key].values, ores[key]), key
    for lk in ['h', 's', 'w', 'c']:
        for i, col in enumerate(['p1', 'p2', 'beta']):
            assert np.array_equal(sim.people.contacts[lk][col], osim.layers[lk][i]), (lk, col)
    held = ivs[1].contacts['h']
    assert np.array_equal(held[0].cpu().numpy(), oivs[1].contacts['h'][0])
