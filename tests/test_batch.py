"""The batched many-sim path: every sim of a batch equals the same sim run alone, bit for bit.

GPU tests compare BatchSim against Sim.run on uneven sims (different sizes, betas, seeds,
import schedules), across launch shapes, and graph against eager launches.
"""

import numpy as np
import pytest

from tests.helpers import COMPARE_KEYS, make_population

BYVAR = ['new_infections_by_variant', 'new_symptomatic_by_variant', 'new_severe_by_variant',
         'new_infectious_by_variant', 'n_exposed_by_variant', 'n_infectious_by_variant']
STATE = ['susceptible', 'exposed', 'infectious', 'symptomatic', 'severe', 'critical', 'recovered', 'dead',
         'date_exposed', 'date_infectious', 'date_symptomatic', 'date_severe', 'date_critical', 'date_recovered',
         'date_dead', 'dur_disease', 'rel_trans', 'nab', 'peak_nab', 'sus_imm', 'n_infections', 'exposed_variant',
         'recovered_variant']


def make_sims(spec, n_days, pop_type, **common):
    import covasim_amd as cv
    sims = []
    for i, (n, beta, seed, variants) in enumerate(spec):
        pop = make_population(n=n, seed=100 + i, pop_type=pop_type)
        gv = [cv.variant(dict(rel_beta=v['rel_beta']), days=v['days'], n_imports=v['n_imports']) for v in variants]
        sims.append(cv.Sim(pop_size=n, pop_type=pop_type, n_days=n_days, rand_seed=seed, beta=beta, popdict=pop,
                           variants=gv, **common))
    return sims


def assert_same_results(a, b):
    for key in COMPARE_KEYS + ['n_exposed', 'n_infectious', 'n_recovered', 'n_dead', 'n_susceptible']:
        assert np.array_equal(a.results[key].values, b.results[key].values), key
    for key in BYVAR:
        assert np.array_equal(a.results['variant'][key].values, b.results['variant'][key].values), key
    np.testing.assert_allclose(a.results['pop_nabs'].values, b.results['pop_nabs'].values, rtol=1e-12, equal_nan=True)


def bits(x):
    x = np.asarray(x)
    return x.view(np.uint8) if x.dtype != bool else x


def assert_same_state(batch, si, sim):
    for key in STATE:
        assert np.array_equal(bits(batch.get(key, si)), bits(sim.people.get(key))), (si, key)


def run_alone(spec, n_days, pop_type, **common):
    alone = make_sims(spec, n_days, pop_type, **common)
    for sim in alone:
        sim.run()
    return alone


HYBRID_SPEC = [(1500, 0.03, 3, []), (2600, 0.02, 4, []), (803, 0.05, 5, []), (4000, 0.025, 6, []),
               (3101, 0.04, 7, [])]


@pytest.fixture
def batch_generate(request):
    """CVK_TUNE_GENERATE_DYNAMIC for the batch only (1: the dynamic layer's endpoints are generated
    inside the batched passes, written once at the end); restored to the default (2) after."""
    from covasim_amd import _lib
    yield request.param
    _lib.check(_lib.lib().cvk_set_tuning(_lib.TUNE_GENERATE_DYNAMIC, 2), 'tuning')


@pytest.mark.gpu
@pytest.mark.parametrize('batch_generate', [2, 1], indirect=True)
def test_batch_equals_individual_runs_hybrid_dynamic(gpu, batch_generate):
    import covasim_amd as cv
    from covasim_amd import _lib
    common = dict(use_waning=False, dynam_layer=dict(c=1))
    alone = run_alone(HYBRID_SPEC, 35, 'hybrid', **common)  # written layers (below the automatic switch)
    _lib.check(_lib.lib().cvk_set_tuning(_lib.TUNE_GENERATE_DYNAMIC, batch_generate), 'tuning')
    batch = cv.BatchSim(make_sims(HYBRID_SPEC, 35, 'hybrid', **common)).run()
    for si, sim in enumerate(batch.sims):
        assert sim.results['cum_infections'].values[-1] > 30
        assert_same_results(sim, alone[si])
        assert_same_state(batch, si, alone[si])
        # the dynamic layer after the run holds the last day's contacts, weights 1
        li = batch.layer_keys.index('c')
        a, c = int(batch.edge_starts[li, si]), int(batch.edge_counts[li, si])
        for col in ('p1', 'p2', 'beta'):
            got = batch._layers['c'][col][a:a + c].cpu().numpy()
            assert np.array_equal(got, alone[si].people.contacts['c'][col]), (si, col)


@pytest.fixture
def source_lists(request):
    """CVK_TUNE_SOURCE_LISTS for the batch (1: the listed layers are always walked, 0: always streamed);
    restored to the default (2: walked on quiet days) after."""
    from covasim_amd import _lib
    yield request.param
    _lib.check(_lib.lib().cvk_set_tuning(_lib.TUNE_SOURCE_LISTS, 2), 'tuning')


@pytest.mark.gpu
@pytest.mark.parametrize('source_lists', [1, 0], indirect=True)
def test_batch_source_lists_walked_or_streamed(gpu, source_lists):
    import covasim_amd as cv
    from covasim_amd import _lib
    common = dict(use_waning=True, dynam_layer=dict(c=1))
    alone = run_alone(HYBRID_SPEC, 30, 'hybrid', **common)  # the default rule (walked on quiet days)
    _lib.check(_lib.lib().cvk_set_tuning(_lib.TUNE_SOURCE_LISTS, source_lists), 'tuning')
    batch = cv.BatchSim(make_sims(HYBRID_SPEC, 30, 'hybrid', **common))
    assert batch.sims[0]._sources  # the static layers have lists
    batch.run()
    for si, sim in enumerate(batch.sims):
        assert sim.results['cum_infections'].values[-1] > 30
        assert_same_results(sim, alone[si])
        assert_same_state(batch, si, alone[si])


@pytest.mark.gpu
def test_batch_with_change_beta_schedules(gpu):
    import covasim_amd as cv
    common = dict(use_waning=False, dynam_layer=dict(c=1))

    def ivs(i):  # per-sim schedules: overall beta, a layer's beta, both, none
        return [[cv.change_beta(days=[8, 20], changes=[0.4, 1.2])],
                [cv.change_beta(days=5, changes=0.2, layers='c')],
                [cv.change_beta(days=12, changes=0.5), cv.change_beta(days=3, changes=2.0, layers=['h', 's'])],
                [], []][i]

    def make():
        sims = make_sims(HYBRID_SPEC, 30, 'hybrid', **common)
        for i, sim in enumerate(sims):
            sim.pars['interventions'] = ivs(i)
        return sims
    alone = [s.run() for s in make()]
    batch = cv.BatchSim(make()).run()
    for si, sim in enumerate(batch.sims):
        assert_same_results(sim, alone[si])
        assert_same_state(batch, si, alone[si])
    plain = cv.BatchSim(make_sims(HYBRID_SPEC, 30, 'hybrid', **common)).run()
    assert plain.sims[0].results['cum_infections'].values[-1] != batch.sims[0].results['cum_infections'].values[-1]
    with pytest.raises(NotImplementedError):
        sims = make_sims(HYBRID_SPEC[:2], 10, 'hybrid', **common)
        sims[0].pars['interventions'] = [cv.test_prob(symp_prob=0.1)]
        cv.BatchSim(sims)


@pytest.mark.gpu
def test_batch_equals_individual_runs_variants_waning(gpu):
    import covasim_amd as cv
    spec = [(2000, 0.03, 11, [dict(rel_beta=1.7, days=[5], n_imports=20)]),
            (1200, 0.04, 12, [dict(rel_beta=1.7, days=[9, 12], n_imports=7)]),
            (2500, 0.02, 13, [dict(rel_beta=1.7, days=[5], n_imports=40)])]
    common = dict(use_waning=True)
    alone = run_alone(spec, 30, 'random', **common)
    batch = cv.BatchSim(make_sims(spec, 30, 'random', **common), log_capacity=20000).run()
    for si, sim in enumerate(batch.sims):
        assert_same_results(sim, alone[si])
        assert_same_state(batch, si, alone[si])
    # infection log: the batch's records of each sim are that sim's records
    blog = batch.infection_log()
    for si, sim in enumerate(alone):
        def rec(r):
            return (r['date'], r['target'], -1 if r['source'] is None else r['source'], r['layer'], r['variant'])
        mine = sorted(rec(r) for r in blog if r['sim'] == si)
        theirs = sorted(rec(r) for r in sim.infection_log())
        assert mine == theirs


@pytest.mark.gpu
def test_batch_launch_shapes_and_graph_agree(gpu):
    import covasim_amd as cv
    from covasim_amd.batch import AGENT_SHAPES, EDGE_SHAPES
    common = dict(use_waning=True, dynam_layer=dict(c=1))
    alone = run_alone(HYBRID_SPEC[:3], 25, 'hybrid', **common)
    batch = cv.BatchSim(make_sims(HYBRID_SPEC[:3], 25, 'hybrid', **common))
    snap = batch.snapshot_state()
    shapes = [e + (2,) for e in EDGE_SHAPES] + [EDGE_SHAPES[0] + (a,) for a in AGENT_SHAPES]
    for shape in shapes:
        for use_graph in (True, False):
            batch.restore_state(snap)
            batch.set_config(*shape)
            assert batch.config() == shape
            batch.run(use_graph=use_graph)
            for si, sim in enumerate(batch.sims):
                assert_same_results(sim, alone[si])


@pytest.mark.gpu
def test_batch_autotune_keeps_results(gpu):
    import covasim_amd as cv
    common = dict(use_waning=False, dynam_layer=dict(c=1))
    alone = run_alone(HYBRID_SPEC, 20, 'hybrid', **common)
    batch = cv.BatchSim(make_sims(HYBRID_SPEC, 20, 'hybrid', **common))
    best = batch.autotune(days=4)
    assert batch.config() == best and len(batch.tuning['trials_ms']) >= 7
    batch.run()
    for si, sim in enumerate(batch.sims):
        assert_same_results(sim, alone[si])


@pytest.mark.gpu
def test_batch_rejects_mismatched_sims(gpu):
    import covasim_amd as cv
    a = make_sims([(500, 0.02, 1, [])], 10, 'hybrid')[0]
    b = make_sims([(500, 0.02, 2, [])], 12, 'hybrid')[0]
    with pytest.raises(ValueError):
        cv.BatchSim([a, b])
    c = make_sims([(500, 0.02, 3, [])], 10, 'random')[0]
    with pytest.raises(ValueError):
        cv.BatchSim([a, c])
