"""MultiSim / multi_run (covasim/run.py:36-330, 1326-1519).

CPU: the set-up follows single_run (run.py:1358-1394): copies of one sim get seeds rand_seed + i and
labels 'Sim i' and leave the base sim as it was; iterpars set one value per run; a list of sims keeps
its seeds unless reseeded; noise multiplies beta by 1 + x or 1 / (1 - x), x = noise * N(0, 1).
GPU: every sim of a multi_run equals the same sim run alone, bit for bit, whether it ran inside a
BatchSim (compatible sims, change_beta schedules included) or alone (a sim that tests and traces, one with
other days); reduce()
gives the median and the 10 / 90 % quantiles of the sims' rows (np.quantile), mean() the mean and
2 standard deviations.
"""

import numpy as np
import pytest

from tests.helpers import COMPARE_KEYS, copy_pop, make_population


def test_multi_run_set_up():
    import covasim_amd as cv
    base = cv.Sim(pop_size=1000, n_days=10, rand_seed=5, beta=0.02)
    sims = cv.multi_run(base, n_runs=3, do_run=False)
    assert [s['rand_seed'] for s in sims] == [5, 6, 7] and [s.label for s in sims] == ['Sim 0', 'Sim 1', 'Sim 2']
    assert base['rand_seed'] == 5 and not base.initialized and all(s is not base for s in sims)
    sims = cv.multi_run(base, iterpars=dict(beta=[0.01, 0.03]), do_run=False)
    assert [s['beta'] for s in sims] == [0.01, 0.03] and [s['rand_seed'] for s in sims] == [5, 6]
    listed = [cv.Sim(pop_size=500, n_days=5, rand_seed=9), cv.Sim(pop_size=500, n_days=5, rand_seed=9)]
    out = cv.multi_run(listed, do_run=False)
    assert [s['rand_seed'] for s in out] == [9, 9] and out[0] is listed[0]
    np.random.seed(3)
    x = 0.2 * np.random.normal(size=3)
    np.random.seed(3)
    noisy = cv.multi_run(base, n_runs=3, noise=0.2, do_run=False)
    want = [0.02 * (1 + v if v > 0 else 1 / (1 - v)) for v in x]
    assert np.allclose([s['beta'] for s in noisy], want, rtol=1e-15)
    with pytest.raises(TypeError):
        cv.multi_run('sim')
    with pytest.raises(ValueError):
        cv.multi_run(base, iterpars=dict(beta=[0.01, 0.02], rel_beta=[1.0]))
    with pytest.raises(KeyError):
        cv.multi_run(base, n_runs=2, sim_args=dict(not_a_parameter=1), do_run=False)


def same(a, b):
    for key in COMPARE_KEYS + ['n_infectious', 'n_recovered']:
        assert np.array_equal(a.results[key].values, b.results[key].values), key


@pytest.mark.gpu
def test_multi_run_equals_single_runs(gpu):
    import covasim_amd as cv
    pop = make_population(n=3000, pop_type='hybrid')
    kw = dict(pop_size=3000, pop_type='hybrid', n_days=30, beta=0.03, dynam_layer=dict(c=1))
    base = cv.Sim(rand_seed=11, popdict=copy_pop(pop), **kw)
    msim = cv.MultiSim(base, n_runs=4).run()
    assert len(msim) == 4
    for i, sim in enumerate(msim.sims):
        alone = cv.Sim(rand_seed=11 + i, popdict=copy_pop(pop), **kw).run()
        assert sim['rand_seed'] == 11 + i
        same(sim, alone)
    assert len({int(s.results['cum_infections'].values[-1]) for s in msim.sims}) > 1  # different seeds
    R = msim.reduce(output=True)
    stack = np.stack([s.results['new_infections'].values for s in msim.sims])
    assert np.array_equal(R['new_infections'].values, np.quantile(stack, 0.5, axis=0))
    assert np.array_equal(R['new_infections'].low, np.quantile(stack, 0.1, axis=0))
    assert np.array_equal(R['new_infections'].high, np.quantile(stack, 0.9, axis=0))
    msim.mean()
    assert np.allclose(msim.results['new_infections'].values, stack.mean(axis=0), rtol=1e-12)
    assert np.allclose(msim.results['new_infections'].high, stack.mean(axis=0) + 2 * stack.std(axis=0), rtol=1e-9)


@pytest.mark.gpu
def test_multi_run_mixed_list(gpu):
    import covasim_amd as cv
    pop = make_population(n=2500, pop_type='hybrid')
    kw = dict(pop_size=2500, pop_type='hybrid', beta=0.03, dynam_layer=dict(c=1))

    def make():
        return [cv.Sim(rand_seed=1, n_days=25, popdict=copy_pop(pop), **kw),
                cv.Sim(rand_seed=2, n_days=25, popdict=copy_pop(pop), **kw),
                cv.Sim(rand_seed=3, n_days=25, popdict=copy_pop(pop),
                       interventions=[cv.change_beta(days=10, changes=0.3)], **kw),
                cv.Sim(rand_seed=4, n_days=18, popdict=copy_pop(pop), **kw),
                cv.Sim(rand_seed=5, n_days=25, popdict=copy_pop(pop), **kw),
                cv.Sim(rand_seed=6, n_days=25, popdict=copy_pop(pop),
                       interventions=[cv.test_prob(symp_prob=0.2), cv.contact_tracing(trace_probs=0.5)], **kw)]
    sims = cv.multi_run(make())
    for sim, alone in zip(sims, [s.run() for s in make()]):
        same(sim, alone)
