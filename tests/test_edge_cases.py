"""Boundary cases of whole runs, the HIP path against the oracle (sim.py:558-685): no seed infections, zero days
(the seeds and day 0 only), a population without any contact, an empty static layer beside a daily-resampled
one, and a population of seven agents. Counts and every agent array are compared exactly, in graph and eager
runs."""

import numpy as np
import pytest

from tests.helpers import copy_pop, make_population

pytestmark = pytest.mark.gpu


def _case(name):
    if name == 'no_seeds':
        return make_population(n=2000, pop_type='hybrid'), dict(n_days=20, pop_infected=0, beta=0.05,
                                                                 pop_type='hybrid')
    if name == 'zero_days':
        return make_population(n=2000, pop_type='hybrid'), dict(n_days=0, beta=0.05, pop_type='hybrid')
    if name == 'no_contacts':
        pop = make_population(n=1500, pop_type='random', contacts={'a': 0})
        assert len(pop['contacts']['a']['p1']) == 0
        return pop, dict(n_days=25, beta=0.05, pop_type='random')
    if name == 'empty_static_layer':
        pop = make_population(n=2000, pop_type='hybrid')
        for c in ('p1', 'p2'):
            pop['contacts']['s'][c] = np.zeros(0, np.int32)
        pop['contacts']['s']['beta'] = np.zeros(0, np.float32)
        return pop, dict(n_days=30, beta=0.05, pop_type='hybrid', dynam_layer=dict(c=1))
    if name == 'seven_agents':
        return make_population(n=7, pop_type='random', contacts={'a': 4}), dict(n_days=15, beta=0.5,
                                                                               pop_type='random', pop_infected=2)
    raise KeyError(name)


@pytest.mark.parametrize('use_graph', [True, False])
@pytest.mark.parametrize('name', ['no_seeds', 'zero_days', 'no_contacts', 'empty_static_layer', 'seven_agents'])
def test_boundary_run_matches_oracle(gpu, name, use_graph):
    from tests.test_gpu_parity import assert_people_equal, assert_results_equal, run_pair
    print(f'boundary case {name} graph={use_graph}', flush=True)
    pop, pars = _case(name)
    n_days = pars.pop('n_days')
    sim, osim, ores = run_pair(copy_pop(pop), n_days=n_days, seed=3, use_graph=use_graph, **pars)
    assert len(sim.results['new_infections'].values) == n_days + 1
    assert_results_equal(sim, ores)
    assert_people_equal(sim, osim)
    if name in ('no_seeds', 'no_contacts'):
        assert sim.results['new_infections'].values.sum() == 0
