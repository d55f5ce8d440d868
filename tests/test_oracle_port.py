"""The compiled C restatement (oracle/cport) and the numpy oracle agree draw-for-draw (CPU)."""

import numpy as np
import pytest

from oracle import covasim_oracle as co
from oracle.cport.port import ROW_KEYS, PortSim
from tests.helpers import PERSON_KEYS, make_population, oracle_sim

CASES = [
    ('random', dict(use_waning=False, beta=0.05)),
    ('hybrid', dict(use_waning=True, beta=0.03, dynam_layer=dict(c=1))),
    ('hybrid', dict(use_waning=True, beta=0.02, variants=[dict(rel_beta=1.7, days=[10], n_imports=30)])),
    ('random', dict(use_waning=False, beta=0.08, n_beds_hosp=2, n_beds_icu=1, rel_severe_prob=20.0)),
]


@pytest.mark.parametrize('case', range(len(CASES)))
@pytest.mark.parametrize('omp', [False, True])
def test_port_equals_numpy_oracle(case, omp):
    pop_type, kw = CASES[case]
    pop = make_population(n=2500, pop_type=pop_type)
    o = oracle_sim(pop, seed=11, n_days=40, pop_type=pop_type, **kw).run()
    r = o.finalize()
    opars = co.make_pars(pop_size=2500, rand_seed=11, n_days=40, pop_type=pop_type, **kw)
    init = dict(age=pop['age'], **{k: pop[k] for k in PERSON_KEYS})
    layers = {k: (pop['contacts'][k]['p1'], pop['contacts'][k]['p2'], pop['contacts'][k]['beta'])
              for k in opars['beta_layer']}
    p = PortSim(opars, init, layers, 11, omp=omp).run()
    for k in ROW_KEYS:
        assert np.array_equal(p.result(k), r[k]), k
    assert r['cum_infections'][-1] > 100
    for k in ['date_infectious', 'date_recovered', 'date_dead', 'rel_trans', 'nab', 'peak_nab']:
        assert np.array_equal(p.a[k].view(np.uint32), getattr(o.people, k).view(np.uint32)), k
