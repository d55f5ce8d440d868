"""Device-state snapshots (covasim/base.py:682-741 save / load; tests/test_resume.py in the reference).

CPU: the parameter encoding round-trips (arrays, nested dicts, variants, interventions).
GPU: a run saved mid-way and loaded into a new Sim finishes exactly as the uninterrupted run, with
waning, two variants, a dynamic layer, change_beta and clip_edges; a finished run reloads with the
same results.
"""

import json
import os

import numpy as np
import pytest

from tests.helpers import COMPARE_KEYS, STOCK_KEYS, copy_pop, make_population


def test_parameter_encoding_round_trips():
    import covasim_amd as cv
    from covasim_amd import checkpoint as cp
    from covasim_amd import parameters as cvpar
    pars = cvpar.make_pars(pop_size=1000, pop_type='hybrid')
    pars['prognoses'] = cvpar.get_prognoses()
    pars['variants'] = [cv.variant('delta', days=[5, 9], n_imports=7)]
    pars['interventions'] = [cv.change_beta([3, 8], [0.5, 1.0], layers='c'), cv.clip_edges(4, 0.3)]
    enc = cp._encode(pars)
    dec = cp._decode(json.loads(json.dumps(enc)))
    assert json.dumps(cp._encode(dec)) == json.dumps(enc)
    assert np.array_equal(dec['prognoses']['death_probs'], pars['prognoses']['death_probs'])
    assert dec['prognoses']['death_probs'].dtype == pars['prognoses']['death_probs'].dtype
    v = dec['variants'][0]
    assert v.p == pars['variants'][0].p and v.days == [5, 9] and v.n_imports == 7
    assert [type(iv).__name__ for iv in dec['interventions']] == ['change_beta', 'clip_edges']
    with pytest.raises(TypeError):
        cp._encode({'f': lambda x: x})


def full_state(sim):
    out = {k: v.cpu().numpy().copy() for k, v in sim.people._d.items()}
    for lk in sim.layer_keys():
        for c in ('p1', 'p2', 'beta'):
            out[f'{lk}.{c}'] = sim.people.contacts[lk][c]
    return out


@pytest.fixture
def generate_dynamic(request):
    """Runs the test under CVK_TUNE_GENERATE_DYNAMIC = request.param (restores the default, 2)."""
    from covasim_amd import _lib
    _lib.check(_lib.lib().cvk_set_tuning(_lib.TUNE_GENERATE_DYNAMIC, request.param), 'tuning')
    yield request.param
    _lib.check(_lib.lib().cvk_set_tuning(_lib.TUNE_GENERATE_DYNAMIC, 2), 'tuning')


@pytest.mark.gpu
@pytest.mark.parametrize('generate_dynamic', [2, 1], indirect=True)
@pytest.mark.parametrize('until', [0, 17, 32])
def test_resume_from_snapshot_equals_uninterrupted(tmp_path, until, generate_dynamic):
    """generate_dynamic 1: the dynamic layer's endpoints are generated inside each run and written
    once when the run (here: the run up to the snapshot) closes, so the snapshot carries them."""
    import covasim_amd as cv
    print(f'snapshot until={until} generate_dynamic={generate_dynamic}', flush=True)
    pop = make_population(n=15000, pop_type='hybrid')

    def make():
        return cv.Sim(pop_size=15000, pop_type='hybrid', n_days=50, rand_seed=12, beta=0.03, use_waning=True,
                      popdict=copy_pop(pop), dynam_layer=dict(c=1),
                      variants=[cv.variant(dict(rel_beta=1.7), days=[10], n_imports=25)],
                      interventions=[cv.change_beta(15, 0.5, layers='w'), cv.clip_edges([20, 35], [0.4, 1.0])])
    a = make()
    a.run()
    b = make()
    b.run(until=until)
    fn = str(tmp_path / f'resume_{until}.sim')
    b.save(fn)
    del b
    c = cv.load(fn)
    assert c.t == until + 1 and not c.complete
    c.run()
    for key in COMPARE_KEYS + [f'n_{k}' for k in STOCK_KEYS] + ['cum_infections', 'r_eff']:
        assert np.array_equal(a.results[key].values, c.results[key].values, equal_nan=True), key
    sa, sc = full_state(a), full_state(c)
    for k in sa:
        assert np.array_equal(sa[k].view(np.uint8), sc[k].view(np.uint8)), k
    def records(sim):  # the log is appended by many workgroups: compare it as a set of records
        return sorted((r['date'], r['target'], -1 if r['source'] is None else r['source'], str(r['layer']),
                       r['variant']) for r in sim.infection_log())
    assert records(a) == records(c)


@pytest.mark.gpu
def test_snapshot_of_finished_run(tmp_path):
    import covasim_amd as cv
    pop = make_population(n=8000, pop_type='random')
    a = cv.Sim(pop_size=8000, n_days=30, rand_seed=3, beta=0.03, popdict=copy_pop(pop))
    a.run()
    fn = str(tmp_path / 'done.sim')
    a.save(fn)
    b = cv.load(fn)
    assert b.complete and b.results_ready
    for key in COMPARE_KEYS + ['cum_infections', 'doubling_time']:
        assert np.array_equal(a.results[key].values, b.results[key].values, equal_nan=True), key
    with pytest.raises(cv.AlreadyRunError):
        b.run()
    os.remove(fn)
