"""Known-answer fixtures of the transmission kernels (tests/golden/kernel_vectors.json, written by
tests/golden/make_kernel_vectors.py from closed forms and plain loops).

CPU: the oracle reproduces every fixture. GPU: the device mirrors (cvk_viral_load, cvk_trans_sus,
cvk_infections with supplied uniforms) reproduce them bit for bit.
"""

import json
import os

import numpy as np
import pytest

from oracle import covasim_oracle as co

F32 = np.float32
G = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'kernel_vectors.json')))


def f32(bits_list):
    return np.asarray(bits_list, dtype=np.uint32).view(F32)


def bits(x):
    return np.asarray(x, dtype=F32).view(np.uint32)


def viral_load_fn(kind):
    if kind == 'oracle':
        return co.compute_viral_load
    import covasim_amd as cv
    return cv.utils.compute_viral_load


def check_viral_load(kind):
    V = G['viral_load']
    fn = viral_load_fn(kind)
    for c in V['cases']:
        got = fn(c['t'], f32([c['date_infectious']]), f32([c['date_recovered']]), f32([c['date_dead']]),
                 V['frac_time'], V['load_ratio'], V['high_cap'])
        assert int(bits(np.asarray(got))[0]) == c['expected'], c


def check_trans_sus(kind):
    T = G['trans_sus']
    if kind == 'oracle':
        fn = co.compute_trans_sus
    else:
        import covasim_amd as cv
        fn = cv.utils.compute_trans_sus
    rt, rs, vl, imm = f32(T['rel_trans']), f32(T['rel_sus']), f32(T['viral_load']), f32(T['sus_imm'])
    n = len(rt)
    for c in T['cases']:
        b = lambda k: np.full(n, bool(c[k]))  # noqa: E731
        got_rt, got_rs = fn(rt, rs, b('inf'), b('sus'), c['beta_layer'], vl, b('symp'), b('iso'), b('quar'),
                            c['asymp_factor'], c['iso_factor'], c['quar_factor'], imm)
        assert bits(np.asarray(got_rt)).tolist() == c['rel_trans_eff'], c['name']
        assert bits(np.asarray(got_rs)).tolist() == c['rel_sus_eff'], c['name']


def check_infections(kind):
    I = G['infections']
    p1, p2 = np.array(I['p1'], np.int32), np.array(I['p2'], np.int32)
    lb, rt, rs = f32(I['layer_betas']), f32(I['rel_trans']), f32(I['rel_sus'])
    U = np.array(I['uniforms'])
    if kind == 'oracle':
        s, t, p = co.compute_infections(I['beta'], p1, p2, lb, rt, rs, lambda d, idx: U[d][idx])
    else:
        import covasim_amd as cv
        s, t, p = cv.utils.compute_infections(I['beta'], p1, p2, lb, rt, rs, uniforms=U, return_probs=True)
    assert [[int(a), int(b)] for a, b in zip(s, t)] == I['pairs']
    assert len(I['pairs']) > 0


def test_oracle_viral_load_fixture():
    check_viral_load('oracle')


def test_oracle_trans_sus_fixture():
    check_trans_sus('oracle')


def test_oracle_infections_fixture():
    check_infections('oracle')


def test_oracle_state_transitions_fixture():
    X = G['transitions']
    n = len(X['dates'])
    pars = co.make_pars(pop_size=n, use_waning=False)
    progs = co.get_prognoses()
    age = np.full(n, 40.0, F32)
    arr = co.set_prognoses(age, progs)
    people = co.People(pars, age, np.ones(n, F32), arr['rel_sus'], arr['symp_prob'], arr['severe_prob'],
                       arr['crit_prob'], arr['death_prob'])
    dates = np.array([f32(d) for d in X['dates']])
    people.susceptible[:] = False
    people.naive[:] = False
    people.exposed[:] = True
    people.date_exposed[:] = 0
    for k, name in enumerate(['date_infectious', 'date_symptomatic', 'date_severe', 'date_critical',
                              'date_recovered', 'date_dead']):
        getattr(people, name)[:] = dates[:, k]
    for t, day in enumerate(X['flags']):
        people.update_states_pre(t)
        for i, want in enumerate(day):
            for key, v in want.items():
                assert bool(getattr(people, key)[i]) == v, (t, i, key)


@pytest.mark.gpu
def test_device_mirrors_reproduce_fixtures():
    check_viral_load('device')
    check_trans_sus('device')
    check_infections('device')
