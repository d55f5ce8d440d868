"""Host logic: parameters, tables and population generation agree with the oracle and the reference's rules."""

import numpy as np
import pytest

from covasim_amd import immunity as cvimm
from covasim_amd import parameters as cvpar
from covasim_amd import population as cvpop
from oracle import covasim_oracle as co


def test_make_pars_matches_oracle_on_path_keys():
    for pop_type in ['random', 'hybrid']:
        a = cvpar.make_pars(pop_type=pop_type, beta=0.02)
        b = co.make_pars(pop_type=pop_type, beta=0.02)
        for k in ['pop_size', 'pop_infected', 'n_days', 'rand_seed', 'beta', 'asymp_factor', 'use_waning',
                  'nab_boost', 'trans_redux', 'rel_beta', 'rel_symp_prob', 'rel_severe_prob', 'rel_crit_prob',
                  'rel_death_prob', 'no_hosp_factor', 'no_icu_factor', 'viral_dist', 'dur', 'nab_init', 'nab_eff',
                  'rel_imm_symp', 'beta_layer', 'iso_factor', 'quar_factor', 'contacts', 'dynam_layer']:
            assert a[k] == b[k], k
        assert list(a['beta_layer']) == (['a'] if pop_type == 'random' else ['h', 's', 'w', 'c'])


def test_prognoses_match_oracle_and_are_conditional():
    a, b = cvpar.get_prognoses(), co.get_prognoses()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    # conditional probabilities recover the absolute table (parameters.py:285-294)
    abs_death = a['death_probs'] * a['crit_probs'] * a['severe_probs'] * a['symp_probs']
    assert np.allclose(abs_death, [0.00002, 0.00002, 0.00010, 0.00032, 0.00098, 0.00265, 0.00766, 0.02439,
                                   0.08292, 0.16190])


def test_nab_kinetics_table_matches_oracle():
    pars = cvpar.make_pars()
    for npts in [61, 181, 400]:
        assert np.array_equal(cvimm.precompute_waning(npts, pars['nab_decay']),
                              co.precompute_waning(npts, co.make_pars()['nab_decay']))


def test_random_contacts_counts():
    rng = np.random.default_rng(0)
    c = cvpop.make_random_contacts(100_000, 20, rng)
    assert c['p1'].dtype == np.int32 and len(c['p1']) == len(c['p2'])
    assert abs(len(c['p1']) / 100_000 - 10.0) < 0.1  # round(Poisson(20)/2) per agent, mean 10
    assert np.all(np.diff(c['p1']) >= 0)  # sources in agent order, as the reference builds them
    assert c['p2'].min() >= 0 and c['p2'].max() < 100_000


def test_households_categorical():
    rng = np.random.default_rng(1)
    ages = cvpop.draw_ages(200_000, rng)
    new_ages, h = cvpop.make_households(200_000, ages, rng)
    assert np.array_equal(np.sort(new_ages), np.sort(ages))  # ages only permuted
    sizes = np.bincount(np.concatenate([h['p1'], h['p2']]), minlength=200_000)  # degree = size - 1
    assert abs((sizes + 1).mean() - 1.0) > 0  # sanity
    assert np.all(h['p1'] < h['p2'])
    # household of agent i = connected block; mean size ~2.6: E[(k-1)] over agents = E[k^2]/E[k] - 1
    p = cvpop.HOUSEHOLD_SIZE_PROBS
    k = np.arange(1, 8)
    assert abs(sizes.mean() - ((k * k * p).sum() / (k * p).sum() - 1)) < 0.05
    # age correlation inside households
    a1, a2 = new_ages[h['p1']], new_ages[h['p2']]
    assert np.corrcoef(a1, a2)[0, 1] > 0.5


def test_hybrid_layers_and_ages():
    pars = cvpar.make_pars(pop_size=50_000, pop_type='hybrid')
    pop = cvpop.make_randpop(pars, np.random.default_rng(2))
    assert list(pop['contacts']) == ['h', 's', 'w', 'c']
    age = pop['age']
    s = pop['contacts']['s']
    assert np.all((age[s['p1']] >= 6) & (age[s['p1']] < 22)) and np.all((age[s['p2']] >= 6) & (age[s['p2']] < 22))
    w = pop['contacts']['w']
    assert np.all((age[w['p1']] >= 22) & (age[w['p1']] < 65))
    # age pyramid marginal (defaults.py:223-243)
    lo = cvpop.DEFAULT_AGE_DATA[:, 0]
    prob = cvpop.DEFAULT_AGE_DATA[:, 2] / cvpop.DEFAULT_AGE_DATA[:, 2].sum()
    counts = np.bincount(np.searchsorted(lo, age, side='right') - 1, minlength=len(lo))
    assert np.all(np.abs(counts / len(age) - prob) < 0.01)


def test_people_arrays_match_oracle_prognoses():
    pars = cvpar.make_pars()
    pars['prognoses'] = cvpar.get_prognoses()
    age = np.array([0.5, 9.99, 10.0, 35, 64.9, 89, 99.5])
    arr = cvpop.make_people_arrays(age, pars, np.random.default_rng(0))
    ref = co.set_prognoses(age, co.get_prognoses())
    for k in ['symp_prob', 'severe_prob', 'crit_prob', 'death_prob', 'rel_sus']:
        assert np.array_equal(arr[k], ref[k]), k
    big = cvpop.make_people_arrays(np.full(200_000, 30.0), pars, np.random.default_rng(1))['rel_trans']
    assert abs(big.mean() - 1.0) < 0.02  # NegBin mean 1 (parameters.py:60)
    # 0.01 * NegBin(mean 100, k 0.45): variance 0.01^2 * (100 + 100^2 / 0.45)
    assert abs(big.var() - (0.01 + 1 / 0.45)) < 0.15


def test_variant_parsing():
    v = cvimm.variant('alpha', days=10)
    assert v.label == 'alpha' and v.p['rel_beta'] == 1.67
    v = cvimm.variant(dict(rel_beta=1.7), days=[30], n_imports=100)
    assert v.label == 'custom' and v.p['rel_severe_prob'] == 1.0 and v.day_list(121) == [30]
    with pytest.raises(KeyError):
        cvimm.variant(dict(rel_betta=2), days=1)
    m = cvimm.init_immunity_matrix(['wild', 'alpha', 'custom'])
    assert m[1, 0] == np.float32(0.5) and m[2, 2] == 1 and m[0, 2] == 1
    # aliases of known variants (parameters.py:317-330), after the same clean-up of the name
    for name, label in [('B.1.1.7', 'alpha'), ('UK', 'alpha'), ('b1351', 'beta'), ('P1', 'gamma'),
                        ('Delta variant', 'delta'), ('india', 'delta'), ('original', 'wild')]:
        assert cvimm.variant(name, days=5).label == label, name
    with pytest.raises(NotImplementedError):
        cvimm.variant('united kingdom', days=5)  # the space is removed before the lookup, as in the reference
    with pytest.raises(NotImplementedError):
        cvimm.variant('omicron', days=5)


def test_variant_labels_set_the_indices():
    """variant.initialize (immunity.py:107-114): parameters stored by label, index = the label's position; two
    dicts without a label are both 'custom', so the second replaces the first's parameters and shares index 1;
    the oracle follows the same rule."""
    import covasim_amd as cv
    vs = [cv.variant(dict(rel_beta=1.7), days=8), cv.variant('alpha', days=9), cv.variant(dict(rel_beta=1.3), days=15),
          cv.variant(dict(rel_beta=2.0, label='x'), days=20)]
    sim = cv.Sim(pop_size=1000, n_days=30, variants=vs)
    sim.init_variants()
    assert sim.pars['n_variants'] == 4 and [v.index for v in vs] == [1, 2, 1, 3]
    assert sim.pars['variant_map'] == {0: 'wild', 1: 'custom', 2: 'alpha', 3: 'x'}
    assert sim.pars['variant_pars']['custom']['rel_beta'] == 1.3
    opars = co.make_pars(variants=[dict(rel_beta=1.7, days=[8]), dict(rel_beta=1.3, days=[15]),
                                   dict(label='x', rel_beta=2.0, days=[20])])
    assert opars['n_variants'] == 3 and opars['variant_index'] == [1, 1, 2]
    assert [vp['rel_beta'] for vp in opars['variant_pars']] == [1.0, 1.3, 2.0]


def test_sim_rejects_out_of_path_parameters():
    import covasim_amd as cv
    with pytest.raises(NotImplementedError):
        cv.Sim(pop_scale=10).validate_pars()
    with pytest.raises(NotImplementedError):
        cv.Sim(interventions=[object()]).validate_pars()


def test_bench_refuses_more_than_one_gpu():
    """The bench runs on one GPU: --gpus 2 prints one JSON line with the reason and exits with status 2."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, os.path.join(root, 'bench.py'), '--gpus', '2'], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 2
    line = json.loads(out.stdout.strip().splitlines()[-1])
    assert 'one GPU' in line['error']
