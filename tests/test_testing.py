"""Testing, contact tracing, quarantine and isolation (covasim/interventions.py:682-1145,
people.py:189-196, 315-375, 589-641).

CPU: the oracle's restatement is pinned by closed forms of whole runs (with certain tests every
symptomatic agent is diagnosed on its symptom day, or test_delay days later, and isolated then;
tests with zero sensitivity or certain loss diagnose nobody), by hand-worked quarantine cases of
check_quar and by a brute-force find_contacts.
GPU: runs with test_prob and contact_tracing (per-layer trace times, presumptive tracing, the
quarantine testing policies, waning and a second variant, clip_edges days between, a snapshot) equal
the oracle bit for bit: results rows, dates and flags, in a captured graph, eagerly and stepped day by
day; the contact scan equals a numpy scan at 20k to 5M agents.
"""

import numpy as np
import pytest

from oracle import covasim_oracle as co
from oracle import interventions_ref as oiv
from tests.helpers import COMPARE_KEYS, STOCK_KEYS, copy_pop, make_population, oracle_sim

nan = np.nan


def run_oracle(pop, ivs, **pars):
    osim = oracle_sim(pop, **pars)
    osim.interventions = [iv.initialize(osim) for iv in ivs]
    osim.run()
    return osim, osim.finalize()


def reached(osim, key):
    """Agents whose date `key` fell inside the run."""
    d = getattr(osim.people, key)
    return ~np.isnan(d) & (d <= osim.npts - 1)


def test_certain_tests_diagnose_on_the_symptom_day():
    pop = make_population(n=4000)
    osim, res = run_oracle(pop, [oiv.test_prob(symp_prob=1.0)], seed=3, n_days=50, beta=0.03, use_waning=False)
    p = osim.people
    sym = reached(osim, 'date_symptomatic')
    assert sym.sum() > 100
    assert np.array_equal(p.date_diagnosed[sym], p.date_symptomatic[sym])
    assert np.isnan(p.date_diagnosed[~sym]).all()
    assert np.array_equal(p.date_end_isolation[sym], p.date_recovered[sym], equal_nan=True)
    assert np.array_equal(res['new_tests'], res['new_symptomatic'])
    assert np.array_equal(res['new_diagnoses'], res['new_symptomatic'])
    assert np.array_equal(res['new_isolated'], res['new_symptomatic'])
    assert res['new_quarantined'].sum() == 0


def test_delayed_results_and_repeat_tests():
    pop = make_population(n=4000)
    delay = 3
    osim, res = run_oracle(pop, [oiv.test_prob(symp_prob=1.0, test_delay=delay)], seed=3, n_days=50, beta=0.03,
                           use_waning=False)
    p = osim.people
    sym = reached(osim, 'date_symptomatic')
    assert np.array_equal(p.date_diagnosed[sym], p.date_symptomatic[sym] + delay)
    assert np.array_equal(res['new_diagnoses'], res['new_symptomatic'])  # counted on the test day
    # the symptomatic test every day until their result arrives, that day included (the tests come
    # before update_states_post sets the flag), or until they stop being symptomatic
    ds, dr, dd = p.date_symptomatic, p.date_recovered, p.date_dead
    stop = np.fmin(np.where(np.isnan(dr), np.inf, dr), np.where(np.isnan(dd), np.inf, dd))
    for t in range(osim.npts):
        expect = np.count_nonzero(sym & (ds <= t) & (t <= ds + delay) & (t < stop))
        assert res['new_tests'][t] == expect, t
    isol = (~np.isnan(p.date_diagnosed)) & (p.date_diagnosed <= osim.npts - 1)
    assert res['new_isolated'].sum() == isol.sum()


def test_no_sensitivity_or_all_lost_diagnoses_nobody():
    pop = make_population(n=3000)
    for iv in (oiv.test_prob(symp_prob=1.0, asymp_prob=0.2, sensitivity=0.0),
               oiv.test_prob(symp_prob=1.0, asymp_prob=0.2, loss_prob=1.0)):
        osim, res = run_oracle(pop, [iv], seed=2, n_days=30, beta=0.03)
        assert res['new_tests'].sum() > 1000
        assert res['new_diagnoses'].sum() == 0 and np.isnan(osim.people.date_diagnosed).all()


def test_test_rates_follow_probabilities():
    pop = make_population(n=20000)
    osim, res = run_oracle(pop, [oiv.test_prob(symp_prob=0.0, asymp_prob=0.05)], seed=5, n_days=5, beta=0.0)
    n = 20000
    for t in range(osim.npts):  # binomial(n, 0.05): within 5 standard deviations
        assert abs(res['new_tests'][t] - 0.05 * n) < 5 * np.sqrt(n * 0.05 * 0.95)


def small_people(n):
    pars = co.make_pars(pop_size=n, use_waning=False)
    z = np.zeros(n, np.float32)
    return co.People(pars, age=np.full(n, 30.0, np.float32), rel_trans=z + 1, rel_sus=z + 1, symp_prob=z,
                     severe_prob=z, crit_prob=z, death_prob=z)


def test_check_quar_cases():
    """people.py:335-358 worked by hand over 8 agents on day 5."""
    p = small_people(8)
    p.t = 5
    p.quarantined[1] = True
    p.date_quarantined[1], p.date_end_quarantine[1] = 2, 9
    p.dead[2] = True
    p.recovered[3] = True
    p.diagnosed[4] = True
    p.isolated[5] = True
    p.quarantined[6] = True
    p.date_quarantined[6], p.date_end_quarantine[6] = 1, 5  # ends today
    p.quarantined[7] = True
    p.date_quarantined[7], p.date_end_quarantine[7], p.date_diagnosed[7] = 3, 12, 5  # diagnosed today
    for ind, end in [(0, 15), (0, 19), (0, 17), (1, 8), (1, 11), (2, 19), (3, 19), (4, 19), (5, 19)]:
        p._pending_quarantine[5].append((ind, end))
    p._pending_quarantine[6].append((1, 30))  # another day: untouched
    n = p.check_quar()
    assert n == 1
    assert p.quarantined.tolist() == [True, True, False, False, False, False, False, False]
    assert p.date_quarantined[0] == 5 and p.date_end_quarantine[0] == 19  # started, extended to the latest end
    assert p.date_end_quarantine[1] == 11  # extended, never shortened
    assert np.isnan(p.date_quarantined[2:6]).all()
    assert p.date_end_quarantine[7] == 5  # diagnosis ends quarantine (isolation takes over)
    assert p._pending_quarantine[6] == [(1, 30)] and 5 not in p._pending_quarantine


def test_check_diagnosed_and_enter_iso():
    p = small_people(4)
    p.t = 7
    p.date_pos_test[:] = [7, 5, nan, 8]
    p.date_diagnosed[:] = [7, 9, 6, 8]
    p.date_recovered[:] = [20, 21, nan, 22]
    flows = p.check_diagnosed()
    assert sorted(flows.tolist()) == [0, 1]  # positive tests taken up to today, counted once
    assert np.isnan(p.date_pos_test[:2]).all() and p.date_pos_test[3] == 8
    assert p.diagnosed.tolist() == [True, False, True, False]
    iso = p.check_enter_iso()
    assert iso.tolist() == [0] and p.isolated.tolist() == [True, False, False, False]
    assert p.date_end_isolation[0] == 20


def test_find_contacts_brute_force():
    rng = np.random.default_rng(4)
    for _ in range(20):
        n = int(rng.integers(2, 30))
        E = int(rng.integers(0, 60))
        p1 = rng.integers(0, n, E).astype(np.int32)
        p2 = rng.integers(0, n, E).astype(np.int32)
        inds = rng.choice(n, int(rng.integers(0, n)), replace=False)
        want = sorted({int(b) for a, b in zip(p1, p2) if a in inds} | {int(a) for a, b in zip(p1, p2) if b in inds})
        assert oiv.find_contacts(p1, p2, inds).tolist() == want


def test_certain_tracing_quarantines_household_contacts():
    """trace_probs 1 and trace_time 0: on the day an agent is diagnosed, its household contacts who are
    alive, undiagnosed, not recovered and not isolated enter quarantine that same day."""
    pop = make_population(n=3000, pop_type='hybrid')
    ivs = [oiv.test_prob(symp_prob=0.5), oiv.contact_tracing(trace_probs=dict(h=1.0, s=0, w=0, c=0), trace_time=0)]
    osim, res = run_oracle(pop, ivs, seed=4, n_days=40, beta=0.03, pop_type='hybrid')
    p = osim.people
    assert res['new_quarantined'].sum() > 20
    h1, h2, _ = osim.layers['h']
    diag_day = p.date_diagnosed
    for a, b in zip(h1.tolist(), h2.tolist()):
        for i, j in ((a, b), (b, a)):
            d = diag_day[i]
            if not np.isnan(d) and d <= osim.npts - 1:
                assert p.date_known_contact[j] <= d or p.dead[j]


def test_host_ops():
    import covasim_amd as cv
    from covasim_amd import _lib

    class FakeSim:
        pars = dict(quar_period=14)

        def layer_keys(self):
            return ['h', 's', 'w', 'c']

    tp = cv.test_prob(0.2, asymp_prob=0.01, quar_policy=[0, 3], loss_prob=0.25, test_delay=2, end_day=30)
    tp.initialize(FakeSim())
    o = tp.op(3, FakeSim().layer_keys())
    assert (o.kind, o.slot, o.end_day, o.quar_policy, o.n_quar_days) == (_lib.OP_TEST, 3, 30, _lib.QP_DAYS, 2)
    assert o.keep_prob == 0.75 and o.symp_quar_prob == 0.2 and o.test_delay == 2.0
    ct = cv.contact_tracing(trace_probs=dict(h=1.0, c=0.3), trace_time=dict(h=0, c=2.5))
    ct.initialize(FakeSim())
    o = ct.op(4, FakeSim().layer_keys())
    assert list(o.trace_prob)[:4] == [1.0, 0.0, 0.0, 0.3] and o.trace_period[3] == 11 and ct.max_delay() == 2
    o = cv.contact_tracing(capacity=10.7)
    o.initialize(FakeSim())
    assert o.op(1, FakeSim().layer_keys()).capacity == 10
    assert cv.contact_tracing().op.__func__  # no capacity: -1
    ct0 = cv.contact_tracing()
    ct0.initialize(FakeSim())
    assert ct0.op(0, FakeSim().layer_keys()).capacity == -1
    with pytest.raises(ValueError):
        cv.contact_tracing(capacity=-1)
    with pytest.raises(NotImplementedError):
        cv.test_prob(0.1, quar_policy=lambda sim: []).op(0, [])
    with pytest.raises(NotImplementedError):
        cv.test_prob(0.1, swab_delay=dict(dist='gamma', par1=1, par2=1))


def test_choice_by_keys_is_exact_and_uniform():
    from scipy import stats
    items = np.arange(3, 3000, 7)
    for k in (0, 1, 50, len(items) - 1, len(items), len(items) + 5):
        got = oiv.choose_items(11, 4, 0xC0000001, items, k)
        assert len(got) == min(k, len(items)) and np.isin(got, items).all() and len(np.unique(got)) == len(got)
    n, k, days = 200, 30, 600
    hits = np.zeros(n)
    for t in range(days):
        hits[oiv.choose_items(5, t, 0xD0000002, np.arange(n), k)] += 1
    exp = days * k / n
    chi2 = ((hits - exp) ** 2 / (exp * (1 - k / n))).sum()
    assert stats.chi2.sf(chi2, n - 1) > 1e-3


def test_ili_agents_are_tested_as_symptomatic():
    # certain tests, no true positives: on a day without COVID symptoms exactly int(ili_prev * n) agents
    # test (the day's ILI agents); later days test the union of the symptomatic and the ILI agents
    pop = make_population(n=3000)
    iv = oiv.test_prob(symp_prob=1.0, asymp_prob=0.0, sensitivity=0.0, ili_prev=0.05, start_day=2)
    osim, res = run_oracle(pop, [iv], seed=4, n_days=40, beta=0.03)
    n_ili = int(0.05 * 3000)
    assert res['new_tests'][0] == res['new_tests'][1] == 0
    assert res['new_tests'][2] == n_ili  # nobody is symptomatic on day 2
    sym = res['n_symptomatic']
    assert np.all(res['new_tests'][2:] >= n_ili) and np.all(res['new_tests'][2:] <= n_ili + sym[2:])
    assert sym[-1] > 50 and res['new_tests'][-1] > n_ili
    iv2 = oiv.test_prob(symp_prob=1.0, ili_prev=[0.0, 0.1, 0.2], start_day=5, sensitivity=0.0)
    osim, res = run_oracle(pop, [iv2], seed=4, n_days=12, beta=0.0, pop_infected=0)  # nobody is ever symptomatic
    assert list(res['new_tests'][:9]) == [0, 0, 0, 0, 0, 0, 300, 600, 0]


def test_swab_delay_probabilities():
    from covasim_amd.utils import get_pdf
    # the pdf's parameters: lognormal of mean par1 and variance par2, shifted by -0.5; uniform on
    # [par1, par1 + par2]; the oracle's and the library's are the same distribution
    for pdf in (oiv.get_pdf('lognormal', 10, 170), get_pdf('lognormal', 10, 170)):
        assert abs(pdf.mean() - 9.5) < 1e-9 and abs(pdf.var() - 170) < 1e-6
    for pdf in (oiv.get_pdf('uniform', 2, 5), get_pdf('uniform', 2, 5)):
        assert pdf.pdf(1.99) == 0 and pdf.pdf(2.0) == pdf.pdf(6.9) == 0.2 and pdf.pdf(7.01) == 0
    assert oiv.get_pdf() is None and get_pdf(None) is None
    # per day since onset s the symptomatic agents share pdf(s) * f(s) (f: symp_prob / (1 - s * symp_prob)
    # below 1, else 1) equally: the sum of their probabilities is n_symp * sum over the present s

    class Probe(oiv.test_prob):
        def __call__(self, sim):
            out = super().__call__(sim)
            p = sim.people
            symp = p.symptomatic.nonzero()[0]
            self.checks.append((sim.t, symp, (sim.t - p.date_symptomatic[symp]).astype(np.int32),
                                self.last_probs.copy(), p.diagnosed.copy()))
            return out

    pop = make_population(n=4000)
    sp = 0.2
    iv = Probe(symp_prob=sp, swab_delay=dict(dist='lognormal', par1=3, par2=4), sensitivity=0.0)
    iv.checks = []
    osim, res = run_oracle(pop, [iv], seed=5, n_days=40, beta=0.04)
    assert len(iv.checks) == osim.npts
    checked = 0
    for t, symp, s, tp, diag in iv.checks:
        if not len(symp):
            continue
        assert not diag.any()  # sensitivity 0: nobody diagnosed, so no probability was zeroed
        present = np.unique(s)
        with np.errstate(divide='ignore'):
            f = np.where(present * sp < 1, sp / (1 - present * sp), 1.0)
        want = len(symp) * np.sum(iv.pdf.pdf(present) * f)
        assert np.isclose(tp[symp].sum(), want, rtol=1e-12), t
        # agents at the same day since onset share one probability
        for v in present[:5]:
            assert len(np.unique(tp[symp[s == v]])) == 1
        checked += 1
    assert checked > 20


def test_subtarget_overrides_and_capacity_limits_tracing():
    pop = make_population(n=3000)
    inds = np.array([5, 17, 17, 900, 2999])
    vals = np.array([0.0, 0.5, 1.0, 1.0, 1.0])  # index 17 keeps its last value
    iv = oiv.test_prob(symp_prob=0.0, subtarget=dict(inds=inds, vals=vals), sensitivity=0.0)
    osim, res = run_oracle(pop, [iv], seed=2, n_days=10, beta=0.0, pop_infected=0)
    assert list(res['new_tests']) == [3] * 11  # 17, 900, 2999 every day; 5 never
    assert np.isnan(osim.people.date_tested[5]) and osim.people.date_tested[[17, 900, 2999]].tolist() == [10] * 3
    # a function of the sim, and a scalar value
    iv = oiv.test_prob(symp_prob=0.0, subtarget=lambda sim: dict(inds=np.arange(sim.t), vals=1.0), sensitivity=0.0)
    osim, res = run_oracle(pop, [iv], seed=2, n_days=10, beta=0.0, pop_infected=0)
    assert list(res['new_tests']) == list(range(11))
    # capacity: the traced cases are min(capacity, cases) of them
    tp = oiv.test_prob(symp_prob=0.5, asymp_prob=0.05)
    traced = []

    class Cap(oiv.contact_tracing):
        def __call__(self, sim):
            out = super().__call__(sim)
            if self.last_traced is not None:
                cases = (sim.people.date_diagnosed == sim.t).nonzero()[0]
                traced.append((len(cases), self.last_traced))
                assert np.isin(self.last_traced, cases).all()
            return out
    osim, res = run_oracle(pop, [tp, Cap(trace_probs=1.0, capacity=4)], seed=2, n_days=40, beta=0.04)
    assert max(c for c, _ in traced) > 4
    assert all(len(tr) == min(c, 4) for c, tr in traced)


# ---------------------------------------------------------------------------------------------
# GPU parity
# ---------------------------------------------------------------------------------------------
DIAG_ROWS = ['new_tests', 'new_diagnoses', 'new_quarantined', 'new_isolated']
DIAG_STOCKS = ['n_diagnosed', 'n_quarantined', 'n_isolated', 'n_known_dead']
DIAG_DATES = ['date_tested', 'date_pos_test', 'date_known_contact', 'date_quarantined', 'date_end_quarantine',
              'date_diagnosed', 'date_end_isolation']
DIAG_FLAGS = ['tested', 'diagnosed', 'known_contact', 'quarantined', 'isolated', 'known_dead']

CASES = {
    'tracing': dict(
        ivs=lambda m: [m.test_prob(symp_prob=0.2, asymp_prob=0.005, sensitivity=0.9, loss_prob=0.1, test_delay=2),
                       m.contact_tracing(trace_probs=dict(h=1.0, s=0.5, w=0.5, c=0.1),
                                         trace_time=dict(h=0, s=1, w=1, c=3))],
        pars=dict(n_days=60, beta=0.03, use_waning=False)),
    'presumptive': dict(
        ivs=lambda m: [m.test_prob(symp_prob=0.3, asymp_prob=0.01, symp_quar_prob=0.9, asymp_quar_prob=0.4,
                                   quar_policy='both', start_day=5, end_day=45),
                       m.contact_tracing(trace_probs=0.6, trace_time=1.5, presumptive=True, quar_period=10)],
        pars=dict(n_days=50, beta=0.03, use_waning=True)),
    'options': dict(
        ivs=lambda m: [m.test_prob(symp_prob=0.25, asymp_prob=0.004, ili_prev=0.01, test_delay=1,
                                   swab_delay=dict(dist='lognormal', par1=3, par2=4),
                                   subtarget=dict(inds=np.arange(0, 20000, 7), vals=np.linspace(0, 0.1, 2858))),
                       m.contact_tracing(trace_probs=dict(h=1.0, s=0.4, w=0.4, c=0.1), trace_time=1, capacity=15)],
        pars=dict(n_days=50, beta=0.03, use_waning=False)),
    'options_dynamic': dict(
        ivs=lambda m: [m.test_prob(symp_prob=0.2, ili_prev=[0.0] * 8 + [0.03] * 30, start_day=3,
                                   swab_delay=dict(dist='uniform', par1=0, par2=6),
                                   subtarget=lambda sim: dict(inds=np.flatnonzero(sim.people.age > 70), vals=0.05)),
                       m.contact_tracing(trace_probs=0.8, trace_time=dict(h=0, s=2, w=2, c=2), presumptive=True,
                                         capacity=3)],
        pars=dict(n_days=40, beta=0.035, use_waning=True)),
    'policies': dict(
        ivs=lambda m: [m.test_prob(symp_prob=0.15, quar_policy=[0, 4], test_delay=1),
                       m.contact_tracing(trace_probs=dict(h=0.9, w=0.4), trace_time=dict(h=1, w=2), start_day=10),
                       m.test_prob(symp_prob=0.05, asymp_prob=0.02, quar_policy='daily', sensitivity=0.8)],
        pars=dict(n_days=45, beta=0.035, use_waning=True),
        variants=[dict(rel_beta=1.7, days=[12], n_imports=30)]),
    # quarantine and isolation that raise transmission (factors above 1): the codes testing and tracing
    # rewrite exceed the ones the agent pass wrote
    'factors_above_one': dict(
        ivs=lambda m: [m.test_prob(symp_prob=0.3, asymp_prob=0.02, test_delay=1),
                       m.contact_tracing(trace_probs=dict(h=1.0, s=0.6, w=0.6, c=0.2), trace_time=1)],
        pars=dict(n_days=45, beta=0.03, use_waning=True, quar_factor=dict(h=1.8, s=1.2, w=1.2, c=1.5),
                  iso_factor=dict(h=2.0, s=1.1, w=1.1, c=1.3)),
        variants=[dict(rel_beta=1.5, days=[8], n_imports=40)]),
}


def compare(sim, osim, ores):
    for key in COMPARE_KEYS + DIAG_ROWS:
        assert np.array_equal(sim.results[key].values, ores[key]), key
    for key in STOCK_KEYS[1:]:
        assert np.array_equal(sim.results[f'n_{key}'].values, ores[f'n_{key}']), key
    for key in DIAG_STOCKS:
        assert np.array_equal(sim.results[key].values, ores[key]), key
    for key in DIAG_DATES:
        assert np.array_equal(sim.people[key], getattr(osim.people, key), equal_nan=True), key
    for key in DIAG_FLAGS:
        assert np.array_equal(sim.people[key], getattr(osim.people, key)), key


@pytest.mark.gpu
@pytest.mark.parametrize('case', sorted(CASES))
@pytest.mark.parametrize('mode', ['graph', 'eager', 'step'])
def test_testing_tracing_matches_oracle(case, mode):
    import covasim_amd as cv
    print(f'testing/tracing case={case} mode={mode}', flush=True)
    spec = CASES[case]
    pop = make_population(n=20000, pop_type='hybrid')
    kw = dict(pop_type='hybrid', dynam_layer=dict(c=1), **spec['pars'])
    variants = spec.get('variants', [])
    gvars = [cv.variant(dict(rel_beta=v['rel_beta']), days=v['days'], n_imports=v['n_imports']) for v in variants]
    sim = cv.Sim(pop_size=20000, rand_seed=7, popdict=copy_pop(pop), interventions=spec['ivs'](cv),
                 variants=gvars, **kw)
    if mode == 'step':
        sim.initialize()
        while not sim.complete:
            sim.step()
        sim.finalize()
    else:
        sim.run(use_graph=(mode == 'graph'))
    osim, ores = run_oracle(pop, spec['ivs'](oiv), seed=7, variants=variants, **kw)
    assert ores['new_tests'].sum() > 0 and ores['new_diagnoses'].sum() > 0
    if case != 'presumptive':
        assert ores['new_quarantined'].sum() > 0
    compare(sim, osim, ores)
    for key in ('test_yield', 'rel_test_yield'):
        assert np.all(np.isfinite(sim.results[key].values)), key


@pytest.mark.gpu
@pytest.mark.parametrize('case', ['factors_above_one', 'policies', 'options'])
def test_testing_tracing_target_bits_matches_oracle(case):
    """The fused day loop with the bitmap out of LDS, so the edge pass reads target bits bounded by the day's
    largest target code: that bound must cover the codes testing and tracing rewrite after the agent pass
    (quarantine / isolation factors up to 2)."""
    import covasim_amd as cv
    from covasim_amd import _lib
    print(f'testing/tracing target bits case={case}', flush=True)
    spec = CASES[case]
    pop = make_population(n=20000, pop_type='hybrid')
    kw = dict(pop_type='hybrid', dynam_layer=dict(c=1), **spec['pars'])
    variants = spec.get('variants', [])
    gvars = [cv.variant(dict(rel_beta=v['rel_beta']), days=v['days'], n_imports=v['n_imports']) for v in variants]
    sim = cv.Sim(pop_size=20000, rand_seed=7, popdict=copy_pop(pop), interventions=spec['ivs'](cv),
                 variants=gvars, **kw)
    _lib.check(_lib.lib().cvk_set_tuning(_lib.TUNE_EDGE_LDS, 0), 'tuning')
    try:
        sim.run(use_graph=True)
    finally:
        _lib.lib().cvk_set_tuning(_lib.TUNE_EDGE_LDS, 1)
    osim, ores = run_oracle(pop, spec['ivs'](oiv), seed=7, variants=variants, **kw)
    assert ores['new_diagnoses'].sum() > 0 and ores['new_quarantined'].sum() > 0
    compare(sim, osim, ores)


@pytest.mark.gpu
def test_tracing_with_clip_edges_and_snapshot(tmp_path):
    import covasim_amd as cv
    print('testing/tracing with clip_edges and a snapshot', flush=True)
    pop = make_population(n=20000, pop_type='hybrid')
    kw = dict(pop_type='hybrid', dynam_layer=dict(c=1), n_days=40, beta=0.03)

    def ivs(m):
        return [m.test_prob(symp_prob=0.25, test_delay=1), m.clip_edges([8, 20], [0.5, 1.0], layers=['w', 'c']),
                m.contact_tracing(trace_probs=dict(h=1.0, w=0.7, c=0.2), trace_time=dict(h=0, w=1, c=2))]
    sim = cv.Sim(pop_size=20000, rand_seed=9, popdict=copy_pop(pop), interventions=ivs(cv), **kw)
    sim.run(until=14)
    fn = str(tmp_path / 'mid.safetensors')
    sim.save(fn)
    sim.run()
    osim, ores = run_oracle(pop, ivs(oiv), seed=9, **kw)
    compare(sim, osim, ores)
    resumed = cv.load(fn)
    resumed.run()
    for key in COMPARE_KEYS + DIAG_ROWS + DIAG_STOCKS:
        assert np.array_equal(resumed.results[key].values, ores[key]), key


@pytest.mark.gpu
@pytest.mark.parametrize('n', [20_000, 150_000, 1_000_000, 5_000_000])
def test_trace_scan_finds_every_contact(n):
    """The contact scan at sizes whose coarse case map has 1 to 128 agents a bit: with trace_probs 1
    and trace_time 0, the partners of the agents diagnosed today become known contacts and, unless
    diagnosed themselves, enter quarantine today (a numpy scan of the same edges is the reference)."""
    import covasim_amd as cv
    print(f'trace scan n={n}', flush=True)
    rng = np.random.default_rng(n)
    E = 3 * n
    p1 = rng.integers(0, n, E).astype(np.int32)
    p2 = rng.integers(0, n, E).astype(np.int32)
    age = rng.uniform(0, 90, n).astype(np.float32)
    pop = dict(age=age, contacts={'a': dict(p1=p1, p2=p2, beta=np.ones(E, np.float32))})
    sim = cv.Sim(pop_size=n, pop_type='random', n_days=3, pop_infected=0, rand_seed=1, popdict=pop,
                 interventions=[cv.contact_tracing(trace_probs=1.0, trace_time=0)])
    sim.initialize()
    cases = rng.choice(n, max(1, n // 2000), replace=False)
    dd = np.full(n, np.nan, np.float32)
    dd[cases] = 0
    sim.people.set('date_diagnosed', dd)
    sim.step()
    is_case = np.zeros(n, bool)
    is_case[cases] = True
    contact = np.zeros(n, bool)
    contact[p2[is_case[p1]]] = True
    contact[p1[is_case[p2]]] = True
    assert np.array_equal(sim.people.known_contact, contact)
    assert np.array_equal(sim.people.quarantined, contact & ~is_case)
    assert np.array_equal(sim.people.isolated, is_case) and np.array_equal(sim.people.diagnosed, is_case)
    from covasim_amd import _lib
    row = sim._w['results'][0].cpu().numpy()
    assert row[_lib.R_COL['new_quarantined']] == np.count_nonzero(contact & ~is_case)
    assert row[_lib.R_COL['new_isolated']] == len(cases)
