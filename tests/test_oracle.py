"""Pins the oracle against what the reference's own tests, fixtures and the mathematics fix.

CPU only. Each check is chosen so that a plausible restatement mistake (dropped
factor, swapped operand, wrong index, wrong branch) fails it.
"""

import json
import math
import os

import numpy as np
import pytest

from oracle import covasim_oracle as co
from oracle import philox as px
from tests.helpers import make_population, oracle_sim

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')
F32 = np.float32


# ---------------------------------------------------------------- Philox

def test_philox_matches_rocrand_vectors():
    d = json.load(open(os.path.join(GOLDEN, 'philox_vectors.json')))
    assert len(d['vectors']) == 80
    for r in d['vectors']:
        w = px.philox4x32(*r['counter'], *r['key'])
        assert [int(x) for x in w] == r['out'], r


def test_philox_random123_known_answers():
    # Random123 kat_vectors: philox4x32_10 on (0,0), all-ones and the pi-digits case
    assert [int(x) for x in px.philox4x32(0, 0, 0, 0, 0, 0)] == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert [int(x) for x in px.philox4x32(*[0xffffffff] * 6)] == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert [int(x) for x in px.philox4x32(0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344, 0xa4093822,
                                          0x299f31d0)] == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_uniform_and_integer_transforms():
    assert px.u01(0, 0) == 0.0
    assert px.u01(0xFFFFFFFF, 0xFFFFFFFF) == (2 ** 53 - 1) / 2 ** 53
    assert px.u01(0x80000000, 0) == 0.5
    assert px.below(0, 0, 10) == 0 and px.below(0xFFFFFFFF, 0xFFFFFFFF, 10) == 9
    assert px.below(0x80000000, 0, 10) == 5
    # uniformity of the counter streams: mean and variance of 200k draws
    u = px.edge_uniforms(5, 3, 0, 1, np.arange(200_000), 0)
    assert abs(u.mean() - 0.5) < 0.003 and abs(u.var() - 1 / 12) < 0.002
    z = px.box_muller(*px.slot_uniforms(9, 2, np.arange(200_000), 0))
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01


def _chi2_pvalue(observed, expected):
    from scipy import stats
    keep = expected > 5
    obs = np.append(observed[keep], observed[~keep].sum())
    exp = np.append(expected[keep], expected[~keep].sum())
    if exp[-1] == 0:
        obs, exp = obs[:-1], exp[:-1]
    exp = exp * obs.sum() / exp.sum()
    return stats.chisquare(obs, exp).pvalue


@pytest.mark.parametrize('rng_cls', [co.PhiloxRNG, co.RefRNG])
def test_lognormal_int_durations_match_scipy(rng_cls):
    """sample('lognormal_int', mean, std) (utils.py:223-231): the mean/std are those of the lognormal itself,
    mapped to (mu, sigma) of the underlying normal, then rounded. Compared with scipy's lognorm."""
    from scipy import stats
    rng = rng_cls(3)
    for par1, par2 in [(4.5, 1.5), (8.0, 2.0), (18.1, 6.3)]:  # exp2inf, sym2sev, crit2die defaults
        x = rng.sample(dict(dist='lognormal_int', par1=par1, par2=par2), np.arange(100_000), 4, 3)
        mu, sigma = co.lognormal_params(par1, par2)
        ref = stats.lognorm(s=sigma, scale=np.exp(mu))
        assert abs(ref.mean() - par1) < 1e-9 and abs(ref.std() - par2) < 1e-9
        k = np.arange(0, int(x.max()) + 2)
        expected = (ref.cdf(k + 0.5) - ref.cdf(k - 0.5)) * len(x)
        observed = np.bincount(x.astype(np.int64), minlength=len(k))[:len(k)].astype(float)
        assert _chi2_pvalue(observed, expected) > 1e-4, (par1, par2)
        xc = rng.sample(dict(dist='lognormal', par1=par1, par2=par2), np.arange(50_000), 5, 2)
        assert stats.kstest(xc, ref.cdf).pvalue > 1e-4


def test_normal_family_durations_match_scipy():
    from scipy import stats
    rng = co.PhiloxRNG(11)
    n = 60_000
    x = rng.sample(dict(dist='normal', par1=2.0, par2=3.0), np.arange(n), 1, 0)
    assert stats.kstest(x, stats.norm(2.0, 3.0).cdf).pvalue > 1e-4
    x = rng.sample(dict(dist='normal_pos', par1=1.0, par2=2.0), np.arange(n), 1, 1)
    assert stats.kstest(x, stats.foldnorm(0.5, scale=2.0).cdf).pvalue > 1e-4
    x = rng.sample(dict(dist='uniform', par1=2.0, par2=5.0), np.arange(n), 1, 2)
    assert stats.kstest(x, stats.uniform(2.0, 3.0).cdf).pvalue > 1e-4
    # par1 = 0 lognormal: all zeros (utils.py:228-229)
    assert np.all(rng.sample(dict(dist='lognormal_int', par1=0.0, par2=1.0), np.arange(10), 1, 3) == 0)


def test_bernoulli_rates_and_independent_halves():
    """binomial_arr (utils.py:343-357) is rand < p; each slot's two halves are independent draws."""
    from scipy import stats
    rng = co.PhiloxRNG(21)
    n = 200_000
    for p in (0.02, 0.3, 0.9):
        for half in (0, 1):
            k = int(rng.bernoulli(np.full(n, p), np.arange(n), 7, 1, half).sum())
            assert stats.binomtest(k, n, p).pvalue > 1e-4
    u0, u1 = px.slot_uniforms(21, 7, np.arange(n), 1)
    assert abs(np.corrcoef(u0, u1)[0, 1]) < 0.01
    # p above 1 always succeeds and p = 0 never does (test_mortality.py:21-22 uses p = 1e6)
    assert rng.bernoulli(np.full(100, 1e6), np.arange(100), 7, 2, 0).all()
    assert not rng.bernoulli(np.zeros(100), np.arange(100), 7, 2, 1).any()


def test_bounded_accept_rule_is_exactly_uniform():
    """Lemire's rule: every value in [0, n) keeps the same number of accepted 32-bit words."""
    two32 = 1 << 32
    for n in (1, 3, 7, 1000, 999_983, 10_000_000, 1_500_000_000, 2**31 - 1):
        thr = (two32 - n) % n
        vs = range(n) if n <= 1000 else [0, 1, n // 3, n // 2, n - 2, n - 1]
        # accepted w with value v: v * 2^32 + thr <= w * n < (v + 1) * 2^32
        counts = {-(-((v + 1) * two32) // n) - -(-(v * two32 + thr) // n) for v in vs}
        assert counts == {(two32 - thr) // n}, n
    # the vectorised restatement applies the same rule
    w = np.array([0, 1, 2**31, two32 - 1, 123456789], dtype=np.uint64)
    for n in (7, 1_500_000_000):
        val, ok = px.bounded(w, n)
        thr = (two32 - n) % n
        assert list(val) == [int(x) * n >> 32 for x in w]
        assert list(ok) == [(int(x) * n) % two32 >= thr for x in w]


def test_resample_retry_stream():
    """n = 1.5e9 rejects ~30 % of first words: retries come from their own stream and stay uniform."""
    n_agents, n_edges = 1_500_000_000, 200_000
    p1, p2 = px.resample_edges(4, 9, 3, n_edges, n_agents)
    assert p1.min() >= 0 and p1.max() < n_agents and p2.min() >= 0 and p2.max() < n_agents
    for p in (p1, p2):
        assert _chi2_pvalue(np.bincount(p // 15_000_000, minlength=100).astype(float),
                            np.full(100, n_edges / 100)) > 1e-4
    # the first word of edge e is word (e % 2) * 2 + j of pair e >> 1; rejected ones were replaced
    k0, k1 = px.seed_key(4)
    q = np.arange(n_edges // 2, dtype=np.uint64)
    w = px.philox4x32(q & px.MASK32, 9, px.STREAM_RESAMPLE | 3, q >> np.uint64(32), k0, k1)
    first = np.empty(n_edges, dtype=np.uint64)
    first[0::2], first[1::2] = w[0], w[2]
    v0, ok0 = px.bounded(first, n_agents)
    assert 0.25 < 1 - ok0.mean() < 0.35
    assert np.array_equal(p1[ok0], v0[ok0])
    assert not np.array_equal(p1[~ok0], v0[~ok0])
    e = np.flatnonzero(~ok0)[:50].astype(np.uint64)
    wr = px.philox4x32(e & px.MASK32, 9, px.STREAM_RESAMPLE_RETRY | 3, (e >> np.uint64(32)) << np.uint64(8), k0, k1)
    vr, okr = px.bounded(wr[0], n_agents)
    assert np.array_equal(p1[e.astype(np.int64)][okr], vr[okr])


def test_resample_endpoints_uniform():
    """Layer.update with frac = 1 (base.py:1849-1876): every edge gets iid uniform endpoints, self-loops allowed."""
    n_agents, n_edges = 1000, 300_000
    p1, p2 = px.resample_edges(4, 9, 3, n_edges, n_agents)
    assert p1.dtype == np.int32 and len(p1) == n_edges and len(p2) == n_edges
    assert p1.min() >= 0 and p2.max() < n_agents
    exp = np.full(n_agents, n_edges / n_agents)
    assert _chi2_pvalue(np.bincount(p1, minlength=n_agents).astype(float), exp) > 1e-4
    assert _chi2_pvalue(np.bincount(p2, minlength=n_agents).astype(float), exp) > 1e-4
    # pairs independent: joint counts over 10 x 10 coarse cells
    joint = np.bincount((p1 // 100) * 10 + p2 // 100, minlength=100).astype(float)
    assert _chi2_pvalue(joint, np.full(100, n_edges / 100)) > 1e-4
    loops = int((p1 == p2).sum())
    assert abs(loops - n_edges / n_agents) < 5 * np.sqrt(n_edges / n_agents)
    q1, _ = px.resample_edges(4, 10, 3, n_edges, n_agents)
    assert (q1 != p1).mean() > 0.99  # a fresh draw each day


def test_choose_k_is_uniform_subset():
    counts = np.zeros(50)
    for s in range(400):
        sel = px.choose_k(s, 0, px.STREAM_CHOOSE, np.arange(50), 10)
        assert len(np.unique(sel)) == 10
        counts[sel] += 1
    # each index chosen with probability 1/5: chi-square-ish bound
    assert np.all(np.abs(counts - 80) < 40)
    assert np.array_equal(px.choose_k(1, 0, px.STREAM_CHOOSE, np.arange(5), 9), np.arange(5))


# ---------------------------------------------------------------- compute_viral_load (utils.py:39-79)

def test_viral_load_closed_form():
    hi = F32(2) / (F32(1) + F32(0.3) * (F32(2) - F32(1)))
    lo = F32(1) / (F32(1) + F32(0.3) * (F32(2) - F32(1)))
    assert hi == F32(2 / 1.3) or abs(hi - 2 / 1.3) < 1e-6
    # (start, recovered, dead, t, expected)
    cases = [
        (np.nan, np.nan, np.nan, 5, lo),   # never infected: NaN comparisons are False
        (0, 10, np.nan, 0, hi),            # (0-0)/10 = 0 < 0.3
        (0, 10, np.nan, 2, hi),            # 0.2 < 0.3
        (0, 10, np.nan, 3, lo),            # 0.3 < 0.3 is False
        (0, 20, np.nan, 3, hi),            # cap: 0.3*20 = 6 > 4 -> switch at 4/20 = 0.2; 0.15 < 0.2
        (0, 20, np.nan, 4, lo),            # 0.2 < 0.2 False
        (5, 50, 12, 6, hi),                # death date overrides recovery: total 7, 1/7 < 0.3
        (5, 50, 12, 8, lo),                # 3/7 >= 0.3
        (3, 3, np.nan, 3, lo),             # zero-length infection: 0/0 -> NaN
    ]
    for s, r, d, t, want in cases:
        got = co.compute_viral_load(t, np.array([s], F32), np.array([r], F32), np.array([d], F32), 0.3, 2, 4)
        assert got.dtype == np.float32 and got[0] == want, (s, r, d, t)


# ---------------------------------------------------------------- compute_trans_sus (utils.py:82-90)

def test_trans_sus_special_cases():
    n = 8
    rt = np.full(n, 2.5, F32)
    rs = np.full(n, 0.67, F32)
    vl = np.full(n, F32(2) / F32(1.3), F32)
    T, Fb = np.ones(n, bool), np.zeros(n, bool)
    imm = np.zeros(n, F32)
    base_rt, base_rs = co.compute_trans_sus(rt, rs, T, T, 1.0, vl, T, Fb, Fb, 1.0, 0.2, 0.3, imm)
    assert np.all(base_rt == rt * vl) and np.all(base_rs == rs)
    z_rt, _ = co.compute_trans_sus(rt, rs, Fb, T, 1.0, vl, T, Fb, Fb, 1.0, 0.2, 0.3, imm)
    assert np.all(z_rt == 0)  # not infectious
    _, z_rs = co.compute_trans_sus(rt, rs, T, Fb, 1.0, vl, T, Fb, Fb, 1.0, 0.2, 0.3, imm)
    assert np.all(z_rs == 0)  # not susceptible
    z_rt, _ = co.compute_trans_sus(rt, rs, T, T, 0.0, vl, T, Fb, Fb, 1.0, 0.2, 0.3, imm)
    assert np.all(z_rt == 0)  # beta_layer = 0
    q_rt, q_rs = co.compute_trans_sus(rt, rs, T, T, 1.0, vl, T, Fb, T, 1.0, 0.2, 0.3, imm)
    assert np.all(q_rt == (rt * F32(0.3)) * vl) and np.all(q_rs == rs * F32(0.3))  # quarantine both sides
    i_rt, i_rs = co.compute_trans_sus(rt, rs, T, T, 1.0, vl, T, T, Fb, 1.0, 0.2, 0.3, imm)
    assert np.all(i_rt == (rt * F32(0.2)) * vl) and np.all(i_rs == rs)  # isolation: source only
    a_rt, _ = co.compute_trans_sus(rt, rs, T, T, 1.0, vl, Fb, Fb, Fb, 0.5, 0.2, 0.3, imm)
    assert np.all(a_rt == (rt * F32(0.5)) * vl)  # asymptomatic factor
    _, m_rs = co.compute_trans_sus(rt, rs, T, T, 1.0, vl, T, Fb, Fb, 1.0, 0.2, 0.3, np.ones(n, F32))
    assert np.all(m_rs == 0)  # full immunity
    _, h_rs = co.compute_trans_sus(rt, rs, T, T, 1.0, vl, T, Fb, Fb, 1.0, 0.2, 0.3, np.full(n, 0.25, F32))
    assert np.all(h_rs == rs * F32(0.75))


# ---------------------------------------------------------------- compute_infections (utils.py:93-128)

def brute_force_infections(beta, p1, p2, lb, rt, rs, U, legacy=False):
    """Scalar double loop over (direction, edge) in the reference's output order."""
    out = []
    dirs = [(p1, p2)] if legacy else [(p1, p2), (p2, p1)]
    for d, (src, tgt) in enumerate(dirs):
        for e in range(len(p1)):
            st = float(rt[src[e]])
            if st == 0:
                continue
            p = F32(F32(F32(F32(beta) * lb[e]) * rt[src[e]]) * rs[tgt[e]])
            if p == 0:
                continue
            if U[d][e] < float(p):
                out.append((int(src[e]), int(tgt[e])))
    return out


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_infections_brute_force(seed):
    rng = np.random.default_rng(seed)
    n, E = 12, 30
    p1 = rng.integers(0, n, E).astype(np.int32)
    p2 = rng.integers(0, n, E).astype(np.int32)
    p1[0] = p2[0]  # a self-loop
    lb = rng.choice([1.0, 0.0, 3.0], E).astype(F32)
    rt = np.where(rng.random(n) < 0.4, rng.random(n) * 2, 0).astype(F32)
    rt[1] = 1e6  # p > 1 always transmits
    rs = np.where(rng.random(n) < 0.8, rng.random(n), 0).astype(F32)
    U = rng.random((2, E))
    for beta in [0.3, 0.0]:
        for legacy in [False, True]:
            s, t, p = co.compute_infections(beta, p1, p2, lb, rt, rs, lambda d, idx: U[d][idx], legacy=legacy)
            want = brute_force_infections(beta, p1, p2, lb, rt, rs, U, legacy)
            assert list(zip(s.tolist(), t.tolist())) == want
            assert np.all(rt[s] > 0) and np.all(rs[t] > 0)  # sources infectious, targets susceptible
            if beta == 0:
                assert len(s) == 0


# ---------------------------------------------------------------- waning kinetics (immunity.py:216-496)

def test_nab_growth_decay_closed_form():
    r1 = np.log(2) / 50
    kin = co.precompute_waning(200, dict(form='nab_growth_decay', growth_time=21, decay_rate1=r1, decay_time1=150,
                                         decay_rate2=np.log(2) / 250, decay_time2=365))
    assert len(kin) == 200  # one value per day since the NAb event (npts of them)
    assert np.allclose(kin[:21], 1 / 21, rtol=0, atol=1e-15)  # linear growth to 1 over 21 days
    i = np.arange(0, 149)
    assert np.allclose(kin[21 + i], np.exp(-(i + 1) * r1) - np.exp(-i * r1), rtol=1e-12, atol=1e-15)
    assert abs(kin[:21].sum() - 1.0) < 1e-12  # telescoping: the peak is reached on day 21


def test_exp_decay_closed_form():
    kin = co.precompute_waning(30, dict(form='exp_decay', init_val=1.0, half_life=10))
    t = np.arange(31)
    assert np.allclose(kin, np.diff(np.exp(-np.log(2) / 10 * t)))


def test_calc_ve_closed_form():
    pars = co.make_pars()['nab_eff']
    assert co.calc_VE(np.array([0.0]), 'sus', pars)[0] == 0.0
    v = co.calc_VE(np.array([1.0]), 'symp', pars)[0]
    assert abs(v - np.exp(-0.739) / (1 + np.exp(-0.739))) < 1e-15
    nab = np.array([0.5, 2.0, 7.0])
    for ax, (a, b) in {'sus': (1.08, 0.967), 'sev': (-0.014, 0.079)}.items():
        ve = co.calc_VE(nab, ax, pars)
        assert np.allclose(np.log(ve / (1 - ve)), a + b * np.log(nab))  # logit-linear in log NAb


# ---------------------------------------------------------------- whole-sim special cases

def run_oracle(n=500, n_days=30, pop_type='random', rng_kind='philox', seed=1, **pars):
    pop = make_population(n=n, pop_type=pop_type)
    sim = oracle_sim(pop, rng_kind=rng_kind, seed=seed, n_days=n_days, pop_type=pop_type, **pars)
    sim.run()
    return sim, sim.finalize()


@pytest.mark.parametrize('rng_kind', ['philox', 'ref'])
def test_beta_zero_no_transmission(rng_kind):
    # tests/unittests/test_transmission.py:16-35
    sim, r = run_oracle(n=500, n_days=30, rng_kind=rng_kind, beta=0, pop_infected=10)
    assert r['new_infections'].sum() == 0
    assert np.all(r['cum_infections'] == 10)


@pytest.mark.parametrize('delay', [1, 5, 15])
def test_exposed_to_infectious_delay_exact(delay):
    # tests/unittests/test_progression.py:17-48: std 0 -> everyone infectious on day `delay`
    dur = co.make_pars()['dur']
    dur['exp2inf'] = dict(dist='normal', par1=delay, par2=0)
    sim, r = run_oracle(n=300, n_days=30, pop_infected=300, dur=dur, rel_symp_prob=0)
    assert r['new_infectious'][:delay].sum() == 0
    assert r['new_infectious'][delay] == 300


@pytest.mark.parametrize('rec', [1, 5, 10])
def test_recovery_duration_exact(rec):
    # tests/unittests/test_progression.py:50-76: infectious on day 1, recover exactly rec days later
    dur = co.make_pars()['dur']
    dur['exp2inf'] = dict(dist='normal', par1=1, par2=0)
    dur['asym2rec'] = dict(dist='normal', par1=rec, par2=0)
    sim, r = run_oracle(n=300, n_days=30, pop_infected=300, dur=dur, rel_symp_prob=0, beta=0)
    assert r['new_recoveries'][1 + rec] == 300


def test_everyone_dies_with_huge_probabilities():
    # tests/unittests/test_mortality.py:13-24 (probabilities > 1 are used unclamped)
    sim, r = run_oracle(n=200, n_days=90, pop_infected=200, rel_symp_prob=1e6, rel_severe_prob=1e6,
                        rel_crit_prob=1e6, rel_death_prob=1e6)
    assert r['cum_deaths'][-1] == 200


def test_nobody_dies_with_zero_death_probability():
    # tests/unittests/test_mortality.py:26-47
    sim, r = run_oracle(n=300, n_days=60, pop_infected=300, rel_symp_prob=1e6, rel_severe_prob=1e6,
                        rel_crit_prob=1e6, rel_death_prob=0.0)
    assert r['new_deaths'].sum() == 0
    assert r['cum_recoveries'][-1] >= 200


@pytest.mark.parametrize('pop_type', ['random', 'hybrid'])
def test_population_conservation(pop_type):
    # tests/unittests/test_pops.py:6-26
    sim, r = run_oracle(n=1000, n_days=10, pop_type=pop_type, pop_infected=50)
    assert r['n_susceptible'][0] + r['cum_infections'][0] == 1000
    assert r['cum_infections'][-1] > r['cum_infections'][0]
    assert r['cum_symptomatic'][-1] > r['cum_symptomatic'][0]


@pytest.mark.parametrize('use_waning', [False, True])
def test_state_diagram_invariants(use_waning):
    # tests/test_immunity.py:25-85 against tests/state_diagram.xlsx
    d = json.load(open(os.path.join(GOLDEN, 'state_diagram.json')))['with_waning' if use_waning else 'without_waning']
    for n_days in [20, 40, 70]:
        sim, r = run_oracle(n=1000, n_days=n_days, pop_infected=20, use_waning=use_waning, beta=0.03,
                            pop_type='hybrid')
        ppl = sim.people
        for s1 in d['states']:
            m1 = getattr(ppl, s1)
            for s2 in d['states']:
                if s1 == s2:
                    continue
                rel = d['relation'][s1][s2]
                v2 = getattr(ppl, s2)[m1]
                if rel == 1:
                    assert v2.all(), (n_days, s1, s2)
                elif rel == -1:
                    assert not v2.any(), (n_days, s1, s2)


def test_step_equals_run():
    # tests/test_resume.py:120-138
    pop = make_population(n=800, pop_type='hybrid')
    a = oracle_sim(pop, seed=4, n_days=25, pop_type='hybrid', beta=0.05).run().finalize()
    s = oracle_sim(pop, seed=4, n_days=25, pop_type='hybrid', beta=0.05)
    for _ in range(26):
        s.step()
    b = s.finalize()
    for k in ['new_infections', 'new_deaths', 'n_exposed']:
        assert np.array_equal(a[k], b[k])


# ---------------------------------------------------------------- fused dedup == layer-sequential infect

def fused_min_key_targets(sim_state, pars, layers, t, seed, viral_load):
    """The GPU's algorithm restated on the CPU: every (variant, layer, dir, edge) slot evaluated with the
    start-of-day susceptibility, winner = smallest key per target."""
    ppl = sim_state
    winner = {}
    sus = ppl.susceptible.copy()
    for v in range(pars['n_variants']):
        beta = F32(pars['beta'] * pars['rel_beta'] * pars['variant_pars'][v]['rel_beta'])
        inf_v = ppl.infectious * (ppl.infectious_variant == v)
        for li, (lk, (p1, p2, lb)) in enumerate(layers.items()):
            rt, rs = co.compute_trans_sus(ppl.rel_trans, ppl.rel_sus, inf_v, sus, pars['beta_layer'][lk], viral_load,
                                          ppl.symptomatic, ppl.isolated, ppl.quarantined, pars['asymp_factor'],
                                          pars['iso_factor'][lk], pars['quar_factor'][lk], ppl.sus_imm[v])
            probs = co.trans_probs(beta, p1, p2, lb, rt, rs)
            for d, (src, tgt) in enumerate([(p1, p2), (p2, p1)]):
                idx = np.flatnonzero(probs[d] != 0)
                u = px.edge_uniforms(seed, t, v, li, idx, d)
                hit = idx[u < probs[d][idx]]
                for e in hit:
                    key = (v, li, d, int(e))
                    tg = int(tgt[e])
                    if tg not in winner or key < winner[tg][0]:
                        winner[tg] = (key, int(src[e]))
    return winner


def test_fused_dedup_equals_layer_sequential():
    pop = make_population(n=1500, pop_type='hybrid')
    seed = 21
    sim = oracle_sim(pop, seed=seed, n_days=40, pop_type='hybrid', beta=0.06)
    checked = 0
    while sim.t < 30:
        t = sim.t
        ppl = sim.people
        if t > 0:
            # state after update_states_pre of day t, before transmission: replay the step in pieces
            import copy
            probe = copy.deepcopy(sim)
            probe.people.update_states_pre(t)
            vd = probe.pars['viral_dist']
            vl = co.compute_viral_load(t, probe.people.date_infectious, probe.people.date_recovered,
                                       probe.people.date_dead, vd['frac_time'], vd['load_ratio'], vd['high_cap'])
            win = fused_min_key_targets(probe.people, probe.pars, probe.layers, t, seed, vl)
            n_log = len(ppl.infection_log)
            sim.step()
            new = [(tg, s) for (s, tg, d, l, v) in ppl.infection_log[n_log:] if d == t]
            assert sorted(new) == sorted((tg, w[1]) for tg, w in win.items())
            checked += len(new)
        else:
            sim.step()
    assert checked > 100
