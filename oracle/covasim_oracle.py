"""CPU restatement (oracle) of Covasim 3.1.7's daily transmission step.

TEST INFRASTRUCTURE ONLY. Only tests/, bench.py's ``cpu_baseline`` leg and
``__graft_entry__.smoke()`` may import this module, and only as the checker.
The product (covasim_amd/) never imports, links or executes anything here.

Every function restates the reference algorithm in numpy float32/float64 with
the same dtypes and the same order of floating-point operations as the
reference, citing the reference file:line it follows (paths relative to the
Covasim checkout). Numba's mixed scalar/array typing is restated explicitly:
an integer scalar combined with a float32 array is computed in float32
(numba's ufunc loop choice), so ``t - time_start`` is ``float32(t) - start``.

Random draws are pluggable (``RefRNG`` mimics the reference's two MT19937
streams and call order; ``PhiloxRNG`` uses the counter-based streams the GPU
kernels use, so whole runs can be compared draw-for-draw).

Pinning (what checks each function independently of this restatement):
  compute_viral_load   closed form (two values 2/1.3, 1/1.3; cap rule; NaN), brute force
  compute_trans_sus    per-element scalar recomputation + special cases (beta_layer=0 ...)
  compute_infections   scalar double-loop brute force over (dir, edge) with supplied uniforms
  infect outcomes      deterministic std-0 progression cases, p=1e6 / p=0 mortality cases,
                       Bernoulli rates and lognormal moments (KS against scipy)
  update_states_pre    tests/state_diagram.xlsx implication matrices + progression cases
  waning kinetics      closed-form nab_growth_decay values; calc_VE logistic closed form
  Philox4x32-10        Random123 known-answer vectors, cross-checked against rocRAND's engine
Seed-for-seed equality with the reference's MT19937 trajectories is parity unpinned
(the reference needs numba/sciris, which are absent); full-run agreement is
distributional.
"""

from collections import defaultdict

import numpy as np

from . import philox as px

F32 = np.float32
F64 = np.float64
NAN = np.float32(np.nan)

# ---------------------------------------------------------------------------
# Parameters (covasim/parameters.py:15-151, 158-227, 230-294; defaults.py:223-243)
# ---------------------------------------------------------------------------

DEFAULT_AGE_DATA = np.array([
    [0, 4, 0.0605], [5, 9, 0.0607], [10, 14, 0.0566], [15, 19, 0.0557],
    [20, 24, 0.0612], [25, 29, 0.0843], [30, 34, 0.0848], [35, 39, 0.0764],
    [40, 44, 0.0697], [45, 49, 0.0701], [50, 54, 0.0681], [55, 59, 0.0653],
    [60, 64, 0.0591], [65, 69, 0.0453], [70, 74, 0.0312], [75, 79, 0.02016],
    [80, 84, 0.01344], [85, 89, 0.01008], [90, 99, 0.00672],
])  # defaults.py:223-243

DUR_KEYS = ['exp2inf', 'inf2sym', 'sym2sev', 'sev2crit', 'asym2rec', 'mild2rec', 'sev2rec', 'crit2rec', 'crit2die']


def get_prognoses():
    """Age-binned prognoses converted to conditional probabilities (parameters.py:230-294)."""
    p = dict(
        age_cutoffs=np.array([0, 10, 20, 30, 40, 50, 60, 70, 80, 90]),
        sus_ORs=np.array([0.34, 0.67, 1.00, 1.00, 1.00, 1.00, 1.24, 1.47, 1.47, 1.47]),
        trans_ORs=np.array([1.00] * 10),
        comorbidities=np.array([1.00] * 10),
        symp_probs=np.array([0.50, 0.55, 0.60, 0.65, 0.70, 0.75, 0.80, 0.85, 0.90, 0.90]),
        severe_probs=np.array([0.00050, 0.00165, 0.00720, 0.02080, 0.03430, 0.07650, 0.13280, 0.20655, 0.24570, 0.24570]),
        crit_probs=np.array([0.00003, 0.00008, 0.00036, 0.00104, 0.00216, 0.00933, 0.03639, 0.08923, 0.17420, 0.17420]),
        death_probs=np.array([0.00002, 0.00002, 0.00010, 0.00032, 0.00098, 0.00265, 0.00766, 0.02439, 0.08292, 0.16190]),
    )
    # relative_prognoses, parameters.py:285-294 (division order matters)
    p['death_probs'] = p['death_probs'] / p['crit_probs']
    p['crit_probs'] = p['crit_probs'] / p['severe_probs']
    p['severe_probs'] = p['severe_probs'] / p['symp_probs']
    return p


LAYER_DEFAULTS = {  # parameters.py:174-189
    'random': dict(beta_layer=dict(a=1.0), contacts=dict(a=20), dynam_layer=dict(a=0),
                   iso_factor=dict(a=0.2), quar_factor=dict(a=0.3)),
    'hybrid': dict(beta_layer=dict(h=3.0, s=0.6, w=0.6, c=0.3), contacts=dict(h=2.0, s=20, w=16, c=20),
                   dynam_layer=dict(h=0, s=0, w=0, c=0), iso_factor=dict(h=0.3, s=0.1, w=0.1, c=0.1),
                   quar_factor=dict(h=0.6, s=0.2, w=0.2, c=0.2)),
}


def make_pars(**kwargs):
    """Default parameters relevant to the transmission step (parameters.py:15-151)."""
    pars = dict(
        pop_size=20e3, pop_infected=20, pop_type='random', n_days=60, rand_seed=1,
        beta_dist=dict(dist='neg_binomial', par1=1.0, par2=0.45, step=0.01),
        viral_dist=dict(frac_time=0.3, load_ratio=2, high_cap=4),
        beta=0.016, asymp_factor=1.0, n_imports=0, n_variants=1,
        use_waning=True,
        nab_init=dict(dist='normal', par1=0, par2=2),
        nab_decay=dict(form='nab_growth_decay', growth_time=21, decay_rate1=np.log(2) / 50, decay_time1=150,
                       decay_rate2=np.log(2) / 250, decay_time2=365),
        nab_boost=1.5,
        nab_eff=dict(alpha_inf=1.08, alpha_inf_diff=1.812, beta_inf=0.967, alpha_symp_inf=-0.739,
                     beta_symp_inf=0.038, alpha_sev_symp=-0.014, beta_sev_symp=0.079),
        rel_imm_symp=dict(asymp=0.85, mild=1, severe=1.5),
        immunity=None, trans_redux=0.59, rel_beta=1.0,
        dur=dict(
            exp2inf=dict(dist='lognormal_int', par1=4.5, par2=1.5),
            inf2sym=dict(dist='lognormal_int', par1=1.1, par2=0.9),
            sym2sev=dict(dist='lognormal_int', par1=6.6, par2=4.9),
            sev2crit=dict(dist='lognormal_int', par1=1.5, par2=2.0),
            asym2rec=dict(dist='lognormal_int', par1=8.0, par2=2.0),
            mild2rec=dict(dist='lognormal_int', par1=8.0, par2=2.0),
            sev2rec=dict(dist='lognormal_int', par1=18.1, par2=6.3),
            crit2rec=dict(dist='lognormal_int', par1=18.1, par2=6.3),
            crit2die=dict(dist='lognormal_int', par1=10.7, par2=4.8),
        ),
        rel_symp_prob=1.0, rel_severe_prob=1.0, rel_crit_prob=1.0, rel_death_prob=1.0,
        n_beds_hosp=None, n_beds_icu=None, no_hosp_factor=2.0, no_icu_factor=2.0, quar_period=14,
        variants=[],  # list of dict(rel_beta=..., days=[...], n_imports=..., ...)
    )
    # variant_pars['wild'] is captured from the defaults before kwargs apply (parameters.py:131-134)
    wild = {k: pars[k] for k in ['rel_beta', 'rel_symp_prob', 'rel_severe_prob', 'rel_crit_prob', 'rel_death_prob']}
    pars.update(kwargs)
    # immunity.py:107-114 (variant.initialize): a variant's parameters are stored under its label (a dict
    # without one is 'custom', immunity.py:78-80) and its index is the label's position, so a later variant
    # with the same label replaces the parameters and shares the index
    vpars = {'wild': wild}
    for var in pars['variants']:
        vp = dict(rel_beta=1.0, rel_symp_prob=1.0, rel_severe_prob=1.0, rel_crit_prob=1.0, rel_death_prob=1.0)
        vp.update({k: v for k, v in var.items() if k in vp})
        vpars[var.get('label') or 'custom'] = vp
    labels = list(vpars)
    pars['variant_index'] = [labels.index(var.get('label') or 'custom') for var in pars['variants']]
    pars['variant_pars'] = [vpars[lab] for lab in labels]
    pars['n_variants'] = len(labels)
    ld = LAYER_DEFAULTS[pars['pop_type']]
    for key in ['beta_layer', 'contacts', 'dynam_layer', 'iso_factor', 'quar_factor']:
        merged = dict(ld[key])
        merged.update(kwargs.get(key, {}) or {})
        pars[key] = merged
    if pars.get('prognoses') is None:
        pars['prognoses'] = get_prognoses()
    return pars


# ---------------------------------------------------------------------------
# Numeric kernels (covasim/utils.py:39-128)
# ---------------------------------------------------------------------------

def compute_viral_load(t, time_start, time_recovered, time_dead, frac_time, load_ratio, high_cap):
    """utils.py:39-79 -- two-level viral load, float32 throughout."""
    frac_time, load_ratio, high_cap = F32(frac_time), F32(load_ratio), F32(high_cap)
    with np.errstate(all='ignore'):
        n = len(time_dead)
        time_stop = np.ones(n, dtype=F32) * time_recovered  # :61
        inds = ~np.isnan(time_dead)  # :62
        time_stop[inds] = time_dead[inds]  # :63
        infect_days_total = time_stop - time_start  # :66
        trans_day = frac_time * infect_days_total  # :67
        inds = trans_day > high_cap  # :68
        cap_frac = high_cap / infect_days_total[inds]  # :69
        trans_point = np.ones(n, dtype=F32) * frac_time  # :72
        trans_point[inds] = cap_frac  # :73
        load = np.ones(n, dtype=F32)  # :76
        early = (F32(t) - time_start) / infect_days_total < trans_point  # :77 (numba: int scalar op f32 array -> f32)
        load = (load_ratio * early + load * ~early) / (load + frac_time * (load_ratio - load))  # :78
    return load.astype(F32)


def compute_trans_sus(rel_trans, rel_sus, inf, sus, beta_layer, viral_load, symp, iso, quar,
                      asymp_factor, iso_factor, quar_factor, immunity_factors):
    """utils.py:82-90 -- strict left-to-right float32 products."""
    asymp_factor, iso_factor, quar_factor, beta_layer = F32(asymp_factor), F32(iso_factor), F32(quar_factor), F32(beta_layer)
    f_asymp = symp + ~symp * asymp_factor  # :85
    f_iso = ~iso + iso * iso_factor  # :86
    f_quar = ~quar + quar * quar_factor  # :87
    rt = rel_trans * inf * f_quar * f_asymp * f_iso * beta_layer * viral_load  # :88
    rs = rel_sus * sus * f_quar * (F32(1) - immunity_factors)  # :89
    return rt.astype(F32), rs.astype(F32)


def compute_infections(beta, p1, p2, layer_betas, rel_trans, rel_sus, uniform_fn, legacy=False):
    """utils.py:93-128 -- who infects whom, dir-major then edge-ascending output.

    ``uniform_fn(direction, edge_inds)`` returns the float64 uniforms for the
    compacted nonzero-probability edges (the reference draws
    ``np.random.random(len(nonzero_betas))`` here, utils.py:123).
    Returns (sources, targets, probs) with probs the float32 transmission probabilities
    of the emitted pairs.
    """
    beta = F32(beta)
    slist, tlist, plist = [], [], []
    pairs = [[p1, p2], [p2, p1]] if not legacy else [[p1, p2]]  # :113
    for d, (sources, targets) in enumerate(pairs):
        source_trans = rel_trans[sources]  # :115
        inf_inds = source_trans.nonzero()[0]  # :116
        betas = beta * layer_betas[inf_inds] * source_trans[inf_inds] * rel_sus[targets[inf_inds]]  # :117
        nonzero_inds = betas.nonzero()[0]  # :118
        nonzero_inf_inds = inf_inds[nonzero_inds]  # :119
        nonzero_betas = betas[nonzero_inds]  # :120
        u = np.asarray(uniform_fn(d, nonzero_inf_inds), dtype=F64)
        transmissions = (u < nonzero_betas).nonzero()[0]  # :123 (float64 < float32 promoted)
        slist.append(sources[nonzero_inf_inds[transmissions]].astype(np.int32))
        tlist.append(targets[nonzero_inf_inds[transmissions]].astype(np.int32))
        plist.append(nonzero_betas[transmissions])
    return np.concatenate(slist), np.concatenate(tlist), np.concatenate(plist).astype(F32)


def trans_probs(beta, p1, p2, layer_betas, rel_trans, rel_sus):
    """Float32 per-edge probabilities for both directions (the p of utils.py:117, 0 where filtered)."""
    beta = F32(beta)
    out = np.zeros((2, len(p1)), dtype=F32)
    for d, (s, tg) in enumerate([[p1, p2], [p2, p1]]):
        st = rel_trans[s]
        inf = st != 0
        out[d, inf] = beta * layer_betas[inf] * st[inf] * rel_sus[tg[inf]]
    return out


# ---------------------------------------------------------------------------
# Sampling (covasim/utils.py:156-236, 343-357, 409-426)
# ---------------------------------------------------------------------------

def lognormal_params(par1, par2):
    """utils.py:223-227: mean/std of the lognormal -> (mu, sigma) of the underlying normal."""
    mean = np.log(par1 ** 2 / np.sqrt(par2 ** 2 + par1 ** 2))
    sigma = np.sqrt(np.log(par2 ** 2 / par1 ** 2 + 1))
    return mean, sigma


def transform_normal(dist, par1, par2, z, u0):
    """Map standard-normal z (and uniform u0) to the sample() distribution (utils.py:215-231)."""
    if dist in ('unif', 'uniform'):
        return par1 + (par2 - par1) * u0
    if dist in ('norm', 'normal'):
        return par1 + par2 * z
    if dist == 'normal_pos':
        return np.abs(par1 + par2 * z)
    if dist == 'normal_int':
        return np.round(np.abs(par1 + par2 * z))
    if dist in ('lognorm', 'lognormal', 'lognorm_int', 'lognormal_int'):
        if par1 > 0:
            mu, sigma = lognormal_params(par1, par2)
            x = np.exp(mu + sigma * z)
        else:
            x = np.zeros(len(z))
        if '_int' in dist:
            x = np.round(x)  # half-to-even, as np.round in utils.py:230-231
        return x
    raise NotImplementedError(f'distribution {dist!r} not supported on the transmission path')


class RefRNG:
    """Reference-like randomness: the numpy global stream and numba's private stream,
    both MT19937 seeded with the sim seed (utils.py:271-298), consumed in the
    reference's call order."""

    kind = 'ref'

    def __init__(self, seed):
        self.np_rs = np.random.RandomState(int(seed))
        self.nb_rs = np.random.RandomState(int(seed))
        self.seed = int(seed)

    def edge_uniform_fn(self, t, v, l):
        return lambda d, inds: self.nb_rs.random(len(inds))  # utils.py:123

    def choose_seeds(self, n, k):
        return self.nb_rs.choice(int(n), int(k), replace=False)  # utils.py:430-443

    def choose_imports(self, sus_inds, k, t, v):
        return self.np_rs.choice(sus_inds, int(k), replace=False)  # immunity.py:127

    def sample(self, durpar, inds, t, slot):
        d = durpar
        n = len(inds)
        if d['dist'] in ('lognorm', 'lognormal', 'lognorm_int', 'lognormal_int'):
            if d['par1'] > 0:
                mu, sigma = lognormal_params(d['par1'], d['par2'])
                x = self.np_rs.lognormal(mean=mu, sigma=sigma, size=n)
            else:
                x = np.zeros(n)
            return np.round(x) if '_int' in d['dist'] else x
        if d['dist'] in ('norm', 'normal'):
            return self.np_rs.normal(loc=d['par1'], scale=d['par2'], size=n)
        if d['dist'] == 'normal_pos':
            return np.abs(self.np_rs.normal(loc=d['par1'], scale=d['par2'], size=n))
        if d['dist'] == 'normal_int':
            return np.round(np.abs(self.np_rs.normal(loc=d['par1'], scale=d['par2'], size=n)))
        if d['dist'] in ('unif', 'uniform'):
            return self.np_rs.uniform(low=d['par1'], high=d['par2'], size=n)
        raise NotImplementedError(d['dist'])

    def bernoulli(self, probs, inds, t, slot, half):
        return self.np_rs.random(len(probs)) < probs  # utils.py:343-357

    def resample(self, t, l, n_edges, n_agents):
        inds = self.nb_rs.choice(n_edges, n_edges, replace=False)  # base.py:1870
        p1 = np.empty(n_edges, dtype=np.int32)
        p2 = np.empty(n_edges, dtype=np.int32)
        p1[inds] = self.nb_rs.choice(n_agents, n_edges, replace=True)  # base.py:1873
        p2[inds] = self.nb_rs.choice(n_agents, n_edges, replace=True)  # base.py:1874
        return p1, p2


class PhiloxRNG:
    """Counter-based randomness identical to the GPU kernels' streams (see oracle/philox.py)."""

    kind = 'philox'

    def __init__(self, seed):
        self.seed = int(seed)

    def edge_uniform_fn(self, t, v, l):
        return lambda d, inds: px.edge_uniforms(self.seed, t, v, l, inds, d)

    def choose_seeds(self, n, k):
        return px.choose_k(self.seed, 0, px.STREAM_CHOOSE, np.arange(int(n)), int(k))

    def choose_imports(self, sus_inds, k, t, v):
        return px.choose_k(self.seed, t, px.STREAM_IMPORT | v, sus_inds, int(k))

    def sample(self, durpar, inds, t, slot):
        u0, u1 = px.slot_uniforms(self.seed, t, inds, slot)
        z = px.box_muller(u0, u1)
        return transform_normal(durpar['dist'], durpar['par1'], durpar['par2'], z, u0)

    def bernoulli(self, probs, inds, t, slot, half):
        u0, u1 = px.slot_uniforms(self.seed, t, inds, slot)
        return (u0 if half == 0 else u1) < probs

    def resample(self, t, l, n_edges, n_agents):
        return px.resample_edges(self.seed, t, l, n_edges, n_agents)


# ---------------------------------------------------------------------------
# Immunity / waning (covasim/immunity.py:138-496)
# ---------------------------------------------------------------------------

def nab_growth_decay(length, growth_time, decay_rate1, decay_time1, decay_rate2, decay_time2):
    """immunity.py:404-448."""
    def f1(t, growth_time):
        return (1 / growth_time) * t

    def f2(t, decay_time1, decay_time2, decay_rate1, decay_rate2):
        decayRate = np.full(len(t), fill_value=decay_rate1)
        decayRate[(t > decay_time2).nonzero()[0]] = decay_rate2
        slowing = (1 / (decay_time2 - decay_time1)) * (decay_rate1 - decay_rate2)
        mid = ((t > decay_time1) * (t <= decay_time2)).nonzero()[0]
        decayRate[mid] = decay_rate1 - slowing * (np.arange(len(mid), dtype=np.int32))
        titre = np.zeros(len(t))
        for i in range(1, len(t)):
            titre[i] = titre[i - 1] + decayRate[i]
        return np.exp(-titre)

    if decay_time2 < decay_time1:
        raise ValueError('decay_time2 must exceed decay_time1')
    length = length + 1
    t1 = np.arange(growth_time, dtype=np.int32)
    t2 = np.arange(length - growth_time, dtype=np.int32)
    y = np.concatenate([f1(t1, growth_time), f2(t2, decay_time1, decay_time2, decay_rate1, decay_rate2)])
    return np.diff(y)[0:length]


def exp_decay(length, init_val, half_life, delay=None):
    """immunity.py:482-496."""
    length = length + 1
    decay_rate = np.log(2) / half_life if ~np.isnan(half_life) else 0.
    if delay is not None:
        t = np.arange(length - delay, dtype=np.int32)
        growth = (init_val / delay) * np.ones(delay)
        decay = init_val * np.exp(-decay_rate * t)
        result = np.concatenate([growth, decay], axis=None)
    else:
        t = np.arange(length, dtype=np.int32)
        result = init_val * np.exp(-decay_rate * t)
    return np.diff(result)


def precompute_waning(length, pars):
    """immunity.py:356-401 (nab_growth_decay and exp_decay forms)."""
    pars = dict(pars)
    form = pars.pop('form')
    if form is None or form == 'nab_growth_decay':
        return nab_growth_decay(length, **pars)
    if form == 'exp_decay':
        if pars.get('half_life') is None:
            pars['half_life'] = np.nan
        return exp_decay(length, **pars)
    raise NotImplementedError(form)


def calc_VE(nab, ax, pars):
    """immunity.py:216-247 -- inverse logit of alpha + beta*log(nab), float64."""
    alpha, beta = {'sus': ('alpha_inf', 'beta_inf'), 'symp': ('alpha_symp_inf', 'beta_symp_inf'),
                   'sev': ('alpha_sev_symp', 'beta_sev_symp')}[ax]
    exp_lo = np.exp(pars[alpha]) * nab ** pars[beta]
    return exp_lo / (1 + exp_lo)


def init_immunity_matrix(nv):
    """immunity.py:269-300 with only the wild label known: custom variants get 1.0."""
    return np.ones((nv, nv), dtype=F32)


# ---------------------------------------------------------------------------
# People (covasim/people.py) -- SoA state in reference dtypes
# ---------------------------------------------------------------------------

STATE_KEYS = ['susceptible', 'naive', 'exposed', 'infectious', 'symptomatic', 'severe', 'critical', 'tested',
              'diagnosed', 'recovered', 'known_dead', 'dead', 'known_contact', 'quarantined', 'isolated',
              'vaccinated']  # defaults.py:54-71
DATE_KEYS = [f'date_{s}' for s in STATE_KEYS] + ['date_pos_test', 'date_end_quarantine', 'date_end_isolation']
DUR_STATE_KEYS = ['dur_exp2inf', 'dur_inf2sym', 'dur_sym2sev', 'dur_sev2crit', 'dur_disease']
FLOW_KEYS = ['new_infections', 'new_reinfections', 'new_infectious', 'new_symptomatic', 'new_severe',
             'new_critical', 'new_recoveries', 'new_deaths', 'new_tests', 'new_diagnoses', 'new_known_deaths',
             'new_quarantined', 'new_isolated', 'new_doses', 'new_vaccinated']  # defaults.py:167-201
VFLOW_KEYS = ['new_infections_by_variant', 'new_symptomatic_by_variant', 'new_severe_by_variant',
              'new_infectious_by_variant']
STOCK_KEYS = ['susceptible', 'exposed', 'infectious', 'symptomatic', 'severe', 'critical', 'recovered', 'dead',
              'diagnosed', 'known_dead', 'quarantined', 'isolated', 'vaccinated']  # defaults.py:145-159


def set_prognoses(age, prognoses):
    """Deterministic part of People.set_prognoses (people.py:151-158)."""
    inds = np.digitize(age, prognoses['age_cutoffs']) - 1
    out = dict(
        symp_prob=prognoses['symp_probs'][inds].astype(F32),
        severe_prob=(prognoses['severe_probs'][inds] * prognoses['comorbidities'][inds]).astype(F32),
        crit_prob=prognoses['crit_probs'][inds].astype(F32),
        death_prob=prognoses['death_probs'][inds].astype(F32),
        rel_sus=prognoses['sus_ORs'][inds].astype(F32),
        trans_OR=prognoses['trans_ORs'][inds],
    )
    return out


class People:
    """Agent state with the reference's keys and dtypes (people.py:47-120)."""

    def __init__(self, pars, age, rel_trans, rel_sus, symp_prob, severe_prob, crit_prob, death_prob):
        n = len(age)
        nv = pars['n_variants']
        self.pars = pars
        self.n = n
        self.t = 0
        self.age = np.asarray(age, dtype=F32)
        self.rel_trans = np.asarray(rel_trans, dtype=F32).copy()
        self.rel_sus = np.asarray(rel_sus, dtype=F32).copy()
        self.symp_prob = np.asarray(symp_prob, dtype=F32)
        self.severe_prob = np.asarray(severe_prob, dtype=F32)
        self.crit_prob = np.asarray(crit_prob, dtype=F32)
        self.death_prob = np.asarray(death_prob, dtype=F32)
        self.n_infections = np.zeros(n, dtype=np.int32)
        self.n_breakthroughs = np.zeros(n, dtype=np.int32)
        for key in STATE_KEYS:
            setattr(self, key, np.full(n, key in ('susceptible', 'naive'), dtype=bool))
        for key in ['exposed_variant', 'infectious_variant', 'recovered_variant']:
            setattr(self, key, np.full(n, np.nan, dtype=F32))
        self.exposed_by_variant = np.zeros((nv, n), dtype=bool)
        self.infectious_by_variant = np.zeros((nv, n), dtype=bool)
        self.sus_imm = np.zeros((nv, n), dtype=F32)
        self.symp_imm = np.zeros((nv, n), dtype=F32)
        self.sev_imm = np.zeros((nv, n), dtype=F32)
        self.peak_nab = np.zeros(n, dtype=F32)
        self.nab = np.zeros(n, dtype=F32)
        self.t_nab_event = np.zeros(n, dtype=np.int32)
        for key in DATE_KEYS + DUR_STATE_KEYS:
            setattr(self, key, np.full(n, np.nan, dtype=F32))
        self.infection_log = []
        self._pending_quarantine = defaultdict(list)  # {start day: [(agent, end day)]}, people.py:116
        self.init_flows()

    def init_flows(self):
        """people.py:123-129."""
        nv = self.pars['n_variants']
        self.flows = {k: 0 for k in FLOW_KEYS}
        self.flows_variant = {k: np.zeros(nv, dtype=F32) for k in VFLOW_KEYS}

    # -- update_states_pre and helpers (people.py:164-312, 368-374) --

    def check_inds(self, current, date, filter_inds=None):
        """people.py:211-219."""
        if filter_inds is None:
            not_current = np.logical_not(current).nonzero()[0]
        else:
            not_current = filter_inds[np.logical_not(current[filter_inds])]
        has_date = not_current[~np.isnan(date[not_current])]
        return has_date[self.t >= date[has_date]]

    def update_states_pre(self, t):
        """people.py:164-186."""
        self.t = t
        self.is_exp = self.exposed.nonzero()[0]
        self.init_flows()
        self.flows['new_infectious'] += len(self.check_infectious())
        self.flows['new_symptomatic'] += len(self.check_symptomatic())
        self.flows['new_severe'] += len(self.check_severe())
        self.flows['new_critical'] += len(self.check_critical())
        self.flows['new_recoveries'] += len(self.check_recovery())
        self.check_exit_iso()
        _, new_deaths, new_known_deaths = self.check_death()
        self.flows['new_deaths'] += new_deaths
        self.flows['new_known_deaths'] += new_known_deaths
        if self.pars['use_waning']:
            self.check_immunity()

    def check_infectious(self):
        """people.py:222-232."""
        inds = self.check_inds(self.infectious, self.date_infectious, filter_inds=self.is_exp)
        self.infectious[inds] = True
        self.infectious_variant[inds] = self.exposed_variant[inds]
        for variant in range(self.pars['n_variants']):
            this = inds[self.infectious_variant[inds] == variant]
            self.flows_variant['new_infectious_by_variant'][variant] += len(this)
            self.infectious_by_variant[variant, this] = True
        return inds

    def check_symptomatic(self):
        inds = self.check_inds(self.symptomatic, self.date_symptomatic, filter_inds=self.is_exp)  # people.py:235-239
        self.symptomatic[inds] = True
        return inds

    def check_severe(self):
        inds = self.check_inds(self.severe, self.date_severe, filter_inds=self.is_exp)  # people.py:242-246
        self.severe[inds] = True
        return inds

    def check_critical(self):
        inds = self.check_inds(self.critical, self.date_critical, filter_inds=self.is_exp)  # people.py:249-253
        self.critical[inds] = True
        return inds

    def check_recovery(self):
        """people.py:256-291."""
        inds = self.check_inds(self.recovered, self.date_recovered, filter_inds=self.is_exp)
        self.exposed[inds] = False
        self.infectious[inds] = False
        self.symptomatic[inds] = False
        self.severe[inds] = False
        self.critical[inds] = False
        self.recovered[inds] = True
        self.recovered_variant[inds] = self.exposed_variant[inds]
        self.infectious_variant[inds] = np.nan
        self.exposed_variant[inds] = np.nan
        self.exposed_by_variant[:, inds] = False
        self.infectious_by_variant[:, inds] = False
        if self.pars['use_waning']:
            self.susceptible[inds] = True
            self.diagnosed[inds] = False
        return inds

    def check_death(self):
        """people.py:294-312."""
        inds = self.check_inds(self.dead, self.date_dead, filter_inds=self.is_exp)
        self.dead[inds] = True
        diag_inds = inds[self.diagnosed[inds]]
        self.known_dead[diag_inds] = True
        for key in ['susceptible', 'exposed', 'infectious', 'symptomatic', 'severe', 'critical', 'known_contact',
                    'quarantined', 'recovered']:
            getattr(self, key)[inds] = False
        self.infectious_variant[inds] = np.nan
        self.exposed_variant[inds] = np.nan
        self.recovered_variant[inds] = np.nan
        return inds, len(inds), len(diag_inds)

    def check_exit_iso(self):
        """people.py:368-374."""
        end_inds = self.check_inds(~self.isolated, self.date_end_isolation, filter_inds=None)
        self.isolated[end_inds] = False
        return end_inds

    # -- testing, quarantine and isolation (people.py:189-196, 315-375, 589-641) --

    def update_states_post(self):
        """people.py:189-196."""
        self.flows['new_diagnoses'] += len(self.check_diagnosed())
        self.flows['new_quarantined'] += self.check_quar()
        self.flows['new_isolated'] += len(self.check_enter_iso())

    def check_diagnosed(self):
        """people.py:315-332: positive tests taken are counted; diagnoses that are due are set."""
        test_pos_inds = self.check_inds(self.diagnosed, self.date_pos_test, filter_inds=None)
        self.date_pos_test[test_pos_inds] = np.nan
        diag_inds = self.check_inds(self.diagnosed, self.date_diagnosed, filter_inds=None)
        self.diagnosed[diag_inds] = True
        return test_pos_inds

    def check_quar(self):
        """people.py:335-358: the day's scheduled quarantines in order, then the ends."""
        n_quarantined = 0
        for ind, end_day in self._pending_quarantine.pop(self.t, []):
            if self.quarantined[ind]:
                self.date_end_quarantine[ind] = max(self.date_end_quarantine[ind], end_day)
            elif not (self.dead[ind] or self.recovered[ind] or self.diagnosed[ind] or self.isolated[ind]):
                self.quarantined[ind] = True
                self.date_quarantined[ind] = self.t
                self.date_end_quarantine[ind] = end_day
                n_quarantined += 1
        diag_inds = (self.quarantined & (self.date_diagnosed == self.t)).nonzero()[0]
        self.date_end_quarantine[diag_inds] = self.t
        end_inds = self.check_inds(~self.quarantined, self.date_end_quarantine, filter_inds=None)
        self.quarantined[end_inds] = False
        return n_quarantined

    def check_enter_iso(self):
        """people.py:361-365."""
        iso_inds = (self.date_diagnosed == self.t).nonzero()[0]
        self.isolated[iso_inds] = True
        self.date_end_isolation[iso_inds] = self.date_recovered[iso_inds]
        return iso_inds

    def test(self, inds, pos_u, keep_u, test_sensitivity=1.0, loss_prob=0.0, test_delay=0):
        """people.py:589-617 with the draws given per agent: pos_u[i] < sensitivity is a true
        positive, keep_u[i] < 1 - loss_prob keeps the result."""
        inds = np.unique(inds)
        self.tested[inds] = True
        self.date_tested[inds] = self.t
        is_infectious = inds[self.infectious[inds]]
        is_inf_pos = is_infectious[pos_u[is_infectious] < test_sensitivity]
        not_diagnosed = is_inf_pos[np.isnan(self.date_diagnosed[is_inf_pos])]
        final_inds = not_diagnosed[keep_u[not_diagnosed] < 1.0 - loss_prob]
        self.date_diagnosed[final_inds] = self.t + test_delay
        self.date_pos_test[final_inds] = self.t
        return final_inds

    def schedule_quarantine(self, inds, start_date=None, period=None):
        """people.py:620-641."""
        start_date = self.t if start_date is None else int(start_date)
        period = self.pars['quar_period'] if period is None else int(period)
        for ind in inds:
            self._pending_quarantine[start_date].append((int(ind), start_date + period))

    def check_immunity(self):
        """immunity.py:303-350 (no vaccines on this path: vaccine_imm = 0)."""
        pars = self.pars
        for variant in range(pars['n_variants']):
            natural_imm = np.zeros(self.n)
            with np.errstate(invalid='ignore'):
                was_inf = (self.t >= self.date_recovered).nonzero()[0]
            natural_imm[was_inf] = pars['immunity'][variant, :][self.recovered_variant[was_inf].astype(int)]
            imm = np.maximum(natural_imm, np.zeros(self.n))
            effective_nabs = self.nab * imm
            self.sus_imm[variant, :] = calc_VE(effective_nabs, 'sus', pars['nab_eff'])
            self.symp_imm[variant, :] = calc_VE(effective_nabs, 'symp', pars['nab_eff'])
            self.sev_imm[variant, :] = calc_VE(effective_nabs, 'sev', pars['nab_eff'])

    # -- infect (people.py:435-586) --

    def infect(self, inds, rng, hosp_max=None, icu_max=None, source=None, layer=None, variant=0):
        if len(inds) == 0:
            return inds
        pars = self.pars
        t = self.t
        inds, unique = np.unique(inds, return_index=True)  # :465 first occurrence wins
        if source is not None:
            source = source[unique]
        keep = self.susceptible[inds]  # :470
        inds = inds[keep]
        if source is not None:
            source = source[keep]

        variant_keys = ['rel_symp_prob', 'rel_severe_prob', 'rel_crit_prob', 'rel_death_prob']
        infect_pars = {k: pars[k] for k in variant_keys}  # :477
        if variant:
            for k in variant_keys:
                infect_pars[k] *= pars['variant_pars'][variant][k]
        durpars = pars['dur']

        breakthrough_inds = inds[self.peak_nab[inds].nonzero()[0]]  # :486
        if len(breakthrough_inds):
            no_prior = self.n_breakthroughs[breakthrough_inds] == 0
            new_bt = breakthrough_inds[no_prior]
            self.rel_trans[new_bt] *= F32(pars['trans_redux'])  # :491

        n_infections = len(inds)
        self.susceptible[inds] = False
        self.naive[inds] = False
        self.recovered[inds] = False
        self.diagnosed[inds] = False
        self.exposed[inds] = True
        self.n_infections[inds] += 1
        self.n_breakthroughs[breakthrough_inds] += 1
        self.exposed_variant[inds] = variant
        self.exposed_by_variant[variant, inds] = True
        self.flows['new_infections'] += n_infections
        self.flows['new_reinfections'] += int((~np.isnan(self.date_recovered[inds])).sum())
        self.flows_variant['new_infections_by_variant'][variant] += n_infections

        for i, target in enumerate(inds):  # :509-511
            self.infection_log.append((int(source[i]) if source is not None else -1, int(target), t, layer, variant))

        self.dur_exp2inf[inds] = rng.sample(durpars['exp2inf'], inds, t, px.SLOT_EXP2INF)  # :514
        self.date_exposed[inds] = t
        self.date_infectious[inds] = self.dur_exp2inf[inds] + F32(t)  # f32 + int -> f32
        for key in ['date_symptomatic', 'date_severe', 'date_critical', 'date_diagnosed', 'date_recovered']:
            getattr(self, key)[inds] = np.nan

        symp_probs = F32(infect_pars['rel_symp_prob']) * self.symp_prob[inds] * (F32(1) - self.symp_imm[variant, inds])  # :523
        is_symp = rng.bernoulli(symp_probs, inds, t, px.SLOT_BERN_SYMP_SEV, 0)
        symp_inds = inds[is_symp]
        asymp_inds = inds[~is_symp]
        self.flows_variant['new_symptomatic_by_variant'][variant] += len(symp_inds)

        dur_asym2rec = rng.sample(durpars['asym2rec'], asymp_inds, t, px.SLOT_DUR_A)  # :530
        self.date_recovered[asymp_inds] = self.date_infectious[asymp_inds] + dur_asym2rec  # f32 + f64 -> f64 -> f32
        self.dur_disease[asymp_inds] = self.dur_exp2inf[asymp_inds] + dur_asym2rec

        self.dur_inf2sym[symp_inds] = rng.sample(durpars['inf2sym'], symp_inds, t, px.SLOT_DUR_A)  # :536
        self.date_symptomatic[symp_inds] = self.date_infectious[symp_inds] + self.dur_inf2sym[symp_inds]
        sev_probs = F32(infect_pars['rel_severe_prob']) * self.severe_prob[symp_inds] * (F32(1) - self.sev_imm[variant, symp_inds])
        is_sev = rng.bernoulli(sev_probs, symp_inds, t, px.SLOT_BERN_SYMP_SEV, 1)
        sev_inds = symp_inds[is_sev]
        mild_inds = symp_inds[~is_sev]
        self.flows_variant['new_severe_by_variant'][variant] += len(sev_inds)

        dur_mild2rec = rng.sample(durpars['mild2rec'], mild_inds, t, px.SLOT_DUR_B)  # :545
        self.date_recovered[mild_inds] = self.date_symptomatic[mild_inds] + dur_mild2rec
        self.dur_disease[mild_inds] = self.dur_exp2inf[mild_inds] + self.dur_inf2sym[mild_inds] + dur_mild2rec

        self.dur_sym2sev[sev_inds] = rng.sample(durpars['sym2sev'], sev_inds, t, px.SLOT_DUR_B)  # :550
        self.date_severe[sev_inds] = self.date_symptomatic[sev_inds] + self.dur_sym2sev[sev_inds]
        crit_probs = F32(infect_pars['rel_crit_prob']) * self.crit_prob[sev_inds] * F32(pars['no_hosp_factor'] if hosp_max else 1.)
        is_crit = rng.bernoulli(crit_probs, sev_inds, t, px.SLOT_BERN_CRIT_DEATH, 0)
        crit_inds = sev_inds[is_crit]
        non_crit_inds = sev_inds[~is_crit]

        dur_sev2rec = rng.sample(durpars['sev2rec'], non_crit_inds, t, px.SLOT_DUR_C)  # :558
        self.date_recovered[non_crit_inds] = self.date_severe[non_crit_inds] + dur_sev2rec
        self.dur_disease[non_crit_inds] = (self.dur_exp2inf[non_crit_inds] + self.dur_inf2sym[non_crit_inds]
                                           + self.dur_sym2sev[non_crit_inds] + dur_sev2rec)

        self.dur_sev2crit[crit_inds] = rng.sample(durpars['sev2crit'], crit_inds, t, px.SLOT_DUR_C)  # :563
        self.date_critical[crit_inds] = self.date_severe[crit_inds] + self.dur_sev2crit[crit_inds]
        death_probs = F32(infect_pars['rel_death_prob']) * self.death_prob[crit_inds] * F32(pars['no_icu_factor'] if icu_max else 1.)
        is_dead = rng.bernoulli(death_probs, crit_inds, t, px.SLOT_BERN_CRIT_DEATH, 1)
        dead_inds = crit_inds[is_dead]
        alive_inds = crit_inds[~is_dead]

        dur_crit2rec = rng.sample(durpars['crit2rec'], alive_inds, t, px.SLOT_DUR_D)  # :571
        self.date_recovered[alive_inds] = self.date_critical[alive_inds] + dur_crit2rec
        self.dur_disease[alive_inds] = (self.dur_exp2inf[alive_inds] + self.dur_inf2sym[alive_inds]
                                        + self.dur_sym2sev[alive_inds] + self.dur_sev2crit[alive_inds] + dur_crit2rec)

        dur_crit2die = rng.sample(durpars['crit2die'], dead_inds, t, px.SLOT_DUR_D)  # :576
        self.date_dead[dead_inds] = self.date_critical[dead_inds] + dur_crit2die
        self.dur_disease[dead_inds] = (self.dur_exp2inf[dead_inds] + self.dur_inf2sym[dead_inds]
                                       + self.dur_sym2sev[dead_inds] + self.dur_sev2crit[dead_inds] + dur_crit2die)
        self.date_recovered[dead_inds] = np.nan

        if pars['use_waning']:  # :582-584
            self.update_peak_nab(inds, asymp_inds, mild_inds, sev_inds, rng)
        return inds

    def update_peak_nab(self, inds, asymp_inds, mild_inds, sev_inds, rng):
        """immunity.py:138-202 for natural infection."""
        pars = self.pars
        has_nabs = self.nab[inds] > 0
        no_prior_nab_inds = inds[~has_nabs]
        prior_nab_inds = inds[has_nabs]
        if len(prior_nab_inds):
            self.peak_nab[prior_nab_inds] *= F32(pars['nab_boost'])
        if len(no_prior_nab_inds):
            init_nab = rng.sample(pars['nab_init'], no_prior_nab_inds, self.t, px.SLOT_NAB)
            no_prior_nab = (2 ** init_nab)
            prior_symp = np.full(self.n, np.nan)
            prior_symp[asymp_inds] = pars['rel_imm_symp']['asymp']
            prior_symp[mild_inds] = pars['rel_imm_symp']['mild']
            prior_symp[sev_inds] = pars['rel_imm_symp']['severe']
            prior_symp[prior_nab_inds] = np.nan
            prior_symp = prior_symp[~np.isnan(prior_symp)]
            norm_factor = 1 + pars['nab_eff']['alpha_inf_diff']
            no_prior_nab = no_prior_nab * prior_symp * norm_factor
            self.peak_nab[no_prior_nab_inds] = no_prior_nab
        self.t_nab_event[inds] = self.t

    def update_nab(self, inds):
        """immunity.py:205-213."""
        t_since_boost = self.t - self.t_nab_event[inds]
        self.nab[inds] += self.pars['nab_kin'][t_since_boost] * self.peak_nab[inds]
        self.nab[inds] = np.where(self.nab[inds] < 0, 0, self.nab[inds])
        self.nab[inds] = np.minimum(self.nab[inds], self.peak_nab[inds])


# ---------------------------------------------------------------------------
# Sim loop (covasim/sim.py:505-532, 558-685, 764-787)
# ---------------------------------------------------------------------------

class Sim:
    """Oracle sim over a fixed population: ``layers`` is an ordered dict lkey -> (p1, p2, beta)."""

    def __init__(self, pars, people_init, layers, rng):
        self.pars = pars
        self.npts = int(pars['n_days']) + 1
        if pars['use_waning']:
            if pars.get('immunity') is None:
                pars['immunity'] = init_immunity_matrix(pars['n_variants'])
            pars['nab_kin'] = precompute_waning(self.npts, pars['nab_decay'])
        self.people = People(pars, **people_init)
        self.layers = {k: [np.asarray(v[0], np.int32).copy(), np.asarray(v[1], np.int32).copy(),
                           np.asarray(v[2], F32).copy()] for k, v in layers.items()}
        self.rng = rng
        self.interventions = []  # called each day after the variant imports, as sim.py:591-593
        # per dynamic layer key, optionally, the update of a Layer subclass (examples/t11_dynamic_layers.py:9-26):
        # fn(t, layer index, (p1, p2, beta), n_agents) -> the day's (p1, p2, beta), of any length
        self.layer_updates = {}
        self.t = 0
        nv = pars['n_variants']
        self.results = {k: np.zeros(self.npts) for k in FLOW_KEYS}
        for k in STOCK_KEYS:
            self.results[f'n_{k}'] = np.zeros(self.npts)
        for k in VFLOW_KEYS:
            self.results[k] = np.zeros((nv, self.npts))
        for k in ['n_exposed_by_variant', 'n_infectious_by_variant']:
            self.results[k] = np.zeros((nv, self.npts))
        for k in ['n_imports', 'pop_nabs', 'pop_protection', 'pop_symp_protection']:
            self.results[k] = np.zeros(self.npts)

    def step(self):
        pars, people, rng = self.pars, self.people, self.rng
        t = self.t
        nv = pars['n_variants']
        if t == 0 and pars['pop_infected']:  # sim.py:570-572 -> init_infections :505-532
            inds = rng.choose_seeds(people.n, pars['pop_infected'])
            people.infect(inds=inds, rng=rng, layer='seed_infection')
        people.update_states_pre(t=t)  # sim.py:577
        for li, (lkey, lay) in enumerate(self.layers.items()):  # people.update_contacts, people.py:199-206
            if not pars['dynam_layer'].get(lkey, 0):
                continue
            if lkey in self.layer_updates:  # a Layer subclass's update(people): new (p1, p2, beta)
                new = self.layer_updates[lkey](t, li, tuple(lay), people.n)
                lay[0], lay[1], lay[2] = (np.asarray(new[0], np.int32), np.asarray(new[1], np.int32),
                                          np.asarray(new[2], F32))
            else:
                lay[0], lay[1] = rng.resample(t, li, len(lay[0]), people.n)
                lay[2] = np.ones(len(lay[0]), dtype=F32)
        hosp_max = people.severe.sum() > pars['n_beds_hosp'] if pars['n_beds_hosp'] is not None else False
        icu_max = people.critical.sum() > pars['n_beds_icu'] if pars['n_beds_icu'] is not None else False
        for var, vi in zip(pars['variants'], pars['variant_index']):  # immunity.py:117-130
            if t in var.get('days', []):
                sus_inds = people.susceptible.nonzero()[0]
                n_imports = int(var.get('n_imports', 1))
                imp = rng.choose_imports(sus_inds, n_imports, t, vi)
                people.infect(inds=imp, rng=rng, layer='importation', variant=vi)
                self.results['n_imports'][t] += n_imports
        for i, intervention in enumerate(self.interventions):  # sim.py:591-593
            self.iv_slot = i  # an intervention's stream slot is its index in the list
            intervention(self)
        people.update_states_post()  # sim.py:596

        vd = pars['viral_dist']
        viral_load = compute_viral_load(t, people.date_infectious, people.date_recovered, people.date_dead,
                                        vd['frac_time'], vd['load_ratio'], vd['high_cap'])  # sim.py:602-608
        sus = people.susceptible
        for variant in range(nv):  # sim.py:622-649
            beta = F32(pars['beta'] * pars['rel_beta'] * pars['variant_pars'][variant]['rel_beta'])
            inf_variant = people.infectious * (people.infectious_variant == variant)
            if ~inf_variant.any():
                continue
            for li, (lkey, (p1, p2, betas)) in enumerate(self.layers.items()):
                sus_imm = people.sus_imm[variant, :]
                rel_trans, rel_sus = compute_trans_sus(people.rel_trans, people.rel_sus, inf_variant, sus,
                                                       pars['beta_layer'][lkey], viral_load, people.symptomatic,
                                                       people.isolated, people.quarantined, pars['asymp_factor'],
                                                       pars['iso_factor'][lkey], pars['quar_factor'][lkey], sus_imm)
                src, tgt, _ = compute_infections(beta, p1, p2, betas, rel_trans, rel_sus,
                                                 rng.edge_uniform_fn(t, variant, li))
                people.infect(inds=tgt, rng=rng, hosp_max=hosp_max, icu_max=icu_max, source=src, layer=lkey,
                              variant=variant)

        res = self.results  # sim.py:651-663
        for key in STOCK_KEYS:
            res[f'n_{key}'][t] = np.count_nonzero(getattr(people, key))
        for v in range(nv):
            res['n_exposed_by_variant'][v, t] = np.count_nonzero(people.exposed_by_variant[v])
            res['n_infectious_by_variant'][v, t] = np.count_nonzero(people.infectious_by_variant[v])
        for key, count in people.flows.items():
            res[key][t] += count
        for key, count in people.flows_variant.items():
            res[key][:, t] += count

        if pars['use_waning']:  # sim.py:666-669
            has_nabs = people.peak_nab.nonzero()[0]
            if len(has_nabs):
                people.update_nab(inds=has_nabs)
        inds_alive = (~people.dead).nonzero()[0]  # sim.py:671-674
        nab_alive = people.nab[inds_alive]
        res['pop_nabs'][t] = np.sum(nab_alive[nab_alive.nonzero()[0]], dtype=F64) / len(inds_alive)
        res['pop_protection'][t] = np.nanmean(people.sus_imm)
        res['pop_symp_protection'][t] = np.nanmean(people.symp_imm)
        self.t += 1

    def run(self):
        while self.t < self.npts:
            self.step()
        return self

    def finalize(self):
        """Cumulative flows (sim.py:781-787)."""
        res = self.results
        for k in FLOW_KEYS:
            res['cum_' + k[4:]] = np.cumsum(res[k])
        res['cum_infections'] = res['cum_infections'] + self.pars['pop_infected']
        return res
