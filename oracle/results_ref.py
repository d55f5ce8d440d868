"""Results finalisation and multi-sim reduction restated in numpy (TEST INFRASTRUCTURE).

Restates covasim/sim.py:764-850 (finalize: cumulative flows, compute_states), :872-895
(compute_doubling), :898-1001 (compute_r_eff, method 'daily') and covasim/run.py:220-303
(MultiSim.reduce) on the per-day count table of one sim. The checker of the device finalisation
(cvk_finalize) and reduction (cvk_reduce).

compute_r_eff smooths with sciris' ``sc.smooth`` (sciris is not installed here): it is restated
from sciris' published source as ``repeats`` passes of ``np.convolve(x, [0.25, 0.5, 0.25],
mode='same')``. The restatement is pinned by closed forms in the tests (an impulse spreads to the
binomial weights), not against sciris itself.
"""

import numpy as np

FLOWS = ['infections', 'reinfections', 'infectious', 'symptomatic', 'severe', 'critical', 'recoveries', 'deaths',
         'tests', 'diagnoses', 'known_deaths', 'quarantined', 'isolated', 'doses', 'vaccinated']  # defaults.py:167-183
STOCKS = ['susceptible', 'exposed', 'infectious', 'symptomatic', 'severe', 'critical', 'recovered', 'dead',
          'diagnosed', 'known_dead', 'quarantined', 'isolated', 'vaccinated']  # defaults.py:145-159
VARIANT_FLOWS = ['infections_by_variant', 'symptomatic_by_variant', 'severe_by_variant', 'infectious_by_variant']


def smooth(data, repeats):
    """sc.smooth of a 1-D array: repeated convolution with (0.25, 0.5, 0.25), zero padded."""
    out = np.array(data, dtype=float)
    kernel = np.array([0.25, 0.5, 0.25])
    for _ in range(repeats):
        out = np.convolve(out, kernel, mode='same')
    return out


def compute_doubling(cum_infections, window=3, max_doubling_time=30):
    """sim.py:872-895."""
    now = cum_infections[window:]
    prev = cum_infections[:-window]
    use = (prev > 0) & (now > prev)
    out = np.full(len(cum_infections), np.nan)
    dt = window * np.log(2) / np.log(now[use] / prev[use])
    out[window:][use] = np.minimum(dt, max_doubling_time)
    return out


def compute_r_eff(new_infections, n_imports, n_infectious, date_infectious, date_recovered, date_dead,
                  initial_period_par, smoothing=2):
    """sim.py:898-952 ('daily'): float32 date arrays, as People holds them."""
    npts = len(new_infections)
    recov = np.flatnonzero(~np.isnan(date_recovered))
    dead = np.flatnonzero(~np.isnan(date_dead))
    date_outcome = np.concatenate((date_recovered[recov], date_dead[dead]))
    inds = np.concatenate((recov, dead))
    date_inf = date_infectious[inds]
    mean_inf = date_outcome.mean() - date_inf.mean() if len(date_outcome) else 0
    new = new_infections - n_imports
    raw = mean_inf * np.divide(new, n_infectious, out=np.zeros(npts), where=n_infectious > 0)
    if npts >= 3:
        ip = int(min(npts, initial_period_par))
        for ind in range(ip):
            raw[ind] = raw[ind:ip].mean()
        values = smooth(raw, smoothing)
        values[:smoothing] = raw[:smoothing]
        values[-smoothing:] = raw[-smoothing:]
    else:
        values = raw
    return values


def finalize(counts, counts_f, col, byvar_base, max_variants, n_variants, pop_size, pop_infected, use_waning,
             imports, date_infectious, date_recovered, date_dead, initial_period_par, smoothing=2):
    """Rows of one sim's finalised results from its [npts][cols] count table (sim.py:764-850)."""
    counts = np.asarray(counts, dtype=np.float64)
    npts = counts.shape[0]
    R = {}
    for f in FLOWS:
        name = f'new_{f}'
        R[name] = counts[:, col[name]].copy() if name in col else np.zeros(npts)
    for k in STOCKS:
        R[f'n_{k}'] = counts[:, col[f'n_{k}']].copy()
    for f in FLOWS:
        R[f'cum_{f}'] = np.cumsum(R[f'new_{f}'])
    R['cum_infections'] = R['cum_infections'] + pop_infected
    V = {}
    for k, name in enumerate(['new_infections_by_variant', 'new_symptomatic_by_variant', 'new_severe_by_variant',
                              'new_infectious_by_variant', 'n_exposed_by_variant', 'n_infectious_by_variant']):
        base = byvar_base + k * max_variants
        V[name] = counts[:, base:base + n_variants].T.copy()
    for f in VARIANT_FLOWS:
        V[f'cum_{f}'] = np.cumsum(V[f'new_{f}'], axis=1)
    V['cum_infections_by_variant'] += pop_infected  # every variant row, sim.py:785-786 (base.py:146-147)
    count_recov = 1 - int(bool(use_waning))  # sim.py:836-850
    with np.errstate(divide='ignore', invalid='ignore'):
        R['n_alive'] = pop_size - R['cum_deaths']
        R['n_naive'] = pop_size - R['cum_deaths'] - R['n_recovered'] - R['n_exposed']
        R['n_susceptible'] = R['n_alive'] - R['n_exposed'] - count_recov * R['cum_recoveries']
        R['n_preinfectious'] = R['n_exposed'] - R['n_infectious']
        R['n_removed'] = count_recov * R['cum_recoveries'] + R['cum_deaths']
        R['prevalence'] = R['n_exposed'] / R['n_alive']
        R['incidence'] = R['new_infections'] / R['n_susceptible']
        R['frac_vaccinated'] = R['n_vaccinated'] / R['n_alive']
        V['incidence_by_variant'] = np.einsum('ji,i->ji', V['new_infections_by_variant'], 1 / R['n_susceptible'])
        V['prevalence_by_variant'] = np.einsum('ji,i->ji', V['new_infections_by_variant'], 1 / R['n_alive'])
        R['n_imports'] = np.zeros(npts) if imports is None else np.asarray(imports, dtype=float)
        R['pop_nabs'] = counts_f[:, 0] / counts[:, col['n_alive']]
        R['pop_protection'] = counts_f[:, 1] / (n_variants * pop_size)
        R['pop_symp_protection'] = counts_f[:, 2] / (n_variants * pop_size)
    R['test_yield'] = np.zeros(npts)  # compute_yield, sim.py:853-868
    inds = R['new_tests'].nonzero()[0]
    R['test_yield'][inds] = R['new_diagnoses'][inds] / R['new_tests'][inds]
    R['rel_test_yield'] = np.zeros(npts)
    inds = R['n_infectious'].nonzero()[0]
    with np.errstate(divide='ignore', invalid='ignore'):
        denom = R['n_infectious'][inds] / (R['n_alive'][inds] - R['cum_diagnoses'][inds])
        R['rel_test_yield'][inds] = R['test_yield'][inds] / denom
    R['doubling_time'] = compute_doubling(R['cum_infections'])
    R['r_eff'] = compute_r_eff(R['new_infections'], R['n_imports'], R['n_infectious'], date_infectious,
                               date_recovered, date_dead, initial_period_par, smoothing)
    R['variant'] = V
    return R


def reduce(values, quantiles=(0.1, 0.9), use_mean=False, bounds=2):
    """run.py:266-290 over values[sim][...]: (center, low, high) with numpy's own routines, on the
    reference's layout (series x sims, sims contiguous, reduced along the last axis)."""
    values = np.asarray(values, dtype=np.float64)
    shape = values.shape[1:]
    raw = np.ascontiguousarray(values.reshape(values.shape[0], -1).T)
    if use_mean:
        mean = np.mean(raw, axis=1)
        std = np.std(raw, axis=1)
        out = (mean, mean - bounds * std, mean + bounds * std)
    else:
        out = (np.quantile(raw, 0.5, axis=1), np.quantile(raw, quantiles[0], axis=1),
               np.quantile(raw, quantiles[1], axis=1))
    return tuple(o.reshape(shape) for o in out)
