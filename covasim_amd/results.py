"""Results on the device: finalisation (covasim/sim.py:764-1001) and reduction over sims
(covasim/run.py:220-303) with the cvk_finalize / cvk_reduce kernels.

``finalize_tables`` turns the day loop's per-day count tables of one sim or a batch into the
finalised result rows (cumulative flows, compute_states, doubling time, r_eff) on the device;
``to_results`` wraps one sim's rows as ``Result`` objects; ``reduce_tables`` gives the median and
quantiles (or mean and standard-deviation bounds) of every row over the sims.
"""

import ctypes as C

import numpy as np

from . import _lib
from .base import Result

__all__ = ['FLOWS', 'STOCKS', 'F_ROW_NAMES', 'FV_ROW_NAMES', 'finalize_tables', 'to_results', 'reduce_tables',
           'initial_period']

FLOWS = ['infections', 'reinfections', 'infectious', 'symptomatic', 'severe', 'critical', 'recoveries', 'deaths',
         'tests', 'diagnoses', 'known_deaths', 'quarantined', 'isolated', 'doses', 'vaccinated']  # defaults.py:167-183
STOCKS = ['susceptible', 'exposed', 'infectious', 'symptomatic', 'severe', 'critical', 'recovered', 'dead',
          'diagnosed', 'known_dead', 'quarantined', 'isolated', 'vaccinated']  # defaults.py:145-159
F_ROW_NAMES = ([f'new_{f}' for f in FLOWS] + [f'cum_{f}' for f in FLOWS] + [f'n_{k}' for k in STOCKS] +
               ['n_alive', 'n_naive', 'n_preinfectious', 'n_removed', 'prevalence', 'incidence', 'frac_vaccinated',
                'n_imports', 'pop_nabs', 'pop_protection', 'pop_symp_protection', 'doubling_time', 'r_eff',
                'test_yield', 'rel_test_yield'])
FV_ROW_NAMES = ['new_infections_by_variant', 'new_symptomatic_by_variant', 'new_severe_by_variant',
                'new_infectious_by_variant', 'cum_infections_by_variant', 'cum_symptomatic_by_variant',
                'cum_severe_by_variant', 'cum_infectious_by_variant', 'n_exposed_by_variant', 'n_infectious_by_variant',
                'incidence_by_variant', 'prevalence_by_variant']
UNSCALED = {'prevalence', 'incidence', 'frac_vaccinated', 'pop_nabs', 'pop_protection', 'pop_symp_protection',
            'doubling_time', 'r_eff', 'test_yield', 'rel_test_yield', 'incidence_by_variant', 'prevalence_by_variant'}


def initial_period(pars, npts):
    """int(min(npts, exp2inf par1 + asym2rec par1)): the seed-infection window r_eff averages over (sim.py:943-944)."""
    dur = pars['dur'][0] if isinstance(pars['dur'], list) else pars['dur']
    return int(min(npts, dur['exp2inf']['par1'] + dur['asym2rec']['par1']))


def finalize_tables(results, results_f, npts, n_variants, use_waning, pop_infected, init_period, agent_offsets,
                    imports, date_infectious, date_recovered, date_dead, smoothing=2, window=3, max_doubling=30.0,
                    stream=None):
    """Device rows [S][len(F_ROW_NAMES)][npts] and [S][len(FV_ROW_NAMES)][n_variants][npts] (float64).

    results / results_f: device pointers (ints) or tensors of the [S][npts][cols] count tables;
    pop_infected, init_period, agent_offsets: host arrays of S, S and S + 1 entries; imports: host
    [S][npts] imported agents per day (or None); dates: device float32 tensors over all agents.
    """
    import torch
    L = _lib.lib()
    S = len(pop_infected)
    dev = date_infectious.device
    pi = np.ascontiguousarray(pop_infected, dtype=np.float64)
    ip = np.ascontiguousarray(init_period, dtype=np.int32)
    off = np.ascontiguousarray(agent_offsets, dtype=np.int64)
    d = _lib.FinalizeDesc(n_sims=S, npts=int(npts), n_variants=int(n_variants), use_waning=int(bool(use_waning)),
                          r_eff_smoothing=int(smoothing), doubling_window=int(window),
                          max_doubling_time=float(max_doubling), pop_infected_host=pi.ctypes.data,
                          initial_period_host=ip.ctypes.data, agent_offsets_host=off.ctypes.data)
    nbytes = int(L.cvk_finalize_work_bytes(C.byref(d)))
    if nbytes <= 0:
        raise ValueError('cvk_finalize: ' + L.cvk_last_error().decode(errors='replace'))
    buf = torch.empty(nbytes + 256, dtype=torch.uint8, device=dev)
    work = C.c_void_p(buf.data_ptr() + (-buf.data_ptr()) % 256)
    out = torch.empty((S, len(F_ROW_NAMES), int(npts)), dtype=torch.float64, device=dev)
    out_v = torch.empty((S, len(FV_ROW_NAMES), int(n_variants), int(npts)), dtype=torch.float64, device=dev)
    imp = None
    if imports is not None:
        imp = torch.as_tensor(np.ascontiguousarray(imports, dtype=np.float64).reshape(S, int(npts)), device=dev)

    def p(x):
        return C.c_void_p(x if isinstance(x, int) else x.data_ptr())
    _lib.check(L.cvk_finalize(C.byref(d), p(results), p(results_f), _lib.ptr(imp), _lib.ptr(date_infectious),
                              _lib.ptr(date_recovered), _lib.ptr(date_dead), _lib.ptr(out), _lib.ptr(out_v), work,
                              nbytes, _lib.stream_ptr(stream)), 'cvk_finalize')
    del buf
    return out, out_v


def to_results(rows, vrows):
    """Result objects of one sim from its host rows [F_ROWS][npts] and [FV_ROWS][nv][npts]."""
    R = {name: Result(name, scale=name not in UNSCALED, values=rows[i].copy()) for i, name in enumerate(F_ROW_NAMES)}
    R['variant'] = {name: Result(name, scale=name not in UNSCALED, values=vrows[i].copy())
                    for i, name in enumerate(FV_ROW_NAMES)}
    return R


def reduce_tables(values, quantiles=None, use_mean=False, bounds=None, stream=None):
    """(center, low, high) over the leading (sim) axis of a device float64 tensor (run.py:220-303):
    median and quantiles (default 0.1 / 0.9, numpy's 'linear' method) or mean -/+ bounds * std."""
    import torch
    L = _lib.lib()
    v = values.to(torch.float64).contiguous()
    S = int(v.shape[0])
    shape = tuple(v.shape[1:])
    n_cols = int(np.prod(shape)) if shape else 1
    if use_mean:
        bounds = 2.0 if bounds is None else float(bounds)
        q = (0.0, 0.0)
    else:
        if quantiles is None:
            quantiles = {'low': 0.1, 'high': 0.9}
        if not isinstance(quantiles, dict):
            quantiles = {'low': float(quantiles[0]), 'high': float(quantiles[1])}
        q = (float(quantiles['low']), float(quantiles['high']))
        bounds = 0.0
    outs = [torch.empty(shape, dtype=torch.float64, device=v.device) for _ in range(3)]
    _lib.check(L.cvk_reduce(S, n_cols, _lib.ptr(v), int(bool(use_mean)), q[0], q[1], bounds, _lib.ptr(outs[0]),
                            _lib.ptr(outs[1]), _lib.ptr(outs[2]), _lib.stream_ptr(stream)), 'cvk_reduce')
    return tuple(outs)
