"""Variants and NAb kinetics tables consumed by the kernels (covasim/immunity.py restated).

The per-agent immunity work (check_immunity, update_peak_nab, update_nab) runs in the
HIP kernels; this module builds the host-side constants they read: the NAb kinetics
table (precompute_waning, immunity.py:356-496) and the cross-immunity matrix
(init_immunity, immunity.py:269-300).
"""

import numpy as np

from . import parameters as cvpar

__all__ = ['variant', 'precompute_waning', 'nab_growth_decay', 'exp_decay', 'init_immunity_matrix']


class variant:
    """A variant introduced on given days with n_imports seed infections (immunity.py:18-130).

    ``variant`` is a known label ('alpha', 'beta', 'gamma', 'delta') or a dict of
    variant parameters (rel_beta, rel_symp_prob, rel_severe_prob, rel_crit_prob,
    rel_death_prob; missing keys default to the wild type).
    """

    def __init__(self, variant, days, label=None, n_imports=1, rescale=True):
        self.days = days
        self.n_imports = int(n_imports)
        self.rescale = rescale
        self.index = None
        if isinstance(variant, str):  # a known variant or an alias of one (immunity.py:53-67)
            label = variant.lower()
            for txt in ['.', ' ', 'variant', 'voc']:
                label = label.replace(txt, '')
            choices, mapping = cvpar.get_variant_choices()
            if label not in mapping:
                raise NotImplementedError(f'The selected variant "{variant}" is not implemented; choices are {choices}')
            label = mapping[label]
            self.p = dict(cvpar.get_variant_pars()[label])
        elif isinstance(variant, dict):
            p = dict(variant)
            label = p.pop('label', label) or 'custom'
            default = cvpar.get_variant_pars(default=True)
            invalid = [k for k in p if k not in default]
            if invalid:
                raise KeyError(f'Could not parse variant keys {invalid}; valid keys are {list(default)}')
            for k, v in default.items():
                p.setdefault(k, v)
            self.p = p
        else:
            raise ValueError(f'Could not understand {type(variant)}; specify a dict or a known variant label')
        self.label = label

    def day_list(self, npts):
        days = self.days
        if np.isscalar(days):
            days = [days]
        return [int(d) for d in days if 0 <= int(d) < npts]


def nab_growth_decay(length, growth_time, decay_rate1, decay_time1, decay_rate2, decay_time2):
    """Linear growth then exponential decay with a slowing rate (immunity.py:404-448)."""
    def f1(t):
        return (1 / growth_time) * t

    def f2(t):
        decay_rate = np.full(len(t), fill_value=decay_rate1)
        decay_rate[np.flatnonzero(t > decay_time2)] = decay_rate2
        slowing = (1 / (decay_time2 - decay_time1)) * (decay_rate1 - decay_rate2)
        mid = np.flatnonzero((t > decay_time1) * (t <= decay_time2))
        decay_rate[mid] = decay_rate1 - slowing * np.arange(len(mid), dtype=np.int32)
        titre = np.concatenate([[0.0], np.cumsum(decay_rate[1:])]) if len(t) else np.zeros(0)
        return np.exp(-titre)

    if decay_time2 < decay_time1:
        raise ValueError(f'Decay time 2 must be larger than decay time 1, but you supplied {decay_time2} < {decay_time1}')
    length = length + 1
    t1 = np.arange(growth_time, dtype=np.int32)
    t2 = np.arange(length - growth_time, dtype=np.int32)
    y = np.concatenate([f1(t1), f2(t2)])
    return np.diff(y)[0:length]


def exp_decay(length, init_val, half_life, delay=None):
    """Exponential decay from init_val (immunity.py:482-496)."""
    length = length + 1
    decay_rate = np.log(2) / half_life if not np.isnan(half_life) else 0.
    if delay is not None:
        t = np.arange(length - delay, dtype=np.int32)
        growth = (init_val / delay) * np.ones(delay)
        result = np.concatenate([growth, init_val * np.exp(-decay_rate * t)], axis=None)
    else:
        t = np.arange(length, dtype=np.int32)
        result = init_val * np.exp(-decay_rate * t)
    return np.diff(result)


def precompute_waning(length, pars):
    """NAb kinetics table of the given length (immunity.py:356-401)."""
    pars = dict(pars)
    form = pars.pop('form', None)
    if form is None or form == 'nab_growth_decay':
        return nab_growth_decay(length, **pars)
    if form == 'exp_decay':
        if pars.get('half_life') is None:
            pars['half_life'] = np.nan
        return exp_decay(length, **pars)
    if callable(form):
        return np.asarray(form(length, **pars), dtype=float)
    raise NotImplementedError(f'The selected functional form "{form}" is not implemented')


def init_immunity_matrix(variant_labels):
    """Cross-immunity matrix: 1.0 by default, known pairs from get_cross_immunity (immunity.py:283-296)."""
    nv = len(variant_labels)
    immunity = np.ones((nv, nv), dtype=np.float32)
    cross = cvpar.get_cross_immunity()
    for i in range(nv):
        for j in range(nv):
            li, lj = variant_labels[i], variant_labels[j]
            if li in cross and lj in cross:
                immunity[j][i] = cross[lj][li]
    return immunity
