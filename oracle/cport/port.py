"""ctypes driver of the compiled CPU restatement (oracle/cport/covasim_port.c).

TEST INFRASTRUCTURE ONLY (CPU baseline and cross-check of the numpy oracle).
``PortSim`` takes the same parameter dict, initial per-agent arrays and layers as
``oracle.covasim_oracle.Sim`` and produces the same per-day counts.
"""

import ctypes as C
import os

import numpy as np

from .. import covasim_oracle as co

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(os.path.dirname(HERE), '_build')
MAXV, MAXL = 8, 8
DIST = {'uniform': 0, 'unif': 0, 'normal': 1, 'norm': 1, 'normal_pos': 2, 'normal_int': 3, 'lognormal': 4,
        'lognorm': 4, 'lognormal_int': 5, 'lognorm_int': 5}
ROW_KEYS = ['new_infections', 'new_reinfections', 'new_infectious', 'new_symptomatic', 'new_severe', 'new_critical',
            'new_recoveries', 'new_deaths', 'new_known_deaths', 'n_susceptible', 'n_exposed', 'n_infectious',
            'n_symptomatic', 'n_severe', 'n_critical', 'n_recovered', 'n_dead']


class Dur(C.Structure):
    _fields_ = [('dist', C.c_int32), ('pad', C.c_int32), ('par1', C.c_double), ('par2', C.c_double),
                ('mu', C.c_double), ('sigma', C.c_double)]


class Params(C.Structure):
    _fields_ = [('n', C.c_int64), ('n_variants', C.c_int32), ('n_layers', C.c_int32), ('use_waning', C.c_int32),
                ('nab_kin_len', C.c_int32), ('seed', C.c_uint64), ('frac_time', C.c_float),
                ('load_ratio', C.c_float), ('high_cap', C.c_float), ('asymp_factor', C.c_float),
                ('beta_v', C.c_float * MAXV), ('beta_layer', C.c_float * MAXL), ('iso_factor', C.c_float * MAXL),
                ('quar_factor', C.c_float * MAXL), ('dynamic', C.c_int32 * MAXL),
                ('rel_symp', C.c_float * MAXV), ('rel_sev', C.c_float * MAXV), ('rel_crit', C.c_float * MAXV),
                ('rel_death', C.c_float * MAXV), ('no_hosp_factor', C.c_float), ('no_icu_factor', C.c_float),
                ('trans_redux', C.c_float), ('nab_boost', C.c_float), ('n_beds_hosp', C.c_int64),
                ('n_beds_icu', C.c_int64), ('dur', Dur * 9), ('nab_init', Dur), ('rel_imm_symp', C.c_double * 3),
                ('nab_norm', C.c_double), ('ve_exp_alpha', C.c_double * 3), ('ve_beta', C.c_double * 3),
                ('immunity', (C.c_float * MAXV) * MAXV), ('nab_kin', C.c_void_p)]


U8 = ['susceptible', 'naive', 'exposed', 'infectious', 'symptomatic', 'severe', 'critical', 'diagnosed', 'recovered',
      'known_dead', 'dead', 'known_contact', 'quarantined', 'isolated']
F32V = ['exposed_variant', 'infectious_variant', 'recovered_variant']
BYV = ['exposed_by_variant', 'infectious_by_variant']
F32A = ['date_exposed', 'date_infectious', 'date_symptomatic', 'date_severe', 'date_critical', 'date_recovered',
        'date_dead', 'date_diagnosed', 'date_end_isolation', 'dur_exp2inf', 'dur_inf2sym', 'dur_sym2sev',
        'dur_sev2crit', 'dur_disease', 'rel_trans', 'rel_sus', 'symp_prob', 'severe_prob', 'crit_prob', 'death_prob']
I32A = ['n_infections', 'n_breakthroughs', 't_nab_event']
F32B = ['nab', 'peak_nab']
IMM = ['sus_imm', 'symp_imm', 'sev_imm']
PEOPLE_FIELDS = U8 + F32V + BYV + F32A + I32A + F32B + IMM


class People(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in PEOPLE_FIELDS]


class Layer(C.Structure):
    _fields_ = [('p1', C.c_void_p), ('p2', C.c_void_p), ('beta', C.c_void_p), ('n_edges', C.c_int64)]


_LIBS = {}


def load(omp=False):
    name = 'libcovasim_port_omp.so' if omp else 'libcovasim_port.so'
    if name not in _LIBS:
        path = os.path.join(BUILD, name)
        if not os.path.exists(path):
            raise RuntimeError(f'{path} missing: run make -C oracle')
        L = C.CDLL(path)
        L.port_step.restype = C.c_int
        L.port_step.argtypes = [C.POINTER(Params), C.POINTER(People), C.POINTER(Layer), C.c_int32, C.c_int64,
                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.port_threads.restype = C.c_int
        _LIBS[name] = L
    return _LIBS[name]


def _dur(d):
    s = Dur()
    s.dist = DIST[d['dist']]
    s.par1, s.par2 = float(d['par1']), float(d.get('par2', 0.0))
    if s.dist in (4, 5) and s.par1 > 0:
        s.mu, s.sigma = co.lognormal_params(s.par1, s.par2)
    return s


class PortSim:
    """Compiled-restatement sim with the oracle's inputs (see oracle.covasim_oracle.Sim)."""

    def __init__(self, pars, people_init, layers, seed, omp=False):
        self.lib = load(omp)
        self.pars = pars
        self.npts = int(pars['n_days']) + 1
        n = len(people_init['age'])
        nv = pars['n_variants']
        self.n = n
        if pars['use_waning']:
            if pars.get('immunity') is None:
                pars['immunity'] = co.init_immunity_matrix(nv)
            pars['nab_kin'] = co.precompute_waning(self.npts, pars['nab_decay'])
        a = {}
        for k in U8:
            a[k] = np.full(n, 1 if k in ('susceptible', 'naive') else 0, np.uint8)
        for k in F32V:
            a[k] = np.full(n, np.nan, np.float32)
        for k in BYV:
            a[k] = np.zeros((nv, n), np.uint8)
        for k in F32A:
            a[k] = np.full(n, np.nan, np.float32)
        for k in ['rel_trans', 'rel_sus', 'symp_prob', 'severe_prob', 'crit_prob', 'death_prob']:
            a[k] = np.ascontiguousarray(people_init[k], dtype=np.float32).copy()
        for k in I32A:
            a[k] = np.zeros(n, np.int32)
        for k in F32B:
            a[k] = np.zeros(n, np.float32)
        for k in IMM:
            a[k] = np.zeros((nv, n), np.float32)
        self.a = a
        self.P = People(**{k: a[k].ctypes.data for k in PEOPLE_FIELDS})
        self.layers_np = [[np.ascontiguousarray(v[0], np.int32).copy(), np.ascontiguousarray(v[1], np.int32).copy(),
                           np.ascontiguousarray(v[2], np.float32).copy()] for v in layers.values()]
        self.L = (Layer * max(1, len(self.layers_np)))()
        for li, (p1, p2, b) in enumerate(self.layers_np):
            self.L[li].p1, self.L[li].p2, self.L[li].beta = p1.ctypes.data, p2.ctypes.data, b.ctypes.data
            self.L[li].n_edges = len(p1)
        e_max = max([len(l[0]) for l in self.layers_np] + [1])
        self.scratch_f = np.zeros(3 * n, np.float32)
        self.scratch_i = np.zeros(n + 2 * e_max + 1, np.int32)
        self.prm = self._params(list(layers.keys()), seed)
        days, vars_, counts = [], [], []
        for var, vi in zip(pars['variants'], pars['variant_index']):
            for d in var.get('days', []):
                days.append(d)
                vars_.append(vi)
                counts.append(int(var.get('n_imports', 1)))
        self.imp = (np.asarray(days, np.int32), np.asarray(vars_, np.int32), np.asarray(counts, np.int64))
        self.results = np.zeros((self.npts, len(ROW_KEYS)))
        self.t = 0

    def _params(self, lkeys, seed):
        pars = self.pars
        p = Params()
        p.n = self.n
        p.n_variants = pars['n_variants']
        p.n_layers = len(lkeys)
        p.use_waning = 1 if pars['use_waning'] else 0
        p.seed = int(seed)
        vd = pars['viral_dist']
        p.frac_time, p.load_ratio, p.high_cap = (np.float32(vd['frac_time']), np.float32(vd['load_ratio']),
                                                 np.float32(vd['high_cap']))
        p.asymp_factor = np.float32(pars['asymp_factor'])
        for v in range(pars['n_variants']):
            vp = pars['variant_pars'][v]
            p.beta_v[v] = np.float32(pars['beta'] * pars['rel_beta'] * vp['rel_beta'])
            for k, f in [('rel_symp_prob', 'rel_symp'), ('rel_severe_prob', 'rel_sev'), ('rel_crit_prob', 'rel_crit'),
                         ('rel_death_prob', 'rel_death')]:
                getattr(p, f)[v] = np.float32(pars[k] * vp[k] if v else pars[k])
        for li, lk in enumerate(lkeys):
            p.beta_layer[li] = np.float32(pars['beta_layer'][lk])
            p.iso_factor[li] = np.float32(pars['iso_factor'][lk])
            p.quar_factor[li] = np.float32(pars['quar_factor'][lk])
            p.dynamic[li] = 1 if pars['dynam_layer'].get(lk, 0) else 0
        p.no_hosp_factor, p.no_icu_factor = np.float32(pars['no_hosp_factor']), np.float32(pars['no_icu_factor'])
        p.trans_redux, p.nab_boost = np.float32(pars['trans_redux']), np.float32(pars['nab_boost'])
        p.n_beds_hosp = -1 if pars['n_beds_hosp'] is None else int(pars['n_beds_hosp'])
        p.n_beds_icu = -1 if pars['n_beds_icu'] is None else int(pars['n_beds_icu'])
        for j, k in enumerate(co.DUR_KEYS):
            p.dur[j] = _dur(pars['dur'][k])
        p.nab_init = _dur(pars['nab_init'])
        if pars['use_waning']:
            self._kin = np.ascontiguousarray(pars['nab_kin'], np.float64)
            p.nab_kin = self._kin.ctypes.data
            p.nab_kin_len = len(self._kin)
            ris = pars['rel_imm_symp']
            p.rel_imm_symp[:] = [ris['asymp'], ris['mild'], ris['severe']]
            ne = pars['nab_eff']
            p.nab_norm = 1 + ne['alpha_inf_diff']
            for j, (a_, b_) in enumerate([('alpha_inf', 'beta_inf'), ('alpha_symp_inf', 'beta_symp_inf'),
                                          ('alpha_sev_symp', 'beta_sev_symp')]):
                p.ve_exp_alpha[j] = float(np.exp(ne[a_]))
                p.ve_beta[j] = float(ne[b_])
            imm = np.asarray(pars['immunity'], np.float32)
            for i in range(pars['n_variants']):
                for j in range(pars['n_variants']):
                    p.immunity[i][j] = imm[i, j]
        return p

    def step(self):
        row = np.zeros(len(ROW_KEYS))
        d, v, c = self.imp
        rc = self.lib.port_step(self.prm, self.P, self.L, self.t, int(self.pars['pop_infected']),
                                d.ctypes.data if len(d) else None, v.ctypes.data if len(v) else None,
                                c.ctypes.data if len(c) else None, len(d), row.ctypes.data,
                                self.scratch_f.ctypes.data, self.scratch_i.ctypes.data)
        if rc != 0:
            raise RuntimeError('port_step failed')
        self.results[self.t] = row
        self.t += 1

    def run(self):
        while self.t < self.npts:
            self.step()
        return self

    def result(self, key):
        return self.results[:, ROW_KEYS.index(key)]
