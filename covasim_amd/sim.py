"""Sim: orchestration of the daily transmission step on the GPU (covasim/sim.py).

``Sim(pars).run()`` builds the population on the host, uploads agent state and
contact layers to the GPU once, and drives the day loop through the C ABI
(``cvk_run_days``: agent pass, dynamic-layer resample, variant imports, fused
all-layer edge pass, infection pass, counts), optionally captured as one
hipGraph. Results come back as float64 per-day arrays with the reference's
keys (sim.py:284-351) and are finalised as in sim.py:764-850.
"""

import ctypes as C
import time

import numpy as np

from . import _lib
from . import results as cvres
from . import interventions as cviv
from . import immunity as cvimm
from . import parameters as cvpar
from . import population as cvpop
from .base import Result, Layer
from .people import People

__all__ = ['Sim', 'AlreadyRunError']

RESULT_FLOWS = ['infections', 'reinfections', 'infectious', 'symptomatic', 'severe', 'critical', 'recoveries',
                'deaths', 'tests', 'diagnoses', 'known_deaths', 'quarantined', 'isolated', 'doses', 'vaccinated']
RESULT_STOCKS = ['susceptible', 'exposed', 'infectious', 'symptomatic', 'severe', 'critical', 'recovered', 'dead',
                 'diagnosed', 'known_dead', 'quarantined', 'isolated', 'vaccinated']
RESULT_FLOWS_BY_VARIANT = ['infections_by_variant', 'symptomatic_by_variant', 'severe_by_variant',
                           'infectious_by_variant']
UNSUPPORTED_NONDEFAULTS = dict(pop_scale=1, frac_susceptible=1.0, n_imports=0, location=None)


class AlreadyRunError(RuntimeError):
    """Raised when stepping a completed sim (sim.py:1550)."""


def _dur_struct(d):
    s = _lib.Dur()
    dist = d['dist']
    if dist not in _lib.DIST_CODES:
        raise NotImplementedError(f'Duration distribution "{dist}" is not supported on the GPU path; choices are '
                                  f'{sorted(set(_lib.DIST_CODES))}')
    s.dist = _lib.DIST_CODES[dist]
    s.par1 = float(d['par1'])
    s.par2 = float(d.get('par2', 0.0))
    if s.dist in (4, 5) and s.par1 > 0:  # utils.py:223-227
        p1, p2 = s.par1, s.par2
        s.mu = float(np.log(p1 ** 2 / np.sqrt(p2 ** 2 + p1 ** 2)))
        s.sigma = float(np.sqrt(np.log(p2 ** 2 / p1 ** 2 + 1)))
    return s


class Sim:
    """A single simulation on one GPU (sim.py:49-80)."""

    def __init__(self, pars=None, label=None, people=None, popdict=None, households='poisson', device=None,
                 log_capacity=None, popgen='host', source_lists=True, **kwargs):
        merged = dict(pars or {})
        merged.update(kwargs)
        variants = merged.pop('variants', [])
        self.pars = cvpar.make_pars(**merged)
        self.pars['variants'] = list(variants) if isinstance(variants, (list, tuple)) else [variants]
        self.label = label
        self.people = people
        self.popdict = popdict
        self.households = households
        if popgen not in ('host', 'device'):
            raise ValueError(f'popgen must be "host" or "device", not "{popgen}"')
        self.popgen = popgen  # 'device': build the population with the cvk_pop_* kernels
        # static layers get source lists (cvk_sources): the edge pass then walks the infectious agents'
        # contacts instead of streaming the layer (16 bytes per stored edge of device memory)
        self.source_lists = bool(source_lists)
        self._sources = {}
        self.device = device
        self.log_capacity = log_capacity
        self.initialized = False
        self.results_ready = False
        self.complete = False
        self.t = 0
        self.results = {}

    # -- dict-like parameter access (base.py:85-110) --
    def __getitem__(self, key):
        return self.pars[key]

    def __setitem__(self, key, value):
        self.pars[key] = value
        if self.initialized:
            self._build_params()

    def update_pars(self, pars=None, **kwargs):
        p = dict(pars or {})
        p.update(kwargs)
        self.pars.update(p)
        if any(k in p for k in ('pop_type',)):
            cvpar.reset_layer_pars(self.pars, force=True)

    @property
    def npts(self):
        return int(self.pars['n_days']) + 1

    # -- initialisation --
    def validate_pars(self):
        for k, v in UNSUPPORTED_NONDEFAULTS.items():
            if self.pars.get(k, v) != v:
                raise NotImplementedError(f'Parameter {k}={self.pars[k]!r} is outside the GPU transmission path '
                                          f'(supported value: {v!r})')
        if self.pars['analyzers']:
            raise NotImplementedError('Analyzers are outside the GPU transmission path')
        for iv in self.pars['interventions']:
            if not isinstance(iv, cviv.SUPPORTED):
                raise NotImplementedError(f'Intervention {iv!r} is outside the GPU transmission path (supported: '
                                          f'{", ".join(c.__name__ for c in cviv.SUPPORTED)})')
        if int(self.pars['n_days']) < 0:
            raise ValueError('n_days must be non-negative')

    def init_variants(self):
        """sim.py:477-495 with variant.initialize (immunity.py:107-114): each variant's parameters are stored
        under its label and its index is that label's position, so a later variant with the same label (two
        dicts without one are both 'custom') replaces the parameters and shares the index."""
        vpars = {'wild': dict(self.pars['variant_pars']['wild'])}
        for i, var in enumerate(self.pars['variants']):
            if not isinstance(var, cvimm.variant):
                raise TypeError(f'Variant {i} ({var}) is not a covasim_amd.variant object')
            vpars[var.label] = dict(var.p)
        labels = list(vpars)
        for var in self.pars['variants']:
            var.index = labels.index(var.label)
        if len(labels) > _lib.MAX_VARIANTS:
            raise ValueError(f'At most {_lib.MAX_VARIANTS} variants are supported')
        self.pars['variant_map'] = dict(enumerate(labels))
        self.pars['variant_pars'] = vpars
        self.pars['n_variants'] = len(labels)

    def initialize(self, reset=False):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError('covasim_amd.Sim needs a ROCm GPU (torch.cuda.is_available() is False)')
        _lib.lib()
        self.validate_pars()
        self.init_variants()
        if self.pars['prognoses'] is None:
            self.pars['prognoses'] = cvpar.get_prognoses(self.pars['prog_by_age'])
        if self.pars['use_waning']:
            if self.pars['immunity'] is None:
                self.pars['immunity'] = cvimm.init_immunity_matrix(list(self.pars['variant_map'].values()))
            self.pars['nab_kin'] = cvimm.precompute_waning(self.npts, self.pars['nab_decay'])
        self.init_people()
        self.init_workspace()
        for iv in self.pars['interventions']:
            iv.initialize(self)
        self._init_diag()
        self._build_params()
        self.t = 0
        self.complete = False
        self.results_ready = False
        self.initialized = True
        return self

    def init_people(self):
        rng = np.random.default_rng(int(self.pars['rand_seed']))
        if isinstance(self.people, People):
            self.people.pars = self.pars
            return
        popdict = self.popdict
        if popdict is None and self.popgen == 'device':
            from . import popgen as cvpg
            popdict = cvpg.make_population(self.pars, households=self.households, device=self.device)
            popdict['age'] = popdict['age'].cpu().numpy()
            popdict['sex'] = popdict['sex'].cpu().numpy()
        if popdict is None:
            popdict = cvpop.make_randpop(self.pars, rng, households=self.households)
        arrays = {k: popdict[k] for k in ('symp_prob', 'severe_prob', 'crit_prob', 'death_prob', 'rel_sus',
                                            'rel_trans') if k in popdict}
        if len(arrays) < 6:
            arrays = cvpop.make_people_arrays(np.asarray(popdict['age']), self.pars, rng)
        if len(popdict['age']) != int(self.pars['pop_size']):
            raise ValueError(f'popdict has {len(popdict["age"])} agents but pop_size is {self.pars["pop_size"]}')
        contacts = popdict['contacts']
        keys = list(self.pars['beta_layer'].keys())
        ordered = {k: contacts[k] for k in keys if k in contacts}
        ordered.update({k: v for k, v in contacts.items() if k not in ordered})
        self.people = People(self.pars, popdict['age'], sex=popdict.get('sex'), contacts=ordered, device=self.device,
                             **arrays)
        self.popdict = None

    def init_workspace(self):
        n, npts = self.people.n, self.npts
        cap = self.log_capacity
        if cap is None:
            cap = max(1024, 2 * n + int(self.pars['pop_infected']))
        self._w, self._work = _lib.make_workspace(n, npts, int(cap), self.people.device)

    def layer_keys(self):
        return list(self.people.contacts.keys())

    def _init_diag(self):
        """Testing and quarantine state when the sim has test_prob / contact_tracing interventions:
        the five dates beside the agents, the ring of scheduled quarantines (one row per day of the
        longest tracing delay, plus one), and the tracing scratch (cvk_diag)."""
        import torch
        ivs = [iv for iv in self.pars['interventions'] if isinstance(iv, cviv.DIAG_TYPES)]
        self._diag = None
        if not ivs:
            return
        people, dev = self.people, self.people.device
        n = people.n
        for k in _lib.DIAG_DATE_FIELDS:
            people._d[k] = torch.full((n,), float('nan'), dtype=torch.float32, device=dev)
        ring = 2 + max([iv.max_delay() for iv in ivs if isinstance(iv, cviv.contact_tracing)] + [0])
        w = self._w
        w['diag_pending'] = torch.full((ring, n), _lib.PENDING_NONE, dtype=torch.int32, device=dev)
        w['diag_cases'] = torch.zeros(2 * ((n + 63) // 64), dtype=torch.int32, device=dev)
        w['diag_coarse'] = torch.zeros(int(_lib.lib().cvk_diag_coarse_words(n)), dtype=torch.int32, device=dev)
        w['diag_contacts'] = torch.zeros(n, dtype=torch.int32, device=dev)
        w['diag_part'] = torch.zeros(4 * _lib.DIAG_MAX_BLOCKS, dtype=torch.int32, device=dev)
        w['diag_select'] = torch.zeros(130, dtype=torch.int64, device=dev)
        bins = self.npts if any(isinstance(iv, cviv.test_prob) and iv.pdf for iv in ivs) else 0
        w['diag_swab_hist'] = torch.zeros(bins + 1, dtype=torch.int32, device=dev)
        w['diag_swab_prob'] = torch.zeros(max(1, bins), dtype=torch.float64, device=dev)
        d = _lib.DiagS()
        for k in _lib.DIAG_DATE_FIELDS:
            setattr(d, k, people._d[k].data_ptr())
        d.pending = w['diag_pending'].data_ptr()
        d.cases = w['diag_cases'].data_ptr()
        d.coarse = w['diag_coarse'].data_ptr()
        d.contacts = w['diag_contacts'].data_ptr()
        d.part = w['diag_part'].data_ptr()
        d.ring = ring
        d.swab_bins = bins
        d.select = w['diag_select'].data_ptr()
        d.swab_hist = w['diag_swab_hist'].data_ptr()
        d.swab_prob = w['diag_swab_prob'].data_ptr()
        self._diag = d

    def _build_params(self):
        import torch
        pars = self.pars
        p = _lib.Params()
        people = self.people
        lkeys = self.layer_keys()
        if len(lkeys) > _lib.MAX_LAYERS:
            raise ValueError(f'At most {_lib.MAX_LAYERS} layers are supported')
        p.n_agents = people.n
        p.n_variants = int(pars['n_variants'])
        p.n_layers = len(lkeys)
        p.seed = int(pars['rand_seed']) & 0xFFFFFFFFFFFFFFFF
        f32 = np.float32
        vd = pars['viral_dist']
        ft, lr, hc = f32(vd['frac_time']), f32(vd['load_ratio']), f32(vd['high_cap'])
        one, zero = f32(1), f32(0)
        den = one + ft * (lr - one)  # utils.py:78 with load = 1
        p.frac_time, p.load_ratio, p.high_cap = ft, lr, hc
        p.vl_hi = (lr * one + one * zero) / den
        p.vl_lo = (lr * zero + one * one) / den
        p.asymp_factor = f32(pars['asymp_factor'])
        labels = [pars['variant_map'][v] for v in range(p.n_variants)]
        beta_v, beta_layer = self._rates(pars)
        for v, lab in enumerate(labels):
            vp = pars['variant_pars'][lab]
            p.beta_v[v] = beta_v[v]
            for k in ('rel_symp_prob', 'rel_severe_prob', 'rel_crit_prob', 'rel_death_prob'):
                val = pars[k] * vp[k] if v else pars[k]  # people.py:477-481
                getattr(p, k)[v] = f32(val)
        for li, lk in enumerate(lkeys):
            p.beta_layer[li] = beta_layer[li]
            p.iso_factor[li] = f32(pars['iso_factor'].get(lk, 0.2))
            p.quar_factor[li] = f32(pars['quar_factor'].get(lk, 0.3))
        p.no_hosp_factor = f32(pars['no_hosp_factor'])
        p.no_icu_factor = f32(pars['no_icu_factor'])
        p.trans_redux = f32(pars['trans_redux'])
        p.nab_boost = f32(pars['nab_boost'])
        p.n_beds_hosp = -1 if pars['n_beds_hosp'] is None else int(pars['n_beds_hosp'])
        p.n_beds_icu = -1 if pars['n_beds_icu'] is None else int(pars['n_beds_icu'])
        for j, k in enumerate(_lib.DUR_ORDER):
            p.dur[j] = _dur_struct(pars['dur'][k])
        p.nab_init = _dur_struct(pars['nab_init'])
        p.use_waning = 1 if pars['use_waning'] else 0
        if pars['use_waning']:
            kin = np.asarray(pars['nab_kin'], dtype=np.float64)
            self._nab_kin = torch.from_numpy(kin.copy()).to(people.device)
            p.nab_kin = self._nab_kin.data_ptr()
            p.nab_kin_len = len(kin)
            ris = pars['rel_imm_symp']
            p.rel_imm_symp[0], p.rel_imm_symp[1], p.rel_imm_symp[2] = ris['asymp'], ris['mild'], ris['severe']
            ne = pars['nab_eff']
            p.nab_norm = 1 + ne['alpha_inf_diff']
            for j, (a, b) in enumerate([('alpha_inf', 'beta_inf'), ('alpha_symp_inf', 'beta_symp_inf'),
                                        ('alpha_sev_symp', 'beta_sev_symp')]):
                p.ve_exp_alpha[j] = float(np.exp(ne[a]))
                p.ve_beta[j] = float(ne[b])
            imm = np.asarray(pars['immunity'], dtype=np.float32)
            for i in range(p.n_variants):
                for j in range(p.n_variants):
                    p.immunity[i][j] = imm[i, j]
        self._params = p
        people._params = p
        layers = (_lib.LayerS * max(1, len(lkeys)))()
        dyn = pars.get('dynam_layer') or {}
        for li, lk in enumerate(lkeys):
            lay = people.contacts[lk]
            lay.validate()
            layers[li].p1 = lay.tensor('p1').data_ptr()
            layers[li].p2 = lay.tensor('p2').data_ptr()
            layers[li].beta = lay.tensor('beta').data_ptr()
            layers[li].n_edges = len(lay)
            layers[li].dynamic = 1 if (dyn.get(lk, 0) and type(lay).update is Layer.update) else 0
            layers[li].beta_ones = 1 if (layers[li].dynamic or lay.beta_all_ones()) else 0
        self._layers = layers
        self._attach_sources()
        # import schedule of the variants (immunity.py:117-130)
        days, vars_, nimp = [], [], []
        for var in pars['variants']:
            for d in var.day_list(self.npts):
                days.append(d)
                vars_.append(var.index)
                nimp.append(var.n_imports)
        self._imp_arrays = (np.asarray(days, np.int32), np.asarray(vars_, np.int32), np.asarray(nimp, np.int64))
        imp = _lib.ImportsS()
        imp.n = len(days)
        imp.days = self._imp_arrays[0].ctypes.data_as(C.POINTER(C.c_int32))
        imp.variants = self._imp_arrays[1].ctypes.data_as(C.POINTER(C.c_int32))
        imp.n_imports = self._imp_arrays[2].ctypes.data_as(C.POINTER(C.c_int64))
        # per-day rates of change_beta interventions (interventions.py:530-586)
        self._beta_days = cviv.beta_schedule(self, self._rates)
        if self._beta_days is not None:
            imp.n_beta_days = self.npts
            imp.beta_v_days = self._beta_days[0].ctypes.data
            imp.beta_layer_days = self._beta_days[1].ctypes.data
        # testing and tracing ops, in the sim's order; an op's stream slot is its intervention's index
        self._diag_ops = [(iv, iv.op(i, lkeys)) for i, iv in enumerate(pars['interventions'])
                          if isinstance(iv, cviv.DIAG_TYPES)]
        if self._diag is not None:
            self._diag_op_array = (_lib.DiagOp * max(1, len(self._diag_ops)))(*[o for _, o in self._diag_ops])
            imp.n_diag_ops = len(self._diag_ops)
            imp.diag_ops = self._diag_op_array
            imp.diag = C.pointer(self._diag)
        self._imports = imp
        self._edge_interventions = [iv for iv in pars['interventions'] if isinstance(iv, cviv.clip_edges)]

    @property
    def _custom_update(self):
        """Dynamic layers whose update() is not the default resample (Layer subclasses, which may be
        put in place after initialize(), examples/t11_dynamic_layers.py:41-43): days then run
        through the individual entry points, with update(people) on the host."""
        dyn = self.pars.get('dynam_layer') or {}
        return [lk for lk in self.layer_keys() if dyn.get(lk, 0) and
                type(self.people.contacts[lk]).update is not Layer.update]

    # -- running --
    def _rates(self, pars):
        """float32 beta per variant (sim.py:627) and per layer (compute_trans_sus's beta_layer)."""
        f32 = np.float32
        beta_v = [f32(pars['beta'] * pars['rel_beta'] * pars['variant_pars'][pars['variant_map'][v]]['rel_beta'])
                  for v in range(int(pars['n_variants']))]
        beta_layer = [f32(pars['beta_layer'].get(lk, 1.0)) for lk in self.layer_keys()]
        return beta_v, beta_layer

    def _apply_schedule(self, t):
        """Day t's rates from the interventions' per-day table, into the parameters of a manual day."""
        if self._beta_days is None:
            return
        row = min(t, len(self._beta_days[0]) - 1)
        for v in range(_lib.MAX_VARIANTS):
            self._params.beta_v[v] = self._beta_days[0][row, v]
        for li in range(_lib.MAX_LAYERS):
            self._params.beta_layer[li] = self._beta_days[1][row, li]

    def _refresh_layers(self):
        for li, lk in enumerate(self.layer_keys()):
            lay = self.people.contacts[lk]
            self._layers[li].p1 = lay.tensor('p1').data_ptr()
            self._layers[li].p2 = lay.tensor('p2').data_ptr()
            self._layers[li].beta = lay.tensor('beta').data_ptr()
            self._layers[li].n_edges = len(lay)
            self._layers[li].beta_ones = 1 if lay.beta_all_ones() else 0
        self._attach_sources()

    def _attach_sources(self):
        """Source lists of the static layers (built on the device, rebuilt when a layer's endpoint tensors
        change, e.g. after clip_edges); dynamic and custom-updated layers are streamed."""
        import torch
        L = _lib.lib()
        dyn = self.pars.get('dynam_layer') or {}
        for li, lk in enumerate(self.layer_keys()):
            lay = self.people.contacts[lk]
            static = not dyn.get(lk, 0)
            if not (self.source_lists and static and len(lay) > 0):
                self._layers[li].sources = None
                self._sources.pop(lk, None)
                continue
            p1, p2 = lay.tensor('p1'), lay.tensor('p2')
            key = (p1.data_ptr(), p2.data_ptr(), p1.numel(), p1._version, p2._version)
            have = self._sources.get(lk)
            if have is None or have[0] != key:
                n, E = self.people.n, len(lay)
                dev = p1.device
                indptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
                target = torch.empty(2 * E, dtype=torch.int32, device=dev)
                edge_dir = torch.empty(2 * E, dtype=torch.int32, device=dev)
                nb = int(L.cvk_sources_work_bytes(n))
                work = torch.empty((nb + 3) // 4, dtype=torch.int32, device=dev)
                slots = C.c_int64(0)
                _lib.check(L.cvk_sources_build(C.byref(self._layers[li]), n, indptr.data_ptr(), target.data_ptr(),
                                               edge_dir.data_ptr(), 2 * E, C.byref(slots), work.data_ptr(), nb,
                                               _lib.stream_ptr()), f'source lists of layer {lk}')
                del work
                s = _lib.SourcesS(indptr.data_ptr(), target.data_ptr(), edge_dir.data_ptr(), slots.value)
                have = (key, s, indptr, target, edge_dir)
                self._sources[lk] = have
            self._layers[li].sources = C.pointer(have[1])

    def _manual_day(self, t):
        """One day through the individual entry points (used by step() and custom layer updates)."""
        L = _lib.lib()
        P = self.people.struct()
        st = _lib.stream_ptr()
        pr, w = self._params, self._work
        if t == 0 and self.pars['pop_infected']:
            _lib.check(L.cvk_choose(P, pr, w, 0, _lib.STREAM_CHOOSE, int(self.pars['pop_infected']), 0,
                                    _lib.LAYER_SEED, 0, st), 'seed infections')
            _lib.check(L.cvk_infect_pass(P, self._layers, pr, w, 0, 0, st), 'seed infections')
        self.people.t = t
        _lib.check(L.cvk_agent_pass(P, pr, w, t, st), 'agent pass')
        for li, lk in enumerate(self.layer_keys()):
            if (self.pars.get('dynam_layer') or {}).get(lk, 0):
                self.people.contacts[lk].update(self.people)
        self._refresh_layers()
        days, vars_, nimp = self._imp_arrays
        for j in range(len(days)):
            if days[j] == t:
                _lib.check(L.cvk_choose(P, pr, w, t, _lib.STREAM_IMPORT | int(vars_[j]), int(nimp[j]), 1,
                                        _lib.LAYER_IMPORT, int(vars_[j]), st), 'variant imports')
        if self._edge_interventions or self._diag is not None:  # after the contact update and the imports,
            self.t = t                                          # in the sim's order (sim.py:591-596)
            ops = {id(iv): o for iv, o in self._diag_ops}
            for iv in self.pars['interventions']:
                if isinstance(iv, cviv.clip_edges):
                    iv(self)
                    self._refresh_layers()
                elif id(iv) in ops:
                    if isinstance(iv, cviv.test_prob):
                        iv.prepare_day(self, ops[id(iv)])
                    _lib.check(L.cvk_diag_apply(P, self._layers, pr, w, C.byref(self._diag), C.byref(ops[id(iv)]), t,
                                                st), iv.label)
            if self._diag is not None:
                _lib.check(L.cvk_states_post(P, pr, w, C.byref(self._diag), t, st), 'update_states_post')
        _lib.check(L.cvk_edge_pass(P, self._layers, pr, w, t, None, st), 'edge pass')
        _lib.check(L.cvk_infect_pass(P, self._layers, pr, w, t, 1, st), 'infect pass')
        _lib.check(L.cvk_count_pass(P, pr, w, t, st), 'count pass')

    def step(self):
        """Advance one day (sim.py:558-685)."""
        if not self.initialized:
            self.initialize()
        if self.complete:
            raise AlreadyRunError('Simulation already complete (call sim.initialize() to re-run)')
        # parameters edited in place between days (e.g. beta_layer[k] *= f, interventions.py:575-586) apply
        self._build_params()
        self._apply_schedule(self.t)
        self._manual_day(self.t)
        self.t += 1
        self.people.t = self.t
        if self.t == self.npts:
            self.complete = True

    def run(self, do_plot=False, until=None, restore_pars=True, reset_seed=True, verbose=None, use_graph=None):
        """Run the day loop (sim.py:688-761); ``until`` is the last day to run (inclusive)."""
        import torch
        if not self.initialized:
            self.initialize()
        if self.complete:
            raise AlreadyRunError('Simulation is already complete (call sim.initialize() to re-run)')
        if self.results_ready:
            raise AlreadyRunError('Simulation has already been finalized')
        t1 = self.npts if until is None else min(int(until) + 1, self.npts)
        if t1 <= self.t:
            raise AlreadyRunError(f'Simulation is already at day {self.t}; cannot run until day {until}')
        T = time.perf_counter()
        stream = torch.cuda.current_stream()
        graph = bool(use_graph) if use_graph is not None else True
        if graph and stream.cuda_stream == 0:
            graph = False
        L = _lib.lib()
        # days on which an edge intervention acts run through the individual passes (the layers
        # change between the contact update and transmission); the days between them run as one
        # day loop (a captured graph)
        manual_days = set(range(self.t, t1)) if self._custom_update else {
            d for iv in self._edge_interventions for d in iv.days if self.t <= d < t1}
        for iv in self.pars['interventions']:  # subtargets given by functions of the sim: evaluated each day
            if isinstance(iv, cviv.test_prob) and iv.dynamic_subtarget:
                manual_days |= {d for d in range(self.t, t1) if iv.active(d)}
        t = self.t
        while t < t1:
            if t in manual_days:
                self._apply_schedule(t)
                self._manual_day(t)
                t += 1
                continue
            nxt = min([d for d in manual_days if d > t] + [t1])
            if manual_days:
                self._refresh_layers()
            _lib.check(L.cvk_run_days(self.people.struct(), self._layers, self._params, self._work, t, nxt,
                                      int(self.pars['pop_infected']), self._imports, 1 if graph else 0,
                                      C.c_void_p(stream.cuda_stream)), 'run')
            t = nxt
        torch.cuda.synchronize()
        self.elapsed = time.perf_counter() - T
        self.t = t1
        self.people.t = t1
        if self.t == self.npts:
            self.complete = True
            self.finalize()
        return self

    # -- results --
    def raw_results(self):
        """Per-day device counters as host arrays (before finalisation)."""
        res = self._w['results'].cpu().numpy().view(np.uint32).astype(np.float64)
        resf = self._w['results_f'].cpu().numpy()
        return res, resf

    def finalize(self, rows=None):
        """Per-day results, cumulative flows, derived states, doubling time and r_eff
        (sim.py:764-1001), computed on the device by cvk_finalize.

        rows: optional (rows [F_ROWS][npts], variant rows [FV_ROWS][nv][npts], raw counts
        [npts][cols]) already finalised for this sim by a batch (BatchSim).
        """
        if self.results_ready:
            raise AlreadyRunError('Simulation has already been finalized')
        if rows is None:
            d = self.people._d
            out, out_v = cvres.finalize_tables(
                self._w['results'], self._w['results_f'], self.npts, int(self.pars['n_variants']),
                self.pars['use_waning'], [self.pars['pop_infected']], [cvres.initial_period(self.pars, self.npts)],
                [0, self.people.n], self.import_schedule()[None, :], d['date_infectious'], d['date_recovered'],
                d['date_dead'])
            counts = self._w['results'].cpu().numpy().view(np.uint32)
            rows = (out[0].cpu().numpy(), out_v[0].cpu().numpy(), counts)
        R = cvres.to_results(rows[0], rows[1])
        self.raw = {'n_susceptible_agents': rows[2][:, _lib.R_COL['n_susceptible']].astype(np.float64)}
        self.results = R
        self.results_ready = True
        self.t = self.npts - 1

    def import_schedule(self):
        """Imported agents per day, from the variants' import events (immunity.py:117-130)."""
        imp = np.zeros(self.npts)
        days, vars_, nimp = self._imp_arrays
        for j in range(len(days)):
            imp[days[j]] += nimp[j]
        return imp

    def compute_states(self):
        """Prevalence, incidence and the recomputed susceptible count (sim.py:828-850): part of
        finalize, kept for callers of the reference's method name."""
        if not self.results_ready:
            self.finalize()
        return self.results

    def save(self, filename):
        """Write a device-state snapshot that ``covasim_amd.load`` resumes (base.py:682-721)."""
        from . import checkpoint
        return checkpoint.save(self, filename)

    @property
    def summary(self):
        return {k: v.values[-1] for k, v in self.results.items() if isinstance(v, Result) and np.ndim(v.values) == 1}

    def infection_log(self):
        """Infection records (source, target, date, layer, variant) (people.py:509-511)."""
        ctl = self._w['ctl'].cpu().numpy().view(np.uint32)
        count = int(ctl[_lib.CTL_LOG_COUNT]) | (int(ctl[_lib.CTL_LOG_COUNT + 1]) << 32)
        count = min(count, self._w['log'].shape[0])
        recs = self._w['log'][:count].cpu().numpy()
        lkeys = self.layer_keys()
        names = {_lib.LAYER_SEED: 'seed_infection', _lib.LAYER_IMPORT: 'importation'}
        out = []
        for s, tg, d, lv in recs:
            l, v = lv & 0xFF, (lv >> 8) & 0xF
            out.append(dict(source=None if s < 0 else int(s), target=int(tg), date=int(d),
                            layer=names.get(l, lkeys[l] if l < len(lkeys) else l),
                            variant=self.pars['variant_map'][v]))
        return out

    def log_overflowed(self):
        return bool(self._w['ctl'][_lib.CTL_LOG_OVERFLOW].item())
