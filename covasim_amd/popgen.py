"""Population generation on the device (covasim/population.py:143-364, people.py:139-160).

``make_population(pars)`` builds what ``population.make_randpop`` + ``make_people_arrays``
build on the host -- ages, sexes, contact layers and the per-agent prognosis arrays -- with the
``cvk_pop_*`` kernels, directly into device tensors. Every draw is a Philox word addressed by an
agent, household, member or edge index, so the population is a pure function of the seed and
the tables below, independent of launch shapes.

The generator follows the reference's construction (population.py): ages from the 19-row age
table, uniform within the bin; household clusters drawn one after another until they cover the
population (the last one cut to fit); school / work layers among the agents of an age band and a
community layer among all agents, each with round(Poisson(n) / 2) edges per member to uniformly
drawn partners. Categorical household sizes (1..7, mean about 2.6) are age-correlated: agents
are ordered by age plus sort-key noise (Irwin-Hall of four uniforms, standard deviation 15
years, 1-year buckets) before being cut into households, and households are then shuffled by a
Feistel permutation. Poisson-sized households are consecutive agents, as in the reference.
"""

import ctypes as C

import numpy as np

from . import _lib
from . import parameters as cvpar
from .population import DEFAULT_AGE_DATA, HOUSEHOLD_SIZE_PROBS

__all__ = ['pop_tables', 'layer_plan', 'count_dist', 'make_population', 'HYBRID_SLOTS']

HYBRID_SLOTS = {'h': 0, 's': 1, 'w': 2, 'c': 3}
SCHOOL_AGES = (6.0, 22.0)
WORK_AGES = (22.0, 65.0)
AGE_NOISE_SD = 15.0
POISSON_HH_SIZES = 64


def _poisson_cdf(lam, n):
    p = np.empty(n)
    p[0] = np.exp(-lam)
    for k in range(1, n):
        p[k] = p[k - 1] * lam / k
    return np.cumsum(p)


def pop_tables(pars, households='categorical', age_data=DEFAULT_AGE_DATA, size_probs=HOUSEHOLD_SIZE_PROBS,
               age_noise_sd=AGE_NOISE_SD):
    """Tables of the generator for a parameter dict (pop_size, rand_seed, contacts, prognoses, beta_dist)."""
    age_data = np.asarray(age_data, dtype=float)
    prob = age_data[:, 2] / age_data[:, 2].sum()
    if len(prob) > _lib.POP_MAX_AGE_BINS:
        raise ValueError(f'at most {_lib.POP_MAX_AGE_BINS} age bins are supported')
    t = dict(seed=int(pars['rand_seed']) & 0xFFFFFFFFFFFFFFFF,
             age_cdf=np.cumsum(prob), age_lo=age_data[:, 0].copy(), age_width=age_data[:, 1] + 1 - age_data[:, 0],
             age_noise=float(age_noise_sd) * np.sqrt(3.0))
    if households == 'categorical':
        sp = np.asarray(size_probs, dtype=float)
        t.update(hh_cdf=np.cumsum(sp / sp.sum()), hh_size_min=1, correlate=1)
    elif households == 'poisson':
        lam = float((pars.get('contacts') or {}).get('h', 2.0))
        t.update(hh_cdf=_poisson_cdf(lam, POISSON_HH_SIZES), hh_size_min=0, correlate=0)
    else:
        raise ValueError(f'households must be "categorical" or "poisson", not "{households}"')
    progs = pars.get('prognoses') or cvpar.get_prognoses(pars.get('prog_by_age', True))
    cut = np.asarray(progs['age_cutoffs'], dtype=float)
    if len(cut) > _lib.POP_MAX_PROG_BINS:
        raise ValueError(f'at most {_lib.POP_MAX_PROG_BINS} prognosis bins are supported')
    t.update(prog_cutoff=cut,
             symp_prob=np.asarray(progs['symp_probs'], dtype=float).astype(np.float32),
             severe_prob=(np.asarray(progs['severe_probs'], dtype=float) *
                          np.asarray(progs['comorbidities'], dtype=float)).astype(np.float32),
             crit_prob=np.asarray(progs['crit_probs'], dtype=float).astype(np.float32),
             death_prob=np.asarray(progs['death_probs'], dtype=float).astype(np.float32),
             rel_sus=np.asarray(progs['sus_ORs'], dtype=float).astype(np.float32),
             trans_or=np.asarray(progs['trans_ORs'], dtype=float))
    bd = pars['beta_dist']
    if bd['dist'] != 'neg_binomial':
        raise NotImplementedError('only the neg_binomial beta_dist of the reference defaults is supported')
    step = float(bd.get('step', 1))
    r = float(bd['par2'])
    p = r / (float(bd['par1']) / step + r)  # utils.py:423-426
    t.update(nb_p0=p ** r, nb_q=1.0 - p, nb_r=r, nb_step=step)
    return t


def count_dist(mean, dispersion=None):
    """Contact-count distribution of a random layer: Poisson(mean), or negative binomial
    (population.py:259-262, utils.py:409-426 with step 1) when dispersion is given."""
    mean = float(mean)
    if dispersion is None:
        if not 0 <= mean <= 700:
            raise ValueError('Poisson contact means must lie in [0, 700]')
        return dict(kind=_lib.POP_COUNT_POISSON, p0=float(np.exp(-mean)), lam=mean, r=0.0, q=0.0)
    r = float(dispersion)
    p = r / (mean + r)
    return dict(kind=_lib.POP_COUNT_NEGBIN, p0=p ** r, lam=0.0, r=r, q=1.0 - p)


def layer_plan(pars, dispersion=None):
    """Layers of the population: (key, slot, band) with band None for all agents, 'households'
    for the household layer, or (lo, hi) for an age band (population.py:332-364)."""
    contacts = dict(pars.get('contacts') or {})
    if pars['pop_type'] == 'random':
        return [(k, slot, None, count_dist(contacts[k], dispersion)) for slot, k in enumerate(contacts)]
    if pars['pop_type'] == 'hybrid':
        c = {'h': 4, 's': 20, 'w': 20, 'c': 20}
        c.update(contacts)
        return [('h', HYBRID_SLOTS['h'], 'households', None),
                ('s', HYBRID_SLOTS['s'], SCHOOL_AGES, count_dist(c['s'], dispersion)),
                ('w', HYBRID_SLOTS['w'], WORK_AGES, count_dist(c['w'], dispersion)),
                ('c', HYBRID_SLOTS['c'], None, count_dist(c['c'], dispersion))]
    raise NotImplementedError(f'Population type "{pars["pop_type"]}" not supported; choices are random or hybrid')


def popgen_struct(t):
    g = _lib.PopGen()
    g.seed = t['seed']
    g.n_age_bins = len(t['age_cdf'])
    g.n_hh_sizes = len(t.get('hh_cdf', [1.0]))
    g.hh_size_min = int(t.get('hh_size_min', 1))
    g.correlate = int(t.get('correlate', 0))
    g.n_prog_bins = len(t['prog_cutoff'])
    for name in ('age_cdf', 'age_lo', 'age_width', 'hh_cdf', 'prog_cutoff', 'trans_or', 'symp_prob', 'severe_prob',
                 'crit_prob', 'death_prob', 'rel_sus'):
        vals = t.get(name, [1.0] if name == 'hh_cdf' else [])
        arr = getattr(g, name)
        for j, v in enumerate(vals):
            arr[j] = v
    g.age_noise = t['age_noise']
    g.nb_p0, g.nb_q, g.nb_r, g.nb_step = t['nb_p0'], t['nb_q'], t['nb_r'], t['nb_step']
    return g


def counts_struct(d):
    c = _lib.PopCounts()
    c.kind, c.p0, c.lam, c.r, c.q = d['kind'], d['p0'], d['lam'], d['r'], d['q']
    return c


def _household_cap(n, t):
    sizes = np.arange(len(t['hh_cdf'])) + t['hh_size_min']
    pmf = np.diff(np.concatenate([[0.0], t['hh_cdf']]))
    mean = max(float((sizes * pmf).sum()), 0.5)
    return int(n / mean * 1.02 + 8 * np.sqrt(n + 1)) + 1024


def make_population(pars, households='categorical', device=None, stream=None, tables=None, dispersion=None):
    """Ages, sexes, contact layers and prognosis arrays of pars['pop_size'] agents, built on the
    device. Returns a dict of device tensors: age, sex, contacts {key: dict(p1, p2, beta)} and
    symp_prob, severe_prob, crit_prob, death_prob, rel_sus, rel_trans."""
    import torch
    L = _lib.lib()
    dev = torch.device(device) if device is not None else torch.device('cuda')
    st = _lib.stream_ptr(stream)
    n = int(pars['pop_size'])
    t = tables if tables is not None else pop_tables(pars, households=households)
    plan = layer_plan(pars, dispersion)
    uses_households = any(band == 'households' for _, _, band, _ in plan)
    if not uses_households:
        t = dict(t, correlate=0)
    g = popgen_struct(t)
    gp = C.byref(g)
    i64 = lambda size: torch.empty(size, dtype=torch.int64, device=dev)  # noqa: E731

    h_cap = _household_cap(n, t) if uses_households else 1
    age_raw = torch.empty(n, dtype=torch.float32, device=dev)
    key = torch.empty(n, dtype=torch.uint8, device=dev) if g.correlate else None
    _lib.check(L.cvk_pop_ages(gp, n, _lib.ptr(age_raw), _lib.ptr(key), st), 'cvk_pop_ages')

    def workspace(cap):
        nbytes = int(L.cvk_pop_work_bytes(n, cap))
        buf = torch.empty(nbytes + 256, dtype=torch.uint8, device=dev)
        off = (-buf.data_ptr()) % 256
        return buf, C.c_void_p(buf.data_ptr() + off), nbytes

    contacts = {}
    age = age_raw
    buf, work, wbytes = workspace(h_cap)
    for lkey, slot, band, counts in plan:
        if band == 'households':
            totals = i64(2)
            while True:
                _lib.check(L.cvk_pop_households_plan(gp, n, _lib.ptr(key), h_cap, work, wbytes, _lib.ptr(totals), st),
                           'cvk_pop_households_plan')
                H, n_edges = (int(v) for v in totals.cpu())
                if H >= 0:
                    break
                h_cap *= 2
                buf, work, wbytes = workspace(h_cap)
            p1 = torch.empty(n_edges, dtype=torch.int32, device=dev)
            p2 = torch.empty(n_edges, dtype=torch.int32, device=dev)
            beta = torch.empty(n_edges, dtype=torch.float32, device=dev)
            age = torch.empty(n, dtype=torch.float32, device=dev) if g.correlate else age_raw
            _lib.check(L.cvk_pop_households_fill(gp, n, _lib.ptr(age_raw), h_cap, work, wbytes, _lib.ptr(age),
                                                 _lib.ptr(p1), _lib.ptr(p2), _lib.ptr(beta), n_edges, st),
                       'cvk_pop_households_fill')
            contacts[lkey] = dict(p1=p1, p2=p2, beta=beta)
    for lkey, slot, band, counts in plan:
        if band == 'households':
            continue
        mapping = None
        n_members = n
        if band is not None:
            mapping = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
            total = i64(1)
            _lib.check(L.cvk_pop_members(n, _lib.ptr(age), float(band[0]), float(band[1]), _lib.ptr(mapping),
                                         _lib.ptr(total), work, wbytes, st), 'cvk_pop_members')
            n_members = int(total.cpu()[0])
            mapping = mapping[:n_members]
        offsets = i64(n_members + 1)
        cs = counts_struct(counts)
        _lib.check(L.cvk_pop_contacts_plan(gp, slot, n_members, C.byref(cs), _lib.ptr(offsets), work, wbytes, st),
                   'cvk_pop_contacts_plan')
        n_edges = int(offsets[n_members].cpu()) if n_members else 0
        p1 = torch.empty(n_edges, dtype=torch.int32, device=dev)
        p2 = torch.empty(n_edges, dtype=torch.int32, device=dev)
        beta = torch.empty(n_edges, dtype=torch.float32, device=dev)
        _lib.check(L.cvk_pop_contacts_fill(gp, slot, n_members, _lib.ptr(mapping), _lib.ptr(offsets), _lib.ptr(p1),
                                           _lib.ptr(p2), _lib.ptr(beta), n_edges, st), 'cvk_pop_contacts_fill')
        contacts[lkey] = dict(p1=p1, p2=p2, beta=beta)
    out = dict(age=age, sex=torch.empty(n, dtype=torch.uint8, device=dev))
    for k in ('symp_prob', 'severe_prob', 'crit_prob', 'death_prob', 'rel_sus', 'rel_trans'):
        out[k] = torch.empty(n, dtype=torch.float32, device=dev)
    _lib.check(L.cvk_pop_people(gp, n, _lib.ptr(age), _lib.ptr(out['sex']), _lib.ptr(out['symp_prob']),
                                _lib.ptr(out['severe_prob']), _lib.ptr(out['crit_prob']), _lib.ptr(out['death_prob']),
                                _lib.ptr(out['rel_sus']), _lib.ptr(out['rel_trans']), work, wbytes, st),
               'cvk_pop_people')
    out['contacts'] = {k: contacts[k] for k, _, _, _ in plan}
    (stream if stream is not None else torch.cuda.current_stream()).synchronize()  # before `buf` is released
    del buf
    return out
