"""ctypes driver of the population generator's C restatement (oracle/cport/popgen_port.c).

TEST INFRASTRUCTURE ONLY: the checker of the device generator (cvk_pop_*) and the CPU baseline
of population building. ``make_population(tables, plan)`` takes the same tables and layer plan
as the device generator and returns numpy arrays.
"""

import ctypes as C

import numpy as np

from .port import BUILD

import os

MAX_AGE, MAX_HH, MAX_PROG = 32, 64, 16


class Tables(C.Structure):
    _fields_ = [('seed', C.c_uint64), ('n_age_bins', C.c_int32), ('n_hh_sizes', C.c_int32),
                ('hh_size_min', C.c_int32), ('correlate', C.c_int32), ('n_prog_bins', C.c_int32), ('pad', C.c_int32),
                ('age_cdf', C.c_double * MAX_AGE), ('age_lo', C.c_double * MAX_AGE),
                ('age_width', C.c_double * MAX_AGE), ('hh_cdf', C.c_double * MAX_HH), ('age_noise', C.c_double),
                ('prog_cutoff', C.c_double * MAX_PROG), ('trans_or', C.c_double * MAX_PROG),
                ('symp_prob', C.c_float * MAX_PROG), ('severe_prob', C.c_float * MAX_PROG),
                ('crit_prob', C.c_float * MAX_PROG), ('death_prob', C.c_float * MAX_PROG),
                ('rel_sus', C.c_float * MAX_PROG), ('nb_p0', C.c_double), ('nb_q', C.c_double), ('nb_r', C.c_double),
                ('nb_step', C.c_double)]


class Counts(C.Structure):
    _fields_ = [('kind', C.c_int32), ('pad', C.c_int32), ('p0', C.c_double), ('lam', C.c_double), ('r', C.c_double),
                ('q', C.c_double)]


_L = None


def lib():
    global _L
    if _L is None:
        L = C.CDLL(os.path.join(BUILD, 'libcovasim_port.so'))
        vp, i64 = C.c_void_p, C.c_int64
        T, K = C.POINTER(Tables), C.POINTER(Counts)
        L.pop_ages.argtypes = [T, i64, vp, vp]
        L.pop_households.argtypes = [T, i64, vp, vp, vp, vp, vp, i64, C.POINTER(C.c_int64)]
        L.pop_households.restype = i64
        L.pop_members.argtypes = [i64, vp, C.c_double, C.c_double, vp]
        L.pop_members.restype = i64
        L.pop_contacts_plan.argtypes = [T, C.c_int32, i64, K, vp]
        L.pop_contacts_plan.restype = i64
        L.pop_contacts_fill.argtypes = [T, C.c_int32, i64, vp, vp, vp, vp]
        L.pop_people.argtypes = [T, i64, vp, vp, vp, vp, vp, vp, vp, vp]
        _L = L
    return _L


def tables_struct(t):
    g = Tables()
    g.seed = t['seed']
    g.n_age_bins = len(t['age_cdf'])
    g.n_hh_sizes = len(t.get('hh_cdf', [1.0]))
    g.hh_size_min = int(t.get('hh_size_min', 1))
    g.correlate = int(t.get('correlate', 0))
    g.n_prog_bins = len(t['prog_cutoff'])
    for name in ('age_cdf', 'age_lo', 'age_width', 'hh_cdf', 'prog_cutoff', 'trans_or', 'symp_prob', 'severe_prob',
                 'crit_prob', 'death_prob', 'rel_sus'):
        for j, v in enumerate(t.get(name, [1.0] if name == 'hh_cdf' else [])):
            getattr(g, name)[j] = v
    g.age_noise = t['age_noise']
    g.nb_p0, g.nb_q, g.nb_r, g.nb_step = t['nb_p0'], t['nb_q'], t['nb_r'], t['nb_step']
    return g


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def make_population(tables, plan, n):
    """plan: [(key, slot, band, counts)] with band None / 'households' / (lo, hi); counts a dict
    (kind, p0, lam, r, q). Returns the same keys as the device generator, as numpy arrays."""
    L = lib()
    uses_households = any(band == 'households' for _, _, band, _ in plan)
    t = dict(tables) if uses_households else dict(tables, correlate=0)
    g = tables_struct(t)
    gp = C.byref(g)
    age_raw = np.empty(n, np.float32)
    key = np.empty(n, np.uint8) if g.correlate else None
    L.pop_ages(gp, n, _p(age_raw), _p(key))
    age = age_raw
    contacts = {}
    H = C.c_int64(0)
    for lkey, slot, band, counts in plan:
        if band != 'households':
            continue
        age = np.empty(n, np.float32)
        dummy = np.empty(1, np.int32)
        n_edges = L.pop_households(gp, n, _p(age_raw), _p(key), _p(age), _p(dummy), _p(dummy), 0, C.byref(H))
        p1 = np.empty(max(n_edges, 1), np.int32)
        p2 = np.empty(max(n_edges, 1), np.int32)
        L.pop_households(gp, n, _p(age_raw), _p(key), _p(age), _p(p1), _p(p2), n_edges, C.byref(H))
        contacts[lkey] = dict(p1=p1[:n_edges], p2=p2[:n_edges], beta=np.ones(n_edges, np.float32))
    for lkey, slot, band, counts in plan:
        if band == 'households':
            continue
        mapping = None
        n_members = n
        if band is not None:
            mapping = np.empty(max(n, 1), np.int32)
            n_members = int(L.pop_members(n, _p(age), float(band[0]), float(band[1]), _p(mapping)))
            mapping = np.ascontiguousarray(mapping[:n_members])
        c = Counts(kind=counts['kind'], pad=0, p0=counts['p0'], lam=counts['lam'], r=counts['r'], q=counts['q'])
        offsets = np.empty(n_members + 1, np.int64)
        n_edges = int(L.pop_contacts_plan(gp, slot, n_members, C.byref(c), _p(offsets)))
        p1 = np.empty(max(n_edges, 1), np.int32)
        p2 = np.empty(max(n_edges, 1), np.int32)
        L.pop_contacts_fill(gp, slot, n_members, _p(mapping), _p(offsets), _p(p1), _p(p2))
        contacts[lkey] = dict(p1=p1[:n_edges], p2=p2[:n_edges], beta=np.ones(n_edges, np.float32))
    out = dict(age=age, sex=np.empty(n, np.uint8), n_households=int(H.value))
    for k in ('symp_prob', 'severe_prob', 'crit_prob', 'death_prob', 'rel_sus', 'rel_trans'):
        out[k] = np.empty(n, np.float32)
    L.pop_people(gp, n, _p(age), _p(out['sex']), _p(out['symp_prob']), _p(out['severe_prob']),
                 _p(out['crit_prob']), _p(out['death_prob']), _p(out['rel_sus']), _p(out['rel_trans']))
    out['contacts'] = {k: contacts[k] for k, _, _, _ in plan}
    return out
