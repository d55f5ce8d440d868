"""Synthetic populations (ages, sexes and contact layers) built with vectorised numpy.

Restates covasim/population.py:143-364 (make_randpop, make_random_contacts,
make_microstructured_contacts, make_hybrid_contacts) without per-agent Python
loops, so 1M-10M agent populations build in seconds. Draws come from a numpy
``Generator`` seeded with the sim seed; the reference's MT19937 stream order is
not reproduced (population building is an input to the transmission step).

Besides the reference's Poisson-sized household clusters, ``make_hybrid_contacts``
offers categorical household sizes over 1..7 (mean about 2.6) whose members have
correlated ages: agents are ordered by age plus N(0, 15 years) noise before being
cut into households, and household blocks are then shuffled, so index order does
not follow age.
"""

import numpy as np

__all__ = ['make_randpop', 'make_random_contacts', 'make_microstructured_contacts', 'make_hybrid_contacts',
           'make_people_arrays', 'DEFAULT_AGE_DATA', 'HOUSEHOLD_SIZE_PROBS']

DEFAULT_AGE_DATA = np.array([
    [0, 4, 0.0605], [5, 9, 0.0607], [10, 14, 0.0566], [15, 19, 0.0557], [20, 24, 0.0612], [25, 29, 0.0843],
    [30, 34, 0.0848], [35, 39, 0.0764], [40, 44, 0.0697], [45, 49, 0.0701], [50, 54, 0.0681], [55, 59, 0.0653],
    [60, 64, 0.0591], [65, 69, 0.0453], [70, 74, 0.0312], [75, 79, 0.02016], [80, 84, 0.01344], [85, 89, 0.01008],
    [90, 99, 0.00672],
])  # defaults.py:223-243

# household sizes 1..7, mean 2.62
HOUSEHOLD_SIZE_PROBS = np.array([0.25, 0.33, 0.16, 0.14, 0.07, 0.03, 0.02])


def draw_ages(pop_size, rng, age_data=DEFAULT_AGE_DATA):
    """Age bins by multinomial draw, uniform within the bin (population.py:198-204)."""
    lo = age_data[:, 0]
    rng_w = age_data[:, 1] + 1 - lo
    prob = age_data[:, 2] / age_data[:, 2].sum()
    bins = np.searchsorted(np.cumsum(prob), rng.random(pop_size))
    bins = np.minimum(bins, len(prob) - 1)
    return lo[bins] + rng_w[bins] * rng.random(pop_size)


def make_random_contacts(pop_size, n, rng, mapping=None):
    """round(Poisson(n)/2) edges per agent to uniformly chosen partners (population.py:241-284)."""
    pop_size = int(pop_size)
    if pop_size == 0:
        return dict(p1=np.zeros(0, np.int32), p2=np.zeros(0, np.int32))
    p_count = np.round(rng.poisson(n, pop_size) / 2.0).astype(np.int64)
    total = int(p_count.sum())
    p1 = np.repeat(np.arange(pop_size, dtype=np.int32), p_count)
    p2 = rng.integers(0, pop_size, size=total, dtype=np.int64).astype(np.int32)
    if mapping is not None:
        mapping = np.asarray(mapping, dtype=np.int32)
        p1 = mapping[p1]
        p2 = mapping[p2]
    return dict(p1=p1, p2=p2)


def _clique_edges(starts, sizes):
    """All pairs (source < target) inside each cluster [start, start + size)."""
    p1s, p2s = [], []
    for k in np.unique(sizes):
        if k < 2:
            continue
        s = starts[sizes == k].astype(np.int64)
        a, b = np.triu_indices(int(k), 1)
        p1s.append((s[:, None] + a[None, :]).ravel())
        p2s.append((s[:, None] + b[None, :]).ravel())
    if not p1s:
        return dict(p1=np.zeros(0, np.int32), p2=np.zeros(0, np.int32))
    p1 = np.concatenate(p1s)
    p2 = np.concatenate(p2s)
    order = np.lexsort((p2, p1))
    return dict(p1=p1[order].astype(np.int32), p2=p2[order].astype(np.int32))


def _cluster_sizes(pop_size, draw):
    sizes = []
    total = 0
    while total < pop_size:
        chunk = draw(max(1024, (pop_size - total) // 2 + 16))
        sizes.append(chunk)
        total += int(chunk.sum())
    sizes = np.concatenate(sizes)
    csum = np.cumsum(sizes)
    last = int(np.searchsorted(csum, pop_size))
    sizes = sizes[:last + 1].copy()
    sizes[-1] -= int(csum[last] - pop_size)
    return sizes


def make_microstructured_contacts(pop_size, cluster_size, rng):
    """Poisson-sized clusters of consecutive agents, fully connected (population.py:287-329)."""
    sizes = _cluster_sizes(int(pop_size), lambda m: rng.poisson(cluster_size, m))
    starts = np.cumsum(sizes) - sizes
    return _clique_edges(starts, sizes)


def make_households(pop_size, ages, rng, size_probs=HOUSEHOLD_SIZE_PROBS, age_noise=15.0):
    """Categorical household sizes with age-correlated members; returns (ages, contacts)."""
    pop_size = int(pop_size)
    sizes = _cluster_sizes(pop_size, lambda m: rng.choice(np.arange(1, len(size_probs) + 1), size=m, p=size_probs))
    order = np.argsort(ages + rng.normal(0, age_noise, pop_size), kind='stable')
    sorted_ages = ages[order]
    n_h = len(sizes)
    perm = rng.permutation(n_h)
    old_starts = np.cumsum(sizes) - sizes
    new_sizes = sizes[perm]
    new_starts = np.cumsum(new_sizes) - new_sizes
    offs = np.arange(pop_size) - np.repeat(new_starts, new_sizes)
    src_pos = np.repeat(old_starts[perm], new_sizes) + offs
    new_ages = sorted_ages[src_pos]
    return new_ages, _clique_edges(new_starts, new_sizes)


def make_hybrid_contacts(pop_size, ages, contacts, rng, households='poisson', school_ages=None, work_ages=None):
    """Households + school + work + community layers (population.py:332-364).

    Returns (ages, contacts) -- ages are reordered when households='categorical'.
    """
    c = {'h': 4, 's': 20, 'w': 20, 'c': 20}
    c.update(contacts or {})
    school_ages = school_ages or [6, 22]
    work_ages = work_ages or [22, 65]
    out = {}
    if households == 'categorical':
        ages, out['h'] = make_households(pop_size, ages, rng)
    else:
        out['h'] = make_microstructured_contacts(pop_size, c['h'], rng)
    s_inds = np.flatnonzero((ages >= school_ages[0]) * (ages < school_ages[1]))
    w_inds = np.flatnonzero((ages >= work_ages[0]) * (ages < work_ages[1]))
    out['s'] = make_random_contacts(len(s_inds), c['s'], rng, mapping=s_inds)
    out['w'] = make_random_contacts(len(w_inds), c['w'], rng, mapping=w_inds)
    out['c'] = make_random_contacts(pop_size, c['c'], rng)
    return ages, out


def make_randpop(pars, rng, households='poisson'):
    """Ages, sexes and contacts for 'random' and 'hybrid' populations (population.py:143-226)."""
    pop_size = int(pars['pop_size'])
    sexes = rng.binomial(1, 0.5, pop_size)
    ages = draw_ages(pop_size, rng)
    if pars['pop_type'] == 'random':
        contacts = {lkey: make_random_contacts(pop_size, n, rng) for lkey, n in pars['contacts'].items()}
    elif pars['pop_type'] == 'hybrid':
        ages, contacts = make_hybrid_contacts(pop_size, ages, pars['contacts'], rng, households=households)
    else:
        raise NotImplementedError(f'Population type "{pars["pop_type"]}" not supported; choices are random or hybrid')
    return dict(uid=np.arange(pop_size, dtype=np.int32), age=ages.astype(np.float32), sex=sexes, contacts=contacts)


def make_people_arrays(age, pars, rng):
    """Per-agent prognoses and relative transmissibility/susceptibility (people.py:139-160)."""
    progs = pars['prognoses']
    inds = np.digitize(age, progs['age_cutoffs']) - 1
    bd = pars['beta_dist']
    if bd['dist'] != 'neg_binomial':
        raise NotImplementedError('only the neg_binomial beta_dist of the reference defaults is supported')
    step = bd.get('step', 1)
    nbn_n = bd['par2']
    nbn_p = bd['par2'] / (bd['par1'] / step + bd['par2'])  # utils.py:423-426
    draws = rng.negative_binomial(n=nbn_n, p=nbn_p, size=len(age)) * step
    return dict(
        symp_prob=progs['symp_probs'][inds].astype(np.float32),
        severe_prob=(progs['severe_probs'][inds] * progs['comorbidities'][inds]).astype(np.float32),
        crit_prob=progs['crit_probs'][inds].astype(np.float32),
        death_prob=progs['death_probs'][inds].astype(np.float32),
        rel_sus=progs['sus_ORs'][inds].astype(np.float32),
        rel_trans=(progs['trans_ORs'][inds] * draws).astype(np.float32),
    )
