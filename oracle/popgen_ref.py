"""Population generator in plain Python loops (TEST INFRASTRUCTURE, small populations only).

A second restatement of the draw rules of the population generator (covasim/population.py:
143-364, people.py:139-160), written independently of oracle/cport/popgen_port.c so that each
checks the other's indexing: households come from a list of drawn sizes, the age order from
Python's stable ``sorted``, layer members from a comprehension, contact counts from
scipy-free CDF inversion. Only for populations of a few hundred agents.
"""

import math

import numpy as np

from . import philox as ph

POP_AGE = 0x70000000
POP_AGE_NOISE = 0x70000100
POP_PERSON = 0x70000200
POP_HH_SIZE = 0x70000300
POP_HH_PERM = 0x70000400
POP_COUNT = 0x70000500
POP_EDGE = 0x70000600
POP_EDGE_RETRY = 0x70000700
KMAX = 1 << 16


def words(seed, c0, c1, c2, c3=0):
    k0, k1 = ph.seed_key(seed)
    w = ph.philox4x32(c0, c1, c2, c3, k0, k1)
    return [int(x) for x in w]


def unit(a, b):
    return float(ph.u01(a, b))


def first_reaching(cdf, u):
    for j, c in enumerate(cdf):
        if c >= u:
            return j
    return len(cdf) - 1


def count_of(c, u):
    k, p = 0, c['p0']
    F = p
    while u > F and k < KMAX:
        if c['kind'] == 0:
            k += 1
            p = p * c['lam'] / k
        else:
            p = p * ((k + c['r']) / (k + 1.0)) * c['q']
            k += 1
        F = F + p
    return k


def draw_age(t, i):
    w = words(t['seed'], i, 0, POP_AGE)
    b = first_reaching(t['age_cdf'], unit(w[0], w[1]))
    a = float(t['age_lo'][b]) + float(t['age_width'][b]) * unit(w[2], w[3])
    z = words(t['seed'], i, 0, POP_AGE_NOISE)
    s = sum(float(x) for x in z) / 4294967296.0
    key = min(255, max(0, math.floor(a + (s - 2.0) * t['age_noise']) + 64))
    return a, key


def feistel_position(seed, h, H):
    nb = (H - 1).bit_length() if H > 1 else 0
    b = max(1, (nb + 1) // 2)
    mask = (1 << b) - 1
    y = h
    while True:
        L, R = y >> b, y & mask
        for r in range(4):
            F = words(seed, R, r, POP_HH_PERM)[0] & mask
            L, R = R, L ^ F
        y = (L << b) | R
        if y < H:
            return y


def make_population(t, plan, n):
    """Same inputs and outputs as oracle.cport.popgen.make_population."""
    uses_households = any(band == 'households' for _, _, band, _ in plan)
    correlate = bool(t.get('correlate', 0)) and uses_households
    draws = [draw_age(t, i) for i in range(n)]
    age_raw = [np.float32(a) for a, _ in draws]
    age = list(age_raw)
    contacts = {}
    n_households = 0
    for lkey, slot, band, counts in plan:
        if band != 'households':
            continue
        sizes = []
        while sum(sizes) < n:
            w = words(t['seed'], len(sizes) & 0xFFFFFFFF, len(sizes) >> 32, POP_HH_SIZE)
            sizes.append(int(t['hh_size_min']) + first_reaching(t['hh_cdf'], unit(w[0], w[1])))
        if sizes:
            sizes[-1] -= sum(sizes) - n
        H = len(sizes)
        n_households = H
        order = sorted(range(n), key=lambda i: draws[i][1]) if correlate else list(range(n))
        members, pos = [], 0
        for m in sizes:
            members.append(order[pos:pos + m])
            pos += m
        at = [None] * H
        for h in range(H):
            at[feistel_position(t['seed'], h, H) if correlate else h] = h
        age, p1, p2, first = [], [], [], 0
        for j in range(H):
            grp = members[at[j]]
            age.extend(age_raw[i] for i in grp)
            for a in range(len(grp)):
                for b in range(a + 1, len(grp)):
                    p1.append(first + a)
                    p2.append(first + b)
            first += len(grp)
        contacts[lkey] = dict(p1=np.array(p1, np.int32), p2=np.array(p2, np.int32),
                              beta=np.ones(len(p1), np.float32))
    for lkey, slot, band, counts in plan:
        if band == 'households':
            continue
        if band is None:
            mapping = list(range(n))
        else:
            mapping = [i for i in range(n) if band[0] <= float(age[i]) < band[1]]
        nm = len(mapping)
        p1, p2 = [], []
        for m in range(nm):
            w = words(t['seed'], m & 0xFFFFFFFF, m >> 32, POP_COUNT | slot)
            k = count_of(counts, unit(w[0], w[1]))
            p1.extend([mapping[m]] * int(round(k / 2.0)))  # Python's round: ties to even
        thr = (2 ** 32 - nm) % nm if nm else 0
        for e in range(len(p1)):
            w = words(t['seed'], (e >> 2) & 0xFFFFFFFF, e >> 34, POP_EDGE | slot)[e & 3]
            m = w * nm
            k = 0
            while (m & 0xFFFFFFFF) < thr and k < 255:
                m = words(t['seed'], e & 0xFFFFFFFF, e >> 32, POP_EDGE_RETRY | slot, k)[0] * nm
                k += 1
            p2.append(mapping[m >> 32])
        contacts[lkey] = dict(p1=np.array(p1, np.int32), p2=np.array(p2, np.int32),
                              beta=np.ones(len(p1), np.float32))
    out = dict(age=np.array(age, np.float32), n_households=n_households)
    cols = {k: [] for k in ('sex', 'symp_prob', 'severe_prob', 'crit_prob', 'death_prob', 'rel_sus', 'rel_trans')}
    nb = dict(kind=1, p0=t['nb_p0'], lam=0.0, r=t['nb_r'], q=t['nb_q'])
    for i in range(n):
        w = words(t['seed'], i, 0, POP_PERSON)
        a = float(age[i])
        b = max(0, sum(1 for c in t['prog_cutoff'] if a >= c) - 1)
        cols['sex'].append(w[0] & 1)
        for k in ('symp_prob', 'severe_prob', 'crit_prob', 'death_prob', 'rel_sus'):
            cols[k].append(t[k][b])
        cols['rel_trans'].append(float(t['trans_or'][b]) * (count_of(nb, unit(w[1], w[2])) * float(t['nb_step'])))
    out['sex'] = np.array(cols['sex'], np.uint8)
    for k in ('symp_prob', 'severe_prob', 'crit_prob', 'death_prob', 'rel_sus', 'rel_trans'):
        out[k] = np.array(cols[k], np.float64).astype(np.float32)
    out['contacts'] = {k: contacts[k] for k, _, _, _ in plan}
    return out
