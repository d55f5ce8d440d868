"""Interventions restated for the oracle Sim (TEST INFRASTRUCTURE).

change_beta follows covasim/interventions.py:530-586: on each of its days it sets the overall beta
(layers=None) or a layer's beta to the value it had when the intervention was initialised times
that day's change, by editing the sim's parameters, which the next transmission step reads.
"""


class change_beta:
    def __init__(self, days, changes, layers=None):
        self.days = [int(days)] if isinstance(days, (int, float)) else [int(d) for d in days]
        self.changes = ([float(changes)] * len(self.days) if isinstance(changes, (int, float))
                        else [float(c) for c in changes])
        self.layers = [None] if layers is None else ([layers] if isinstance(layers, str) else list(layers))
        self.orig_betas = None

    def initialize(self, sim):
        self.orig_betas = {}
        for lkey in self.layers:
            if lkey is None:
                self.orig_betas['overall'] = sim.pars['beta']
            else:
                self.orig_betas[lkey] = sim.pars['beta_layer'][lkey]
        return self

    def __call__(self, sim):
        for ind, day in enumerate(self.days):
            if day == sim.t:
                for lkey, beta in self.orig_betas.items():
                    if lkey == 'overall':
                        sim.pars['beta'] = beta * self.changes[ind]
                    else:
                        sim.pars['beta_layer'][lkey] = beta * self.changes[ind]


def item_keys(seed, t, stream, items):
    """Choice keys of items: (philox(seed; i lo, t, stream, i hi).x << 32) | i lo, distinct per item."""
    import numpy as np
    from . import philox as ph
    e = np.asarray(items, dtype=np.uint64)
    k0, k1 = ph.seed_key(seed)
    w = ph.philox4x32(e & np.uint64(0xFFFFFFFF), t, stream, e >> np.uint64(32), k0, k1)[0]
    return (np.asarray(w, dtype=np.uint64) << np.uint64(32)) | (e & np.uint64(0xFFFFFFFF))


def choose_items(seed, t, stream, items, k):
    """The k of items with the smallest keys (all of them when k >= len(items)): a uniform choice
    without replacement, as cvu.choose / np.random.choice(replace=False)."""
    import numpy as np
    items = np.asarray(items, dtype=np.int64)
    if k <= 0:
        return items[:0]
    if k >= len(items):
        return items
    keys = item_keys(seed, t, stream, items)
    kth = np.partition(keys, k - 1)[k - 1]
    return items[keys <= kth]


def chosen_edges(seed, t, slot, n, k):
    """Mask of the k edges of n with the smallest keys in stream 0x80000000 | slot (clip_edges)."""
    import numpy as np
    mask = np.zeros(n, dtype=bool)
    mask[choose_items(seed, t, 0x80000000 | slot, np.arange(n), k)] = True
    return mask


class clip_edges:
    """covasim/interventions.py:589-673: on each of its days, a layer keeps the share changes[i] of
    its edges (sim's plus the intervention's store); the edges moved out or back are appended, in
    their order, to the other side, and both sides keep their remaining edges in order
    (Layer.pop_inds / append, base.py:1739-1768)."""

    def __init__(self, days, changes, layers=None):
        self.days = [int(days)] if isinstance(days, (int, float)) else [int(d) for d in days]
        self.changes = ([float(changes)] * len(self.days) if isinstance(changes, (int, float))
                        else [float(c) for c in changes])
        self.layers = layers
        self.contacts = None

    def initialize(self, sim):
        import numpy as np
        keys = list(sim.layers.keys())
        self.layers = keys if self.layers is None else ([self.layers] if isinstance(self.layers, str)
                                                         else list(self.layers))
        self.contacts = {k: [np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32)]
                         for k in self.layers}
        return self

    def __call__(self, sim):
        import numpy as np
        keys = list(sim.layers.keys())
        for ind, day in enumerate(self.days):
            if day != sim.t:
                continue
            for lkey in self.layers:
                lay, held = sim.layers[lkey], self.contacts[lkey]
                n_sim, n_int = len(lay[0]), len(held[0])
                n = n_sim + n_int
                if not n:
                    continue
                n_to_move = int((n_sim / n - self.changes[ind]) * n)
                slot = keys.index(lkey)
                if n_to_move > 0:
                    mask = chosen_edges(sim.pars['rand_seed'], sim.t, slot, n_sim, n_to_move)
                    for i in range(3):
                        held[i] = np.concatenate([held[i], lay[i][mask]])
                        lay[i] = lay[i][~mask]
                elif n_to_move < 0:
                    mask = chosen_edges(sim.pars['rand_seed'], sim.t, slot, n_int, -n_to_move)
                    for i in range(3):
                        lay[i] = np.concatenate([lay[i], held[i][mask]])
                        held[i] = held[i][~mask]


# ---------------------------------------------------------------------------------------------
# Testing and contact tracing (interventions.py:682-710, 858-1145)
# ---------------------------------------------------------------------------------------------
def _active(iv, t):
    return t >= iv.start_day and (iv.end_day is None or t <= iv.end_day)


def get_quar_inds(quar_policy, sim):
    """interventions.py:682-710 (the policies given as a string or numbers)."""
    import numpy as np
    t, people = sim.t, sim.people
    if quar_policy is None:
        return np.array([], dtype=np.int64)
    if quar_policy == 'start':
        return (people.date_quarantined == t - 1).nonzero()[0]
    if quar_policy == 'end':
        return (people.date_end_quarantine == t + 1).nonzero()[0]
    if quar_policy == 'both':
        return np.concatenate([(people.date_quarantined == t - 1).nonzero()[0],
                               (people.date_end_quarantine == t + 1).nonzero()[0]])
    if quar_policy == 'daily':
        return people.quarantined.nonzero()[0]
    q = np.atleast_1d(np.asarray(quar_policy))
    return np.unique(np.concatenate([(people.date_quarantined == t - 1 - v).nonzero()[0] for v in q]))


def get_pdf(dist=None, par1=None, par2=None):
    """utils.py:240-268: scipy's uniform on [par1, par1 + par2], or the lognormal of mean par1 and
    variance par2 shifted by -0.5; None for no distribution."""
    import numpy as np
    import scipy.stats as sps
    if dist in ('None', 'none', None):
        return None
    if dist == 'uniform':
        return sps.uniform(loc=par1, scale=par2)
    if dist == 'lognormal':
        mean = np.log(par1 ** 2 / np.sqrt(par2 + par1 ** 2))
        sigma = np.sqrt(np.log(par2 / par1 ** 2 + 1))
        return sps.lognorm(sigma, loc=-0.5, scale=np.exp(mean))
    raise NotImplementedError(dist)


def get_subtargets(subtarget, sim):
    """interventions.py:150-187."""
    if callable(subtarget):
        subtarget = subtarget(sim)
    inds = subtarget['inds'](sim) if callable(subtarget['inds']) else subtarget['inds']
    vals = subtarget['vals'](sim) if callable(subtarget['vals']) else subtarget['vals']
    return inds, vals


class test_prob:
    """interventions.py:858-981: test probabilities set piece by piece in the reference's order, one
    Bernoulli draw per agent, then People.test. The draws of agent i are philox(i, t, 0x90000000 |
    slot, 0) (tested, true positive) and philox(i, t, 0x90000000 | slot, 1) (not lost), slot = the
    intervention's index in the sim. Options: subtarget (interventions.py:958-960), ili_prev (the
    day's int(ili_prev[t - start_day] * pop_size) agents with the smallest keys in stream 0xC0000000 |
    slot, interventions.py:936-942) and swab_delay (interventions.py:924-933, with the pdf taken from
    a table of its values at whole days 0..npts-1, which is where it is evaluated)."""

    def __init__(self, symp_prob, asymp_prob=0.0, symp_quar_prob=None, asymp_quar_prob=None, quar_policy=None,
                 subtarget=None, ili_prev=None, sensitivity=1.0, loss_prob=0.0, test_delay=0, start_day=0,
                 end_day=None, swab_delay=None):
        self.symp_prob = symp_prob
        self.asymp_prob = asymp_prob
        self.symp_quar_prob = symp_quar_prob if symp_quar_prob is not None else symp_prob
        self.asymp_quar_prob = asymp_quar_prob if asymp_quar_prob is not None else asymp_prob
        self.quar_policy = quar_policy if quar_policy else 'start'
        self.subtarget = subtarget
        self.ili_prev = ili_prev
        self.sensitivity = sensitivity
        self.loss_prob = loss_prob
        self.test_delay = test_delay
        self.start_day = start_day
        self.end_day = end_day
        self.pdf = get_pdf(**(swab_delay or {}))
        self.last_ili = None
        self.last_probs = None

    def initialize(self, sim):
        import numpy as np
        if self.ili_prev is not None and np.isscalar(self.ili_prev):
            self.ili_prev = np.array([self.ili_prev] * sim.npts)
        self.pdf_table = None
        if self.pdf:
            self.pdf_table = np.asarray(self.pdf.pdf(np.arange(sim.npts, dtype=np.int32)), dtype=np.float64)
        return self

    def __call__(self, sim):
        import numpy as np
        from . import philox as px
        t, people = sim.t, sim.people
        if not _active(self, t):
            return
        n = people.n
        seed = sim.pars['rand_seed']
        symp_inds = people.symptomatic.nonzero()[0]
        symp_prob = self.symp_prob
        if self.pdf:
            symp_time = (t - people.date_symptomatic[symp_inds]).astype(np.int32)
            inv_count = np.bincount(symp_time) / len(symp_time)
            count = np.nan * np.ones(inv_count.shape)
            count[inv_count != 0] = 1 / inv_count[inv_count != 0]
            symp_prob = np.ones(len(symp_time))
            inds = 1 > (symp_time * self.symp_prob)
            symp_prob[inds] = self.symp_prob / (1 - symp_time[inds] * self.symp_prob)
            symp_prob = self.pdf_table[symp_time] * symp_prob * count[symp_time]
        ili_inds = np.array([], dtype=np.int64)
        if self.ili_prev is not None:
            rel_t = t - self.start_day
            if rel_t < len(self.ili_prev):
                n_ili = int(self.ili_prev[rel_t] * n)
                ili_inds = choose_items(seed, t, px.STREAM_ILI | sim.iv_slot, np.arange(n), n_ili)
                ili_inds = np.setdiff1d(ili_inds, symp_inds)
        self.last_ili = ili_inds
        asymp_inds = np.setdiff1d(np.setdiff1d(np.arange(n), symp_inds), ili_inds)
        quar_test_inds = get_quar_inds(self.quar_policy, sim)
        symp_quar_inds = np.intersect1d(quar_test_inds, symp_inds)
        asymp_quar_inds = np.intersect1d(quar_test_inds, asymp_inds)
        diag_inds = people.diagnosed.nonzero()[0]
        test_probs = np.zeros(n)
        test_probs[symp_inds] = symp_prob
        test_probs[ili_inds] = self.symp_prob
        test_probs[asymp_inds] = self.asymp_prob
        test_probs[symp_quar_inds] = self.symp_quar_prob
        test_probs[asymp_quar_inds] = self.asymp_quar_prob
        if self.subtarget is not None:
            sub_inds, sub_vals = get_subtargets(self.subtarget, sim)
            test_probs[sub_inds] = sub_vals
        test_probs[diag_inds] = 0.0
        self.last_probs = test_probs
        stream = px.STREAM_TEST | sim.iv_slot
        u_test, u_pos = px.stream_uniforms(seed, t, np.arange(n), stream, 0)
        u_keep, _ = px.stream_uniforms(seed, t, np.arange(n), stream, 1)
        test_inds = (u_test < test_probs).nonzero()[0]
        people.test(test_inds, u_pos, u_keep, test_sensitivity=self.sensitivity, loss_prob=self.loss_prob,
                    test_delay=self.test_delay)
        sim.results['new_tests'][t] += len(test_inds)
        return test_inds


def find_contacts(p1, p2, inds):
    """utils.py:131-147 + Layer.find_contacts (base.py): the partners of inds, sorted and unique."""
    import numpy as np
    inds = set(int(i) for i in inds)
    partners = set()
    for a, b in zip(p1.tolist(), p2.tolist()):
        if a in inds:
            partners.add(b)
        if b in inds:
            partners.add(a)
    return np.array(sorted(partners), dtype=np.int64)


class contact_tracing:
    """interventions.py:984-1151: the agents diagnosed today (or, presumptive, exposed and tested
    today) are traced through each layer with trace_prob > 0 -- under a capacity, when there are more
    of them, the int(capacity) with the smallest keys in stream 0xD0000000 | slot (select_cases,
    :1078-1083); each contact found in layer l is kept when philox(i, t, 0xA0000000 | slot << 8 | l,
    0) < trace_prob[l], and the living ones are notified and scheduled to quarantine on day
    int(t + trace_time[l])."""

    def __init__(self, trace_probs=None, trace_time=None, start_day=0, end_day=None, presumptive=False,
                 quar_period=None, capacity=None):
        self.trace_probs = trace_probs
        self.trace_time = trace_time
        self.start_day = start_day
        self.end_day = end_day
        self.presumptive = presumptive
        self.quar_period = quar_period
        self.capacity = capacity
        self.last_traced = None

    def initialize(self, sim):
        keys = list(sim.layers.keys())
        if self.trace_probs is None:
            self.trace_probs = 1.0
        if self.trace_time is None:
            self.trace_time = 0.0
        if self.quar_period is None:
            self.quar_period = sim.pars['quar_period']
        if isinstance(self.trace_probs, (int, float)):
            self.trace_probs = {k: self.trace_probs for k in keys}
        if isinstance(self.trace_time, (int, float)):
            self.trace_time = {k: self.trace_time for k in keys}
        return self

    def __call__(self, sim):
        import numpy as np
        from . import philox as px
        t, people = sim.t, sim.people
        if not _active(self, t):
            return
        if not self.presumptive:
            trace_inds = (people.date_diagnosed == t).nonzero()[0]
        else:
            just_tested = (people.date_tested == t).nonzero()[0]
            trace_inds = just_tested[people.exposed[just_tested]]
        if self.capacity is not None:
            capacity = int(self.capacity / 1.0)  # rescale_vec[t] is 1 in unscaled sims
            if len(trace_inds) > capacity:
                trace_inds = np.sort(choose_items(sim.pars['rand_seed'], t, px.STREAM_TRACE_CAP | sim.iv_slot,
                                                  trace_inds, capacity))
        self.last_traced = trace_inds
        if not len(trace_inds):
            return {}
        keys = list(sim.layers.keys())
        contacts = {}
        for lkey, prob in self.trace_probs.items():
            if prob == 0:
                continue
            lay = sim.layers[lkey]
            traceable = find_contacts(lay[0], lay[1], trace_inds)
            if len(traceable):
                u, _ = px.stream_uniforms(sim.pars['rand_seed'], t, traceable,
                                          px.STREAM_TRACE | (sim.iv_slot << 8) | keys.index(lkey), 0)
                contacts.setdefault(self.trace_time[lkey], []).extend(traceable[u < prob].tolist())
        is_dead = people.dead.nonzero()[0]
        for trace_time, inds in contacts.items():  # notify_contacts, interventions.py:1125-1145
            inds = np.setdiff1d(np.array(inds, dtype=np.int64), is_dead)
            people.known_contact[inds] = True
            people.date_known_contact[inds] = np.fmin(people.date_known_contact[inds], t + trace_time)
            people.schedule_quarantine(inds, start_date=t + trace_time, period=self.quar_period - trace_time)
        return contacts
