"""Interventions that act on the transmission step (covasim/interventions.py).

``change_beta`` (interventions.py:530-586) rescales the overall beta or a layer's beta from given
days on. ``clip_edges`` (interventions.py:589-673) moves a share of a layer's edges out of the sim
into the intervention's store, and back, with the cvk_move_edges kernels; runs with it go day by
day (the layer sizes change between days). The GPU day loop takes its effect as a per-day table of the float32 rates the kernels
use (``cvk_imports.beta_v_days`` / ``beta_layer_days``), built once when the sim initialises by
replaying the interventions over the days on a copy of the parameters. A run therefore stays one
captured hipGraph with the rates of each day baked into that day's launches, and the sim's own
parameters are left as they were (the reference restores them after a run, sim.py:803-808).
"""

import numpy as np

from . import _lib

__all__ = ['Intervention', 'change_beta', 'clip_edges', 'test_prob', 'contact_tracing', 'beta_schedule']


class Intervention:
    """Base class (interventions.py:75-260): a label and an ``apply(sim)`` run on each day."""

    def __init__(self, label=None, **kwargs):
        self.label = label if label is not None else type(self).__name__
        self.initialized = False

    def initialize(self, sim):
        self.initialized = True

    def __call__(self, sim):
        if not self.initialized:
            raise RuntimeError(f'Intervention {self.label} (label={self.label}) has not been initialized')
        return self.apply(sim)

    def apply(self, sim):
        raise NotImplementedError

    def __repr__(self):
        return f'{type(self).__name__}({self.label})'


def process_days(sim, days):
    """Days as integers (interventions.py:38-57): ints pass through, dates go through sim.day."""
    days = [days] if np.isscalar(days) or isinstance(days, str) else list(days)
    out = []
    for d in days:
        if isinstance(d, str):
            if not hasattr(sim, 'day'):
                raise ValueError(f'Cannot convert the date "{d}" to a day on this sim')
            d = sim.day(d)
        out.append(int(d))
    return out


def process_changes(changes, days):
    """One change per day (interventions.py:60-72)."""
    if np.isscalar(changes):
        return [float(changes)] * len(days)
    changes = [float(c) for c in changes]
    if len(changes) != len(days):
        raise ValueError(f'Number of days supplied ({len(days)}) does not match number of changes ({len(changes)})')
    return changes


class change_beta(Intervention):
    """Scale beta by ``changes[i]`` from ``days[i]`` on, overall (layers=None) or in given layers
    (interventions.py:530-586)."""

    def __init__(self, days, changes, layers=None, **kwargs):
        super().__init__(**kwargs)
        self.days = days
        self.changes = changes
        self.layers = layers
        self.orig_betas = None

    def initialize(self, sim):
        super().initialize(sim)
        self.days = process_days(sim, self.days)
        self.changes = process_changes(self.changes, self.days)
        layers = self.layers
        self.layers = [None] if layers is None else ([layers] if isinstance(layers, str) else list(layers))
        pars = sim.pars
        self.orig_betas = {}
        for lkey in self.layers:
            if lkey is None:
                self.orig_betas['overall'] = pars['beta']
            else:
                if lkey not in pars['beta_layer']:
                    raise ValueError(f'change_beta: layer "{lkey}" not in {list(pars["beta_layer"])}')
                self.orig_betas[lkey] = pars['beta_layer'][lkey]

    def apply(self, sim):
        pars = sim.pars
        for ind, day in enumerate(self.days):
            if day != sim.t:
                continue
            for lkey, beta in self.orig_betas.items():
                new_beta = beta * self.changes[ind]
                if lkey == 'overall':
                    pars['beta'] = new_beta
                else:
                    pars['beta_layer'][lkey] = new_beta


def move_edges(cols, n, k, dest, seed, t, slot):
    """Move k of the n edges in cols (p1, p2, beta tensors), chosen by cvk_move_edges, onto the end of
    dest (p1, p2, beta tensors). Returns (kept columns, extended destination columns)."""
    import ctypes as C
    import torch
    L = _lib.lib()
    dev = cols[0].device
    m = int(dest[0].numel())
    keep = [torch.empty(n - k, dtype=c.dtype, device=dev) for c in cols]
    out = [torch.empty(m + k, dtype=c.dtype, device=dev) for c in cols]
    for o, d in zip(out, dest):
        o[:m].copy_(d)
    nbytes = int(L.cvk_move_edges_work_bytes(n))
    buf = torch.empty(nbytes + 256, dtype=torch.uint8, device=dev)
    work = C.c_void_p(buf.data_ptr() + (-buf.data_ptr()) % 256)
    tail = [C.c_void_p(o.data_ptr() + m * o.element_size()) for o in out]
    _lib.check(L.cvk_move_edges(_lib.ptr(cols[0]), _lib.ptr(cols[1]), _lib.ptr(cols[2]), n, k,
                                int(seed) & 0xFFFFFFFFFFFFFFFF, int(t), int(slot), _lib.ptr(keep[0]), _lib.ptr(keep[1]),
                                _lib.ptr(keep[2]), tail[0], tail[1], tail[2], work, nbytes, _lib.stream_ptr()),
               'cvk_move_edges')
    torch.cuda.current_stream().synchronize()  # before `buf` is released
    return keep, out


class clip_edges(Intervention):
    """From ``days[i]`` on, keep the share ``changes[i]`` of each layer's edges in the sim: edges move
    to the intervention's store or back, chosen uniformly without replacement
    (interventions.py:589-673). Layers default to all of the sim's."""

    def __init__(self, days, changes, layers=None, **kwargs):
        super().__init__(**kwargs)
        self.days = days
        self.changes = changes
        self.layers = layers
        self.contacts = None

    def initialize(self, sim):
        import torch
        super().initialize(sim)
        self.days = process_days(sim, self.days)
        self.changes = process_changes(self.changes, self.days)
        keys = sim.layer_keys() if hasattr(sim, 'layer_keys') else list(sim.pars['beta_layer'])
        self.layers = list(keys) if self.layers is None else (
            [self.layers] if isinstance(self.layers, str) else list(self.layers))
        dev = getattr(sim, 'device', None) or 'cuda'
        self.contacts = {k: (torch.empty(0, dtype=torch.int32, device=dev), torch.empty(0, dtype=torch.int32, device=dev),
                             torch.empty(0, dtype=torch.float32, device=dev)) for k in self.layers}

    def apply(self, sim):
        for ind, day in enumerate(self.days):
            if day != sim.t:
                continue
            keys = sim.layer_keys()
            for lkey in self.layers:
                layer = sim.people.contacts[lkey]
                held = self.contacts[lkey]
                n_sim, n_int = len(layer), int(held[0].numel())
                n = n_sim + n_int
                if not n:
                    continue
                n_to_move = int((n_sim / n - self.changes[ind]) * n)
                cols = (layer.tensor('p1'), layer.tensor('p2'), layer.tensor('beta'))
                slot = keys.index(lkey)
                if n_to_move > 0:  # out of the sim
                    keep, self.contacts[lkey] = move_edges(cols, n_sim, n_to_move, held, sim.pars['rand_seed'],
                                                           sim.t, slot)
                    new = keep
                elif n_to_move < 0:  # back into the sim
                    kept_held, new = move_edges(held, n_int, -n_to_move, cols, sim.pars['rand_seed'], sim.t, slot)
                    self.contacts[lkey] = tuple(kept_held)
                else:
                    continue
                layer['p1'], layer['p2'], layer['beta'] = new[0], new[1], new[2]


def _day_or_none(sim, day):
    if day is None:
        return None
    return process_days(sim, day)[0]


def process_daily_data(daily_data, sim, start_day):
    """Per-day values from the intervention's start (interventions.py:111-147): a number is repeated
    over the sim's days, a pandas Series / DataFrame with a date index is reindexed from the start
    date (missing days 0), a sequence is used as it is. Named columns of loaded data are not held by
    this sim."""
    if daily_data is None:
        return None
    if isinstance(daily_data, str):
        raise NotImplementedError('per-day data by column name needs data loaded into the sim')
    if np.isscalar(daily_data):
        return np.array([daily_data] * sim.npts)
    try:
        import pandas as pd
    except ImportError:  # pragma: no cover
        pd = None
    if pd is not None and isinstance(daily_data, (pd.Series, pd.DataFrame)):
        start_date = pd.Timestamp(sim['start_day']) + pd.Timedelta(days=start_day)
        dateindex = pd.date_range(start_date, daily_data.index[-1])
        out = daily_data.reindex(dateindex, fill_value=0).to_numpy()
        return out.reshape(len(out), -1)[:, 0] if out.ndim == 2 and out.shape[1] == 1 else out
    return np.asarray(daily_data)


def get_subtargets(subtarget, sim):
    """Indices and values of a subtarget (interventions.py:150-187): a dict with 'inds' and 'vals',
    or a function of the sim returning one; 'inds' and 'vals' may be functions of the sim."""
    if callable(subtarget):
        subtarget = subtarget(sim)
    if 'inds' not in subtarget:
        raise ValueError(f'The subtarget dict must have keys "inds" and "vals", but you supplied {subtarget}')
    inds = subtarget['inds'](sim) if callable(subtarget['inds']) else subtarget['inds']
    vals = subtarget['vals'](sim) if callable(subtarget['vals']) else subtarget['vals']
    if np.iterable(vals) and not isinstance(vals, str) and len(vals) != len(inds):
        raise ValueError(f'Length of subtargeting indices ({len(inds)}) does not match length of values ({len(vals)})')
    return inds, vals


def subtarget_arrays(subtarget, sim, device):
    """The subtarget of a day as device arrays: a bitmap of the agents ([(n + 31) / 32] words) and
    their probabilities (float64 [n], assigned as test_probs[inds] = vals is, so a repeated index
    keeps its last value), or the one value when vals is a number."""
    import torch
    n = sim.people.n
    inds, vals = get_subtargets(subtarget, sim)
    inds = np.asarray(inds)
    if inds.dtype == bool:
        inds = inds.nonzero()[0]
    inds = inds.astype(np.int64).ravel()
    if len(inds) and (inds.min() < -n or inds.max() >= n):
        raise IndexError('subtarget indices out of range')
    mask = np.zeros(n, dtype=bool)
    mask[inds] = True
    words = np.zeros((n + 31) // 32 * 4, dtype=np.uint8)
    packed = np.packbits(mask, bitorder='little')
    words[:len(packed)] = packed
    mask_t = torch.from_numpy(words.view(np.int32).copy()).to(device)
    if np.isscalar(vals):
        return mask_t, None, float(vals)
    probs = np.zeros(n, dtype=np.float64)
    probs[inds] = np.asarray(vals, dtype=np.float64).ravel()
    return mask_t, torch.from_numpy(probs).to(device), 0.0


class test_prob(Intervention):
    """Test each agent with a probability set by its symptoms and quarantine state; true positives
    among the infectious are diagnosed test_delay days later unless lost to follow-up
    (interventions.py:858-981, People.test people.py:589-617). Runs on the device (cvk_diag_apply),
    with the subtarget, ili_prev and swab_delay options: a subtarget of index / value arrays is
    uploaded once, one given by functions of the sim is evaluated on the host on each of the
    intervention's days (those days then run one by one); the ILI agents of a day are a uniform
    choice on the device; the swab delay's pdf is tabulated by whole days since onset."""

    def __init__(self, symp_prob, asymp_prob=0.0, symp_quar_prob=None, asymp_quar_prob=None, quar_policy=None,
                 subtarget=None, ili_prev=None, sensitivity=1.0, loss_prob=0.0, test_delay=0, start_day=0,
                 end_day=None, swab_delay=None, **kwargs):
        super().__init__(**kwargs)
        from .utils import get_pdf
        self._kwargs = dict(symp_prob=symp_prob, asymp_prob=asymp_prob, symp_quar_prob=symp_quar_prob,
                            asymp_quar_prob=asymp_quar_prob, quar_policy=quar_policy, subtarget=subtarget,
                            ili_prev=ili_prev, sensitivity=sensitivity, loss_prob=loss_prob, test_delay=test_delay,
                            start_day=start_day, end_day=end_day, swab_delay=swab_delay, label=kwargs.get('label'))
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
        self.swab_delay = swab_delay
        self.pdf = get_pdf(**(swab_delay or {}))
        self._dev = {}

    @property
    def dynamic_subtarget(self):
        """Whether the subtarget is given by functions of the sim (evaluated on each active day)."""
        st = self.subtarget
        return st is not None and (callable(st) or callable(st.get('inds')) or callable(st.get('vals')))

    def active(self, t):
        return t >= self.start_day and (self.end_day is None or t <= self.end_day)

    def initialize(self, sim):
        super().initialize(sim)
        self.start_day = _day_or_none(sim, self.start_day)
        self.end_day = _day_or_none(sim, self.end_day)
        self.days = [self.start_day, self.end_day]
        dev = sim.people.device if hasattr(sim, 'people') else None
        self._dev = {}
        self._ili_counts = None
        if self.ili_prev is not None:
            daily = process_daily_data(self.ili_prev, sim, self.start_day)
            pop_size = int(sim['pop_size'])
            counts = np.array([int(v * pop_size) for v in daily], dtype=np.int64)
            if len(counts) and (counts.min() < 0 or counts.max() > pop_size):
                raise ValueError('ili_prev must give between 0 and pop_size agents per day')
            self._ili_counts = counts
        self._swab_base = None
        if self.pdf:
            # interventions.py:925-933 by whole days since onset s: pdf(s) * (symp_prob / (1 - s * symp_prob)
            # where s * symp_prob < 1, else 1); the day's counts complete it on the device
            s = np.arange(sim.npts, dtype=np.int32)
            sp = np.ones(len(s))
            inds = 1 > (s * self.symp_prob)
            sp[inds] = self.symp_prob / (1 - s[inds] * self.symp_prob)
            self._swab_base = np.asarray(self.pdf.pdf(s), dtype=np.float64) * sp
            if dev is not None:
                import torch
                self._dev['swab_base'] = torch.from_numpy(self._swab_base).to(dev)
        if self.subtarget is not None and not self.dynamic_subtarget and dev is not None:
            self._dev['sub'] = subtarget_arrays(self.subtarget, sim, dev)

    def prepare_day(self, sim, op):
        """A subtarget given by functions of the sim: evaluated for today, into the op."""
        if self.dynamic_subtarget and self.active(sim.t):
            self._dev['sub'] = subtarget_arrays(self.subtarget, sim, sim.people.device)
            self._set_sub(op)

    def _set_sub(self, o):
        sub = self._dev.get('sub')
        o.sub_mask = sub[0].data_ptr() if sub is not None else None
        o.sub_vals = sub[1].data_ptr() if sub is not None and sub[1] is not None else None
        o.sub_val = sub[2] if sub is not None else 0.0

    def op(self, slot, layer_keys):
        """This intervention as a cvk_diag_op (slot: its index among the sim's interventions)."""
        o = _lib.DiagOp()
        o.kind = _lib.OP_TEST
        o.slot = slot
        o.start_day = self.start_day
        o.end_day = -1 if self.end_day is None else self.end_day
        o.symp_prob, o.asymp_prob = float(self.symp_prob), float(self.asymp_prob)
        o.symp_quar_prob, o.asymp_quar_prob = float(self.symp_quar_prob), float(self.asymp_quar_prob)
        o.sensitivity = float(self.sensitivity)
        o.keep_prob = 1.0 - float(self.loss_prob)
        o.test_delay = float(self.test_delay)
        qp = self.quar_policy
        named = {None: _lib.QP_NONE, 'start': _lib.QP_START, 'end': _lib.QP_END, 'both': _lib.QP_BOTH,
                 'daily': _lib.QP_DAILY}
        if isinstance(qp, str) or qp is None:
            if qp not in named:
                raise ValueError(f'Quarantine policy "{qp}" not recognized')
            o.quar_policy = named[qp]
        elif callable(qp):
            raise NotImplementedError('test_prob: a callable quar_policy is not supported by the device implementation')
        else:
            qd = [int(q) for q in np.atleast_1d(qp)]
            if not 0 < len(qd) <= _lib.MAX_QUAR_DAYS:
                raise ValueError(f'test_prob: between 1 and {_lib.MAX_QUAR_DAYS} quarantine test days are supported')
            o.quar_policy = _lib.QP_DAYS
            o.n_quar_days = len(qd)
            for j, q in enumerate(qd):
                o.quar_days[j] = q
        self._set_sub(o)
        if self._ili_counts is not None:
            o.ili_counts = self._ili_counts.ctypes.data
            o.n_ili_days = len(self._ili_counts)
        if self._swab_base is not None:
            o.n_swab_bins = len(self._swab_base)
            o.swab_base = self._dev['swab_base'].data_ptr() if 'swab_base' in self._dev else None
        return o

    def apply(self, sim):
        raise RuntimeError('test_prob runs inside the sim\'s day loop')


class contact_tracing(Intervention):
    """Trace the contacts of the agents diagnosed today (presumptive: exposed and tested today) in
    each layer with probability trace_probs[layer]; the living ones become known contacts and are
    scheduled to quarantine trace_time[layer] days later (interventions.py:984-1151). Runs on the
    device (cvk_diag_apply); with a capacity, at most int(capacity) of the day's cases, chosen
    uniformly on the device, are traced (the sims here run unscaled: rescale_vec is 1)."""

    def __init__(self, trace_probs=None, trace_time=None, start_day=0, end_day=None, presumptive=False,
                 quar_period=None, capacity=None, **kwargs):
        super().__init__(**kwargs)
        self._kwargs = dict(trace_probs=trace_probs, trace_time=trace_time, start_day=start_day, end_day=end_day,
                            presumptive=presumptive, quar_period=quar_period, capacity=capacity,
                            label=kwargs.get('label'))
        if capacity is not None and not capacity >= 0:
            raise ValueError('contact_tracing: capacity must be >= 0')
        self.capacity = capacity
        self.trace_probs = trace_probs
        self.trace_time = trace_time
        self.start_day = start_day
        self.end_day = end_day
        self.presumptive = presumptive
        self.quar_period = quar_period

    def initialize(self, sim):
        super().initialize(sim)
        self.start_day = _day_or_none(sim, self.start_day)
        self.end_day = _day_or_none(sim, self.end_day)
        self.days = [self.start_day, self.end_day]
        keys = sim.layer_keys()
        if self.trace_probs is None:
            self.trace_probs = 1.0
        if self.trace_time is None:
            self.trace_time = 0.0
        if self.quar_period is None:
            self.quar_period = sim.pars['quar_period']
        if np.isscalar(self.trace_probs):
            self.trace_probs = {k: self.trace_probs for k in keys}
        if np.isscalar(self.trace_time):
            self.trace_time = {k: self.trace_time for k in keys}
        for lk, prob in self.trace_probs.items():
            if lk not in keys:
                raise ValueError(f'contact_tracing: layer "{lk}" not in {keys}')
            if prob != 0 and lk not in self.trace_time:
                raise ValueError(f'contact_tracing: no trace_time for layer "{lk}"')
            if prob != 0 and not float(self.trace_time[lk]) >= 0:
                raise ValueError('contact_tracing: trace_time must be >= 0')

    def max_delay(self):
        """The largest whole-day delay from tracing to the start of quarantine."""
        return max([int(float(self.trace_time[lk])) for lk, p in self.trace_probs.items() if p != 0] + [0])

    def op(self, slot, layer_keys):
        o = _lib.DiagOp()
        o.kind = _lib.OP_TRACE
        o.slot = slot
        o.start_day = self.start_day
        o.end_day = -1 if self.end_day is None else self.end_day
        o.presumptive = 1 if self.presumptive else 0
        o.capacity = -1 if self.capacity is None else max(0, int(self.capacity / 1.0))  # rescale_vec[t] == 1
        for li, lk in enumerate(layer_keys):
            prob = float(self.trace_probs.get(lk, 0.0))
            o.trace_prob[li] = prob
            if prob != 0:
                tt = float(self.trace_time[lk])
                o.trace_time[li] = tt
                o.trace_period[li] = int(self.quar_period - tt)  # schedule_quarantine's int(period)
        return o

    def apply(self, sim):
        raise RuntimeError('contact_tracing runs inside the sim\'s day loop')


DIAG_TYPES = (test_prob, contact_tracing)


class _Replay:
    """What an intervention sees while the schedule is replayed: a day and a copy of the rates."""

    def __init__(self, pars):
        self.pars = dict(pars)
        self.pars['beta_layer'] = dict(pars['beta_layer'])
        self.t = 0


def beta_schedule(sim, rates):
    """Per-day float32 rates [npts][MAX_VARIANTS] and [npts][MAX_LAYERS] after the sim's
    change_beta interventions; ``rates(pars)`` gives (beta_v, beta_layer) for a parameter dict as
    the day loop uses them. None when no intervention changes a rate."""
    ivs = [iv for iv in sim.pars['interventions'] if isinstance(iv, change_beta)]
    if not ivs:
        return None
    view = _Replay(sim.pars)
    bv = np.zeros((sim.npts, _lib.MAX_VARIANTS), np.float32)
    bl = np.zeros((sim.npts, _lib.MAX_LAYERS), np.float32)
    for t in range(sim.npts):
        view.t = t
        for iv in ivs:
            iv(view)
        v, l = rates(view.pars)
        bv[t, :len(v)] = v
        bl[t, :len(l)] = l
    return bv, bl


SUPPORTED = (change_beta, clip_edges, test_prob, contact_tracing)
