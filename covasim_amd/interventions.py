"""Interventions that act on the transmission step (covasim/interventions.py).

``change_beta`` (interventions.py:530-586) rescales the overall beta or a layer's beta from given
days on. The GPU day loop takes its effect as a per-day table of the float32 rates the kernels
use (``cvk_imports.beta_v_days`` / ``beta_layer_days``), built once when the sim initialises by
replaying the interventions over the days on a copy of the parameters. A run therefore stays one
captured hipGraph with the rates of each day baked into that day's launches, and the sim's own
parameters are left as they were (the reference restores them after a run, sim.py:803-808).
"""

import numpy as np

from . import _lib

__all__ = ['Intervention', 'change_beta', 'beta_schedule']


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


SUPPORTED = (change_beta,)
