"""Several sims at once (covasim/run.py:36-330 MultiSim, 1326-1519 single_run / multi_run).

The reference runs the copies of a sim (or a given list of sims) one per process with
``sc.parallelize``. Here they share the GPU: ``multi_run`` groups the sims a ``BatchSim`` can advance
together (same days, layer keys, variant count, waning setting and dynamic layers, no interventions
other than change_beta and no custom ``Layer.update``) and runs each group with one launch per phase per day; any other sim
runs on its own (``Sim.run``). Every sim's results are those of running it alone, whichever way it
ran. ``MultiSim.reduce`` / ``mean`` / ``median`` summarise the finished sims on the device
(``cvk_reduce``).
"""

import numpy as np

from . import results as cvres
from .batch import BatchSim
from .interventions import change_beta
from .sim import Sim

__all__ = ['MultiSim', 'single_run', 'multi_run']


def _prepare(sim, ind=0, reseed=True, noise=0.0, noisepar=None, sim_args=None):
    """single_run's set-up (run.py:1358-1394): the label, the seed rand_seed + ind, the
    normally distributed fractional noise on noisepar (default beta), extra parameters."""
    if sim.initialized:
        raise ValueError('multi_run takes sims that have not been initialised (their copies get new seeds)')
    if not sim.label:
        sim.label = f'Sim {ind}'
    if reseed:
        sim['rand_seed'] = sim['rand_seed'] + ind
    if noisepar is None:
        noisepar = 'beta'
    if noisepar not in sim.pars:
        raise KeyError(f'Noise parameter {noisepar} was not found in sim parameters')
    noiseval = noise * np.random.normal()
    factor = 1 + noiseval if noiseval > 0 else 1 / (1 - noiseval)
    if noise:
        sim[noisepar] = sim[noisepar] * factor
    for key, val in (sim_args or {}).items():
        if key not in sim.pars:
            raise KeyError(f'Could not set key {key}: not a valid parameter name')
        sim[key] = val
    return sim


def _copy(sim):
    """A fresh, uninitialised copy of an uninitialised sim (its own parameters; the population
    dictionary, which initialisation only reads, is shared)."""
    import copy
    pars = copy.deepcopy({k: v for k, v in sim.pars.items() if k not in ('interventions', 'variants')})
    pars['interventions'] = copy.deepcopy(sim.pars['interventions'])
    pars['variants'] = copy.deepcopy(sim.pars['variants'])
    popdict = sim.popdict
    if popdict is not None:  # the containers are the copy's own; the arrays are only read
        popdict = dict(popdict)
        popdict['contacts'] = {k: dict(v) for k, v in popdict['contacts'].items()}
    new = Sim(pars=pars, label=sim.label, popdict=popdict, households=sim.households, device=sim.device,
              log_capacity=sim.log_capacity, popgen=sim.popgen, source_lists=sim.source_lists)
    new.pars = pars  # exactly the source's parameters (make_pars has already filled them in)
    return new


def single_run(sim, ind=0, reseed=True, noise=0.0, noisepar=None, sim_args=None, run_args=None, do_run=True):
    """run.py:1326-1403: prepare one sim (seed + ind, noise) and run it."""
    _prepare(sim, ind, reseed, noise, noisepar, sim_args)
    if do_run:
        sim.run(**(run_args or {}))
    return sim


def batchable(sim):
    """Whether a sim can run inside a BatchSim."""
    if not sim.initialized:
        sim.initialize()
    return all(isinstance(iv, change_beta) for iv in sim.pars['interventions']) and not sim._custom_update \
        and sim.t == 0


def batch_key(sim):
    """Sims with equal keys can share a BatchSim."""
    return (sim.npts, tuple(sim.layer_keys()), int(sim._params.n_variants), bool(sim.pars['use_waning']),
            tuple(bool(sim._layers[i].dynamic) for i in range(len(sim.layer_keys()))))


def run_sims(sims, batch=True, min_batch=2):
    """Run every sim to the end: compatible sims together (BatchSim), the others one by one."""
    groups = {}
    alone = []
    for sim in sims:
        if batch and batchable(sim):
            groups.setdefault(batch_key(sim), []).append(sim)
        else:
            alone.append(sim)
    for key, group in groups.items():
        if len(group) < min_batch:
            alone.extend(group)
            continue
        for start in range(0, len(group), 4096):  # BatchSim's limit per batch
            BatchSim(group[start:start + 4096]).run()
    for sim in alone:
        sim.run()
    return sims


def multi_run(sim, n_runs=4, reseed=None, noise=0.0, noisepar=None, iterpars=None, sim_args=None, run_args=None,
              do_run=True, batch=True, **kwargs):
    """run.py:1406-1519: copies of one sim with seeds rand_seed + i (reseed, default for one sim) or a
    given list of sims (reseed default off), with optional noise and per-run parameters (iterpars: a
    dict of equal-length lists), all run on this GPU -- together where a BatchSim can take them."""
    sim_args = dict(sim_args or {}, **kwargs)
    iterpars = dict(iterpars or {})
    if iterpars:
        lens = {len(v) for v in iterpars.values()}
        if len(lens) != 1:
            raise ValueError('Each entry in iterpars must have the same length')
        n_runs = lens.pop()
    if isinstance(sim, Sim):
        if reseed is None:
            reseed = True
        sims = [_copy(sim) for _ in range(n_runs)]
    elif isinstance(sim, list):
        if reseed is None:
            reseed = False
        sims = list(sim)
        if iterpars and len(sims) != n_runs:
            raise ValueError('iterpars must have one entry per sim')
    else:
        raise TypeError(f'Must be Sim object or list, not {type(sim)}')
    for i, s in enumerate(sims):
        extra = dict(sim_args, **{k: v[i] for k, v in iterpars.items()})
        _prepare(s, i, reseed, noise, noisepar, extra)
    if do_run:
        if run_args:
            for s in sims:
                s.run(**run_args)
        else:
            run_sims(sims, batch=batch)
    return sims


class MultiSim:
    """run.py:36-330 (running and reducing): copies of a base sim (n_runs, noise, reseed) or a list of
    sims; ``run()`` runs them all (multi_run), ``reduce()`` / ``mean()`` / ``median()`` put the
    median and quantiles (or mean and standard-deviation bounds) of every result row in
    ``results``, each Result carrying ``low`` and ``high``."""

    def __init__(self, sims=None, base_sim=None, label=None, n_runs=4, noise=0.0, noisepar=None, **kwargs):
        if base_sim is None:
            if isinstance(sims, Sim):
                base_sim, sims = sims, None
            elif isinstance(sims, list) and sims:
                base_sim = sims[0]
            else:
                raise ValueError('MultiSim needs a sim or a list of sims')
        self.base_sim = base_sim
        self.sims = sims
        self.label = label if label is not None else base_sim.label
        self.run_args = dict(n_runs=n_runs, noise=noise, noisepar=noisepar, **kwargs)
        self.results = None
        self.which = None

    def __len__(self):
        return len(self.sims) if self.sims is not None else 0

    def run(self, reduce=False, **kwargs):
        args = dict(self.run_args, **kwargs)
        if self.sims is None:
            self.sims = multi_run(self.base_sim, **args)
        else:
            for s, sim in enumerate(self.sims):
                if sim.label is None:
                    sim.label = f'Sim {s}'
            args.pop('n_runs', None)
            self.sims = multi_run(self.sims, **args)
        if reduce:
            self.reduce()
        return self

    def _tables(self):
        import torch
        if not self.sims or not all(s.results_ready for s in self.sims):
            raise RuntimeError('MultiSim: run the sims before reducing them')
        dev = self.sims[0].people.device if self.sims[0].people is not None else 'cuda'
        rows = np.stack([[s.results[k].values for k in cvres.F_ROW_NAMES] for s in self.sims]).astype(np.float64)
        vrows = np.stack([[s.results['variant'][k].values for k in cvres.FV_ROW_NAMES] for s in self.sims])
        return (torch.as_tensor(rows, device=dev), torch.as_tensor(vrows.astype(np.float64), device=dev))

    def reduce(self, quantiles=None, use_mean=False, bounds=None, output=False):
        """run.py:220-303 on the device (cvk_reduce)."""
        rows, vrows = self._tables()
        parts = [cvres.reduce_tables(t, quantiles=quantiles, use_mean=use_mean, bounds=bounds) for t in (rows, vrows)]
        (c, lo, hi), (cv, lov, hiv) = [[x.cpu().numpy() for x in p] for p in parts]
        R = cvres.to_results(c, cv)
        for i, name in enumerate(cvres.F_ROW_NAMES):
            R[name].low, R[name].high = lo[i], hi[i]
        for i, name in enumerate(cvres.FV_ROW_NAMES):
            R['variant'][name].low, R['variant'][name].high = lov[i], hiv[i]
        self.results = R
        self.which = 'reduced'
        return R if output else None

    def mean(self, bounds=None, **kwargs):
        return self.reduce(use_mean=True, bounds=bounds, **kwargs)

    def median(self, quantiles=None, **kwargs):
        return self.reduce(use_mean=False, quantiles=quantiles, **kwargs)
