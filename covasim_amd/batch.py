"""Many independent sims advanced together on one GPU (the cvk_batch_* C ABI).

The reference runs independent sims one per process (``MultiSim.run`` ->
``run.py:1477`` ``sc.parallelize`` over ``Sim.run``, ``sim.py:688-761``). ``BatchSim``
takes the same list of ``Sim`` objects and advances all of them with one launch per
phase per day:
- agents of all sims are concatenated;
- each layer's edge lists are concatenated, each sim's run starting on a 4-edge boundary;
- a table of per-sim parameters and offsets drives the kernels.

Each sim's draws and transmission keys depend only on its own seed and its own agent and
edge indices, so every sim's results equal those of running it alone (``Sim.run``).
After ``run()`` each sim's ``results`` are finalised exactly as ``Sim.finalize`` does.

``autotune()`` is the launch-shape autotuner. It times a few days of the batch for each
candidate shape, from the initial state and from the state a third of the way through the
run (restored between trials), and keeps the fastest.
The shapes are the edge-pass workgroup size, the tile depth, the workgroups per CU and
the agents per thread in the agent pass.
"""

import ctypes as C
import time

import numpy as np

from . import _lib
from . import results as cvres
from .interventions import change_beta

__all__ = ['BatchSim']

# launch shapes the autotuner tries: (edge_block, edge_chunks, edge_blocks_per_cu, agents_per_thread)
EDGE_SHAPES = [(1024, 4, 1), (1024, 2, 1), (512, 4, 1), (512, 2, 1), (512, 4, 2), (512, 2, 2)]
AGENT_SHAPES = [1, 2, 4]


class BatchSim:
    """A batch of ``Sim`` objects sharing n_days, layer keys, variants count, waning and dynamic layers."""

    def __init__(self, sims, log_capacity=0, device=None):
        import torch
        sims = list(sims)
        if not sims:
            raise ValueError('BatchSim needs at least one sim')
        if len(sims) > _lib.BATCH_MAX_SIMS:
            raise ValueError(f'At most {_lib.BATCH_MAX_SIMS} sims per batch')
        for sim in sims:
            if not sim.initialized:
                sim.initialize()
            if sim.t != 0 or sim.complete:
                raise ValueError('BatchSim takes sims that have not started running')
            if sim._custom_update:
                raise NotImplementedError('custom Layer.update subclasses are not supported in a batch')
            if any(not isinstance(iv, change_beta) for iv in sim.pars['interventions']):
                raise NotImplementedError('a batch takes change_beta interventions; other interventions run on '
                                          'single sims (multi_run does this by itself)')
        s0 = sims[0]
        lkeys = s0.layer_keys()
        for sim in sims[1:]:
            if sim.npts != s0.npts:
                raise ValueError('all sims of a batch need the same n_days')
            if sim.layer_keys() != lkeys:
                raise ValueError('all sims of a batch need the same layer keys')
            if sim._params.n_variants != s0._params.n_variants or bool(sim.pars['use_waning']) != bool(s0.pars['use_waning']):
                raise ValueError('all sims of a batch need the same number of variants and waning setting')
            if [sim._layers[i].dynamic for i in range(len(lkeys))] != [s0._layers[i].dynamic for i in range(len(lkeys))]:
                raise ValueError('all sims of a batch need the same dynamic layers')
        self.sims = sims
        self.S = len(sims)
        self.npts = s0.npts
        self.layer_keys = lkeys
        self.n_variants = int(s0._params.n_variants)
        self.device = s0.people.device
        ns = np.array([sim.people.n for sim in sims], dtype=np.int64)
        self.agent_offsets = np.zeros(self.S + 1, dtype=np.int64)
        self.agent_offsets[1:] = np.cumsum(ns)
        self.n_total = int(self.agent_offsets[-1])
        # concatenated agent state (copies of each sim's initial device state)
        d = {}
        for k, t0 in s0.people._d.items():
            parts = [sim.people._d[k] for sim in sims]
            d[k] = torch.cat(parts, dim=1 if t0.dim() == 2 else 0).contiguous()
        self._d = d
        # concatenated layers, each sim's run padded to a multiple of 4 edges
        L = len(lkeys)
        self.edge_starts = np.zeros((L, self.S), dtype=np.int64)
        self.edge_counts = np.zeros((L, self.S), dtype=np.int64)
        self._layers = {}
        for li, lk in enumerate(lkeys):
            cols = {c: [] for c in ('p1', 'p2', 'beta')}
            pos = 0
            for si, sim in enumerate(sims):
                lay = sim.people.contacts[lk]
                n = len(lay)
                pad = (-n) % 4
                self.edge_starts[li, si] = pos
                self.edge_counts[li, si] = n
                for c in cols:
                    cols[c].append(lay.tensor(c))
                    if pad:
                        z = torch.zeros(pad, dtype=lay.tensor(c).dtype, device=self.device)
                        cols[c].append(z)
                pos += n + pad
            self._layers[lk] = {c: (torch.cat(v).contiguous() if v else torch.zeros(0, device=self.device))
                                for c, v in cols.items()}
        self.dynamic = np.array([s0._layers[i].dynamic for i in range(L)], dtype=np.int32)
        beta_ones = [int(all(sim._layers[i].beta_ones for sim in sims)) for i in range(L)]
        self._params = (_lib.Params * self.S)(*[sim._params for sim in sims])
        desc = _lib.BatchDesc()
        desc.n_sims, desc.n_layers, desc.n_variants = self.S, L, self.n_variants
        desc.use_waning = 1 if s0.pars['use_waning'] else 0
        desc.n_agents_total = self.n_total
        desc.agent_offsets_host = self.agent_offsets.ctypes.data
        desc.edge_starts_host = self.edge_starts.ctypes.data
        desc.edge_counts_host = self.edge_counts.ctypes.data
        desc.params_host = C.addressof(self._params)
        desc.dynamic_host = self.dynamic.ctypes.data
        desc.people = self.struct()
        for li, lk in enumerate(lkeys):
            desc.p1[li] = self._layers[lk]['p1'].data_ptr()
            desc.p2[li] = self._layers[lk]['p2'].data_ptr()
            desc.beta[li] = self._layers[lk]['beta'].data_ptr()
            desc.beta_ones[li] = beta_ones[li]
        # the sims' source lists of their static layers (built when each sim initialised; local indices)
        self._sources = (_lib.SourcesS * max(1, self.S * L))()
        for si, sim in enumerate(sims):
            for li in range(L):
                ptr = sim._layers[li].sources
                if ptr:
                    self._sources[si * L + li] = ptr.contents
        desc.sources = C.cast(self._sources, C.POINTER(_lib.SourcesS))
        # per-day rates of the sims with change_beta schedules (the others repeat their own rates)
        if any(sim._beta_days is not None for sim in sims):
            bv = np.zeros((self.S, self.npts, _lib.MAX_VARIANTS), np.float32)
            bl = np.zeros((self.S, self.npts, _lib.MAX_LAYERS), np.float32)
            for si, sim in enumerate(sims):
                if sim._beta_days is not None:
                    bv[si], bl[si] = sim._beta_days[0][:self.npts], sim._beta_days[1][:self.npts]
                else:
                    bv[si] = np.asarray(sim._params.beta_v, np.float32)
                    bl[si] = np.asarray(sim._params.beta_layer, np.float32)
            self._rate_tables = (np.ascontiguousarray(bv), np.ascontiguousarray(bl))
            desc.n_beta_days = self.npts
            desc.beta_v_days = self._rate_tables[0].ctypes.data
            desc.beta_layer_days = self._rate_tables[1].ctypes.data
        self._desc = desc
        self.log_capacity = int(log_capacity)
        handle = C.c_void_p()
        _lib.check(_lib.lib().cvk_batch_create(C.byref(desc), self.npts, self.log_capacity, C.byref(handle)),
                   'cvk_batch_create')
        self._handle = handle
        self._initial = None
        # seeds and import events (union of the sims' schedules)
        self.pop_infected = np.array([int(sim.pars['pop_infected']) for sim in sims], dtype=np.int64)
        # one event per (day, variant, occurrence): the single-sim loop selects each import separately
        events = {}
        for si, sim in enumerate(sims):
            days, vars_, nimp = sim._imp_arrays
            seen = {}
            for j in range(len(days)):
                dv = (int(days[j]), int(vars_[j]))
                occ = seen.get(dv, 0)
                seen[dv] = occ + 1
                events.setdefault(dv + (occ,), np.zeros(self.S, dtype=np.int64))[si] = int(nimp[j])
        keys = sorted(events)
        self._ev_days = np.array([k[0] for k in keys], dtype=np.int32)
        self._ev_vars = np.array([k[1] for k in keys], dtype=np.int32)
        self._ev_counts = np.ascontiguousarray(np.stack([events[k] for k in keys]) if keys
                                               else np.zeros((0, self.S), dtype=np.int64))
        imp = _lib.BatchImports()
        imp.n_events = len(keys)
        imp.days = self._ev_days.ctypes.data if len(keys) else None
        imp.variants = self._ev_vars.ctypes.data if len(keys) else None
        imp.counts = self._ev_counts.ctypes.data if len(keys) else None
        self._imports = imp
        self.t = 0
        self.complete = False
        self.tuning = None
        self.reset()

    # -- plumbing --
    def __del__(self):
        h = getattr(self, '_handle', None)
        if h is not None and h.value:
            try:
                _lib.lib().cvk_batch_destroy(h)
            except Exception:
                pass
            self._handle = None

    def struct(self):
        s = _lib.PeopleS()
        s.n = self.n_total
        for k in _lib.PEOPLE_PTR_FIELDS:
            setattr(s, k, self._d[k].data_ptr())
        return s

    def reset(self, stream=None):
        """Clear keys, counters, results and log (the agent state is left as it is)."""
        _lib.check(_lib.lib().cvk_batch_reset(self._handle, _lib.stream_ptr(stream)), 'cvk_batch_reset')

    def snapshot_state(self):
        """Device copies of the agent state and layers (for repeated timed runs)."""
        return ({k: v.clone() for k, v in self._d.items()},
                {lk: {c: v.clone() for c, v in cols.items()} for lk, cols in self._layers.items()})

    def restore_state(self, snap):
        state, layers = snap
        for k, v in state.items():
            self._d[k].copy_(v)
        for lk, cols in layers.items():
            for c, v in cols.items():
                self._layers[lk][c].copy_(v)
        self.reset()
        self.t = 0
        self.complete = False

    def config(self):
        c = _lib.BatchConfig()
        _lib.check(_lib.lib().cvk_batch_get_config(self._handle, C.byref(c)), 'cvk_batch_get_config')
        return c.as_tuple()

    def set_config(self, edge_block, edge_chunks, edge_blocks_per_cu, agents_per_thread):
        c = _lib.BatchConfig(edge_block, edge_chunks, edge_blocks_per_cu, agents_per_thread)
        _lib.check(_lib.lib().cvk_batch_set_config(self._handle, C.byref(c)), 'cvk_batch_set_config')

    # -- running --
    def run(self, use_graph=True, stream=None):
        """Run every sim to the end (days [t, npts)), then finalise each sim's results."""
        import torch
        if self.complete:
            raise RuntimeError('BatchSim already complete (restore_state() to run again)')
        st = stream if stream is not None else torch.cuda.current_stream()
        graph = bool(use_graph) and st.cuda_stream != 0
        _lib.check(_lib.lib().cvk_batch_run_days(self._handle, self.t, self.npts, self.pop_infected.ctypes.data,
                                                 C.byref(self._imports), 1 if graph else 0, C.c_void_p(st.cuda_stream)),
                   'cvk_batch_run_days')
        torch.cuda.synchronize()
        self.t = self.npts
        self.complete = True
        self.finalize()
        return self

    def graph_create(self, stream):
        """An instantiated hipGraph of the whole run (replay with _lib cvk_graph_launch)."""
        exec_ = C.c_void_p()
        _lib.check(_lib.lib().cvk_batch_graph_create(self._handle, 0, self.npts, self.pop_infected.ctypes.data,
                                                     C.byref(self._imports), C.c_void_p(stream.cuda_stream),
                                                     C.byref(exec_)), 'cvk_batch_graph_create')
        return exec_

    def raw_results(self):
        res = np.zeros((self.S, self.npts, _lib.R_COLS), dtype=np.uint32)
        resf = np.zeros((self.S, self.npts, _lib.RF_COLS), dtype=np.float64)
        _lib.check(_lib.lib().cvk_batch_copy_results(self._handle, res.ctypes.data, resf.ctypes.data),
                   'cvk_batch_copy_results')
        return res, resf

    def device_results(self):
        """Finalised rows of every sim on the device (cvk_finalize over the batch's tables):
        [S][F_ROWS][npts] and [S][FV_ROWS][n_variants][npts] float64 tensors."""
        L = _lib.lib()
        rp, rfp = C.c_void_p(), C.c_void_p()
        _lib.check(L.cvk_batch_results(self._handle, C.byref(rp), C.byref(rfp)), 'cvk_batch_results')
        imports = np.stack([sim.import_schedule() for sim in self.sims])
        d = self._d
        return cvres.finalize_tables(rp.value, rfp.value, self.npts, self.n_variants, self.sims[0].pars['use_waning'],
                                     [sim.pars['pop_infected'] for sim in self.sims],
                                     [cvres.initial_period(sim.pars, self.npts) for sim in self.sims],
                                     self.agent_offsets, imports, d['date_infectious'], d['date_recovered'],
                                     d['date_dead'])

    def finalize(self):
        out, out_v = self.device_results()
        out, out_v = out.cpu().numpy(), out_v.cpu().numpy()
        res, _ = self.raw_results()
        for si, sim in enumerate(self.sims):
            sim.results_ready = False
            sim.finalize(rows=(out[si], out_v[si], res[si]))
            sim.complete = True

    def reduce(self, quantiles=None, use_mean=False, bounds=None):
        """MultiSim.reduce (run.py:220-303) over the batch, on the device: for every result row, the
        median and quantiles (default 0.1 / 0.9) or the mean -/+ bounds (default 2) standard
        deviations across sims. Returns results like a Sim's, each Result carrying low and high."""
        out, out_v = self.device_results()
        parts = [cvres.reduce_tables(t, quantiles=quantiles, use_mean=use_mean, bounds=bounds) for t in (out, out_v)]
        (c, lo, hi), (cv, lov, hiv) = [[x.cpu().numpy() for x in p] for p in parts]
        R = cvres.to_results(c, cv)
        for i, name in enumerate(cvres.F_ROW_NAMES):
            R[name].low, R[name].high = lo[i], hi[i]
        for i, name in enumerate(cvres.FV_ROW_NAMES):
            R['variant'][name].low, R['variant'][name].high = lov[i], hiv[i]
        return R

    def infection_log(self):
        """Records (sim, source, target, date, layer, variant); source/target are indices within the sim."""
        count = C.c_int64()
        buf = np.zeros((max(1, self.log_capacity), 4), dtype=np.int32)
        _lib.check(_lib.lib().cvk_batch_copy_log(self._handle, buf.ctypes.data, len(buf), C.byref(count)),
                   'cvk_batch_copy_log')
        names = {_lib.LAYER_SEED: 'seed_infection', _lib.LAYER_IMPORT: 'importation'}
        out = []
        for s, tg, d, lv in buf[:count.value]:
            l, v, si = lv & 0xFF, (lv >> 8) & 0xF, lv >> 12
            out.append(dict(sim=int(si), source=None if s < 0 else int(s), target=int(tg), date=int(d),
                            layer=names.get(l, self.layer_keys[l] if l < len(self.layer_keys) else l),
                            variant=self.sims[si].pars['variant_map'][v]))
        return out

    def get(self, key, sim_index):
        """Host copy of one sim's state array (decoded like People.get)."""
        a, b = int(self.agent_offsets[sim_index]), int(self.agent_offsets[sim_index + 1])
        sim = self.sims[sim_index]
        saved = sim.people._d
        sim.people._d = {k: (v[:, a:b] if v.dim() == 2 else v[a:b]) for k, v in self._d.items()}
        try:
            return sim.people.get(key)
        finally:
            sim.people._d = saved

    # -- autotuner --
    def autotune(self, days=8, repeats=1, verbose=False):
        """Pick the fastest launch shape by timing, per candidate, `days` days from the initial state
        (quiet days) plus `days` days from the state a third of the way through the run (busier days)."""
        import torch
        snap = self.snapshot_state()
        st = torch.cuda.Stream()
        trials = {}
        lib = _lib.lib()
        mid = min(max(self.npts // 3, days), max(self.npts - days, 0))

        def run(t0, t1):
            _lib.check(lib.cvk_batch_run_days(self._handle, t0, t1, self.pop_infected.ctypes.data,
                                              C.byref(self._imports), 0, C.c_void_p(st.cuda_stream)), 'autotune')

        run(0, mid)  # the busier state, with the shape in place now
        st.synchronize()
        snap_mid = self.snapshot_state() if mid > 0 else None

        def timed(state, t0):
            self.restore_state(state)
            torch.cuda.synchronize()
            c0 = time.perf_counter()
            run(t0, min(t0 + days, self.npts))
            st.synchronize()
            return time.perf_counter() - c0

        def trial(shape):
            self.set_config(*shape)
            best = None
            for _ in range(repeats):
                el = timed(snap, 0) + (timed(snap_mid, mid) if snap_mid is not None else 0.0)
                best = el if best is None else min(best, el)
            trials[shape] = best
            if verbose:
                print(f'autotune {shape}: {best * 1e3:.2f} ms', flush=True)
            return best

        apt0 = self.config()[3]
        edge = min(EDGE_SHAPES, key=lambda e: trial(e + (apt0,)))
        apt = min(AGENT_SHAPES, key=lambda a: trial(edge + (a,)) if edge + (a,) not in trials else trials[edge + (a,)])
        best = edge + (apt,)
        self.set_config(*best)
        self.restore_state(snap)
        del snap, snap_mid
        self.tuning = dict(best=best, days=days, from_days=[0, mid], trials_ms={'x'.join(map(str, k)): round(v * 1e3, 3)
                                                              for k, v in trials.items()})
        return best

