"""Device-state snapshots of a sim (covasim/base.py:682-741 Sim.save / load, misc.py:94-147).

The reference pickles the whole Sim object. Here a snapshot is one safetensors file: every device
array of the agents (state flags, dates, durations, immunity), the contact layers as they stand
(resampled and clipped), the day loop's workspace (per-day result tables, infection log, control
words, scheduled quarantines) and the stores of clip_edges interventions, plus the day and the parameters as JSON in the
file's metadata. Randomness is counter-based (Philox keyed by the seed and the day), so the seed
and the day are the whole random state: a loaded sim continues exactly as the saved one would
have. Loading executes nothing from the file (no pickle).
"""

import json

import numpy as np

__all__ = ['save', 'load', 'FORMAT']

FORMAT = 'covasim_amd.sim/1'


def _encode(obj):
    from . import immunity as cvimm
    from . import interventions as cviv
    if isinstance(obj, np.ndarray):
        return {'__ndarray__': obj.tolist(), 'dtype': str(obj.dtype), 'shape': list(obj.shape)}
    if isinstance(obj, np.generic):
        return obj.item()
    if isinstance(obj, dict):
        return {'__dict__': [[_encode(k), _encode(v)] for k, v in obj.items()]}
    if isinstance(obj, (list, tuple)):
        return [_encode(v) for v in obj]
    if isinstance(obj, cvimm.variant):
        return {'__variant__': dict(p=_encode(obj.p), label=obj.label, days=_encode(obj.days),
                                    n_imports=obj.n_imports, rescale=obj.rescale)}
    if isinstance(obj, cviv.change_beta) or isinstance(obj, cviv.clip_edges):
        layers = obj.layers
        if isinstance(obj, cviv.change_beta) and layers == [None]:
            layers = None
        return {'__intervention__': type(obj).__name__, 'days': _encode(obj.days), 'changes': _encode(obj.changes),
                'layers': _encode(layers), 'label': obj.label}
    if isinstance(obj, cviv.DIAG_TYPES):  # rebuilt from the arguments it was made with
        return {'__intervention__': type(obj).__name__, 'kwargs': _encode(obj._kwargs)}
    if obj is None or isinstance(obj, (bool, int, float, str)):
        return obj
    raise TypeError(f'cannot save a parameter of type {type(obj).__name__} in a snapshot')


def _decode(obj):
    from . import immunity as cvimm
    from . import interventions as cviv
    if isinstance(obj, list):
        return [_decode(v) for v in obj]
    if isinstance(obj, dict):
        if '__ndarray__' in obj:
            return np.array(obj['__ndarray__'], dtype=obj['dtype']).reshape(obj['shape'])
        if '__dict__' in obj:
            return {_decode(k): _decode(v) for k, v in obj['__dict__']}
        if '__variant__' in obj:
            d = obj['__variant__']
            return cvimm.variant(dict(_decode(d['p']), label=d['label']), _decode(d['days']),
                                 n_imports=d['n_imports'], rescale=d['rescale'])
        if '__intervention__' in obj:
            if 'kwargs' in obj:
                cls = {'test_prob': cviv.test_prob, 'contact_tracing': cviv.contact_tracing}[obj['__intervention__']]
                return cls(**_decode(obj['kwargs']))
            cls = {'change_beta': cviv.change_beta, 'clip_edges': cviv.clip_edges}[obj['__intervention__']]
            return cls(_decode(obj['days']), _decode(obj['changes']), layers=_decode(obj['layers']),
                       label=obj['label'])
        raise ValueError('unknown object in a snapshot')
    return obj


def save(sim, filename):
    """Write the sim's device state (at its current day) to ``filename``."""
    import torch
    from safetensors.torch import save_file
    from . import interventions as cviv
    if not sim.initialized:
        sim.initialize()
    torch.cuda.synchronize()
    tensors = {}
    for k, v in sim.people._d.items():
        tensors[f'people.{k}'] = v.detach().contiguous().cpu()
    tensors['host.age'] = torch.from_numpy(np.ascontiguousarray(sim.people.age, dtype=np.float32))
    tensors['host.sex'] = torch.from_numpy(np.ascontiguousarray(sim.people.sex, dtype=np.float32))
    keys = sim.layer_keys()
    for li, lk in enumerate(keys):
        for c in ('p1', 'p2', 'beta'):
            tensors[f'layer.{li}.{c}'] = sim.people.contacts[lk].tensor(c).contiguous().cpu()
    for k, v in sim._w.items():
        if isinstance(v, torch.Tensor):
            tensors[f'work.{k}'] = v.detach().contiguous().cpu()
    for i, iv in enumerate(sim.pars['interventions']):
        if isinstance(iv, cviv.clip_edges) and iv.contacts is not None:
            for lk, cols in iv.contacts.items():
                for c, col in zip(('p1', 'p2', 'beta'), cols):
                    tensors[f'iv.{i}.{keys.index(lk)}.{c}'] = col.contiguous().cpu()
    pars = {k: v for k, v in sim.pars.items() if k not in ('variant_map', 'variant_pars', 'n_variants', 'nab_kin')}
    meta = {'format': FORMAT, 't': str(int(sim.t)), 'complete': str(int(bool(sim.complete))),
            'layer_keys': json.dumps(keys), 'households': str(sim.households), 'label': json.dumps(sim.label),
            'log_capacity': str(int(sim._w['log'].shape[0])), 'pars': json.dumps(_encode(pars))}
    save_file(tensors, filename, metadata=meta)
    return filename


def load(filename, device=None):
    """A Sim restored from a snapshot, at the saved day (finalised again if the saved run was complete)."""
    import torch
    from safetensors import safe_open
    from safetensors.torch import load_file
    from . import interventions as cviv
    from .sim import Sim
    with safe_open(filename, 'pt') as f:
        meta = f.metadata()
    if not meta or meta.get('format') != FORMAT:
        raise ValueError(f'{filename} is not a {FORMAT} snapshot')
    T = load_file(filename)
    keys = json.loads(meta['layer_keys'])
    pars = _decode(json.loads(meta['pars']))
    popdict = dict(age=T['host.age'].numpy(), sex=T['host.sex'].numpy(),
                   contacts={lk: {c: T[f'layer.{li}.{c}'] for c in ('p1', 'p2', 'beta')} for li, lk in enumerate(keys)})
    for k in ('symp_prob', 'severe_prob', 'crit_prob', 'death_prob', 'rel_sus', 'rel_trans'):
        popdict[k] = T[f'people.{k}']
    sim = Sim(pars, label=json.loads(meta['label']), popdict=popdict, households=meta['households'], device=device,
              log_capacity=int(meta['log_capacity']))
    sim.initialize()
    dev = sim.people.device
    for k, v in sim.people._d.items():
        v.copy_(T[f'people.{k}'].to(dev))
    for k, v in sim._w.items():
        if isinstance(v, torch.Tensor) and f'work.{k}' in T:
            v.copy_(T[f'work.{k}'].to(dev))
    for i, iv in enumerate(sim.pars['interventions']):
        if isinstance(iv, cviv.clip_edges):
            iv.contacts = {lk: tuple(T[f'iv.{i}.{keys.index(lk)}.{c}'].to(dev) for c in ('p1', 'p2', 'beta'))
                           for lk in iv.layers}
    sim._build_params()
    sim.t = int(meta['t'])
    sim.people.t = sim.t
    if int(meta['complete']):
        sim.complete = True
        sim.finalize()
    torch.cuda.synchronize()
    return sim
