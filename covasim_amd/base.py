"""Containers: Result, Layer and Contacts (covasim/base.py:117-192, 1509-1876).

A ``Layer`` keeps its int32 ``p1``/``p2`` and float32 ``beta`` columns as device
tensors; item access returns host numpy copies and item assignment uploads (with
the dtype conversion and length validation of base.py:1716-1736, plus a range
check of the endpoints, since the kernels index agent arrays with them).
"""

import numpy as np

__all__ = ['Result', 'Layer', 'Contacts']


class Result:
    """One result time series (base.py:117-192)."""

    def __init__(self, name=None, npts=None, scale=True, n_variants=0, values=None):
        self.name = name
        self.scale = scale
        if values is None:
            shape = (n_variants, npts) if n_variants > 0 else npts
            values = np.zeros(shape, dtype=np.float64)
        self.values = values
        self.low = None
        self.high = None

    def __getitem__(self, key):
        return self.values[key]

    def __setitem__(self, key, value):
        self.values[key] = value

    def __len__(self):
        return len(self.values)

    def __repr__(self):
        return f'Result({self.name!r}, {self.values!r})'

    @property
    def npts(self):
        return self.values.shape[-1]


class Layer:
    """A single layer of undirected contact edges (base.py:1610-1876)."""

    meta = {'p1': np.int32, 'p2': np.int32, 'beta': np.float32}
    basekey = 'p1'

    def __init__(self, *args, label=None, device=None, **kwargs):
        import torch
        self.label = label
        self.device = torch.device(device) if device is not None else torch.device('cuda')
        data = {}
        for a in args:
            data.update(a)
        data.update(kwargs)
        self._d = {k: torch.empty(0, dtype=torch.int32 if k != 'beta' else torch.float32, device=self.device)
                   for k in self.meta}
        self.n_agents = None
        for k, v in data.items():
            self[k] = v
        if 'beta' not in data:
            self['beta'] = np.ones(len(self), dtype=np.float32)

    def _columns(self):
        """The device columns; created empty when a subclass's __init__ filled the layer item by item
        without calling Layer.__init__ (as examples/t11_dynamic_layers.py:12-17 does)."""
        d = self.__dict__.get('_d')
        if d is None:
            import torch
            self.__dict__.setdefault('label', None)
            self.__dict__.setdefault('device', torch.device('cuda'))
            self.__dict__.setdefault('n_agents', None)
            d = {k: torch.empty(0, dtype=torch.int32 if k != 'beta' else torch.float32, device=self.device)
                 for k in self.meta}
            self.__dict__['_d'] = d
        return d

    # -- dict-like interface --
    def keys(self):
        return list(self.meta.keys())

    def items(self):
        return [(k, self[k]) for k in self.meta]

    def __len__(self):
        return int(self._columns()['p1'].numel())

    def __contains__(self, item):
        p1, p2 = self['p1'], self['p2']
        return bool((item == p1).any() or (item == p2).any())

    def __getitem__(self, key):
        d = self._columns()
        if key not in d:
            raise KeyError(key)
        return d[key].cpu().numpy()

    def __setitem__(self, key, value):
        import torch
        self._columns()
        if key not in self.meta:
            raise KeyError(f'Layer keys are {list(self.meta)}, not "{key}"')
        if isinstance(value, torch.Tensor):
            t = value.to(device=self.device, dtype=torch.int32 if key != 'beta' else torch.float32).contiguous()
        else:
            t = torch.from_numpy(np.ascontiguousarray(np.asarray(value, dtype=self.meta[key]))).to(self.device)
        if key != 'beta' and t.numel() and self.n_agents is not None:
            lo, hi = int(t.min()), int(t.max())
            if lo < 0 or hi >= self.n_agents:
                raise ValueError(f'Layer "{self.label}" column {key} has indices outside [0, {self.n_agents})')
        self._d[key] = t

    def tensor(self, key):
        """The device tensor backing a column."""
        return self._columns()[key]

    def beta_all_ones(self):
        """True when every edge weight is exactly 1.0 (checked on the device, cached per column tensor)."""
        t = self._columns()['beta']
        key = (t.data_ptr(), t.numel(), t._version)
        if getattr(self, '_ones_key', None) != key:
            self._ones = bool((t == 1.0).all()) if t.numel() else True
            self._ones_key = key
        return self._ones

    def validate(self):
        """Equal lengths for every column (base.py:1716-1736)."""
        n = len(self)
        for k in self.meta:
            if self._columns()[k].numel() != n:
                raise TypeError(f'Expecting length {n} for layer key "{k}"; got {self._columns()[k].numel()}')

    def set_n_agents(self, n):
        self.n_agents = int(n)
        for k in ('p1', 'p2'):
            t = self._columns()[k]
            if t.numel() and (int(t.min()) < 0 or int(t.max()) >= n):
                raise ValueError(f'Layer "{self.label}" column {k} has indices outside [0, {n})')

    def pop_inds(self, inds):
        """Remove edges by index and return them as a dict (base.py:1739-1751)."""
        out = {}
        for k in self.meta:
            arr = self[k]
            out[k] = arr[inds]
            self[k] = np.delete(arr, inds)
        return out

    def append(self, contacts):
        """Append edges (base.py:1754-1768)."""
        for k in self.meta:
            self[k] = np.concatenate([self[k], np.asarray(contacts[k], dtype=self.meta[k])])

    def members(self):
        return np.unique(np.concatenate([self['p1'], self['p2']]))

    def find_contacts(self, inds, as_array=True):
        """Contacts of the given people in either direction (base.py:1790-1833)."""
        inds = np.asarray(inds)
        p1, p2 = self['p1'], self['p2']
        out = set(p2[np.isin(p1, inds)].tolist()) | set(p1[np.isin(p2, inds)].tolist())
        if as_array:
            return np.sort(np.fromiter(out, dtype=np.int32, count=len(out)))
        return out

    def update(self, people, frac=1.0):
        """Regenerate the layer's edges (base.py:1849-1876) on the GPU: round(n_edges * frac) edges,
        a uniform subset (all of them for frac = 1), get new endpoints drawn uniformly with
        replacement and weight 1."""
        import torch
        from . import _lib
        L = _lib.lib()
        n = len(self)
        k = int(np.round(n * frac))  # base.py:1868
        if not 0 <= k <= n:
            raise ValueError(f'frac must be in [0, 1], not {frac}')
        s = _lib.LayerS(_lib.ptr(self._columns()['p1']).value, _lib.ptr(self._columns()['p2']).value,
                        _lib.ptr(self._columns()['beta']).value, n, 1, 0)
        li, t = int(people._layer_index(self.label)), int(people.t)
        if k == n:
            _lib.check(L.cvk_resample_layer(s, li, people._params, t, _lib.stream_ptr()), 'Layer.update')
        elif k > 0:
            nb = int(L.cvk_resample_frac_work_bytes())
            work = torch.empty(nb // 4, dtype=torch.int32, device=self.device)
            _lib.check(L.cvk_resample_layer_frac(s, li, people._params, t, k, _lib.ptr(work), nb, _lib.stream_ptr()),
                       'Layer.update')

    def __repr__(self):
        return f'Layer("{self.label}", n_edges={len(self)})'


class Contacts(dict):
    """Ordered dict of layers (base.py:1509-1607)."""

    def __setitem__(self, lkey, layer):
        if isinstance(layer, Layer):  # a layer put in place after the sim was built (examples/t11_dynamic_layers.py:43)
            layer._columns()
            if layer.label is None:
                layer.label = lkey
            old = self.get(lkey)
            if old is not None and layer.n_agents is None:
                layer.n_agents = old.n_agents
        super().__setitem__(lkey, layer)

    def __repr__(self):
        return 'Contacts(' + ', '.join(f'{k}: {len(v)}' for k, v in self.items()) + ')'

    def add_layer(self, **kwargs):
        for lkey, layer in kwargs.items():
            if not isinstance(layer, Layer):
                layer = Layer(layer, label=lkey)
            layer.validate()
            self[lkey] = layer

    def pop_layer(self, *args):
        for lkey in args:
            self.pop(lkey)
