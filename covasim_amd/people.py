"""People: agent state as device structure-of-arrays (covasim/people.py, base.py:1094-1370).

Sixteen boolean states are bits of one uint32 ``flags`` word per agent, and the
variant states (exposed/recovered variant, exposed/infectious by variant) are
packed in a second uint32 ``vstate`` word; dates, durations and per-agent
factors are float32 arrays; immunity rows are float32 [n_variants, n]. Reading
an attribute (``people.susceptible``, ``people.date_infectious`` ...) returns a
host numpy array in the reference's dtype (bool, float32 with NaN = undefined);
``people.set(key, values)`` uploads.
"""

import numpy as np

from . import _lib

__all__ = ['People', 'STATE_KEYS', 'DATE_KEYS', 'DUR_KEYS']

STATE_KEYS = list(_lib.STATE_BITS.keys())  # defaults.py:54-71
DEVICE_DATE_KEYS = ['date_exposed', 'date_infectious', 'date_symptomatic', 'date_severe', 'date_critical',
                    'date_recovered', 'date_dead', 'date_diagnosed', 'date_end_isolation']
DATE_KEYS = [f'date_{s}' for s in STATE_KEYS] + ['date_pos_test', 'date_end_quarantine', 'date_end_isolation']
DUR_KEYS = ['dur_exp2inf', 'dur_inf2sym', 'dur_sym2sev', 'dur_sev2crit', 'dur_disease']
F32_KEYS = DEVICE_DATE_KEYS + DUR_KEYS + ['rel_trans', 'rel_sus', 'symp_prob', 'severe_prob', 'crit_prob',
                                         'death_prob', 'nab', 'peak_nab']
I32_KEYS = ['n_infections', 'n_breakthroughs', 't_nab_event']
IMM_KEYS = ['sus_imm', 'symp_imm', 'sev_imm']
VARIANT_KEYS = ['exposed_variant', 'infectious_variant', 'recovered_variant']
BYVAR_KEYS = ['exposed_by_variant', 'infectious_by_variant']


class People:
    """Agent state on the GPU. Usually created by ``Sim.initialize``."""

    def __init__(self, pars, age, sex=None, contacts=None, device=None, **person_arrays):
        import torch
        self.pars = pars
        self.device = torch.device(device) if device is not None else torch.device('cuda')
        n = int(len(age))
        self.n = n
        self.t = 0
        nv = int(pars['n_variants'])
        self.n_variants = nv
        dev = self.device
        self.age = np.asarray(age, dtype=np.float32)
        self.sex = np.asarray(sex if sex is not None else np.zeros(n), dtype=np.float32)
        self.uid = np.arange(n, dtype=np.int32)
        d = {}
        d['flags'] = torch.full((n,), _lib.SUSCEPTIBLE | _lib.NAIVE | (_lib.VAR_NONE << _lib.INFVAR_SHIFT),
                                dtype=torch.int32, device=dev)
        d['vstate'] = torch.full((n,), _lib.VAR_NONE | (_lib.VAR_NONE << 4), dtype=torch.int32, device=dev)
        for k in DEVICE_DATE_KEYS + DUR_KEYS:
            d[k] = torch.full((n,), float('nan'), dtype=torch.float32, device=dev)
        for k in ['rel_trans', 'rel_sus', 'symp_prob', 'severe_prob', 'crit_prob', 'death_prob', 'nab', 'peak_nab']:
            d[k] = torch.zeros(n, dtype=torch.float32, device=dev)
        for k in I32_KEYS:
            d[k] = torch.zeros(n, dtype=torch.int32, device=dev)
        for k in IMM_KEYS:
            d[k] = torch.zeros((nv, n), dtype=torch.float32, device=dev)
        self._d = d
        for k, v in person_arrays.items():
            self.set(k, v)
        from .base import Contacts
        self.contacts = Contacts()
        if contacts is not None:
            self.add_contacts(contacts)
        self._params = None
        self._layer_keys = []

    # -- structure for the C ABI --
    def struct(self):
        s = _lib.PeopleS()
        s.n = self.n
        for k in _lib.PEOPLE_PTR_FIELDS:
            setattr(s, k, self._d[k].data_ptr())
        return s

    def _layer_index(self, lkey):
        return list(self.contacts.keys()).index(lkey)

    def add_contacts(self, contacts):
        """Add layers from a dict lkey -> dict(p1, p2[, beta]) or Layer (base.py:1371-1415)."""
        from .base import Layer
        for lkey, lay in contacts.items():
            if isinstance(lay, Layer):
                layer = lay
            else:
                layer = Layer(lay, label=lkey, device=self.device)
            layer.validate()
            layer.set_n_agents(self.n)
            self.contacts[lkey] = layer

    # -- attribute access --
    def __len__(self):
        return self.n

    def keys(self):
        return (['uid', 'age', 'sex'] + STATE_KEYS + VARIANT_KEYS + BYVAR_KEYS + IMM_KEYS + ['peak_nab', 'nab',
                't_nab_event'] + DATE_KEYS + DUR_KEYS + ['rel_trans', 'rel_sus', 'symp_prob', 'severe_prob',
                'crit_prob', 'death_prob', 'n_infections', 'n_breakthroughs'])

    def flags(self):
        return self._d['flags'].cpu().numpy().view(np.uint32)

    def vstate(self):
        return self._d['vstate'].cpu().numpy().view(np.uint32)

    def __getitem__(self, key):
        return self.get(key)

    def __getattr__(self, key):
        if key.startswith('_') or key in ('pars', 'device', 'n', 't', 'contacts', 'age', 'sex', 'uid'):
            raise AttributeError(key)
        try:
            return self.get(key)
        except KeyError as E:
            raise AttributeError(key) from E

    def get(self, key):
        """Host copy of a state array in the reference's dtype."""
        if key in ('age', 'sex', 'uid'):
            return getattr(self, key)
        if key in _lib.STATE_BITS:
            return (self.flags() & _lib.STATE_BITS[key]) != 0
        if key in self._d and key not in ('flags', 'vstate'):
            return self._d[key].cpu().numpy()
        if key in DATE_KEYS:  # dates the transmission path never sets
            return np.full(self.n, np.nan, dtype=np.float32)
        if key in VARIANT_KEYS:
            if key == 'infectious_variant':
                code = (self.flags() >> _lib.INFVAR_SHIFT) & 0xF
            else:
                code = (self.vstate() >> (0 if key == 'exposed_variant' else 4)) & 0xF
            out = code.astype(np.float32)
            out[code == _lib.VAR_NONE] = np.nan
            return out
        if key in BYVAR_KEYS:
            shift = 8 if key == 'exposed_by_variant' else 16
            vs = self.vstate()
            return np.stack([((vs >> (shift + v)) & 1).astype(bool) for v in range(self.n_variants)])
        raise KeyError(f'People has no key "{key}"')

    def set(self, key, value):
        """Upload values for a key (states, dates, factors)."""
        import torch
        if key in ('age', 'sex', 'uid'):
            setattr(self, key, np.asarray(value))
            return
        if key in _lib.STATE_BITS:
            bit = _lib.STATE_BITS[key]
            val = torch.from_numpy(np.asarray(value, dtype=bool).copy()).to(self.device)
            f = self._d['flags']
            self._d['flags'] = torch.where(val, f | bit, f & ~bit)
            return
        if key not in self._d or key in ('flags', 'vstate'):
            raise KeyError(f'Cannot set "{key}" on the device state')
        t = self._d[key]
        if isinstance(value, torch.Tensor):
            t.copy_(value.to(device=t.device, dtype=t.dtype).expand(t.shape))
            return
        arr = np.asarray(value, dtype=np.float32 if t.dtype == torch.float32 else np.int32)
        if arr.shape != tuple(t.shape):
            arr = np.broadcast_to(arr, tuple(t.shape))
        t.copy_(torch.from_numpy(np.ascontiguousarray(arr)))

    def count(self, key):
        """Number of agents with a state (base.py:1133-1135)."""
        return int(np.count_nonzero(self.get(key)))

    def count_not(self, key):
        return self.n - self.count(key)

    def count_by_variant(self, key, variant):
        return int(np.count_nonzero(self.get(key)[variant]))

    def true(self, key):
        return np.flatnonzero(self.get(key))

    def false(self, key):
        return np.flatnonzero(~self.get(key))

    def defined(self, key):
        return np.flatnonzero(~np.isnan(self.get(key)))

    def undefined(self, key):
        return np.flatnonzero(np.isnan(self.get(key)))

    def snapshot(self):
        """All device arrays copied to host (for checkpoints and parity tests)."""
        return {k: v.cpu().numpy().copy() for k, v in self._d.items()}

    def restore(self, snap):
        import torch
        for k, v in snap.items():
            self._d[k].copy_(torch.from_numpy(np.ascontiguousarray(v)))
