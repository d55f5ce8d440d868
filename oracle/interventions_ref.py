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


def chosen_edges(seed, t, slot, n, k):
    """Mask of the k edges of n with the smallest keys (philox(seed; e lo, t, 0x80000000 | slot,
    e hi).x << 32) | e: a uniform choice without replacement, as cvu.choose in clip_edges."""
    import numpy as np
    from . import philox as ph
    mask = np.zeros(n, dtype=bool)
    if k <= 0:
        return mask
    e = np.arange(n, dtype=np.uint64)
    k0, k1 = ph.seed_key(seed)
    w = ph.philox4x32(e & np.uint64(0xFFFFFFFF), t, 0x80000000 | slot, e >> np.uint64(32), k0, k1)[0]
    keys = (np.asarray(w, dtype=np.uint64) << np.uint64(32)) | (e & np.uint64(0xFFFFFFFF))
    kth = np.partition(keys, k - 1)[k - 1]
    mask[keys <= kth] = True
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
