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
