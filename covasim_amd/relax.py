"""Differentiable expected-infection relaxation of the transmission step, for calibration.

``Relaxation(sim)`` takes an initialized ``Sim`` and uses:
- its agents (ages -> prognosis age bins, ``rel_trans``);
- its static contact layers as they stand on day 0, and its daily-resampled layers (``dynam_layer``) in
  expectation: 2 E_l / N contact slots per agent, each with a uniformly drawn source;
- its disease parameters.

``Relaxation.expected(theta)`` returns the expected new infections of each day, E_t, as
a differentiable torch function of theta = [beta, beta_layer for each layer, rel_sus for
the 10 prognosis age bins]. The model and its adjoint are spelled out in
``csrc/cvk_relax.hip``. The forward runs in float32 (precision=32, log-space products
accumulated in float32) or float64 (precision=64). The gradient comes from the
hand-written adjoint kernels, with the state checkpointed every ``checkpoint_every``
days. ``calibrate`` fits theta to an observed daily-incidence curve with Adam on log(theta).

The per-agent relative transmissibility over time since infection,
w_b(d) = E[infectious on lag d x viral load x asymptomatic factor], is tabulated per age
bin from the duration and prognosis model (people.py:513-579, utils.py:39-79). The
table is a Monte Carlo average with a fixed seed.
"""

import ctypes as C

import numpy as np

from . import _lib

__all__ = ['Relaxation', 'infectiousness_table', 'neighbour_lists', 'calibrate', 'N_BINS']

N_BINS = 10


def _sample(rng, d, n):
    """sample() of utils.py:156-236 for the duration distributions (host, float64)."""
    dist, p1, p2 = d['dist'], float(d['par1']), float(d.get('par2', 0.0))
    if dist in ('lognormal', 'lognorm', 'lognormal_int', 'lognorm_int'):
        if p1 <= 0:
            x = np.zeros(n)
        else:
            mu = np.log(p1 ** 2 / np.sqrt(p2 ** 2 + p1 ** 2))
            sigma = np.sqrt(np.log(p2 ** 2 / p1 ** 2 + 1))
            x = rng.lognormal(mu, sigma, n)
        return np.round(x) if dist.endswith('_int') else x
    if dist in ('normal', 'norm'):
        return rng.normal(p1, p2, n)
    if dist == 'normal_pos':
        return np.abs(rng.normal(p1, p2, n))
    if dist == 'normal_int':
        return np.round(np.abs(rng.normal(p1, p2, n)))
    if dist in ('uniform', 'unif'):
        return rng.uniform(p1, p2, n)
    raise NotImplementedError(dist)


def infectiousness_table(pars, n_lags, n_samples=200_000, seed=0):
    """w[b][d] for age bins b and lags d = 0..n_lags-1 days after infection (w[:, 0] = 0).

    An agent infected on day tau is infectious on day tau + d when date_infectious <= tau + d
    < date_recovered (or date_dead), people.py:222-312, with the two-level viral load of
    utils.py:58-78 and asymp_factor while not symptomatic (utils.py:86-88).
    """
    progs = pars['prognoses']
    dur = pars['dur']
    vd = pars['viral_dist']
    ft, lr, cap = float(vd['frac_time']), float(vd['load_ratio']), float(vd['high_cap'])
    vl_hi, vl_lo = lr / (1 + ft * (lr - 1)), 1 / (1 + ft * (lr - 1))
    af = float(pars['asymp_factor'])
    rng = np.random.default_rng(seed)
    w = np.zeros((N_BINS, n_lags))
    lags = np.arange(n_lags, dtype=np.float64)
    m = n_samples
    for b in range(N_BINS):
        p_symp = pars['rel_symp_prob'] * progs['symp_probs'][b]
        p_sev = pars['rel_severe_prob'] * progs['severe_probs'][b]
        p_crit = pars['rel_crit_prob'] * progs['crit_probs'][b]
        p_death = pars['rel_death_prob'] * progs['death_probs'][b]
        e2i = _sample(rng, dur['exp2inf'], m)
        symp = rng.random(m) < p_symp
        sev = symp & (rng.random(m) < p_sev)
        crit = sev & (rng.random(m) < p_crit)
        death = crit & (rng.random(m) < p_death)
        i2s = _sample(rng, dur['inf2sym'], m)
        s2s = _sample(rng, dur['sym2sev'], m)
        s2c = _sample(rng, dur['sev2crit'], m)
        length = np.where(~symp, _sample(rng, dur['asym2rec'], m),
                 np.where(~sev, i2s + _sample(rng, dur['mild2rec'], m),
                 np.where(~crit, i2s + s2s + _sample(rng, dur['sev2rec'], m),
                 np.where(~death, i2s + s2s + s2c + _sample(rng, dur['crit2rec'], m),
                          i2s + s2s + s2c + _sample(rng, dur['crit2die'], m)))))
        switch = np.where(ft * length > cap, cap / np.where(length > 0, length, 1.0), ft)
        sym_from = np.where(symp, e2i + i2s, np.inf)
        for d in range(1, n_lags):
            since = lags[d] - e2i
            on = (since >= 0) & (lags[d] < e2i + length)
            with np.errstate(divide='ignore', invalid='ignore'):
                early = (since / length) < switch
            vl = np.where(early, vl_hi, vl_lo)
            fa = np.where(lags[d] >= sym_from, 1.0, af)
            w[b, d] = float(np.mean(np.where(on, vl * fa, 0.0)))
    return w


def neighbour_lists(p1, p2, n):
    """Both directions of every stored pair, self-loops dropped, sorted by (agent, neighbour):
    (indptr int64 [n+1], nbr int32) as device tensors."""
    import torch
    src = torch.cat([p1, p2]).to(torch.int64)
    tgt = torch.cat([p2, p1]).to(torch.int64)
    keep = src != tgt
    src, tgt = src[keep], tgt[keep]
    key, _ = torch.sort(tgt * n + src)
    tgt_sorted = key // n
    nbr = (key - tgt_sorted * n).to(torch.int32).contiguous()
    counts = torch.bincount(tgt_sorted, minlength=n)
    indptr = torch.zeros(n + 1, dtype=torch.int64, device=p1.device)
    indptr[1:] = torch.cumsum(counts, 0)
    return indptr, nbr


def age_bins(age, pars):
    return (np.digitize(np.asarray(age), pars['prognoses']['age_cutoffs']) - 1).astype(np.uint8)


class Relaxation:
    """Expected-infection relaxation of an initialized Sim (static layers, n_days days)."""

    def __init__(self, sim, n_days=None, checkpoint_every=10, precision=32, w_samples=200_000, w_seed=0):
        import torch
        if not sim.initialized:
            sim.initialize()
        pars = sim.pars
        if int(pars['n_variants']) != 1 or pars['use_waning']:
            raise NotImplementedError('the relaxation covers one variant without waning')
        self.sim = sim
        self.n = sim.people.n
        self.n_days = int(n_days if n_days is not None else pars['n_days'])
        self.layer_keys = sim.layer_keys()
        self.n_layers = len(self.layer_keys)
        self.precision = precision
        dev = sim.people.device
        self.w = infectiousness_table(pars, self.n_days + 1, n_samples=w_samples, seed=w_seed)
        self.bins_np = age_bins(sim.people.age, pars)
        self._bins = torch.from_numpy(self.bins_np).to(dev)
        self._rt = sim.people._d['rel_trans'].clone()
        dyn = pars.get('dynam_layer') or {}
        self.dyn_slots = [2.0 * len(sim.people.contacts[lk]) / self.n if dyn.get(lk, 0) else 0.0
                          for lk in self.layer_keys]
        self._csr = [None if self.dyn_slots[li] else
                     neighbour_lists(sim.people.contacts[lk].tensor('p1'), sim.people.contacts[lk].tensor('p2'), self.n)
                     for li, lk in enumerate(self.layer_keys)]
        self.seed_frac = float(pars['pop_infected']) / self.n
        d = _lib.RelaxDesc()
        d.n_agents, d.n_layers, d.n_days = self.n, self.n_layers, self.n_days
        d.kernel_len = self.w.shape[1]
        d.checkpoint_every = int(checkpoint_every)
        d.precision = int(precision)
        d.seed_frac = self.seed_frac
        for li, lay in enumerate(self._csr):
            if lay is None:
                d.dyn_slots[li] = self.dyn_slots[li]
                continue
            d.indptr[li] = lay[0].data_ptr()
            d.nbr[li] = lay[1].data_ptr()
        d.age_bin = self._bins.data_ptr()
        d.rel_trans = self._rt.data_ptr()
        self._w_host = np.ascontiguousarray(self.w, dtype=np.float64)
        d.w_host = self._w_host.ctypes.data
        self._desc = d
        h = C.c_void_p()
        _lib.check(_lib.lib().cvk_relax_create(C.byref(d), C.byref(h)), 'cvk_relax_create')
        self._handle = h

    def __del__(self):
        h = getattr(self, '_handle', None)
        if h is not None and h.value:
            try:
                _lib.lib().cvk_relax_destroy(h)
            except Exception:
                pass
            self._handle = None

    @property
    def n_params(self):
        return 1 + self.n_layers + N_BINS

    def theta0(self):
        """The sim's own parameters: beta, beta_layer (layer order), prognosis rel_sus (sus_ORs) per age bin."""
        p = self.sim.pars
        return np.array([p['beta']] + [p['beta_layer'][lk] for lk in self.layer_keys]
                        + list(np.asarray(p['prognoses']['sus_ORs'], dtype=np.float64)), dtype=np.float64)

    def csr_host(self):
        """Neighbour lists as host arrays (for the CPU reference); None for a daily-resampled layer."""
        return [None if lay is None else (lay[0].cpu().numpy(), lay[1].cpu().numpy()) for lay in self._csr]

    def forward_np(self, theta):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        if theta.shape != (self.n_params,):
            raise ValueError(f'theta must have {self.n_params} entries')
        out = np.zeros(self.n_days, dtype=np.float64)
        _lib.check(_lib.lib().cvk_relax_forward(self._handle, theta.ctypes.data, out.ctypes.data, _lib.stream_ptr()),
                   'cvk_relax_forward')
        self._last_theta = theta.copy()
        return out

    def backward_np(self, theta, grad_expected):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        ge = np.ascontiguousarray(grad_expected, dtype=np.float64)
        out = np.zeros(self.n_params, dtype=np.float64)
        _lib.check(_lib.lib().cvk_relax_backward(self._handle, theta.ctypes.data, ge.ctypes.data, out.ctypes.data,
                                                 _lib.stream_ptr()), 'cvk_relax_backward')
        return out

    def expected(self, theta):
        """E_t (float64 CPU tensor [n_days]) as a differentiable function of theta (float64 tensor)."""
        return _ExpectedInfections.apply(theta, self)


def _make_function():
    import torch

    class ExpectedInfections(torch.autograd.Function):
        """Forward: cvk_relax_forward; backward: the adjoint kernels (cvk_relax_backward)."""

        @staticmethod
        def forward(ctx, theta, relax):
            th = theta.detach().cpu().to(torch.float64).numpy()
            e = relax.forward_np(th)
            ctx.relax = relax
            ctx.theta = th
            ctx.device, ctx.dtype = theta.device, theta.dtype
            return torch.from_numpy(e).to(device=theta.device, dtype=theta.dtype)

        @staticmethod
        def backward(ctx, grad_e):
            ge = grad_e.detach().cpu().to(torch.float64).numpy()
            relax = ctx.relax
            # the adjoint needs the forward state of this theta (a later forward may have replaced it)
            last = getattr(relax, '_last_theta', None)
            if last is None or not np.array_equal(last, ctx.theta):
                relax.forward_np(ctx.theta)
            g = relax.backward_np(ctx.theta, ge)
            return torch.from_numpy(g).to(device=ctx.device, dtype=ctx.dtype), None

    return ExpectedInfections


class _Lazy:
    fn = None

    @classmethod
    def apply(cls, theta, relax):
        if cls.fn is None:
            cls.fn = _make_function()
        return cls.fn.apply(theta, relax)


_ExpectedInfections = _Lazy


def log_incidence_loss(expected, observed):
    """Mean squared difference of log(1 + daily incidence), days 1..n."""
    import torch
    return torch.mean((torch.log1p(expected[1:]) - torch.log1p(observed[1:])) ** 2)


def calibrate(relax, observed, theta_init, iters=100, lr=0.05, fit_mask=None, verbose=False):
    """Adam on log(theta) against an observed daily-incidence curve. Returns (theta, loss history)."""
    import torch
    obs = torch.as_tensor(np.asarray(observed[:relax.n_days], dtype=np.float64))
    log_theta = torch.tensor(np.log(np.asarray(theta_init, dtype=np.float64)), requires_grad=True)
    mask = torch.ones(relax.n_params, dtype=torch.float64) if fit_mask is None else \
        torch.as_tensor(np.asarray(fit_mask, dtype=np.float64))
    opt = torch.optim.Adam([log_theta], lr=lr)
    history = []
    for it in range(iters):
        opt.zero_grad()
        theta = torch.exp(log_theta)
        loss = log_incidence_loss(relax.expected(theta), obs)
        loss.backward()
        log_theta.grad *= mask
        opt.step()
        history.append(float(loss.item()))
        if verbose and it % 10 == 0:
            print(f'calibrate {it}: loss {history[-1]:.5f}', flush=True)
    return np.exp(log_theta.detach().numpy()), history
