"""Drop-in mirrors of the reference's numba kernels (covasim/utils.py:39-128), on the GPU.

Each function accepts numpy arrays (copied to the GPU and the result copied
back) or CUDA torch tensors (results stay on the device), checks dtypes like
numba's typed dispatch (a mismatch raises TypeError), and calls the matching
C-ABI entry point. There is no CPU fallback.
"""

import ctypes as C

import numpy as np

from . import _lib

__all__ = ['compute_viral_load', 'compute_trans_sus', 'compute_infections', 'set_seed', 'get_seed', 'choose',
           'choose_r', 'default_int', 'default_float', 'get_pdf']

default_int = np.int32    # defaults.py:26-30 (precision 32)
default_float = np.float32

_SEED = [1]


def set_seed(seed=None):
    """Seed of the counter-based streams used by compute_infections (utils.py:271-298)."""
    _SEED[0] = int(seed) if seed is not None else int(np.random.SeedSequence().entropy % (1 << 63))


def get_seed():
    return _SEED[0]


def _dev(x, dtype, name):
    import torch
    want = {np.float32: torch.float32, np.int32: torch.int32, np.bool_: torch.bool, np.float64: torch.float64}[dtype]
    if isinstance(x, torch.Tensor):
        if x.dtype != want:
            raise TypeError(f'No matching definition for argument {name}: expected {want}, got {x.dtype}')
        if x.device.type != 'cuda':
            x = x.cuda()
        return x.contiguous(), True
    a = np.asarray(x)
    if a.dtype != np.dtype(dtype):
        raise TypeError(f'No matching definition for argument {name}: expected {np.dtype(dtype)}, got {a.dtype}')
    return torch.from_numpy(np.ascontiguousarray(a)).cuda(), False


def _out(t, was_tensor):
    return t if was_tensor else t.cpu().numpy()


def compute_viral_load(t, time_start, time_recovered, time_dead, frac_time, load_ratio, high_cap):
    """Viral load of every agent on day t (utils.py:39-79)."""
    import torch
    ts, is_t = _dev(time_start, np.float32, 'time_start')
    tr, _ = _dev(time_recovered, np.float32, 'time_recovered')
    td, _ = _dev(time_dead, np.float32, 'time_dead')
    n = ts.numel()
    if tr.numel() != n or td.numel() != n:
        raise ValueError('time arrays must have equal lengths')
    out = torch.empty(n, dtype=torch.float32, device=ts.device)
    _lib.check(_lib.lib().cvk_viral_load(int(t), _lib.ptr(ts), _lib.ptr(tr), _lib.ptr(td), n,
                                         float(np.float32(frac_time)), float(np.float32(load_ratio)),
                                         float(np.float32(high_cap)), _lib.ptr(out), _lib.stream_ptr()),
               'compute_viral_load')
    return _out(out, is_t)


def compute_trans_sus(rel_trans, rel_sus, inf, sus, beta_layer, viral_load, symp, iso, quar, asymp_factor,
                      iso_factor, quar_factor, immunity_factors):
    """Relative transmissibility and susceptibility (utils.py:82-90)."""
    import torch
    rt, is_t = _dev(rel_trans, np.float32, 'rel_trans')
    rs, _ = _dev(rel_sus, np.float32, 'rel_sus')
    vl, _ = _dev(viral_load, np.float32, 'viral_load')
    imf, _ = _dev(immunity_factors, np.float32, 'immunity_factors')
    bools = []
    for name, b in [('inf', inf), ('sus', sus), ('symp', symp), ('iso', iso), ('quar', quar)]:
        bt, _ = _dev(b, np.bool_, name)
        bools.append(bt.to(torch.uint8))
    n = rt.numel()
    for x in [rs, vl, imf] + bools:
        if x.numel() != n:
            raise ValueError('all per-agent arrays must have the same length')
    rt_out = torch.empty(n, dtype=torch.float32, device=rt.device)
    rs_out = torch.empty(n, dtype=torch.float32, device=rt.device)
    f = lambda x: float(np.float32(x))
    _lib.check(_lib.lib().cvk_trans_sus(_lib.ptr(rt), _lib.ptr(rs), _lib.ptr(bools[0]), _lib.ptr(bools[1]),
                                        f(beta_layer), _lib.ptr(vl), _lib.ptr(bools[2]), _lib.ptr(bools[3]),
                                        _lib.ptr(bools[4]), f(asymp_factor), f(iso_factor), f(quar_factor),
                                        _lib.ptr(imf), n, _lib.ptr(rt_out), _lib.ptr(rs_out), _lib.stream_ptr()),
               'compute_trans_sus')
    return _out(rt_out, is_t), _out(rs_out, is_t)


def compute_infections(beta, p1, p2, layer_betas, rel_trans, rel_sus, legacy=False, uniforms=None, t=0,
                       variant=0, layer=0, return_probs=False):
    """Who infects whom (utils.py:93-128): (sources, targets) in direction-major, edge-ascending order.

    ``uniforms`` (float64 [2, n_edges]) supplies the per-(direction, edge) uniform
    draws; without it, counter-based draws keyed by (seed, t, variant, layer, edge)
    are used.
    """
    import torch
    a1, is_t = _dev(p1, np.int32, 'p1')
    a2, _ = _dev(p2, np.int32, 'p2')
    lb, _ = _dev(layer_betas, np.float32, 'layer_betas')
    rt, _ = _dev(rel_trans, np.float32, 'rel_trans')
    rs, _ = _dev(rel_sus, np.float32, 'rel_sus')
    E = a1.numel()
    if a2.numel() != E or lb.numel() != E:
        raise ValueError('p1, p2 and layer_betas must have equal lengths')
    n = rt.numel()
    if rs.numel() != n:
        raise ValueError('rel_trans and rel_sus must have equal lengths')
    if E and (int(torch.minimum(a1.min(), a2.min())) < 0 or int(torch.maximum(a1.max(), a2.max())) >= n):
        raise IndexError('contact indices outside [0, len(rel_trans))')
    u = None
    if uniforms is not None:
        u, _ = _dev(uniforms, np.float64, 'uniforms')
        if u.numel() != 2 * E:
            raise ValueError('uniforms must have shape (2, n_edges)')
    cap = (E if legacy else 2 * E)
    src = torch.empty(max(cap, 1), dtype=torch.int32, device=a1.device)
    tgt = torch.empty(max(cap, 1), dtype=torch.int32, device=a1.device)
    prob = torch.empty(max(cap, 1), dtype=torch.float32, device=a1.device)
    count = C.c_int64(0)
    _lib.check(_lib.lib().cvk_infections(float(np.float32(beta)), _lib.ptr(a1), _lib.ptr(a2), _lib.ptr(lb), E,
                                         _lib.ptr(rt), _lib.ptr(rs), n, _lib.ptr(u) if u is not None else None,
                                         get_seed() & 0xFFFFFFFFFFFFFFFF, int(t), int(variant), int(layer),
                                         1 if legacy else 0, _lib.ptr(src), _lib.ptr(tgt), _lib.ptr(prob), cap,
                                         C.byref(count), _lib.stream_ptr()), 'compute_infections')
    k = count.value
    out = (_out(src[:k], is_t), _out(tgt[:k], is_t))
    if return_probs:
        out = out + (_out(prob[:k], is_t),)
    return out


def choose(max_n, n):
    """n distinct integers from [0, max_n), uniformly (utils.py:430-443), on the host: for Layer
    subclasses that build their own contacts (examples/t11_dynamic_layers.py:19-25)."""
    return np.random.choice(max_n, n, replace=False)


def choose_r(max_n, n):
    """n integers from [0, max_n) with replacement (utils.py:446-459), on the host."""
    return np.random.choice(max_n, n, replace=True)


def get_pdf(dist=None, par1=None, par2=None):
    """Frozen scipy distribution of the delay from symptom onset to swab (utils.py:240-268): 'uniform'
    on [par1, par1 + par2], or 'lognormal' with mean par1 and variance par2, shifted by -0.5 days;
    None for no distribution."""
    import scipy.stats as sps
    if dist in ('None', 'none', None):
        return None
    if dist == 'uniform':
        return sps.uniform(loc=par1, scale=par2)
    if dist == 'lognormal':
        mean = np.log(par1 ** 2 / np.sqrt(par2 + par1 ** 2))  # the underlying normal's mean and sigma
        sigma = np.sqrt(np.log(par2 / par1 ** 2 + 1))
        return sps.lognorm(sigma, loc=-0.5, scale=np.exp(mean))
    raise NotImplementedError(f'The selected distribution "{dist}" is not implemented; choices are: none, uniform, '
                              f'lognormal')
