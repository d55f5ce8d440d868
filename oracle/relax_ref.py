"""CPU float64 reference of the expected-infection relaxation, differentiated by torch autograd.

TEST INFRASTRUCTURE ONLY: the checker of the GPU relaxation (covasim_amd/csrc/cvk_relax.hip),
never part of the product path. Same inputs as the GPU: per-layer neighbour lists (indptr,
nbr), age bins, rel_trans, the infectiousness table w[bin][lag], the seed fraction and
theta = [beta, beta_l..., sigma_b (10)]. The model, day by day:

    tr_j(t)  = rel_trans_j * sum_{d=1}^{min(t, K)} D_j(t-d) w[bin_j][d]
    L_i(t)   = sum_l sum_{j in N_l(i)} log1p(-min(beta * beta_l * sigma[bin_i] * tr_j, XMAX))
    inf_i(t) = s_i(t) * (1 - exp(L_i(t)));  s_i(t+1) = s_i(t) exp(L_i(t))
    D_i(t)   = inf_i(t) + (f if t == 0 else 0);  s_i(0) = 1 - f;  E_t = sum_i inf_i(t)

This is the mean-field reading of the transmission step (utils.py:88-123: a susceptible
target escapes every infectious neighbour independently) with indicator states replaced by
probabilities. A layer resampled every day (base.py:1849-1876: E_l pairs with uniform endpoints)
enters in expectation: agent i has 2 E_l / N contact slots on average, each with a uniformly drawn
source, and self-pairs do not transmit, so its term is

    L_i(t) += c_l / N * sum_{j != i} log1p(-min(beta * beta_l * sigma[bin_i] * tr_j, XMAX)),  c_l = 2 E_l / N

(the mean over resampled layers of that layer's term, by linearity).
"""

import numpy as np

XMAX = 0.999999


def log_escape(beta, bl, sig, bins, slots, dyn_slots, tr):
    """L_i of one day (torch float64 [n]) from tr: every held layer's neighbour slots (tgt, src), and the
    expected term of every daily-resampled layer (slots[l] None)."""
    import torch
    n = len(bins)
    sig_i = sig[bins]
    L = torch.zeros(n, dtype=torch.float64)
    for li, sl in enumerate(slots):
        if sl is None:
            xb = (beta * bl[li]) * sig  # [10]
            g = torch.log1p(-torch.clamp(xb[:, None] * tr[None, :], max=XMAX))  # [10][n]
            own = g[bins, torch.arange(n)]
            L = L + dyn_slots[li] * (g.sum(dim=1)[bins] - own) / n
            continue
        tgt, src = sl
        x = ((beta * bl[li]) * sig_i[tgt]) * tr[src]
        x = torch.clamp(x, max=XMAX)
        L = L.index_add(0, tgt, torch.log1p(-x))
    return L


def layer_slots(csr, n):
    """(tgt, src) torch index pairs of each held layer's neighbour lists; None for a resampled layer."""
    import torch
    slots = []
    for lay in csr:
        if lay is None:
            slots.append(None)
            continue
        indptr, nbr = lay
        indptr = np.asarray(indptr, dtype=np.int64)
        tgt = torch.as_tensor(np.repeat(np.arange(n), np.diff(indptr)))
        slots.append((tgt, torch.as_tensor(np.asarray(nbr, dtype=np.int64))))
    return slots


def expected_infections(theta, csr, bins, rel_trans, w, n_days, seed_frac, dyn_slots=None):
    """E (torch float64 [n_days]) as a differentiable function of theta (torch float64). csr[l] is
    (indptr, nbr) for a layer held as it is, or None for a layer resampled every day, whose expected
    slots per agent c_l are dyn_slots[l]."""
    import torch
    bins = torch.as_tensor(np.asarray(bins, dtype=np.int64))
    rt = torch.as_tensor(np.asarray(rel_trans, dtype=np.float64))
    w = torch.as_tensor(np.asarray(w, dtype=np.float64))
    n = len(bins)
    n_layers = len(csr)
    beta = theta[0]
    bl = theta[1:1 + n_layers]
    sig = theta[1 + n_layers:]
    slots = layer_slots(csr, n)
    kmax = min(w.shape[1] - 1, n_days)
    s = torch.full((n,), 1.0 - seed_frac, dtype=torch.float64)
    hist = []
    out = []
    for t in range(n_days):
        inf_t = torch.zeros(n, dtype=torch.float64)
        for d in range(1, min(t, kmax) + 1):
            inf_t = inf_t + hist[t - d] * w[bins, d]
        tr = rt * inf_t
        L = log_escape(beta, bl, sig, bins, slots, dyn_slots, tr)
        inf = s * (-torch.expm1(L))
        s = s * torch.exp(L)
        hist.append(inf + (seed_frac if t == 0 else 0.0))
        out.append(inf.sum())
    return torch.stack(out)
