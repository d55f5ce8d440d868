"""Writes tests/golden/kernel_vectors.json: small known-answer cases of the transmission kernels.

Every expected value comes from the closed form or a plain loop written here, not from the
oracle, so the fixtures check the oracle (CPU tests) and the device mirrors (GPU tests):
  viral_load   compute_viral_load (covasim/utils.py:39-79) on hand-picked dates: two load values,
               the high_cap rule, NaN dates, a death date overriding recovery, zero-length infections
  trans_sus    compute_trans_sus (utils.py:82-90) on 8 agents under each factor on its own
  infections   compute_infections (utils.py:93-128): 12 agents, 30 edges (a self-loop, a duplicate
               edge, p > 1, p = 0) with the uniforms stored; expected pairs from a double loop
  transitions  update_states_pre (people.py:164-312) of 6 agents with fixed dates over 10 days:
               expected flags per day from the date rule "X when t >= date_X"
Float32 values are stored as their bit patterns (uint32).

Run: python tests/golden/make_kernel_vectors.py
"""

import json
import os

import numpy as np

F32 = np.float32
HERE = os.path.dirname(os.path.abspath(__file__))


def bits(x):
    return [int(v) for v in np.asarray(x, dtype=F32).view(np.uint32)]


def viral_load_cases():
    ft, lr, hc = F32(0.3), F32(2.0), F32(4.0)
    hi = (lr * F32(1) + F32(1) * F32(0)) / (F32(1) + ft * (lr - F32(1)))
    lo = (lr * F32(0) + F32(1) * F32(1)) / (F32(1) + ft * (lr - F32(1)))
    nan = float('nan')
    # (date_infectious, date_recovered, date_dead, t, expected)
    rows = [(nan, nan, nan, 5, lo), (0, 10, nan, 0, hi), (0, 10, nan, 2, hi), (0, 10, nan, 3, lo),
            (0, 20, nan, 3, hi), (0, 20, nan, 4, lo), (5, 50, 12, 6, hi), (5, 50, 12, 8, lo),
            (3, 3, nan, 3, lo), (2, 40, nan, 5, hi), (2, 40, nan, 6, lo)]
    return dict(frac_time=0.3, load_ratio=2.0, high_cap=4.0,
                cases=[dict(date_infectious=bits([s])[0], date_recovered=bits([r])[0], date_dead=bits([d])[0],
                            t=t, expected=bits([w])[0]) for s, r, d, t, w in rows])


def trans_sus_cases():
    n = 8
    rt = np.array([2.5, 0.0, 1.0, 0.3, 12.0, 1.0, 0.01, 4.0], F32)
    rs = np.array([0.67, 1.0, 0.34, 1.47, 1.0, 0.0, 1.24, 1.0], F32)
    vl = np.array([F32(2) / F32(1.3), F32(1) / F32(1.3)] * 4, F32)
    imm = np.array([0, 0.25, 0, 0.5, 0, 0, 1.0, 0.1], F32)
    cases = []
    for name, inf, sus, symp, iso, quar, bl, af, isf, qf in [
            ('plain', 1, 1, 1, 0, 0, 1.0, 1.0, 0.2, 0.3), ('quarantine', 1, 1, 1, 0, 1, 1.0, 1.0, 0.2, 0.3),
            ('isolation', 1, 1, 1, 1, 0, 1.0, 1.0, 0.2, 0.3), ('asymptomatic', 1, 1, 0, 0, 0, 0.6, 0.5, 0.2, 0.3),
            ('not_infectious', 0, 1, 1, 0, 0, 3.0, 1.0, 0.2, 0.3), ('beta_layer_zero', 1, 1, 1, 0, 0, 0.0, 1.0, 0.2, 0.3)]:
        I = np.full(n, bool(inf))
        S = np.full(n, bool(sus))
        Y = np.full(n, bool(symp))
        Iso = np.full(n, bool(iso))
        Q = np.full(n, bool(quar))
        f_asymp = Y + ~Y * F32(af)           # utils.py:86-88, left to right in float32
        f_iso = ~Iso + Iso * F32(isf)
        f_quar = ~Q + Q * F32(qf)
        exp_rt = rt * I * f_quar * f_asymp * f_iso * F32(bl) * vl
        exp_rs = rs * S * f_quar * (F32(1) - imm)
        cases.append(dict(name=name, inf=inf, sus=sus, symp=symp, iso=iso, quar=quar, beta_layer=bl,
                          asymp_factor=af, iso_factor=isf, quar_factor=qf,
                          rel_trans_eff=bits(exp_rt), rel_sus_eff=bits(exp_rs)))
    return dict(rel_trans=bits(rt), rel_sus=bits(rs), viral_load=bits(vl), sus_imm=bits(imm), cases=cases)


def infections_case():
    rng = np.random.default_rng(20240611)
    n, E = 12, 30
    p1 = rng.integers(0, n, E).astype(np.int32)
    p2 = rng.integers(0, n, E).astype(np.int32)
    p1[0] = p2[0] = 4          # self-loop
    p1[1], p2[1] = p1[2], p2[2]  # duplicate edge
    lb = np.ones(E, F32)
    lb[5] = 0.0                # p = 0
    lb[6] = 2.5
    rt = np.where(rng.random(n) < 0.5, rng.random(n) * 2, 0).astype(F32)
    rt[3] = 1e6                # p > 1
    rs = rng.random(n).astype(F32)
    rs[7] = 0.0
    U = rng.random((2, E))
    beta = 0.3
    out = []
    probs = []
    for d, (src, tgt) in enumerate([(p1, p2), (p2, p1)]):  # utils.py:114-127: dir-major, edge order
        for e in range(E):
            if float(rt[src[e]]) == 0:
                continue
            p = F32(F32(F32(F32(beta) * lb[e]) * rt[src[e]]) * rs[tgt[e]])
            if p == 0:
                continue
            probs.append(int(np.asarray([p], F32).view(np.uint32)[0]))
            if U[d][e] < float(p):
                out.append([int(src[e]), int(tgt[e])])
    return dict(beta=beta, p1=p1.tolist(), p2=p2.tolist(), layer_betas=bits(lb), rel_trans=bits(rt),
                rel_sus=bits(rs), uniforms=U.tolist(), pairs=out, probs_nonzero_in_order=probs)


def transitions_case():
    nan = float('nan')
    # agents exposed on day 0: (date_infectious, date_symptomatic, date_severe, date_critical,
    # date_recovered, date_dead)
    dates = [(2, nan, nan, nan, 6, nan), (1, 3, nan, nan, 8, nan), (1, 2, 4, nan, 9, nan),
             (2, 4, 5, 7, nan, 9), (3, 5, nan, nan, 10, nan), (0, nan, nan, nan, 3, nan)]
    flags = []
    for t in range(10):
        day = []
        for (di, ds, dv, dc, dr, dd) in dates:
            dead = dd == dd and t >= dd
            rec = dr == dr and t >= dr and not dead
            over = dead or rec
            day.append(dict(infectious=bool(t >= di and not over), symptomatic=bool(ds == ds and t >= ds and not over),
                            severe=bool(dv == dv and t >= dv and not over),
                            critical=bool(dc == dc and t >= dc and not over), recovered=bool(rec), dead=bool(dead)))
        flags.append(day)
    return dict(dates=[bits(list(d)) for d in dates], flags=flags)


def main():
    out = dict(viral_load=viral_load_cases(), trans_sus=trans_sus_cases(), infections=infections_case(),
               transitions=transitions_case())
    with open(os.path.join(HERE, 'kernel_vectors.json'), 'w') as f:
        json.dump(out, f, indent=1)
    print('wrote kernel_vectors.json')


if __name__ == '__main__':
    main()
