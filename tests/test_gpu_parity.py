"""GPU parity: HIP kernels (through the C ABI) against the oracle on the same seeded inputs.

Bit-exact for indices and counts; float32 values compared as bit patterns
(the kernels keep the reference's float32 operation order and are built with
-ffp-contract=off and correctly rounded division).
"""

import numpy as np
import pytest

from oracle import covasim_oracle as co
from oracle import philox as px
from tests.helpers import COMPARE_KEYS, STOCK_KEYS, copy_pop, make_population, oracle_sim

pytestmark = pytest.mark.gpu


def bits(a):
    return np.asarray(a, dtype=np.float32).view(np.uint32)


def random_dates(n, rng):
    start = np.where(rng.random(n) < 0.7, rng.integers(0, 40, n), np.nan).astype(np.float32)
    dur = rng.choice([0, 1, 2, 3, 5, 8, 13, 21, 40], n).astype(np.float32)
    rec = (start + dur).astype(np.float32)
    dead = np.where(rng.random(n) < 0.2, start + rng.integers(0, 30, n), np.nan).astype(np.float32)
    rec[~np.isnan(dead)] = np.nan
    return start, rec, dead


@pytest.mark.parametrize('n_agents', [4000, 1_500_000_000])
def test_resample_layer_matches_oracle(gpu, n_agents):
    """Layer.update via cvk_resample_layer, odd edge count; n = 1.5e9 rejects ~30 % of first words (retry stream)."""
    import ctypes as C
    import torch
    from covasim_amd import _lib
    from oracle import philox as px
    n_edges = 100_001
    p1 = torch.full((n_edges,), -1, dtype=torch.int32, device='cuda')
    p2 = torch.full((n_edges,), -1, dtype=torch.int32, device='cuda')
    beta = torch.zeros(n_edges, dtype=torch.float32, device='cuda')
    lay = _lib.LayerS(p1.data_ptr(), p2.data_ptr(), beta.data_ptr(), n_edges, 1, 1)
    prm = _lib.Params()
    prm.n_agents = n_agents
    prm.seed = 77
    _lib.check(_lib.lib().cvk_resample_layer(C.byref(lay), 3, C.byref(prm), 12, _lib.stream_ptr()), 'resample')
    torch.cuda.synchronize()
    q1, q2 = px.resample_edges(77, 12, 3, n_edges, n_agents)
    assert np.array_equal(p1.cpu().numpy(), q1)
    assert np.array_equal(p2.cpu().numpy(), q2)
    assert bool((beta == 1.0).all())


def test_viral_load_bit_exact(gpu):
    import covasim_amd as cv
    rng = np.random.default_rng(3)
    n = 100_003
    start, rec, dead = random_dates(n, rng)
    for t in [0, 1, 7, 20, 45]:
        for (ft, lr, hc) in [(0.3, 2, 4), (0.5, 3.0, 1.5), (0.0, 1.0, 4.0)]:
            got = cv.utils.compute_viral_load(t, start, rec, dead, ft, lr, hc)
            want = co.compute_viral_load(t, start, rec, dead, ft, lr, hc)
            assert np.array_equal(bits(got), bits(want)), (t, ft, lr, hc)


def test_trans_sus_bit_exact(gpu):
    import covasim_amd as cv
    rng = np.random.default_rng(4)
    n = 50_001
    rt = (rng.negative_binomial(0.45, 0.0045, n) * 0.01).astype(np.float32)
    rs = rng.choice([0.34, 0.67, 1.0, 1.24, 1.47], n).astype(np.float32)
    b = lambda p: rng.random(n) < p
    vl = rng.choice([2 / 1.3, 1 / 1.3], n).astype(np.float32)
    imm = np.where(b(0.3), rng.random(n), 0).astype(np.float32)
    for (bl, af, isf, qf) in [(1.0, 1.0, 0.2, 0.3), (3.0, 0.8, 0.3, 0.6), (0.0, 1.0, 0.1, 0.2)]:
        args = (rt, rs, b(0.2), b(0.6), bl, vl, b(0.5), b(0.1), b(0.1), af, isf, qf, imm)
        got = cv.utils.compute_trans_sus(*args)
        want = co.compute_trans_sus(*args)
        assert np.array_equal(bits(got[0]), bits(want[0]))
        assert np.array_equal(bits(got[1]), bits(want[1]))


def edge_case_inputs(rng, n=5000, E=60_000):
    p1 = rng.integers(0, n, E).astype(np.int32)
    p2 = rng.integers(0, n, E).astype(np.int32)
    p1[:50] = p2[:50]  # self-loops
    p1[50:100] = p1[0]  # a hub
    lb = rng.choice([1.0, 0.0, 2.5, 1e-30], E, p=[0.9, 0.03, 0.05, 0.02]).astype(np.float32)
    rt = np.where(rng.random(n) < 0.15, rng.random(n) * 3, 0).astype(np.float32)
    rt[:10] = 1e6  # p > 1
    rs = np.where(rng.random(n) < 0.8, rng.random(n), 0).astype(np.float32)
    return p1, p2, lb, rt, rs


def test_infections_supplied_uniforms_bit_exact(gpu):
    import covasim_amd as cv
    rng = np.random.default_rng(5)
    p1, p2, lb, rt, rs = edge_case_inputs(rng)
    U = rng.random((2, len(p1)))
    for beta in [0.016, 0.5, 0.0]:
        for legacy in [False, True]:
            s, t, p = cv.utils.compute_infections(beta, p1, p2, lb, rt, rs, legacy=legacy, uniforms=U,
                                                  return_probs=True)
            ws, wt, wp = co.compute_infections(beta, p1, p2, lb, rt, rs, lambda d, idx: U[d][idx], legacy=legacy)
            assert np.array_equal(s, ws) and np.array_equal(t, wt), (beta, legacy)
            assert np.array_equal(bits(p), bits(wp))
            if beta == 0.0:
                assert len(s) == 0  # tests/unittests/test_transmission.py:16-35


def test_infections_philox_bit_exact(gpu):
    import covasim_amd as cv
    rng = np.random.default_rng(6)
    p1, p2, lb, rt, rs = edge_case_inputs(rng)
    cv.utils.set_seed(77)
    for (t, v, l) in [(0, 0, 0), (5, 1, 3)]:
        s, tg = cv.utils.compute_infections(0.3, p1, p2, lb, rt, rs, t=t, variant=v, layer=l)
        ws, wt, _ = co.compute_infections(0.3, p1, p2, lb, rt, rs, lambda d, idx: px.edge_uniforms(77, t, v, l, idx, d))
        assert np.array_equal(s, ws) and np.array_equal(tg, wt)


def run_pair(pop, n_days=40, seed=1, variants=(), use_graph=None, **pars):
    import covasim_amd as cv
    gpars = dict(pars)
    gvars = [cv.variant({k: v[k] for k in ('rel_beta', 'label') if k in v}, days=v['days'], n_imports=v['n_imports'])
             for v in variants]
    sim = cv.Sim(pop_size=len(pop['age']), n_days=n_days, rand_seed=seed, popdict=copy_pop(pop), variants=gvars,
                 **gpars)
    sim.run(use_graph=use_graph)
    osim = oracle_sim(pop, seed=seed, n_days=n_days, variants=list(variants), **pars).run()
    ores = osim.finalize()
    return sim, osim, ores


def assert_results_equal(sim, ores, nv=1):
    for key in COMPARE_KEYS:
        assert np.array_equal(sim.results[key].values, ores[key]), (key, sim.results[key].values, ores[key])
    assert np.array_equal(sim.raw['n_susceptible_agents'], ores['n_susceptible'])
    for key in STOCK_KEYS[1:]:
        assert np.array_equal(sim.results[f'n_{key}'].values, ores[f'n_{key}']), key
    V = sim.results['variant']
    for key in ['new_infections_by_variant', 'new_symptomatic_by_variant', 'new_severe_by_variant',
                'new_infectious_by_variant']:
        assert np.array_equal(V[key].values, ores[key][:nv]), key
    assert np.array_equal(V['n_exposed_by_variant'].values, ores['n_exposed_by_variant'])
    assert np.array_equal(V['n_infectious_by_variant'].values, ores['n_infectious_by_variant'])


def assert_people_equal(sim, osim):
    gp, op = sim.people, osim.people
    for key in ['susceptible', 'naive', 'exposed', 'infectious', 'symptomatic', 'severe', 'critical', 'recovered',
                'dead', 'diagnosed', 'known_dead']:
        assert np.array_equal(gp.get(key), getattr(op, key)), key
    for key in ['date_exposed', 'date_infectious', 'date_symptomatic', 'date_severe', 'date_critical',
                'date_recovered', 'date_dead', 'dur_exp2inf', 'dur_inf2sym', 'dur_sym2sev', 'dur_sev2crit',
                'dur_disease', 'rel_trans', 'nab', 'peak_nab']:
        assert np.array_equal(bits(gp.get(key)), bits(getattr(op, key))), key
    for key in ['exposed_variant', 'infectious_variant', 'recovered_variant']:
        assert np.array_equal(bits(gp.get(key)), bits(getattr(op, key))), key
    for key in ['n_infections', 'n_breakthroughs', 't_nab_event']:
        assert np.array_equal(gp.get(key), getattr(op, key)), key
    for key in ['sus_imm', 'symp_imm', 'sev_imm']:
        assert np.array_equal(bits(gp.get(key)), bits(getattr(op, key))), key
    # contact layers after the run (a dynamic layer holds the last day's resample, weights 1)
    for lk, (p1, p2, beta) in osim.layers.items():
        lay = gp.contacts[lk]
        assert np.array_equal(lay.tensor('p1').cpu().numpy(), p1), lk
        assert np.array_equal(lay.tensor('p2').cpu().numpy(), p2), lk
        assert np.array_equal(bits(lay.tensor('beta').cpu().numpy()), bits(np.asarray(beta, np.float32))), lk


CONFIGS = {
    'random_plain': dict(pop=dict(n=3000, pop_type='random'), pars=dict(use_waning=False, beta=0.05)),
    'random_waning': dict(pop=dict(n=3000, pop_type='random'), pars=dict(use_waning=True, beta=0.05)),
    'hybrid_dynamic': dict(pop=dict(n=4000, pop_type='hybrid'),
                           pars=dict(use_waning=False, beta=0.03, dynam_layer=dict(c=1))),
    'hybrid_categorical_waning': dict(pop=dict(n=4000, pop_type='hybrid', households='categorical'),
                                      pars=dict(use_waning=True, beta=0.03)),
    'variants_waning': dict(pop=dict(n=4000, pop_type='hybrid'),
                            pars=dict(use_waning=True, beta=0.02,
                                      variants=[dict(rel_beta=1.7, days=[10], n_imports=30)])),
    # bitmap not staged: the coarse map in LDS filters the L2 reads of bitmap words
    'weighted_edges_l2_bitmap': dict(pop=dict(n=3000, pop_type='hybrid', weighted=True),
                                     pars=dict(use_waning=True, beta=0.04), tuning=dict(edge_lds=0)),
    'hybrid_dynamic_deep_tiles': dict(pop=dict(n=4000, pop_type='hybrid'),
                                      pars=dict(use_waning=False, beta=0.03, dynam_layer=dict(c=1)),
                                      tuning=dict(edge_chunks=4)),
    'variants_waning_shallow_l2': dict(pop=dict(n=4000, pop_type='hybrid'),
                                       pars=dict(use_waning=True, beta=0.02,
                                                 variants=[dict(rel_beta=1.7, days=[10], n_imports=30)]),
                                       tuning=dict(edge_chunks=2, edge_lds=0, edge_coarse=0)),
    'hybrid_dynamic_prefetch_tiles': dict(pop=dict(n=4000, pop_type='hybrid'),
                                          pars=dict(use_waning=True, beta=0.03, dynam_layer=dict(c=1)),
                                          tuning=dict(edge_chunks=2, edge_prefetch=1)),
    'hybrid_dynamic_resample_in_agent_pass': dict(pop=dict(n=4000, pop_type='hybrid'),
                                                  pars=dict(use_waning=True, beta=0.03, dynam_layer=dict(c=1)),
                                                  tuning=dict(overlap_resample=0)),
    'hybrid_dynamic_generated_contacts': dict(pop=dict(n=4000, pop_type='hybrid'),
                                              pars=dict(use_waning=False, beta=0.03, dynam_layer=dict(c=1)),
                                              tuning=dict(generate_dynamic=1)),
    # p_ub >= 1 for most slots: whole rounds pass the prefilter (exact tests in place and queue drains)
    'dense_exact_tests': dict(pop=dict(n=3000, pop_type='random'), pars=dict(use_waning=False, beta=0.9)),
    # every static layer streamed instead of walked through its source lists
    'hybrid_dynamic_streamed': dict(pop=dict(n=4000, pop_type='hybrid'),
                                    pars=dict(use_waning=False, beta=0.03, dynam_layer=dict(c=1)),
                                    tuning=dict(source_lists=0)),
    'variants_waning_streamed_l2': dict(pop=dict(n=4000, pop_type='hybrid'),
                                        pars=dict(use_waning=True, beta=0.02,
                                                  variants=[dict(rel_beta=1.7, days=[10], n_imports=30)]),
                                        tuning=dict(source_lists=0, edge_lds=0)),
    # static layers walked through their source lists on every day, busy ones included
    'hybrid_waning_always_walked': dict(pop=dict(n=4000, pop_type='hybrid', households='categorical'),
                                        pars=dict(use_waning=True, beta=0.04, dynam_layer=dict(c=1)),
                                        tuning=dict(source_lists=1)),
    # the choice between walking and streaming switching back and forth (threshold at 5 per thousand)
    'hybrid_walk_switching': dict(pop=dict(n=4000, pop_type='hybrid'),
                                  pars=dict(use_waning=False, beta=0.03, dynam_layer=dict(c=1)),
                                  tuning=dict(source_lists=2, source_permille=5)),
    # source lists with the bitmap read from L2 / filtered by the coarse map, and weighted static layers
    'weighted_sources_coarse': dict(pop=dict(n=3000, pop_type='hybrid', weighted=True),
                                    pars=dict(use_waning=True, beta=0.05, dynam_layer=dict(c=1)),
                                    tuning=dict(edge_lds=0, edge_coarse=1, source_lists=1)),
    # three variants (two imported), waning: default tuning, and streamed with the coarse map
    'three_variants_waning': dict(pop=dict(n=4000, pop_type='hybrid'),
                                  pars=dict(use_waning=True, beta=0.02, dynam_layer=dict(c=1),
                                            variants=[dict(label='v1', rel_beta=1.7, days=[8], n_imports=30),
                                                      dict(label='v2', rel_beta=1.3, days=[15], n_imports=30)])),
    'three_variants_streamed_coarse': dict(pop=dict(n=4000, pop_type='hybrid'),
                                           pars=dict(use_waning=True, beta=0.02, dynam_layer=dict(c=1),
                                                     variants=[dict(label='v1', rel_beta=1.7, days=[8], n_imports=30),
                                                               dict(label='v2', rel_beta=1.3, days=[15],
                                                                    n_imports=30)]),
                                           tuning=dict(source_lists=0, edge_lds=0)),
    # target bits (the bitmap not in LDS: each mixed edge reads its target's bit instead of its code, bounded by
    # the day's largest target code): three variants; dense exact tests; walk switching; bitmap in L2; and the
    # codes path in coarse mode with the bits off
    'target_bits_three_variants': dict(pop=dict(n=4000, pop_type='hybrid', households='categorical'),
                                       pars=dict(use_waning=True, beta=0.03, dynam_layer=dict(c=1),
                                                 variants=[dict(label='v1', rel_beta=1.7, days=[8], n_imports=30),
                                                           dict(label='v2', rel_beta=1.3, days=[15],
                                                                n_imports=30)]),
                                       tuning=dict(edge_lds=0, source_lists=0)),
    'target_bits_dense': dict(pop=dict(n=3000, pop_type='random'), pars=dict(use_waning=False, beta=0.9),
                              tuning=dict(edge_lds=0)),
    'target_bits_walk_switching': dict(pop=dict(n=4000, pop_type='hybrid'),
                                       pars=dict(use_waning=True, beta=0.03, dynam_layer=dict(c=1)),
                                       tuning=dict(edge_lds=0, source_lists=2, source_permille=5)),
    'target_bits_l2_weighted': dict(pop=dict(n=3000, pop_type='hybrid', weighted=True),
                                    pars=dict(use_waning=False, beta=0.05, dynam_layer=dict(c=1)),
                                    tuning=dict(edge_lds=0, edge_coarse=0, source_lists=0)),
    'codes_coarse_variants': dict(pop=dict(n=4000, pop_type='hybrid'),
                                  pars=dict(use_waning=True, beta=0.02, dynam_layer=dict(c=1),
                                            variants=[dict(rel_beta=1.7, days=[10], n_imports=30)]),
                                  tuning=dict(edge_lds=0, source_lists=0, target_bits=0)),
    # two variants without labels are both 'custom': one index, the second's parameters (immunity.py:107-114)
    'same_label_variants': dict(pop=dict(n=4000, pop_type='hybrid'),
                                pars=dict(use_waning=True, beta=0.02, dynam_layer=dict(c=1),
                                          variants=[dict(rel_beta=1.7, days=[8], n_imports=30),
                                                    dict(rel_beta=1.3, days=[15], n_imports=30)]),
                                n_variants=2),
    'beds_limited': dict(pop=dict(n=3000, pop_type='random'),
                         pars=dict(use_waning=False, beta=0.08, n_beds_hosp=2, n_beds_icu=1,
                                   rel_severe_prob=20.0, rel_crit_prob=5.0)),
}


@pytest.mark.parametrize('name', list(CONFIGS))
def test_full_sim_matches_oracle(gpu, name):
    from covasim_amd import _lib
    cfg = CONFIGS[name]
    pop = make_population(**cfg['pop'])
    pars = dict(cfg['pars'])
    pars['pop_type'] = cfg['pop']['pop_type']
    variants = pars.pop('variants', [])
    keys = dict(edge_lds=(_lib.TUNE_EDGE_LDS, 1), edge_chunks=(_lib.TUNE_EDGE_CHUNKS, 2),
                generate_dynamic=(_lib.TUNE_GENERATE_DYNAMIC, 2),
                overlap_resample=(_lib.TUNE_OVERLAP_RESAMPLE, 1), edge_coarse=(_lib.TUNE_EDGE_COARSE, 1),
                edge_prefetch=(_lib.TUNE_EDGE_PREFETCH, 0), source_lists=(_lib.TUNE_SOURCE_LISTS, 2),
                source_permille=(_lib.TUNE_SOURCE_PERMILLE, 30), target_bits=(_lib.TUNE_TARGET_BITS, 1))
    for k, v in cfg.get('tuning', {}).items():
        _lib.check(_lib.lib().cvk_set_tuning(keys[k][0], v), 'tuning')
    try:
        sim, osim, ores = run_pair(pop, n_days=45, seed=11, variants=variants, **pars)
    finally:
        for key, default in keys.values():
            _lib.lib().cvk_set_tuning(key, default)
    assert sim.results['cum_infections'].values[-1] > 50  # the epidemic took off
    nv = cfg.get('n_variants', 1 + len(variants))
    assert sim.pars['n_variants'] == osim.pars['n_variants'] == nv
    assert_results_equal(sim, ores, nv=nv)
    assert_people_equal(sim, osim)


def test_graph_and_eager_and_step_agree(gpu):
    import covasim_amd as cv
    pop = make_population(n=2500, pop_type='hybrid')
    runs = []
    for mode in ['graph', 'eager', 'step', 'chunks']:
        sim = cv.Sim(pop_size=2500, pop_type='hybrid', n_days=30, rand_seed=5, beta=0.04, popdict=copy_pop(pop),
                     dynam_layer=dict(c=1))
        if mode == 'step':
            sim.initialize()
            while not sim.complete:
                sim.step()
            sim.finalize()
        elif mode == 'chunks':  # run(until) then resume (tests/test_resume.py:19-52)
            sim.run(until=9)
            sim.run(until=20)
            sim.run()
        else:
            sim.run(use_graph=(mode == 'graph'))
        runs.append(sim)
    for key in COMPARE_KEYS + ['n_exposed', 'n_infectious', 'n_recovered']:
        for other in runs[1:]:
            assert np.array_equal(runs[0].results[key].values, other.results[key].values), key
    snaps = [r.people.snapshot() for r in runs]
    for other in snaps[1:]:
        for k in snaps[0]:
            assert np.array_equal(snaps[0][k].view(np.uint8), other[k].view(np.uint8)), k  # NaN-safe, bitwise


def test_infection_log_attribution(gpu):
    pop = make_population(n=3000)
    sim, osim, ores = run_pair(pop, n_days=25, seed=3, use_waning=False, beta=0.05)
    glog = sorted((r['date'], r['target'], -1 if r['source'] is None else r['source'], r['layer'])
                  for r in sim.infection_log())
    olog = sorted((d, tg, -1 if s < 0 else s, l) for (s, tg, d, l, v) in osim.people.infection_log)
    assert glog == olog


@pytest.mark.parametrize('n, self_loops', [(5000, False), (70_001, True)])
def test_source_lists_hold_both_directions_of_every_edge(gpu, n, self_loops):
    """cvk_sources_build: agent i's slots are exactly (j, e, dir) for every edge e = (i, j) (dir 0) and
    e = (j, i) (dir 1), self-loops left out; offsets are the running degree counts."""
    import ctypes as C
    import torch
    from covasim_amd import _lib
    rng = np.random.default_rng(n)
    E = 3 * n + 7
    p1 = rng.integers(0, n, E).astype(np.int32)
    p2 = rng.integers(0, n, E).astype(np.int32)
    if self_loops:
        p1[::97] = p2[::97]
    d1, d2 = torch.from_numpy(p1).cuda(), torch.from_numpy(p2).cuda()
    beta = torch.ones(E, dtype=torch.float32, device='cuda')
    lay = _lib.LayerS(d1.data_ptr(), d2.data_ptr(), beta.data_ptr(), E, 0, 1)
    L = _lib.lib()
    indptr = torch.empty(n + 1, dtype=torch.int32, device='cuda')
    target = torch.empty(2 * E, dtype=torch.int32, device='cuda')
    edge_dir = torch.empty(2 * E, dtype=torch.int32, device='cuda')
    nb = int(L.cvk_sources_work_bytes(n))
    work = torch.empty((nb + 3) // 4, dtype=torch.int32, device='cuda')
    slots = C.c_int64(-1)
    _lib.check(L.cvk_sources_build(C.byref(lay), n, indptr.data_ptr(), target.data_ptr(), edge_dir.data_ptr(),
                                   2 * E, C.byref(slots), work.data_ptr(), nb, _lib.stream_ptr()), 'sources')
    torch.cuda.synchronize()
    keep = p1 != p2
    assert slots.value == 2 * int(keep.sum())
    ip = indptr.cpu().numpy().astype(np.int64)
    deg = np.bincount(np.concatenate([p1[keep], p2[keep]]), minlength=n)
    assert ip[0] == 0 and np.array_equal(np.diff(ip), deg)
    tg = target.cpu().numpy()[:slots.value]
    ed = edge_dir.cpu().numpy()[:slots.value].view(np.uint32).astype(np.int64)
    src = np.repeat(np.arange(n), deg)
    got = np.lexsort((ed, src))
    e_all = np.flatnonzero(keep)
    want_src = np.concatenate([p1[e_all], p2[e_all]])
    want_tgt = np.concatenate([p2[e_all], p1[e_all]])
    want_ed = np.concatenate([e_all << 1, (e_all << 1) | 1])
    order = np.lexsort((want_ed, want_src))
    assert np.array_equal(src[got], want_src[order])
    assert np.array_equal(tg[got], want_tgt[order])
    assert np.array_equal(ed[got], want_ed[order])
