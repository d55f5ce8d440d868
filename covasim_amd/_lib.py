"""ctypes binding of libcovasim_amd.so (the C ABI declared in include/covasim_amd.h).

The library is loaded from the package folder; there is no fallback: if it is
missing, every entry point raises. Structs below mirror the C layout and are
checked against the library's own ``cvk_abi_layout`` report on load.
"""

import ctypes as C
import os

MAX_LAYERS = 8
MAX_VARIANTS = 8
POP_MAX_AGE_BINS, POP_MAX_HH_SIZES, POP_MAX_PROG_BINS = 32, 64, 16
POP_COUNT_POISSON, POP_COUNT_NEGBIN = 0, 1
N_DURS = 9

OK, E_INVALID, E_CAPACITY, E_HIP, E_ALIGN = 0, -1, -2, -3, -4
INFECTIONS_NOALLOC_MAX_SLOTS = 65536 * 2048  # CVK_INFECTIONS_NOALLOC_MAX_SLOTS

# state bits (flags word)
SUSCEPTIBLE, NAIVE, EXPOSED, INFECTIOUS, SYMPTOMATIC, SEVERE, CRITICAL, TESTED = (1 << k for k in range(8))
DIAGNOSED, RECOVERED, KNOWN_DEAD, DEAD, KNOWN_CONTACT, QUARANTINED, ISOLATED, VACCINATED = (1 << k for k in range(8, 16))
EARLY = 1 << 16
INFVAR_SHIFT = 20
VAR_NONE = 0xF
STATE_BITS = {
    'susceptible': SUSCEPTIBLE, 'naive': NAIVE, 'exposed': EXPOSED, 'infectious': INFECTIOUS,
    'symptomatic': SYMPTOMATIC, 'severe': SEVERE, 'critical': CRITICAL, 'tested': TESTED,
    'diagnosed': DIAGNOSED, 'recovered': RECOVERED, 'known_dead': KNOWN_DEAD, 'dead': DEAD,
    'known_contact': KNOWN_CONTACT, 'quarantined': QUARANTINED, 'isolated': ISOLATED, 'vaccinated': VACCINATED,
}

DIST_CODES = {'uniform': 0, 'unif': 0, 'normal': 1, 'norm': 1, 'normal_pos': 2, 'normal_int': 3,
              'lognormal': 4, 'lognorm': 4, 'lognormal_int': 5, 'lognorm_int': 5}
DUR_ORDER = ['exp2inf', 'inf2sym', 'sym2sev', 'sev2crit', 'asym2rec', 'mild2rec', 'sev2rec', 'crit2rec', 'crit2die']

TUNE_EDGE_LDS = 1
TUNE_EDGE_CHUNKS = 2
TUNE_GENERATE_DYNAMIC = 4
TUNE_OVERLAP_RESAMPLE = 5
TUNE_EDGE_COARSE = 6
TUNE_EDGE_PREFETCH = 7
TUNE_SOURCE_LISTS = 8
TUNE_COUNT_SOURCES = 9
TUNE_SOURCE_PERMILLE = 10
TUNE_TARGET_BITS = 11

LAYER_SEED = 14
LAYER_IMPORT = 15

STREAM_TX = 0x10000000
STREAM_INFECT = 0x20000000
STREAM_CHOOSE = 0x30000000
STREAM_IMPORT = 0x40000000
STREAM_RESAMPLE = 0x50000000
STREAM_CLIP = 0x80000000

# result columns
R_NAMES = ['new_infections', 'new_reinfections', 'new_infectious', 'new_symptomatic', 'new_severe',
           'new_critical', 'new_recoveries', 'new_deaths', 'new_known_deaths',
           'n_susceptible', 'n_exposed', 'n_infectious', 'n_symptomatic', 'n_severe', 'n_critical',
           'n_recovered', 'n_dead', 'n_diagnosed', 'n_known_dead', 'n_quarantined', 'n_isolated',
           'n_vaccinated', 'n_imports', 'mid_severe', 'mid_critical', 'n_alive',
           'new_tests', 'new_diagnoses', 'new_quarantined', 'new_isolated']
R_COL = {k: i for i, k in enumerate(R_NAMES)}
R_BYVAR = len(R_NAMES)
RV_NAMES = ['new_infections_by_variant', 'new_symptomatic_by_variant', 'new_severe_by_variant',
            'new_infectious_by_variant', 'n_exposed_by_variant', 'n_infectious_by_variant']
R_COLS = R_BYVAR + len(RV_NAMES) * MAX_VARIANTS
RF_COLS = 4
CTL_WORDS = 64
CTL_LOG_COUNT, CTL_LOG_OVERFLOW, CTL_SEL_SHORT = 2, 4, 10
CTL_SRC_SLOTS, CTL_SRC_AGENTS, CTL_SRC_STREAMED = 12, 14, 16


class Dur(C.Structure):
    _fields_ = [('dist', C.c_int32), ('pad', C.c_int32), ('par1', C.c_double), ('par2', C.c_double),
                ('mu', C.c_double), ('sigma', C.c_double)]


class Params(C.Structure):
    _fields_ = [
        ('n_agents', C.c_int64), ('n_variants', C.c_int32), ('n_layers', C.c_int32), ('seed', C.c_uint64),
        ('frac_time', C.c_float), ('load_ratio', C.c_float), ('high_cap', C.c_float),
        ('vl_hi', C.c_float), ('vl_lo', C.c_float), ('asymp_factor', C.c_float),
        ('beta_v', C.c_float * MAX_VARIANTS),
        ('beta_layer', C.c_float * MAX_LAYERS), ('iso_factor', C.c_float * MAX_LAYERS),
        ('quar_factor', C.c_float * MAX_LAYERS),
        ('rel_symp_prob', C.c_float * MAX_VARIANTS), ('rel_severe_prob', C.c_float * MAX_VARIANTS),
        ('rel_crit_prob', C.c_float * MAX_VARIANTS), ('rel_death_prob', C.c_float * MAX_VARIANTS),
        ('no_hosp_factor', C.c_float), ('no_icu_factor', C.c_float), ('trans_redux', C.c_float),
        ('nab_boost', C.c_float),
        ('n_beds_hosp', C.c_int64), ('n_beds_icu', C.c_int64),
        ('dur', Dur * N_DURS), ('nab_init', Dur),
        ('use_waning', C.c_int32), ('nab_kin_len', C.c_int32),
        ('rel_imm_symp', C.c_double * 3), ('nab_norm', C.c_double),
        ('ve_exp_alpha', C.c_double * 3), ('ve_beta', C.c_double * 3),
        ('immunity', (C.c_float * MAX_VARIANTS) * MAX_VARIANTS),
        ('nab_kin', C.c_void_p),
    ]


PEOPLE_PTR_FIELDS = ['flags', 'vstate', 'date_exposed', 'date_infectious', 'date_symptomatic', 'date_severe',
                     'date_critical', 'date_recovered', 'date_dead', 'date_diagnosed', 'date_end_isolation',
                     'dur_exp2inf', 'dur_inf2sym', 'dur_sym2sev', 'dur_sev2crit', 'dur_disease', 'rel_trans',
                     'rel_sus', 'symp_prob', 'severe_prob', 'crit_prob', 'death_prob', 'n_infections',
                     'n_breakthroughs', 'nab', 'peak_nab', 't_nab_event', 'sus_imm', 'symp_imm', 'sev_imm']


class PeopleS(C.Structure):
    _fields_ = [('n', C.c_int64)] + [(k, C.c_void_p) for k in PEOPLE_PTR_FIELDS]


class SourcesS(C.Structure):
    _fields_ = [('indptr', C.c_void_p), ('target', C.c_void_p), ('edge_dir', C.c_void_p), ('n_slots', C.c_int64)]


class LayerS(C.Structure):
    _fields_ = [('p1', C.c_void_p), ('p2', C.c_void_p), ('beta', C.c_void_p), ('n_edges', C.c_int64),
                ('dynamic', C.c_int32), ('beta_ones', C.c_int32), ('sources', C.POINTER(SourcesS))]


class WorkS(C.Structure):
    _fields_ = [('winner', C.c_void_p), ('touched', C.c_void_p), ('bitmap', C.c_void_p), ('ctl', C.c_void_p),
                ('results', C.c_void_p), ('results_f', C.c_void_p), ('log', C.c_void_p),
                ('log_capacity', C.c_int64), ('select_hist', C.c_void_p), ('coarse', C.c_void_p),
                ('rec', C.c_void_p), ('partials', C.c_void_p), ('partials_f', C.c_void_p), ('bound', C.c_void_p),
                ('npts', C.c_int32), ('pad', C.c_int32), ('day_clock', C.c_void_p), ('aux', C.c_void_p)]


MAX_REDUCE_BLOCKS = 65536
PARTIAL_COLS = 64
WORK_PTR_FIELDS = ['winner', 'touched', 'bitmap', 'ctl', 'results', 'results_f', 'log', 'select_hist', 'coarse',
                   'rec', 'partials', 'partials_f', 'bound', 'day_clock', 'aux']


def make_workspace(n, npts, log_capacity, device):
    """Device workspace tensors for n agents and npts days, and the cvk_work struct over them."""
    import torch
    words = ((n + 127) // 128) * 4
    w = dict(
        winner=torch.full((n,), -1, dtype=torch.int64, device=device),
        touched=torch.zeros((n + 31) // 32, dtype=torch.int32, device=device),
        bitmap=torch.zeros(words, dtype=torch.int32, device=device),
        coarse=torch.zeros(words, dtype=torch.int32, device=device),
        rec=torch.zeros(n, dtype=torch.int64, device=device),
        ctl=torch.zeros(CTL_WORDS, dtype=torch.int32, device=device),
        results=torch.zeros((npts, R_COLS), dtype=torch.int32, device=device),
        results_f=torch.zeros((npts, RF_COLS), dtype=torch.float64, device=device),
        log=torch.zeros((max(1, log_capacity), 4), dtype=torch.int32, device=device),
        select_hist=torch.zeros(256, dtype=torch.int32, device=device),
        partials=torch.zeros((MAX_REDUCE_BLOCKS, PARTIAL_COLS), dtype=torch.int32, device=device),
        partials_f=torch.zeros((MAX_REDUCE_BLOCKS, RF_COLS), dtype=torch.float64, device=device),
        bound=torch.zeros(n, dtype=torch.int16, device=device),
        day_clock=torch.zeros(3 * npts + 1, dtype=torch.int64, device=device),
        aux=torch.zeros(words, dtype=torch.int32, device=device),
    )
    s = WorkS()
    for k in WORK_PTR_FIELDS:
        setattr(s, k, w[k].data_ptr())
    s.log_capacity = log_capacity
    s.npts = npts
    return w, s


def reset_workspace(w):
    """Clear a workspace for a fresh run (device-side fills, no host copies)."""
    w['winner'].fill_(-1)
    for k in ('touched', 'bitmap', 'coarse', 'rec', 'ctl', 'results', 'results_f', 'select_hist', 'aux'):
        w[k].zero_()


OP_TEST, OP_TRACE = 1, 2
QP_NONE, QP_START, QP_END, QP_BOTH, QP_DAILY, QP_DAYS = range(6)
MAX_QUAR_DAYS = 8
PENDING_NONE = -2**31
DIAG_MAX_BLOCKS = 4096


class DiagOp(C.Structure):
    _fields_ = [('kind', C.c_int32), ('slot', C.c_int32), ('start_day', C.c_int32), ('end_day', C.c_int32),
                ('symp_prob', C.c_double), ('asymp_prob', C.c_double), ('symp_quar_prob', C.c_double),
                ('asymp_quar_prob', C.c_double), ('sensitivity', C.c_double), ('keep_prob', C.c_double),
                ('test_delay', C.c_double), ('quar_policy', C.c_int32), ('n_quar_days', C.c_int32),
                ('quar_days', C.c_int32 * MAX_QUAR_DAYS),
                ('sub_mask', C.c_void_p), ('sub_vals', C.c_void_p), ('sub_val', C.c_double),
                ('ili_counts', C.c_void_p), ('n_ili_days', C.c_int32), ('n_swab_bins', C.c_int32),
                ('swab_base', C.c_void_p), ('trace_prob', C.c_double * MAX_LAYERS),
                ('trace_time', C.c_double * MAX_LAYERS), ('trace_period', C.c_int32 * MAX_LAYERS),
                ('presumptive', C.c_int32), ('pad1', C.c_int32), ('capacity', C.c_int64)]


DIAG_DATE_FIELDS = ['date_tested', 'date_pos_test', 'date_known_contact', 'date_quarantined', 'date_end_quarantine']


class DiagS(C.Structure):
    _fields_ = ([(k, C.c_void_p) for k in DIAG_DATE_FIELDS] +
                [('pending', C.c_void_p), ('cases', C.c_void_p), ('coarse', C.c_void_p), ('contacts', C.c_void_p),
                 ('part', C.c_void_p), ('ring', C.c_int32), ('swab_bins', C.c_int32), ('select', C.c_void_p),
                 ('swab_hist', C.c_void_p), ('swab_prob', C.c_void_p)])


class ImportsS(C.Structure):
    _fields_ = [('n', C.c_int32), ('days', C.POINTER(C.c_int32)), ('variants', C.POINTER(C.c_int32)),
                ('n_imports', C.POINTER(C.c_int64)), ('n_beta_days', C.c_int32), ('beta_v_days', C.c_void_p),
                ('beta_layer_days', C.c_void_p), ('n_diag_ops', C.c_int32), ('diag_ops', C.POINTER(DiagOp)),
                ('diag', C.POINTER(DiagS))]


BATCH_MAX_SIMS = 4096


class BatchDesc(C.Structure):
    _fields_ = [('n_sims', C.c_int32), ('n_layers', C.c_int32), ('n_variants', C.c_int32), ('use_waning', C.c_int32),
                ('n_agents_total', C.c_int64), ('agent_offsets_host', C.c_void_p), ('edge_starts_host', C.c_void_p),
                ('edge_counts_host', C.c_void_p), ('params_host', C.c_void_p), ('dynamic_host', C.c_void_p),
                ('people', PeopleS), ('p1', C.c_void_p * MAX_LAYERS), ('p2', C.c_void_p * MAX_LAYERS),
                ('beta', C.c_void_p * MAX_LAYERS), ('beta_ones', C.c_int32 * MAX_LAYERS),
                ('sources', C.POINTER(SourcesS)), ('n_beta_days', C.c_int32), ('beta_v_days', C.c_void_p),
                ('beta_layer_days', C.c_void_p)]


class BatchImports(C.Structure):
    _fields_ = [('n_events', C.c_int32), ('days', C.c_void_p), ('variants', C.c_void_p), ('counts', C.c_void_p)]


class BatchConfig(C.Structure):
    _fields_ = [('edge_block', C.c_int32), ('edge_chunks', C.c_int32), ('edge_blocks_per_cu', C.c_int32),
                ('agents_per_thread', C.c_int32)]

    def as_tuple(self):
        return (self.edge_block, self.edge_chunks, self.edge_blocks_per_cu, self.agents_per_thread)


class RelaxDesc(C.Structure):
    _fields_ = [('n_agents', C.c_int64), ('n_layers', C.c_int32), ('n_days', C.c_int32), ('kernel_len', C.c_int32),
                ('checkpoint_every', C.c_int32), ('precision', C.c_int32), ('pad', C.c_int32),
                ('seed_frac', C.c_double), ('indptr', C.c_void_p * MAX_LAYERS), ('nbr', C.c_void_p * MAX_LAYERS),
                ('age_bin', C.c_void_p), ('rel_trans', C.c_void_p), ('w_host', C.c_void_p),
                ('dyn_slots', C.c_double * MAX_LAYERS)]


class FinalizeDesc(C.Structure):
    _fields_ = [('n_sims', C.c_int32), ('npts', C.c_int32), ('n_variants', C.c_int32), ('use_waning', C.c_int32),
                ('r_eff_smoothing', C.c_int32), ('doubling_window', C.c_int32), ('max_doubling_time', C.c_double),
                ('pop_infected_host', C.c_void_p), ('initial_period_host', C.c_void_p),
                ('agent_offsets_host', C.c_void_p)]


class PopGen(C.Structure):
    _fields_ = [('seed', C.c_uint64), ('n_age_bins', C.c_int32), ('n_hh_sizes', C.c_int32),
                ('hh_size_min', C.c_int32), ('correlate', C.c_int32), ('n_prog_bins', C.c_int32), ('pad', C.c_int32),
                ('age_cdf', C.c_double * POP_MAX_AGE_BINS), ('age_lo', C.c_double * POP_MAX_AGE_BINS),
                ('age_width', C.c_double * POP_MAX_AGE_BINS), ('hh_cdf', C.c_double * POP_MAX_HH_SIZES),
                ('age_noise', C.c_double), ('prog_cutoff', C.c_double * POP_MAX_PROG_BINS),
                ('trans_or', C.c_double * POP_MAX_PROG_BINS), ('symp_prob', C.c_float * POP_MAX_PROG_BINS),
                ('severe_prob', C.c_float * POP_MAX_PROG_BINS), ('crit_prob', C.c_float * POP_MAX_PROG_BINS),
                ('death_prob', C.c_float * POP_MAX_PROG_BINS), ('rel_sus', C.c_float * POP_MAX_PROG_BINS),
                ('nb_p0', C.c_double), ('nb_q', C.c_double), ('nb_r', C.c_double), ('nb_step', C.c_double)]


class PopCounts(C.Structure):
    _fields_ = [('kind', C.c_int32), ('pad', C.c_int32), ('p0', C.c_double), ('lam', C.c_double), ('r', C.c_double),
                ('q', C.c_double)]


_LIB = None
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'libcovasim_amd.so')


class CovasimAMDError(RuntimeError):
    """Raised when a library call returns a failure status."""


def lib():
    """Load libcovasim_amd.so (once); raises if it was not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f'covasim_amd: HIP library not built ({LIB_PATH} missing); run __graft_entry__.build()')
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, u32, u64, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_float
    L.cvk_last_error.restype = C.c_char_p
    L.cvk_version.restype = C.c_int
    L.cvk_source_hash.restype = C.c_uint64
    L.cvk_wall_clock_rate.argtypes = [C.POINTER(C.c_int64)]
    L.cvk_abi_layout.argtypes = [C.POINTER(C.c_int64), i32]
    L.cvk_device_info.argtypes = [C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.cvk_set_tuning.argtypes = [i32, i32]
    L.cvk_workspace_sizes.argtypes = [i64, i32] + [C.POINTER(C.c_int64)] * 5
    L.cvk_workspace_sizes2.argtypes = [i64] + [C.POINTER(C.c_int64)] * 4
    L.cvk_viral_load.argtypes = [i32, vp, vp, vp, i64, f32, f32, f32, vp, vp]
    L.cvk_trans_sus.argtypes = [vp, vp, vp, vp, f32, vp, vp, vp, vp, f32, f32, f32, vp, i64, vp, vp, vp]
    L.cvk_infections.argtypes = [f32, vp, vp, vp, i64, vp, vp, i64, vp, u64, i32, i32, i32, i32, vp, vp, vp, i64,
                                 C.POINTER(C.c_int64), vp]
    L.cvk_infections_noalloc.argtypes = L.cvk_infections.argtypes
    P, Pa, W, Ly = C.POINTER(PeopleS), C.POINTER(Params), C.POINTER(WorkS), C.POINTER(LayerS)
    L.cvk_choose.argtypes = [P, Pa, W, i32, u32, i64, i32, i32, i32, vp]
    L.cvk_agent_pass.argtypes = [P, Pa, W, i32, vp]
    L.cvk_resample_layer.argtypes = [Ly, i32, Pa, i32, vp]
    L.cvk_resample_frac_work_bytes.restype = i64
    L.cvk_sources_work_bytes.argtypes = [i64]
    L.cvk_sources_work_bytes.restype = i64
    L.cvk_sources_build.argtypes = [Ly, i64, vp, vp, vp, i64, C.POINTER(C.c_int64), vp, i64, vp]
    L.cvk_resample_layer_frac.argtypes = [Ly, i32, Pa, i32, i64, vp, i64, vp]
    L.cvk_edge_pass.argtypes = [P, Ly, Pa, W, i32, C.POINTER(C.c_void_p), vp]
    L.cvk_infect_pass.argtypes = [P, Ly, Pa, W, i32, i32, vp]
    L.cvk_count_pass.argtypes = [P, Pa, W, i32, vp]
    L.cvk_run_days.argtypes = [P, Ly, Pa, W, i32, i32, i64, C.POINTER(ImportsS), i32, vp]
    L.cvk_day_plan.argtypes = [Ly, Pa, C.POINTER(ImportsS), i32, C.POINTER(C.c_int32)]
    L.cvk_graph_create.argtypes = [P, Ly, Pa, W, i32, i32, i64, C.POINTER(ImportsS), vp, C.POINTER(C.c_void_p)]
    L.cvk_graph_launch.argtypes = [vp, vp]
    L.cvk_graph_destroy.argtypes = [vp]
    BD, BI, BC = C.POINTER(BatchDesc), C.POINTER(BatchImports), C.POINTER(BatchConfig)
    L.cvk_batch_time_edges.argtypes = [vp, i32]
    L.cvk_batch_edge_times.argtypes = [vp, C.POINTER(C.c_float), i32]
    L.cvk_batch_create.argtypes = [BD, i32, i64, C.POINTER(C.c_void_p)]
    L.cvk_batch_destroy.argtypes = [vp]
    L.cvk_batch_set_config.argtypes = [vp, BC]
    L.cvk_batch_get_config.argtypes = [vp, BC]
    L.cvk_batch_reset.argtypes = [vp, vp]
    L.cvk_batch_run_days.argtypes = [vp, i32, i32, vp, BI, i32, vp]
    L.cvk_batch_graph_create.argtypes = [vp, i32, i32, vp, BI, vp, C.POINTER(C.c_void_p)]
    L.cvk_batch_copy_results.argtypes = [vp, vp, vp]
    L.cvk_batch_copy_log.argtypes = [vp, vp, i64, C.POINTER(C.c_int64)]
    L.cvk_relax_create.argtypes = [C.POINTER(RelaxDesc), C.POINTER(C.c_void_p)]
    L.cvk_relax_destroy.argtypes = [vp]
    L.cvk_relax_forward.argtypes = [vp, vp, vp, vp]
    L.cvk_relax_backward.argtypes = [vp, vp, vp, vp, vp]
    FD = C.POINTER(FinalizeDesc)
    L.cvk_batch_results.argtypes = [vp, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]
    L.cvk_batch_results.restype = C.c_int
    L.cvk_finalize_work_bytes.argtypes = [FD]
    L.cvk_finalize_work_bytes.restype = i64
    L.cvk_finalize.argtypes = [FD, vp, vp, vp, vp, vp, vp, vp, vp, vp, i64, vp]
    L.cvk_finalize.restype = C.c_int
    L.cvk_move_edges_work_bytes.argtypes = [i64]
    L.cvk_move_edges_work_bytes.restype = i64
    L.cvk_move_edges.argtypes = [vp, vp, vp, i64, i64, u64, i32, i32, vp, vp, vp, vp, vp, vp, vp, i64, vp]
    L.cvk_move_edges.restype = C.c_int
    L.cvk_reduce.argtypes = [i32, i64, vp, i32, C.c_double, C.c_double, C.c_double, vp, vp, vp, vp]
    DS, DO = C.POINTER(DiagS), C.POINTER(DiagOp)
    L.cvk_diag_apply.argtypes = [P, Ly, Pa, W, DS, DO, i32, vp]
    L.cvk_diag_apply.restype = C.c_int
    L.cvk_states_post.argtypes = [P, Pa, W, DS, i32, vp]
    L.cvk_states_post.restype = C.c_int
    L.cvk_diag_coarse_words.argtypes = [i64]
    L.cvk_diag_coarse_words.restype = i64
    L.cvk_reduce.restype = C.c_int
    PG, PC = C.POINTER(PopGen), C.POINTER(PopCounts)
    L.cvk_pop_work_bytes.argtypes = [i64, i64]
    L.cvk_pop_work_bytes.restype = i64
    L.cvk_pop_ages.argtypes = [PG, i64, vp, vp, vp]
    L.cvk_pop_households_plan.argtypes = [PG, i64, vp, i64, vp, i64, vp, vp]
    L.cvk_pop_households_fill.argtypes = [PG, i64, vp, i64, vp, i64, vp, vp, vp, vp, i64, vp]
    L.cvk_pop_members.argtypes = [i64, vp, C.c_double, C.c_double, vp, vp, vp, i64, vp]
    L.cvk_pop_contacts_plan.argtypes = [PG, i32, i64, PC, vp, vp, i64, vp]
    L.cvk_pop_contacts_fill.argtypes = [PG, i32, i64, vp, vp, vp, vp, vp, i64, vp]
    L.cvk_pop_people.argtypes = [PG, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, i64, vp]
    for name in ['cvk_pop_ages', 'cvk_pop_households_plan', 'cvk_pop_households_fill', 'cvk_pop_members',
                 'cvk_pop_contacts_plan', 'cvk_pop_contacts_fill', 'cvk_pop_people']:
        getattr(L, name).restype = C.c_int
    for name in ['cvk_relax_create', 'cvk_relax_destroy', 'cvk_relax_forward', 'cvk_relax_backward']:
        getattr(L, name).restype = C.c_int
    for name in ['cvk_batch_create', 'cvk_batch_destroy', 'cvk_batch_set_config', 'cvk_batch_get_config',
                 'cvk_batch_reset', 'cvk_batch_run_days', 'cvk_batch_graph_create', 'cvk_batch_copy_results',
                 'cvk_batch_copy_log']:
        getattr(L, name).restype = C.c_int
    for name in ['cvk_set_tuning', 'cvk_graph_create', 'cvk_graph_launch', 'cvk_graph_destroy', 'cvk_workspace_sizes2',
                 'cvk_workspace_sizes', 'cvk_viral_load', 'cvk_trans_sus', 'cvk_infections', 'cvk_choose',
                 'cvk_agent_pass', 'cvk_resample_layer', 'cvk_edge_pass', 'cvk_infect_pass', 'cvk_count_pass',
                 'cvk_run_days', 'cvk_abi_layout', 'cvk_device_info']:
        getattr(L, name).restype = C.c_int
    _check_layout(L)
    _LIB = L
    return L


def _check_layout(L):
    out = (C.c_int64 * 32)()
    m = L.cvk_abi_layout(out, 32)
    got = list(out[:m])
    want = [C.sizeof(Dur), C.sizeof(Params), C.sizeof(PeopleS), C.sizeof(LayerS), C.sizeof(WorkS),
            C.sizeof(ImportsS), Params.dur.offset, Params.nab_init.offset, Params.rel_imm_symp.offset,
            Params.immunity.offset, Params.nab_kin.offset, PeopleS.sev_imm.offset, WorkS.log_capacity.offset,
            WorkS.npts.offset, R_COLS, RF_COLS, C.sizeof(BatchDesc), BatchDesc.people.offset, C.sizeof(BatchImports),
            C.sizeof(BatchConfig), C.sizeof(RelaxDesc), C.sizeof(PopGen), PopGen.nb_p0.offset, C.sizeof(PopCounts),
            C.sizeof(DiagOp), DiagOp.trace_prob.offset, C.sizeof(DiagS), ImportsS.diag.offset]
    if got != want:
        raise RuntimeError(f'covasim_amd: ctypes struct layout mismatch with the library: {got} != {want}')


def check(rc, what=''):
    """Raise on a negative status, mapping codes to Python exception types like the reference's."""
    if rc >= 0:
        return rc
    msg = lib().cvk_last_error().decode(errors='replace')
    if rc == E_INVALID:
        raise ValueError(f'{what}: {msg}')
    if rc == E_ALIGN:
        raise ValueError(f'{what}: {msg}')
    if rc == E_CAPACITY:
        raise OverflowError(f'{what}: {msg}')
    raise CovasimAMDError(f'{what}: {msg}')


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)
