"""The C-ABI library loads without a GPU and exports every entry point include/covasim_amd.h declares."""

import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'covasim_amd.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(cvk_[a-z0-9_]+)\s*\(', text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ['cvk_viral_load', 'cvk_trans_sus', 'cvk_infections', 'cvk_agent_pass', 'cvk_edge_pass',
                 'cvk_infect_pass', 'cvk_count_pass', 'cvk_run_days', 'cvk_choose', 'cvk_resample_layer',
                 'cvk_pop_ages', 'cvk_pop_households_plan', 'cvk_pop_households_fill', 'cvk_pop_members',
                 'cvk_pop_contacts_plan', 'cvk_pop_contacts_fill', 'cvk_pop_people']:
        assert must in names


def test_library_exports_every_declared_symbol():
    from covasim_amd import _lib
    L = _lib.lib()  # also verifies the ctypes struct layouts against cvk_abi_layout
    for name in declared_functions():
        assert hasattr(L, name), name


def test_struct_layout_report():
    from covasim_amd import _lib
    L = _lib.lib()
    out = (C.c_int64 * 32)()
    m = L.cvk_abi_layout(out, 32)
    assert m == 28
    assert out[0] == C.sizeof(_lib.Dur) == 40
    assert out[21] == C.sizeof(_lib.PopGen) and out[23] == C.sizeof(_lib.PopCounts) == 40
    assert out[24] == C.sizeof(_lib.DiagOp) and out[26] == C.sizeof(_lib.DiagS) == 112
    assert L.cvk_abi_layout(out, 3) < 0  # too-small output is an error, not an overflow
    # the struct sizes above belong to one ABI version: the header's macro and the library agree
    hdr = open(os.path.join(ROOT, 'include', 'covasim_amd.h')).read()
    version = int(re.search(r'#define CVK_ABI_VERSION (\d+)', hdr).group(1))
    assert L.cvk_version() == version == 9
    assert L.cvk_source_hash() != 0  # set by the build script


def test_host_side_argument_errors():
    """Invalid arguments fail before any device work (no GPU needed)."""
    from covasim_amd import _lib
    L = _lib.lib()
    assert L.cvk_viral_load(0, None, None, None, 10, 0.3, 2.0, 4.0, None, None) == _lib.E_INVALID
    assert b'null' in L.cvk_last_error()
    cnt = C.c_int64(7)
    assert L.cvk_infections(0.1, None, None, None, 5, None, None, 5, None, 0, 0, 0, 0, 0, None, None, None, 0,
                            C.byref(cnt), None) == _lib.E_INVALID
    assert cnt.value == 0
    assert L.cvk_agent_pass(None, None, None, 0, None) == _lib.E_INVALID
    w = _lib.WorkS()
    sizes = [C.c_int64() for _ in range(5)]
    assert L.cvk_workspace_sizes(1000, 61, *[C.byref(s) for s in sizes]) == 0
    assert [s.value for s in sizes] == [8000, 128, 128, 61 * _lib.R_COLS * 4, 61 * _lib.RF_COLS * 8]
    assert L.cvk_workspace_sizes(0, 61, *[C.byref(s) for s in sizes]) == _lib.E_INVALID
    assert L.cvk_graph_launch(None, None) == _lib.E_INVALID
    with pytest.raises(ValueError):
        _lib.check(L.cvk_agent_pass(None, None, None, 0, None), 'agent pass')


def test_diag_argument_errors():
    """cvk_diag_apply / cvk_states_post reject missing state, bad ops and days before any device work."""
    from covasim_amd import _lib
    L = _lib.lib()
    P, prm, w, d, op = _lib.PeopleS(), _lib.Params(), _lib.WorkS(), _lib.DiagS(), _lib.DiagOp()
    assert L.cvk_states_post(None, None, None, None, 0, None) == _lib.E_INVALID
    assert L.cvk_diag_apply(C.byref(P), None, C.byref(prm), C.byref(w), C.byref(d), C.byref(op), 0,
                            None) == _lib.E_INVALID
    assert b'people arrays' in L.cvk_last_error()


def test_diag_coarse_map_sizes():
    """The coarse case map is the smallest power-of-two grouping that fits 2048 words (8 KiB of LDS)."""
    from covasim_amd import _lib
    L = _lib.lib()
    for n in [1, 63, 64, 20_000, 65_536, 65_537, 150_000, 1_000_000, 5_000_000, 10_000_000, 2**31 - 1]:
        words = L.cvk_diag_coarse_words(n)
        shift = next(s for s in range(32) if -(-n // (1 << s)) <= 2048 * 32)
        assert words == -(-(-(-n // (1 << shift))) // 32) and 1 <= words <= 2048, n
    assert L.cvk_diag_coarse_words(0) == 0


def test_population_argument_errors():
    """The cvk_pop_* entry points reject bad tables, sizes and workspaces before any device work."""
    from covasim_amd import _lib
    L = _lib.lib()
    g = _lib.PopGen()
    assert L.cvk_pop_ages(C.byref(g), 10, None, None, None) == _lib.E_INVALID  # zero-row tables
    g.n_age_bins, g.n_hh_sizes, g.n_prog_bins, g.hh_size_min, g.correlate = 19, 7, 10, 1, 1
    assert L.cvk_pop_ages(C.byref(g), 10, None, None, None) == _lib.E_INVALID  # null outputs
    assert L.cvk_pop_ages(C.byref(g), -1, None, None, None) == _lib.E_INVALID
    assert L.cvk_pop_work_bytes(1000, 500) > 0 and L.cvk_pop_work_bytes(-1, 5) == 0
    totals = C.c_void_p(256)  # never dereferenced: the call fails on its workspace first
    key = C.c_void_p(512)
    assert L.cvk_pop_households_plan(C.byref(g), 1000, key, 500, C.c_void_p(4096 + 8), 1 << 20, totals,
                                     None) == _lib.E_ALIGN
    assert L.cvk_pop_households_plan(C.byref(g), 1000, key, 500, C.c_void_p(4096), 64, totals,
                                     None) == _lib.E_CAPACITY
    assert L.cvk_pop_households_plan(C.byref(g), 1000, None, 500, C.c_void_p(4096), 1 << 20, totals,
                                     None) == _lib.E_INVALID  # correlate needs sort keys
    c = _lib.PopCounts(kind=_lib.POP_COUNT_POISSON, p0=0.0, lam=800.0)
    assert L.cvk_pop_contacts_plan(C.byref(g), 0, 100, C.byref(c), C.c_void_p(256), C.c_void_p(4096), 1 << 20,
                                   None) == _lib.E_INVALID
    assert b'p0' in L.cvk_last_error()
    assert L.cvk_pop_contacts_fill(C.byref(g), 300, 100, None, C.c_void_p(256), None, None, None, 10,
                                   None) == _lib.E_INVALID  # slot above 255
    g.nb_p0 = 0.0
    assert L.cvk_pop_people(C.byref(g), 10, C.c_void_p(256), *([C.c_void_p(256)] * 7), None, 0, None) == _lib.E_INVALID


def test_day_plan_generate_dynamic_switch():
    """The default generate-dynamic mode (2) generates dynamic-layer endpoints from 2^25 dynamic
    edges up and writes them below; 0 and 1 force either way; testing and tracing always write."""
    from covasim_amd import _lib
    L = _lib.lib()
    prm = _lib.Params()
    prm.n_layers = 2
    layers = (_lib.LayerS * 2)()
    layers[0].n_edges, layers[0].dynamic = 5_000_000, 0
    layers[1].dynamic = 1
    out = (C.c_int32 * 2)()

    def plan(n_dyn, days=10, imports=None):
        layers[1].n_edges = n_dyn
        _lib.check(L.cvk_day_plan(layers, C.byref(prm), imports, days, out), 'plan')
        return tuple(out)
    try:
        assert plan(2 ** 25 - 1) == (0, 1)  # written, with the next day's resample beside the infect pass
        assert plan(2 ** 25) == (1, 0)      # generated
        assert plan(2 ** 25, days=1) == (1, 0)
        assert plan(2 ** 25 - 1, days=1) == (0, 0)  # a single day has no next day to resample
        assert plan(0) == (0, 0)             # no dynamic edges at all
        diag = _lib.DiagS()
        imp = _lib.ImportsS()
        imp.diag = C.pointer(diag)
        assert plan(2 ** 26, imports=C.byref(imp)) == (0, 1)  # tracing reads each day's contacts
        _lib.check(L.cvk_set_tuning(_lib.TUNE_GENERATE_DYNAMIC, 1), 'tuning')
        assert plan(1000) == (1, 0)
        _lib.check(L.cvk_set_tuning(_lib.TUNE_GENERATE_DYNAMIC, 0), 'tuning')
        assert plan(2 ** 27) == (0, 1)
        _lib.check(L.cvk_set_tuning(_lib.TUNE_OVERLAP_RESAMPLE, 0), 'tuning')
        assert plan(2 ** 27) == (0, 0)
    finally:
        _lib.check(L.cvk_set_tuning(_lib.TUNE_GENERATE_DYNAMIC, 2), 'tuning')
        _lib.check(L.cvk_set_tuning(_lib.TUNE_OVERLAP_RESAMPLE, 1), 'tuning')
    assert L.cvk_day_plan(layers, None, None, 1, out) == _lib.E_INVALID


def test_noalloc_limit_matches_header():
    from covasim_amd import _lib
    hdr = open(HEADER).read()
    blocks = int(re.search(r'#define CVK_INFECTIONS_NOALLOC_MAX_BLOCKS (\d+)', hdr).group(1))
    chunk = int(re.search(r'#define CVK_INFECTIONS_NOALLOC_MAX_SLOTS \(CVK_INFECTIONS_NOALLOC_MAX_BLOCKS \* (\d+)LL\)',
                          hdr).group(1))
    assert _lib.INFECTIONS_NOALLOC_MAX_SLOTS == blocks * chunk
