"""The C ABI of libsfft.so without a GPU: it loads, exports every symbol include/sfft.h
declares (and nothing the header lacks among the public entry points), keeps the reference's
POD layouts, validates parameters like the reference's asserts, and its host-side canonical
FFT description (pass split, twiddles) equals the oracle's bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
from helpers import N20_K1000, N27_K16000, SMALL_CONFIGS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sfft.h")


@pytest.fixture(scope="module")
def S():
    import sfft_amd
    return sfft_amd


def _header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text))
    keywords = {"if", "sizeof", "defined", "extern", "typedef", "struct", "return"}
    return {n for n in names if n not in keywords and not n.isupper()}


def test_library_exports_every_declared_symbol(S):
    from sfft_amd import _lib
    declared = _header_functions()
    assert set(_lib.EXPORTED_FUNCTIONS) == declared, declared ^ set(_lib.EXPORTED_FUNCTIONS)
    for name in declared:
        assert hasattr(S.lib, name), name
    for name in _lib.EXPORTED_DATA:
        assert C.c_int.in_dll(S.lib, name) is not None


def test_pod_layouts(S):
    from sfft_amd import _lib
    assert C.sizeof(_lib.complex_t) == 16
    assert C.sizeof(_lib.Node) == 24 and _lib.Node.value.offset == 8
    assert C.sizeof(_lib.Params) == 12 * 4
    assert S.NODE_DTYPE.itemsize == 24


def test_globals_defaults(S):
    assert C.c_int.in_dll(S.lib, "TIMING").value == 1       # outerLoop.cu:51
    assert C.c_int.in_dll(S.lib, "VERBOSE").value == 0
    assert C.c_int.in_dll(S.lib, "ALGORITHM1").value == 1


def test_timesmod(S):
    rng = np.random.default_rng(0)
    for _ in range(100):
        n = 1 << int(rng.integers(4, 31))
        x, a = int(rng.integers(0, n)), int(rng.integers(0, n))
        assert S.lib.timesmod(x, a, n) == (x * a) % n


def test_params_match_oracle(S):
    for cfg in SMALL_CONFIGS + [N20_K1000, N27_K16000, (18, 2048, 1.0, 2, 2, 5, 256, 2)]:
        a = S.derive_params(*cfg)
        b = O.derive_params(*cfg)
        for key in ("n", "B", "B_thresh", "W_RLFC", "w", "b_f", "BB", "lobefrac", "t0", "loops"):
            assert a[key] == b[key], (cfg, key)
        assert bool(a["assert_ok"]) == bool(b["assert_ok"])


def _valid():
    return O.derive_params(*SMALL_CONFIGS[1])


def test_validation_mirrors_reference_asserts(S):
    p = _valid()
    assert S.validate(p) == 0
    bad = []
    q = dict(p); q["B_thresh"] = q["B"]; bad.append(q)            # experiment.c:83
    q = dict(p); q["B_thresh"] = q["W_RLFC"]; bad.append(q)       # experiment.c:82
    q = dict(p); q["loops_thresh"] = q["loops_loc"] + 1; bad.append(q)   # experiment.c:84
    q = dict(p); q["w"] = q["n"] + 1; bad.append(q)               # cuExperiment.cu:283
    q = dict(p); q["n"] = q["n"] + 2; bad.append(q)               # not a power of two
    q = dict(p); q["B"] = 3 * q["B"] // 2; bad.append(q)
    q = dict(p); q["w"] = q["B"] - 1; bad.append(q)               # tiled tail would overrun
    for q in bad:
        assert S.validate(q) == 1, q
    assert S.validate(p, fold_mode=7) == 1
    assert S.validate(p, hit_capacity=-1) == 1
    assert S.validate(p, front_end=4) == 1 and S.validate(p, front_end="stream") == 0
    assert S.validate(p, front_end="columns") == 0


def test_plan_without_device_reports_status(S):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(S.SfftError) as e:
        _plan_raw(S, _valid())
    assert e.value.status == 6


def _plan_raw(S, prm):
    from sfft_amd import _lib
    p = _lib.Params(n=prm["n"], B=prm["B"], B_thresh=prm["B_thresh"], W_RLFC=prm["W_RLFC"],
                    RLFC_loops=prm["RLFC_loops"], loops_thresh=prm["loops_thresh"],
                    loops_loc=prm["loops_loc"], loops=prm["loops"], w=prm["w"], fold_mode=0,
                    hit_capacity=0, front_end=0)
    h = C.c_void_p()
    st = S.lib.sfft_plan_create(C.byref(p), 0, None, C.byref(h))
    if st != 0:
        raise S.SfftError(st)
    S.lib.sfft_plan_destroy(h)


def test_fft_pass_split_matches_oracle(S):
    for l2 in range(0, 31):
        assert S.fft_plan_passes(l2) == O.fft_plan(l2)


def test_twiddles_match_oracle_bitwise(S):
    rng = np.random.default_rng(1)
    for l2 in range(0, 31):
        T = 1 << l2
        es = {0, T - 1, T // 2, T // 4, T // 8} | set(int(v) for v in rng.integers(0, T, 64))
        for e in es:
            if e >= T:
                continue
            for inv in (False, True):
                a = S.twiddle(e, T, inv)
                b = O.twiddle(e, T, inv)
                assert (a.real, a.imag) == (b.real, b.imag) and \
                    np.signbit(a.imag) == np.signbit(b.imag), (e, T, inv)


def test_filter_taps_match(S):
    for cfg in SMALL_CONFIGS + [N20_K1000, N27_K16000]:
        p = O.derive_params(*cfg)
        assert S.filter_taps(p["lobefrac"], p["tolerance"]) == p["w"]


class _RandomData(C.Structure):   # glibc struct random_data
    _fields_ = [("fptr", C.c_void_p), ("rptr", C.c_void_p), ("state", C.c_void_p),
                ("rand_type", C.c_int), ("rand_deg", C.c_int), ("rand_sep", C.c_int),
                ("end_ptr", C.c_void_p)]


def _glibc_random(seed, count):
    """count draws of glibc's random() generator from srandom(seed), on a fresh private
    state (random_r), as a process that never touched the global state would draw them"""
    libc = C.CDLL("libc.so.6")
    buf = C.create_string_buffer(128)
    rd = _RandomData()
    assert libc.initstate_r(C.c_uint(seed), buf, C.c_size_t(128), C.byref(rd)) == 0
    out, r = [], C.c_int32()
    for _ in range(count):
        libc.random_r(C.byref(rd), C.byref(r))
        out.append(r.value)
    return out


def test_permutation_generator_is_glibc_random(S):
    """outer_loop's permutation draws come from glibc's random() generator on a private state
    that starts at seed 1 (a program that never calls srandom()); 1804289383 is the first
    value such a program's random() returns."""
    S.lib.sfft_srandom(1)
    got = [S.lib.sfft_random() for _ in range(50)]
    assert got[0] == 1804289383
    assert got == _glibc_random(1, 50)
    S.lib.sfft_srandom(8765)
    assert [S.lib.sfft_random() for _ in range(50)] == _glibc_random(8765, 50)
    S.lib.sfft_srandom(1)


def test_status_strings(S):
    for s in range(7):
        assert S.lib.sfft_status_string(s)


def test_make_filter_ex_rejects_bad_arguments(S):
    """argument checks of sfft_make_filter_ex run before any device call"""
    import ctypes as C
    d = C.c_void_p(16)                      # never dereferenced: every call fails validation
    L = S.lib
    assert L.sfft_make_filter_ex(1e-3, 1e-8, 1 << 14, 100, 301, 2, d, None, None, None) == 1
    assert L.sfft_make_filter_ex(1e-3, 1e-8, 2048, 100, 301, 1, d, None, None, None) == 1
    assert L.sfft_make_filter_ex(1e-3, 1e-8, 3000, 100, 301, 0, d, None, None, None) == 1
    assert L.sfft_make_filter_ex(1e-3, 1e-8, 1 << 14, 0, 301, 1, d, None, None, None) == 1
    assert L.sfft_make_filter_ex(1e-3, 1e-8, 1 << 14, 100, 2, 1, d, None, None, None) == 1
    assert L.sfft_make_filter_ex(1e-3, 2.0, 1 << 14, 100, 301, 1, d, None, None, None) == 1
    assert L.sfft_make_filter_ex(1e-3, 1e-8, 1 << 14, 100, 301, 1, None, None, None, None) == 1
    assert L.sfft_filter_window_device(None, 1 << 14, 16, d, None) == 1
    assert L.sfft_filter_window_device(d, 1 << 14, 3, d, None) == 1
    with pytest.raises(ValueError):
        S.make_filter(1e-3, 1e-8, 1 << 14, 100, mode="fast")
