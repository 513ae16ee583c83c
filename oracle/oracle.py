"""ctypes wrapper for the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.  Every function takes and returns numpy
arrays; complex data is complex128 (or complex64 for the FP32 helpers), which has the
complex_t / float2 interleaved layout the C code expects.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SFFT_ORACLE_LIBRARY: another build of the same source (the sanitizer run's instrumented one)
_LIB_PATH = os.environ.get("SFFT_ORACLE_LIBRARY") or os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "-B", "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class Params(C.Structure):
    _fields_ = [("q", C.c_int), ("n", C.c_int), ("k", C.c_int), ("Bcst", C.c_double),
                ("loops_loc", C.c_int), ("loops_thresh", C.c_int), ("loops_est", C.c_int),
                ("loops", C.c_int), ("RLFC_cst", C.c_double), ("RLFC_loops", C.c_int),
                ("tolerance", C.c_double), ("BB", C.c_double), ("lobefrac", C.c_double),
                ("b_f", C.c_int), ("B", C.c_int), ("B_thresh", C.c_int), ("W_RLFC", C.c_int),
                ("w", C.c_int), ("t0", C.c_double), ("assert_ok", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Plan(C.Structure):
    _fields_ = [("n", C.c_int), ("B", C.c_int), ("B_thresh", C.c_int), ("W_RLFC", C.c_int),
                ("RLFC_loops", C.c_int), ("loops_thresh", C.c_int), ("loops_loc", C.c_int),
                ("loops", C.c_int), ("w", C.c_int), ("fold_mode", C.c_int),
                ("hit_capacity", C.c_int)]


NODE_DTYPE = np.dtype([("key", np.uint32), ("pad", np.uint32), ("re", np.float64),
                       ("im", np.float64)])

_p = C.c_void_p
_i = C.c_int
_d = C.c_double


def _declare(L):
    L.orc_derive_params.argtypes = [_i, _i, _d, _i, _i, _i, _i, _i, _d, C.POINTER(Params)]
    L.orc_floor_to_pow2.argtypes = [_d]
    L.orc_floor_to_pow2.restype = _i
    L.orc_gcd.argtypes = [_i, _i]
    L.orc_gcd.restype = _i
    L.orc_mod_inverse.argtypes = [_i, _i]
    L.orc_mod_inverse.restype = _i
    L.orc_timesmod.argtypes = [_i, _i, _i]
    L.orc_timesmod.restype = _i
    L.orc_fft_plan.argtypes = [_i, _p]
    L.orc_fft_plan.restype = _i
    L.orc_twiddle64.argtypes = [C.c_uint64, C.c_uint64, _i, _p]
    L.orc_fft64.argtypes = [_p, _i, _i, _p]
    L.orc_fft32.argtypes = [_p, _i, _i, _p]
    L.orc_dft_any64.argtypes = [_p, _i, _i]
    L.orc_filter_window.argtypes = [_d, _d, _i, _p, _p]
    L.orc_filter_window.restype = _i
    L.orc_make_multiple_t.argtypes = [_d, _d, _i, _i, _i, _p, _p]
    L.orc_make_multiple_t_taps.argtypes = [_d, _d, _i, _i, _i, _p, _p, _p, _p, _p]
    L.orc_perm_filter64.argtypes = [_p, _i, _p, _i, _i, _i, _i, _p]
    L.orc_perm_filter32.argtypes = [_p, _i, _p, _i, _i, _i, _i, _p]
    L.orc_samples64.argtypes = [_p, _i, _p]
    L.orc_samples32.argtypes = [_p, _i, _p]
    L.orc_cutoff64.argtypes = [_p, _i, _i, _p]
    L.orc_cutoff32.argtypes = [_p, _i, _i, _p]
    L.orc_rlfc64.argtypes = [_p, _i, _i, _i, _i, _p]
    L.orc_merge_J2.argtypes = [_p, _i]
    L.orc_merge_J2.restype = _i
    L.orc_reverse_hash.argtypes = [_p, _i, _i, _i, _i, _i, _i, _p, _p, _i, _i, _p, _i, _i]
    L.orc_reverse_hash.restype = _i
    L.orc_estimate_values.argtypes = [_p, _i, _p, _i, _i, _p, _i, _p, _p]
    L.orc_outer_loop64.argtypes = [C.POINTER(Plan), _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                   _p, _p, _p]
    L.orc_outer_loop64.restype = _i
    L.orc_outer_loop32.argtypes = [C.POINTER(Plan), _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                   _p, _p]
    L.orc_outer_loop32.restype = _i
    L.orc_rlfc32.argtypes = [_p, _i, _i, _i, _i, _p]
    L.orc_metric_dense.argtypes = [_p, _i, _i, _p, _i, _p, _p, _p, _p]


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.complex64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---- parameters ---------------------------------------------------------------------------
def derive_params(q, k, Bcst, loops_loc, loops_thresh, loops_est, RLFC_cst, RLFC_loops,
                  tolerance=1e-8):
    p = Params()
    lib().orc_derive_params(q, k, Bcst, loops_loc, loops_thresh, loops_est, int(RLFC_cst),
                            RLFC_loops, tolerance, C.byref(p))
    return p.as_dict()


def floor_to_pow2(x):
    return lib().orc_floor_to_pow2(float(x))


def gcd(a, b):
    return lib().orc_gcd(a, b)


def mod_inverse(a, n):
    return lib().orc_mod_inverse(a, n)


def timesmod(x, a, n):
    return lib().orc_timesmod(x, a, n)


# ---- FFT ----------------------------------------------------------------------------------
def fft_plan(log2n):
    out = np.zeros(8, dtype=np.int32)
    np_ = lib().orc_fft_plan(log2n, _ptr(out))
    return [int(v) for v in out[:np_]]


def twiddle(e, T, inverse=False):
    out = np.zeros(1, dtype=np.complex128)
    lib().orc_twiddle64(e, T, int(inverse), _ptr(out))
    return out[0]


def fft64(x, inverse=False):
    x = _c128(x).copy()
    n = x.size
    l2 = n.bit_length() - 1
    assert 1 << l2 == n
    scr = np.empty_like(x)
    lib().orc_fft64(_ptr(x), l2, int(inverse), _ptr(scr))
    return x


def fft32(x, inverse=False):
    x = _c64(x).copy()
    n = x.size
    l2 = n.bit_length() - 1
    assert 1 << l2 == n
    scr = np.empty_like(x)
    lib().orc_fft32(_ptr(x), l2, int(inverse), _ptr(scr))
    return x


def dft_any64(x, inverse=False):
    x = _c128(x).copy()
    lib().orc_dft_any64(_ptr(x), x.size, int(inverse))
    return x


# ---- filter -------------------------------------------------------------------------------
def filter_window(lobefrac, tolerance, n):
    w = C.c_int()
    t0 = C.c_double()
    ok = lib().orc_filter_window(lobefrac, tolerance, n, C.byref(w), C.byref(t0))
    return w.value, t0.value, bool(ok)


def make_multiple_t(lobefrac, tolerance, n, b_f, w):
    ft = np.zeros(w, dtype=np.complex128)
    ff = np.zeros(n, dtype=np.complex128)
    lib().orc_make_multiple_t(lobefrac, tolerance, n, b_f, w, _ptr(ft), _ptr(ff))
    return ft, ff


def make_multiple_t_taps(lobefrac, tolerance, n, b_f, w):
    ft = np.zeros(w, dtype=np.complex128)
    ff = np.zeros(n, dtype=np.complex128)
    v = np.zeros(w, dtype=np.complex128)
    V = np.zeros(w, dtype=np.complex128)
    G = np.zeros(n, dtype=np.complex128)
    lib().orc_make_multiple_t_taps(lobefrac, tolerance, n, b_f, w, _ptr(ft), _ptr(ff), _ptr(v),
                                   _ptr(V), _ptr(G))
    return ft, ff, v, V, G


def filter_window_slice(filter_freq, n, B):
    """win[j] = filter_freq[(j - m/2) mod n], j < m = n/B (the only entries estimation reads)."""
    m = n // B
    idx = (np.arange(m, dtype=np.int64) - m // 2) % n
    return np.ascontiguousarray(filter_freq[idx])


# ---- stages -------------------------------------------------------------------------------
FOLD_REFERENCE = 0
FOLD_ALL_TAPS = 1


def perm_filter64(x, n, f, B, ai, fold_mode=FOLD_REFERENCE):
    x = _c128(x)
    f = _c128(f)
    out = np.zeros(B, dtype=np.complex128)
    lib().orc_perm_filter64(_ptr(x), n, _ptr(f), f.size, B, ai, fold_mode, _ptr(out))
    return out


def perm_filter32(x, n, f, B, ai, fold_mode=FOLD_REFERENCE):
    x = _c64(x)
    f = _c64(f)
    out = np.zeros(B, dtype=np.complex64)
    lib().orc_perm_filter32(_ptr(x), n, _ptr(f), f.size, B, ai, fold_mode, _ptr(out))
    return out


def samples64(X):
    X = _c128(X)
    s = np.empty(X.size, dtype=np.float64)
    lib().orc_samples64(_ptr(X), X.size, _ptr(s))
    return s


def samples32(X):
    X = _c64(X)
    s = np.empty(X.size, dtype=np.float32)
    lib().orc_samples32(_ptr(X), X.size, _ptr(s))
    return s


def cutoff64(s, K):
    s = np.ascontiguousarray(s, dtype=np.float64)
    out = np.empty(K, dtype=np.int32)
    lib().orc_cutoff64(_ptr(s), s.size, K, _ptr(out))
    return out


def cutoff32(s, K):
    s = np.ascontiguousarray(s, dtype=np.float32)
    out = np.empty(K, dtype=np.int32)
    lib().orc_cutoff32(_ptr(s), s.size, K, _ptr(out))
    return out


def rlfc64(x, n, W, offset, B_thresh):
    x = _c128(x)
    out = np.empty(B_thresh, dtype=np.int32)
    lib().orc_rlfc64(_ptr(x), n, W, offset, B_thresh, _ptr(out))
    return out


def merge_J2(J2):
    J2 = _i32(J2).copy()
    c = lib().orc_merge_J2(_ptr(J2), J2.size)
    return J2[:c].copy()


def reverse_hash(J, n, B, a, ai, loop_thresh, vote, hits, hit_count, J2, W,
                 hit_capacity=None):
    """vote (int32[n]) and hits (int32[capacity]) are updated in place; returns new count."""
    J = _i32(J)
    J2 = _i32(J2)
    assert vote.dtype == np.int32 and hits.dtype == np.int32
    cap = hits.size if hit_capacity is None else hit_capacity
    return lib().orc_reverse_hash(_ptr(J), J.size, n, B, a, ai, loop_thresh, _ptr(vote),
                                  _ptr(hits), hit_count, cap, _ptr(J2), J2.size, W)


def estimate_values(hits, bins_f, loops, n, ai, B, filter_win):
    hits = _i32(hits)
    bins_f = _c128(bins_f)
    ai = _i32(ai)
    filter_win = _c128(filter_win)
    out = np.zeros(hits.size, dtype=NODE_DTYPE)
    lib().orc_estimate_values(_ptr(hits), hits.size, _ptr(bins_f), loops, n, _ptr(ai), B,
                              _ptr(filter_win), _ptr(out))
    return out


def outer_loop64(prm, x, filter_time, filter_win, a, ai, offsets, fold_mode=FOLD_REFERENCE,
                 hit_capacity=524288, want_bins=False, want_J=False):
    """The whole transform with explicit draws.  prm: dict with n, B, B_thresh, W_RLFC,
    RLFC_loops, loops_thresh, loops_loc, loops (and w from filter_time)."""
    pl = Plan(n=prm["n"], B=prm["B"], B_thresh=prm["B_thresh"], W_RLFC=prm["W_RLFC"],
              RLFC_loops=prm["RLFC_loops"], loops_thresh=prm["loops_thresh"],
              loops_loc=prm["loops_loc"], loops=prm["loops"], w=int(filter_time.size),
              fold_mode=fold_mode, hit_capacity=hit_capacity)
    x = _c128(x)
    ft = _c128(filter_time)
    fw = _c128(filter_win)
    a = _i32(a)
    ai = _i32(ai)
    offsets = _i32(offsets)
    hits = np.zeros(hit_capacity, dtype=np.int32)
    nodes = np.zeros(hit_capacity, dtype=NODE_DTYPE)
    bins = np.zeros(prm["loops"] * prm["B"], dtype=np.complex128) if want_bins else None
    J = np.zeros(prm["loops_loc"] * prm["B_thresh"], dtype=np.int32) if want_J else None
    J2 = np.zeros(prm["RLFC_loops"] * prm["B_thresh"], dtype=np.int32)
    nj2 = C.c_int()
    st = np.zeros(7, dtype=np.float64)
    I_F = lib().orc_outer_loop64(C.byref(pl), _ptr(x), _ptr(ft), _ptr(fw), _ptr(a), _ptr(ai),
                                 _ptr(offsets), _ptr(hits), _ptr(nodes),
                                 _ptr(bins) if bins is not None else None,
                                 _ptr(J) if J is not None else None, _ptr(J2), C.byref(nj2),
                                 _ptr(st))
    res = {"I_F": I_F, "stage_seconds": st, "J2": J2[:nj2.value].copy()}
    if I_F >= 0:
        res["hits"] = hits[:I_F].copy()
        res["nodes"] = nodes[:I_F].copy()
    if bins is not None:
        res["bins_f"] = bins.reshape(prm["loops"], prm["B"])
    if J is not None:
        res["J"] = J.reshape(prm["loops_loc"], prm["B_thresh"])
    return res


def outer_loop32(prm, x, filter_time, filter_win, a, ai, offsets, fold_mode=FOLD_REFERENCE,
                 hit_capacity=524288, want_bins=False, want_J=False):
    """The FP32 path with explicit draws (x and filter_time complex64)."""
    pl = Plan(n=prm["n"], B=prm["B"], B_thresh=prm["B_thresh"], W_RLFC=prm["W_RLFC"],
              RLFC_loops=prm["RLFC_loops"], loops_thresh=prm["loops_thresh"],
              loops_loc=prm["loops_loc"], loops=prm["loops"], w=int(filter_time.size),
              fold_mode=fold_mode, hit_capacity=hit_capacity)
    x = _c64(x)
    ft = _c64(filter_time)
    fw = _c128(filter_win)
    a, ai, offsets = _i32(a), _i32(ai), _i32(offsets)
    hits = np.zeros(hit_capacity, dtype=np.int32)
    nodes = np.zeros(hit_capacity, dtype=NODE_DTYPE)
    bins = np.zeros(prm["loops"] * prm["B"], dtype=np.complex64) if want_bins else None
    J = np.zeros(prm["loops_loc"] * prm["B_thresh"], dtype=np.int32) if want_J else None
    J2 = np.zeros(prm["RLFC_loops"] * prm["B_thresh"], dtype=np.int32)
    nj2 = C.c_int()
    I_F = lib().orc_outer_loop32(C.byref(pl), _ptr(x), _ptr(ft), _ptr(fw), _ptr(a), _ptr(ai),
                                 _ptr(offsets), _ptr(hits), _ptr(nodes),
                                 _ptr(bins) if bins is not None else None,
                                 _ptr(J) if J is not None else None, _ptr(J2), C.byref(nj2))
    res = {"I_F": I_F, "J2": J2[:nj2.value].copy()}
    if I_F >= 0:
        res["hits"] = hits[:I_F].copy()
        res["nodes"] = nodes[:I_F].copy()
    if bins is not None:
        res["bins_f"] = bins.reshape(prm["loops"], prm["B"])
    if J is not None:
        res["J"] = J.reshape(prm["loops_loc"], prm["B_thresh"])
    return res


def metric_dense(nodes, k, large_freq, n, x_f):
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE).copy()
    lf = _i32(large_freq)
    xf = _c128(x_f)
    me = C.c_int()
    mr = C.c_int()
    err = C.c_double()
    lib().orc_metric_dense(_ptr(nodes), nodes.size, k, _ptr(lf), n, _ptr(xf), C.byref(me),
                           C.byref(mr), C.byref(err))
    return me.value, mr.value, err.value


# ---- value map and its adjoint (project addition) -------------------------------------------
def _plan_of(prm, w, fold_mode):
    return Plan(n=prm["n"], B=prm["B"], B_thresh=prm["B_thresh"], W_RLFC=prm["W_RLFC"],
                RLFC_loops=prm["RLFC_loops"], loops_thresh=prm["loops_thresh"],
                loops_loc=prm["loops_loc"], loops=prm["loops"], w=int(w),
                fold_mode=fold_mode, hit_capacity=524288)


def estimate_select(hits, bins_f, loops, n, ai, B, filter_win):
    """uint8[nh, 2]: the loops whose values the real / imaginary lower medians returned"""
    hits = _i32(hits)
    sel = np.zeros(2 * hits.size, dtype=np.uint8)
    lib().orc_estimate_select(_ptr(hits), hits.size, _ptr(_c128(np.ravel(bins_f))), loops, n,
                              _ptr(_i32(ai)), B, _ptr(_c128(filter_win)), _ptr(sel))
    return sel.reshape(-1, 2)


def value_apply64(prm, x, filter_time, filter_win, ai, hits, sel,
                  fold_mode=FOLD_REFERENCE):
    hits = _i32(hits)
    sel = np.ascontiguousarray(sel, dtype=np.uint8).ravel()
    out = np.zeros(hits.size, dtype=np.complex128)
    pl = _plan_of(prm, np.size(filter_time), fold_mode)
    lib().orc_value_apply64(C.byref(pl), _ptr(_c128(x)), _ptr(_c128(filter_time)),
                            _ptr(_c128(filter_win)), _ptr(_i32(ai)), _ptr(hits), hits.size,
                            _ptr(sel), _ptr(out))
    return out


def value_adjoint64(prm, filter_time, filter_win, ai, hits, sel, g,
                    fold_mode=FOLD_REFERENCE):
    hits = _i32(hits)
    sel = np.ascontiguousarray(sel, dtype=np.uint8).ravel()
    gx = np.zeros(prm["n"], dtype=np.complex128)
    pl = _plan_of(prm, np.size(filter_time), fold_mode)
    lib().orc_value_adjoint64(C.byref(pl), _ptr(_c128(filter_time)), _ptr(_c128(filter_win)),
                              _ptr(_i32(ai)), _ptr(hits), hits.size, _ptr(sel),
                              _ptr(_c128(g)), _ptr(gx))
    return gx
