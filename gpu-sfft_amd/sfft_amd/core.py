"""Host-side mirror of the reference interface over the C ABI.

Device memory is held in torch tensors (complex128 = complex_t, complex64 = float2); only
raw pointers cross into libsfft.  Parameter derivation follows experiment.c:54-84 and
cuExperiment.cu:279-284; the transform follows outerLoop.cu:65-190 with explicit draws.
"""
import ctypes as C
import math

import numpy as np

from . import _lib as L
from ._lib import lib

NODE_DTYPE = np.dtype([("key", np.uint32), ("pad", np.uint32), ("re", np.float64),
                       ("im", np.float64)])
assert NODE_DTYPE.itemsize == C.sizeof(L.Node) == 24

FOLD_REFERENCE = L.SFFT_FOLD_REFERENCE
FOLD_ALL_TAPS = L.SFFT_FOLD_ALL_TAPS
FRONT = {"auto": L.SFFT_FRONT_AUTO, "gather": L.SFFT_FRONT_GATHER, "stream": L.SFFT_FRONT_STREAM,
         "columns": L.SFFT_FRONT_COLUMNS}


class SfftError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = lib.sfft_status_string(status).decode()
        super().__init__(f"{what}: {msg} (status {status})" if what else msg)


def _check(status, what):
    if status != L.SFFT_OK:
        raise SfftError(status, what)


def _torch():
    import torch
    return torch


def _stream_handle(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _np_ptr(a):
    return a.ctypes.data_as(C.c_void_p)


# ---- parameters (experiment.c:54-84) -------------------------------------------------------
def floor_to_pow2(x):
    ans = 1
    while ans <= x:
        ans <<= 1
    return ans // 2


def derive_params(q, k, Bcst, loops_loc, loops_thresh, loops_est, RLFC_cst, RLFC_loops,
                  tolerance=1e-8):
    """The reference's parameter derivation (experiment.c:54-84; w: cuExperiment.cu:279-284)."""
    n = floor_to_pow2(int(math.pow(2, q)))
    RLFC_cst = float(int(RLFC_cst))                      # atoi (experiment.c:62)
    BB = float(int(Bcst * math.sqrt(float(n) * k / math.log2(n))))   # (unsigned) cast
    lobefrac = 0.5 / BB
    b_f = int(1.2 * 1.1 * (float(n) / BB))
    B = floor_to_pow2(BB)
    B_thresh = 2 * k
    W = floor_to_pow2(RLFC_cst * n / B)
    t0 = C.c_double()
    w = lib.sfft_filter_taps(lobefrac, tolerance, C.byref(t0))
    ok = (b_f <= n) and (B_thresh < W) and (B_thresh < B) and (loops_thresh <= loops_loc) \
        and (w <= n)
    return dict(q=q, n=n, k=k, Bcst=Bcst, loops_loc=loops_loc, loops_thresh=loops_thresh,
                loops_est=loops_est, loops=loops_loc + loops_est, RLFC_cst=RLFC_cst,
                RLFC_loops=RLFC_loops, tolerance=tolerance, BB=BB, lobefrac=lobefrac, b_f=b_f,
                B=B, B_thresh=B_thresh, W_RLFC=W, w=w, t0=t0.value, assert_ok=ok)


def draw_permutations(n, loops, RLFC_loops, W, seed):
    """Seeded draws: a odd uniform on [1, n), ai = a^-1 mod n, comb offsets on [0, n/W)."""
    rng = np.random.default_rng(seed)
    a = (rng.integers(0, n // 2, size=loops, dtype=np.int64) * 2 + 1).astype(np.int64)
    ai = np.array([pow(int(v), -1, n) for v in a], dtype=np.int64)
    offs = rng.integers(0, n // W, size=RLFC_loops, dtype=np.int64)
    return a.astype(np.int32), ai.astype(np.int32), offs.astype(np.int32)


def validate(params, fold_mode=FOLD_REFERENCE, hit_capacity=0, front_end="auto"):
    p = _cparams(params, fold_mode, hit_capacity, front_end)
    return lib.sfft_params_validate(C.byref(p))


def _cparams(params, fold_mode, hit_capacity, front_end="auto"):
    fe = FRONT[front_end] if isinstance(front_end, str) else int(front_end)
    return L.Params(n=params["n"], B=params["B"], B_thresh=params["B_thresh"],
                    W_RLFC=params["W_RLFC"], RLFC_loops=params["RLFC_loops"],
                    loops_thresh=params["loops_thresh"], loops_loc=params["loops_loc"],
                    loops=params["loops"], w=params["w"], fold_mode=fold_mode,
                    hit_capacity=hit_capacity, front_end=fe)


# ---- plan ------------------------------------------------------------------------------------
class Plan:
    """One transform configuration on the current device (outer_loop with explicit draws)."""

    def __init__(self, params, precision="fp64", fold_mode=FOLD_REFERENCE, hit_capacity=0,
                 stream=None, front_end="auto"):
        self.params = dict(params)
        self.precision = precision
        self.fp32 = precision == "fp32"
        self.stream = stream
        self._cp = _cparams(params, fold_mode, hit_capacity, front_end)
        self.capacity = hit_capacity if hit_capacity > 0 else 524288
        h = C.c_void_p()
        _check(lib.sfft_plan_create(C.byref(self._cp), L.SFFT_FP32 if self.fp32 else L.SFFT_FP64,
                                    _stream_handle(stream), C.byref(h)), "sfft_plan_create")
        self._h = h
        self._ft = None
        # pinned host buffer: the node copy at the end of execute() is a direct DMA
        torch = _torch()
        self._nodes_buf = torch.empty(self.capacity * NODE_DTYPE.itemsize, dtype=torch.uint8,
                                      pin_memory=True)
        self._nodes = self._nodes_buf.numpy().view(NODE_DTYPE)
        self._nodes_vp = C.c_void_p(self._nodes.ctypes.data)   # fixed for the plan's life
        self._draw_cache = {}

    def close(self):
        if getattr(self, "_h", None):
            lib.sfft_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def workspace_bytes(self):
        return lib.sfft_plan_workspace_bytes(self._h)

    def set_filter(self, filter_time, filter_win):
        """filter_time: device tensor (complex128 for fp64 plans, complex64 for fp32) of w taps.
        filter_win: numpy complex128[m], win[j] = filter_freq[(j - m/2) mod n]."""
        torch = _torch()
        want = torch.complex64 if self.fp32 else torch.complex128
        if filter_time.dtype != want or not filter_time.is_cuda:
            raise TypeError(f"filter_time must be a {want} device tensor")
        if filter_time.numel() != self.params["w"]:
            raise ValueError("filter_time has the wrong number of taps")
        win = np.ascontiguousarray(filter_win, dtype=np.complex128)
        if win.size != self.params["n"] // self.params["B"]:
            raise ValueError("filter window must hold n/B entries")
        self._ft = filter_time
        self._win = win
        _check(lib.sfft_plan_set_filter(self._h, _ptr(filter_time), _np_ptr(win)),
               "sfft_plan_set_filter")

    def set_profiling(self, on=True):
        _check(lib.sfft_plan_set_profiling(self._h, int(on)), "sfft_plan_set_profiling")

    def set_graph(self, on=True):
        """Capture the transform into a hipGraph on first use and replay it afterwards."""
        _check(lib.sfft_plan_set_graph(self._h, int(on)), "sfft_plan_set_graph")

    def stage_times(self):
        out = np.zeros(7, dtype=np.float64)
        _check(lib.sfft_plan_stage_times(self._h, _np_ptr(out)), "sfft_plan_stage_times")
        return dict(zip(["votetable", "restr-loc-index", "perm+filter", "reverse-hash",
                         "estimation", "FFT+Cutoff", "total"], out.tolist()))

    def interval_times(self):
        out = np.zeros(7, dtype=np.float64)
        _check(lib.sfft_plan_interval_times(self._h, _np_ptr(out)), "sfft_plan_interval_times")
        return dict(zip(["clear", "stream", "comb", "fold", "bucket_fft_cutoff", "hash",
                         "estimate"], out.tolist()))

    def _draws(self, a, ai, offsets):
        """The draws struct (pointers to int32 host arrays, read by the library at each call).
        When the caller's own int32 arrays are passed, the struct is kept for those array
        objects (holding them, so their ids stay theirs): taking an array's address costs a
        few microseconds per array and call."""
        key = (id(a), id(ai), id(offsets))
        hit = self._draw_cache.get(key)
        if hit is not None:
            return hit[3]
        self._a = np.ascontiguousarray(a, dtype=np.int32)
        self._ai = np.ascontiguousarray(ai, dtype=np.int32)
        self._off = np.ascontiguousarray(offsets, dtype=np.int32)
        dr = L.Draws(self._a.ctypes.data, self._ai.ctypes.data, self._off.ctypes.data)
        if self._a is a and self._ai is ai and self._off is offsets:
            if len(self._draw_cache) >= 64:
                self._draw_cache.clear()
            self._draw_cache[key] = (a, ai, offsets, dr)
        return dr

    def _check_x(self, x):
        torch = _torch()
        want = torch.complex64 if self.fp32 else torch.complex128
        if x.dtype != want or not x.is_cuda or x.numel() != self.params["n"] or \
                not x.is_contiguous():
            raise TypeError(f"x must be a contiguous {want} device tensor of n entries")

    def execute(self, x, a, ai, offsets, copy=True, sync_inputs=True):
        """Runs one transform and returns (nodes[I_F] numpy structured array, I_F).  With
        copy=False the nodes are a view of the plan's pinned buffer (valid until the next
        execute).  sync_inputs: wait for pending device work first (the plan may run on a
        stream of its own); False when x is known to be ready."""
        self._check_x(x)
        if sync_inputs:
            _torch().cuda.synchronize()
        dr = self._draws(a, ai, offsets)
        cnt = C.c_int()
        st = lib.sfft_execute(self._h, _ptr(x), C.byref(dr), self._nodes_vp, C.byref(cnt))
        _check(st, "sfft_execute")
        nodes = self._nodes[:cnt.value]
        return (nodes.copy() if copy else nodes), cnt.value

    def execute_async(self, x, a, ai, offsets, d_nodes=None, d_count=None):
        """Enqueues one transform on the plan stream; results stay on the device."""
        self._check_x(x)
        dr = self._draws(a, ai, offsets)
        st = lib.sfft_execute_async(self._h, _ptr(x), C.byref(dr),
                                    _ptr(d_nodes) if d_nodes is not None else None,
                                    _ptr(d_count) if d_count is not None else None)
        _check(st, "sfft_execute_async")

    def selection(self, count):
        """Device copies of the last transform's located indices (int32[count]) and
        median-selected loops (uint8[count, 2]: real part, imaginary part)."""
        torch = _torch()
        hits = torch.empty(max(count, 1), dtype=torch.int32, device="cuda")[:count]
        sel = torch.empty((max(count, 1), 2), dtype=torch.uint8, device="cuda")[:count]
        torch.cuda.synchronize()
        for t, which, nb in ((hits, 3, 4 * count), (sel, 7, 2 * count)):
            if nb:
                err = _hip().hipMemcpy(C.c_void_p(t.data_ptr()),
                                       C.c_void_p(lib.sfft_plan_debug_buffer(self._h, which)),
                                       C.c_size_t(nb), 3)
                if err != 0:
                    raise RuntimeError(f"hipMemcpy failed ({err})")
        return hits, sel

    def _value_args(self, hits, sel):
        torch = _torch()
        if hits.dtype != torch.int32 or sel.dtype != torch.uint8 or not hits.is_cuda or \
                not sel.is_cuda or sel.numel() != 2 * hits.numel():
            raise TypeError("hits must be int32 and sel uint8[nh, 2] device tensors")
        return hits.contiguous(), sel.contiguous()

    def value_apply(self, x, a, ai, offsets, hits, sel):
        """The estimated values of the fixed hits/selections as a function of x (FP64 plans):
        complex128[nh] device tensor."""
        torch = _torch()
        self._check_x(x)
        hits, sel = self._value_args(hits, sel)
        out = torch.empty(max(hits.numel(), 1), dtype=torch.complex128, device="cuda")
        dr = self._draws(a, ai, offsets)
        torch.cuda.synchronize()
        _check(lib.sfft_value_apply(self._h, _ptr(x), C.byref(dr), hits.numel(), _ptr(hits),
                                    _ptr(sel), _ptr(out)), "sfft_value_apply")
        torch.cuda.synchronize()
        return out[:hits.numel()]

    def value_adjoint(self, grad, a, ai, offsets, hits, sel):
        """The transpose of value_apply: grad (complex128[nh], dL/dRe + i dL/dIm of each
        value) -> complex128[n] device tensor (dL/dRe x + i dL/dIm x)."""
        torch = _torch()
        hits, sel = self._value_args(hits, sel)
        g = grad.to(torch.complex128).contiguous()
        if g.numel() != hits.numel() or not g.is_cuda:
            raise TypeError("grad must be a device tensor of nh entries")
        gx = torch.empty(self.params["n"], dtype=torch.complex128, device="cuda")
        dr = self._draws(a, ai, offsets)
        torch.cuda.synchronize()
        _check(lib.sfft_value_adjoint(self._h, _ptr(g) if g.numel() else None, C.byref(dr),
                                      hits.numel(), _ptr(hits), _ptr(sel), _ptr(gx)),
               "sfft_value_adjoint")
        torch.cuda.synchronize()
        return gx

    @property
    def front_end(self):
        """The front end this plan runs ("gather", "stream" or "columns"; AUTO resolved)."""
        return {v: k for k, v in FRONT.items()}[lib.sfft_plan_front_end(self._h)]

    def comb_key_order(self):
        """(lo, mid) of the comb keys' storage order (debug buffer 8): position s holds bin
        (s mod 2^lo) | ((s >> lo) mod 2^mid) << (log2 W - mid) | (s >> (lo + mid)) << lo."""
        lo, mid = C.c_int(), C.c_int()
        _check(lib.sfft_plan_comb_key_order(self._h, C.byref(lo), C.byref(mid)),
               "sfft_plan_comb_key_order")
        return lo.value, mid.value

    def debug_array(self, which, count, dtype):
        """Copies `count` items of a plan-owned device buffer to a numpy array
        (0 bins_f, 1 J, 2 J2 lists, 3 hits, 4 bins_t, 5 nodes, 6 count, 7 sel, 8 comb keys).
        bins_t is in bucket order, or in column order (bucket c * a_l mod B at c) after the
        column front end's slab fold."""
        torch = _torch()
        ptr = lib.sfft_plan_debug_buffer(self._h, which)
        out = np.empty(count, dtype=dtype)
        torch.cuda.synchronize()
        err = _hip().hipMemcpy(_np_ptr(out), C.c_void_p(ptr), C.c_size_t(out.nbytes), 2)
        if err != 0:
            raise RuntimeError(f"hipMemcpy failed ({err})")
        return out


# ---- batches of uneven signals -----------------------------------------------------------------
def batch_workload(count, seed, q_range=(18, 24), k_range=(50, 4000), Bcst=1.0, loops_loc=2,
                   loops_thresh=2, loops_est=5, RLFC_cst=256, RLFC_loops=2):
    """Per-signal parameters of a batch of uneven signals: q uniform on q_range (inclusive),
    k log-uniform on k_range.  A draw that violates the reference's asserts (experiment.c:
    77-84; e.g. q=18 gives B=4096, so k >= 2048 breaks B_thresh < B) or this library's plan
    limits is redrawn (k only, q kept)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        q = int(rng.integers(q_range[0], q_range[1] + 1))
        for _try in range(1000):
            k = int(round(math.exp(rng.uniform(math.log(k_range[0]), math.log(k_range[1])))))
            prm = derive_params(q, k, Bcst, loops_loc, loops_thresh, loops_est, RLFC_cst,
                                RLFC_loops)
            if prm["assert_ok"] and validate(prm) == 0:
                break
        else:
            raise ValueError(f"no k in {k_range} gives valid parameters at q={q}")
        out.append(prm)
    return out


class Batch:
    """Several signals of uneven n and k transformed together (sfft_batch_*): one persistent
    permute+filter launch for all signals, the other stages on worker streams, optional
    hipGraph capture and an autotuner over bit-equivalent launch shapes."""

    def __init__(self, params_list, precision="fp32", fold_mode=FOLD_REFERENCE,
                 front_end="gather"):
        """front_end: "gather" or "auto" (the batch always gathers); others are refused."""
        self.params = [dict(p) for p in params_list]
        self.count = len(self.params)
        self.fp32 = precision == "fp32"
        arr = (L.Params * self.count)(*[_cparams(p, fold_mode, 0, front_end)
                                       for p in self.params])
        h = C.c_void_p()
        _check(lib.sfft_batch_create(arr, self.count, L.SFFT_FP32 if self.fp32 else L.SFFT_FP64,
                                     None, C.byref(h)), "sfft_batch_create")
        self._h = h
        self._keep = {}
        self._stage = None     # pinned node staging of all signals (grows when needed)

    def close(self):
        if getattr(self, "_h", None):
            lib.sfft_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_filter(self, i, filter_time, filter_win):
        win = np.ascontiguousarray(filter_win, dtype=np.complex128)
        self._keep[("ft", i)] = filter_time
        _check(lib.sfft_batch_set_filter(self._h, i, _ptr(filter_time), _np_ptr(win)),
               "sfft_batch_set_filter")

    def set_graph(self, on=True):
        _check(lib.sfft_batch_set_graph(self._h, int(on)), "sfft_batch_set_graph")

    def set_shape(self, lpt, block, grid, qmax_bucket=0, qmax_comb=0):
        _check(lib.sfft_batch_set_shape(self._h, lpt, block, grid, qmax_bucket, qmax_comb),
               "sfft_batch_set_shape")

    def _args(self, xs, draws):
        torch = _torch()
        want = torch.complex64 if self.fp32 else torch.complex128
        for x, p in zip(xs, self.params):
            if x.dtype != want or not x.is_cuda or x.numel() != p["n"] or not x.is_contiguous():
                raise TypeError(f"each signal must be a contiguous {want} device tensor of n "
                                "entries")
        ptrs = (C.c_void_p * self.count)(*[x.data_ptr() for x in xs])
        keep, dr = [], (L.Draws * self.count)()
        for i, (a, ai, off) in enumerate(draws):
            a, ai, off = (np.ascontiguousarray(v, dtype=np.int32) for v in (a, ai, off))
            keep.append((a, ai, off))
            dr[i] = L.Draws(a.ctypes.data, ai.ctypes.data, off.ctypes.data)
        return ptrs, dr, keep

    def execute_async(self, xs, draws):
        torch = _torch()
        ptrs, dr, keep = self._args(xs, draws)
        torch.cuda.synchronize()
        _check(lib.sfft_batch_execute_async(self._h, ptrs, dr), "sfft_batch_execute_async")
        self._keep["draws"] = keep

    def execute(self, xs, draws, copy=True):
        """Runs the batch; returns a list of (nodes numpy structured array, I_F) per signal.
        The nodes of all signals come back in one round trip into a pinned staging buffer;
        with copy=False they are views of it (valid until the next execute)."""
        torch = _torch()
        ptrs, dr, keep = self._args(xs, draws)
        counts = np.zeros(self.count, dtype=np.int32)
        torch.cuda.synchronize()
        _check(lib.sfft_batch_execute(self._h, ptrs, dr, None, _np_ptr(counts)),
               "sfft_batch_execute")
        kept = [int(c) for c in counts]   # within capacity (sfft_batch_execute checked)
        total = sum(kept)
        if self._stage is None or self._stage.numel() < max(total, 1) * NODE_DTYPE.itemsize:
            self._stage = torch.empty(max(total, 1) * NODE_DTYPE.itemsize, dtype=torch.uint8,
                                      pin_memory=True)
        stage = self._stage.numpy().view(NODE_DTYPE)
        hip, base, off = _hip(), self._stage.data_ptr(), 0
        for i, cnt in enumerate(kept):
            if cnt:
                src = lib.sfft_plan_debug_buffer(lib.sfft_batch_plan(self._h, i), 5)
                err = hip.hipMemcpyAsync(C.c_void_p(base + off * NODE_DTYPE.itemsize),
                                         C.c_void_p(src), C.c_size_t(cnt * NODE_DTYPE.itemsize),
                                         2, None)
                if err != 0:
                    raise RuntimeError(f"hipMemcpyAsync failed ({err})")
            off += cnt
        if hip.hipStreamSynchronize(None) != 0:
            raise RuntimeError("hipStreamSynchronize failed")
        out, off = [], 0
        for i, cnt in enumerate(kept):
            nodes = stage[off:off + cnt]
            out.append((nodes.copy() if copy else nodes, int(counts[i])))
            off += cnt
        return out

    def tune(self, xs, draws, reps=3):
        """Times the bit-equivalent launch shapes on this batch and keeps the fastest."""
        torch = _torch()
        ptrs, dr, keep = self._args(xs, draws)
        ch = np.zeros(4, dtype=np.int32)
        torch.cuda.synchronize()
        _check(lib.sfft_batch_tune(self._h, ptrs, dr, reps, _np_ptr(ch)), "sfft_batch_tune")
        return {"loops_per_thread": int(ch[0]), "block": int(ch[1]), "grid": int(ch[2]),
                "bucket_fft_columns": int(ch[3] & 0xff), "comb_fft_columns": int(ch[3] >> 8)}


_hip_lib = None


def _hip():
    global _hip_lib
    if _hip_lib is None:
        _hip_lib = C.CDLL("libamdhip64.so")
        _hip_lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip_lib.hipMemcpy.restype = C.c_int
        _hip_lib.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int,
                                            C.c_void_p]
        _hip_lib.hipMemcpyAsync.restype = C.c_int
        _hip_lib.hipStreamSynchronize.argtypes = [C.c_void_p]
        _hip_lib.hipStreamSynchronize.restype = C.c_int
    return _hip_lib


# ---- filter design (make_multiple_t) -------------------------------------------------------
def filter_taps(lobefrac, tolerance=1e-8):
    return lib.sfft_filter_taps(lobefrac, tolerance, None)


DESIGN_MODES = {"faithful": 0, "parallel": 1}


def make_filter(lobefrac, tolerance, n, b_f, w=None, mode="faithful", freq="host"):
    """FP64 filter of make_multiple_t: (filter_time device complex128[w], filter_freq).

    mode "faithful" reproduces the reference's sequential boxcar and phase-ramp recurrences bit
    for bit; "parallel" computes the same filter on the device without recurrences (n >= 4096;
    different rounding, no drift with n).  freq: "host" returns filter_freq as numpy
    complex128[n], "device" as a torch complex128[n] on the GPU, None skips it."""
    torch = _torch()
    if mode not in DESIGN_MODES:
        raise ValueError(f"mode must be one of {sorted(DESIGN_MODES)}")
    if freq not in ("host", "device", None):
        raise ValueError('freq must be "host", "device" or None')
    if w is None:
        w = filter_taps(lobefrac, tolerance)
    ft = torch.empty(w, dtype=torch.complex128, device="cuda")
    ff_h = np.empty(n, dtype=np.complex128) if freq == "host" else None
    ff_d = torch.empty(n, dtype=torch.complex128, device="cuda") if freq == "device" else None
    torch.cuda.synchronize()
    _check(lib.sfft_make_filter_ex(lobefrac, tolerance, n, b_f, w, DESIGN_MODES[mode], _ptr(ft),
                                   _np_ptr(ff_h) if ff_h is not None else None,
                                   _ptr(ff_d) if ff_d is not None else None, _stream_handle()),
           "sfft_make_filter_ex")
    return ft, (ff_h if freq == "host" else ff_d)


def filter_window(filter_freq, n, B):
    """win[j] = filter_freq[(j - m/2) mod n], j < m = n/B, from a host or device response."""
    m = n // B
    if not isinstance(filter_freq, np.ndarray):        # torch tensor on the GPU
        win = np.empty(m, dtype=np.complex128)
        _check(lib.sfft_filter_window_device(_ptr(filter_freq), n, B, _np_ptr(win),
                                             _stream_handle()), "sfft_filter_window_device")
        return win
    idx = (np.arange(m, dtype=np.int64) - m // 2) % n
    return np.ascontiguousarray(filter_freq[idx])


def to_complex64(x):
    torch = _torch()
    out = torch.empty(x.numel(), dtype=torch.complex64, device=x.device)
    _check(lib.sfft_convert_c128_to_c64(_ptr(x), _ptr(out), x.numel(), _stream_handle()),
           "convert")
    return out


# ---- signals, truth, metric ------------------------------------------------------------------
def signal_spectrum(n, k, seed):
    locs = np.empty(k, dtype=np.int32)
    vals = np.empty(k, dtype=np.complex128)
    _check(lib.sfft_signal_spectrum(n, k, seed, _np_ptr(locs), _np_ptr(vals)), "spectrum")
    return locs, vals


def synthesize(n, k, locs, vals, snr_db=0.0, noise_seed=0, fp32=False):
    torch = _torch()
    x = torch.empty(n, dtype=torch.complex64 if fp32 else torch.complex128, device="cuda")
    locs = np.ascontiguousarray(locs, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.complex128)
    torch.cuda.synchronize()
    _check(lib.sfft_signal_synthesize(n, k, _np_ptr(locs), _np_ptr(vals), float(snr_db),
                                      noise_seed, int(fp32), _ptr(x), _stream_handle()),
           "synthesize")
    return x


def truth_spectrum(x):
    torch = _torch()
    n = x.numel()
    X = torch.empty(n, dtype=torch.complex128, device=x.device)
    torch.cuda.synchronize()
    _check(lib.sfft_truth_spectrum(_ptr(x), n, int(x.dtype == torch.complex64), _ptr(X),
                                   _stream_handle()), "truth_spectrum")
    return X


def metric(nodes, k, locs, X):
    """run_experiment's metric against the truth spectrum X (device complex128[n])."""
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    locs = np.ascontiguousarray(locs, dtype=np.int32)
    me, mr = C.c_int(), C.c_int()
    et, es, ma = C.c_double(), C.c_double(), C.c_double()
    _check(lib.sfft_metric(_np_ptr(nodes), nodes.size, k, _np_ptr(locs), _ptr(X), X.numel(),
                           C.byref(me), C.byref(mr), C.byref(et), C.byref(es), C.byref(ma),
                           _stream_handle()), "metric")
    return dict(missed_est=me.value, missed_res=mr.value, mse_k=et.value, mse=et.value / k,
                support_err=es.value, max_abs=ma.value)


# ---- canonical FFT -----------------------------------------------------------------------------
def fft(x, inverse=False, out=None):
    """Canonical power-of-two FFT of the last axis of a contiguous device tensor."""
    torch = _torch()
    n = x.shape[-1]
    l2 = n.bit_length() - 1
    if 1 << l2 != n:
        raise ValueError("length must be a power of two")
    count = x.numel() // n
    if out is None:
        out = torch.empty_like(x)
    prec = L.SFFT_FP32 if x.dtype == torch.complex64 else L.SFFT_FP64
    _check(lib.sfft_fft_device(_ptr(x), _ptr(out), l2, count, int(inverse), prec,
                               _stream_handle()), "sfft_fft_device")
    return out


def dft_any(x, inverse=False):
    """In-place arbitrary-length FP64 DFT (Bluestein) of a device complex128 vector."""
    _check(lib.sfft_dft_any_device(_ptr(x), x.numel(), int(inverse), _stream_handle()),
           "sfft_dft_any_device")
    return x


def twiddle(e, T, inverse=False):
    re, im = C.c_double(), C.c_double()
    lib.sfft_twiddle(e, T, int(inverse), C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def fft_plan_passes(log2n):
    out = (C.c_int * 8)()
    n = lib.sfft_fft_plan_passes(log2n, out)
    return [out[i] for i in range(n)]


# ---- autograd ---------------------------------------------------------------------------------
def _make_function():
    torch = _torch()

    class SparseFFTValues(torch.autograd.Function):
        """One sparse FFT of x (complex128 device tensor) with an FP64 plan.  Returns the
        estimated values (complex128[I_F], differentiable) and the located indices (int64,
        not differentiable).  The backward pass holds the located indices and the
        median-selected loops of the forward transform fixed and applies the adjoint of the
        value map (sfft_value_adjoint) through the C ABI."""

        @staticmethod
        def forward(ctx, x, plan, a, ai, offsets):
            if plan.fp32:
                raise TypeError("the adjoint needs an fp64 plan")
            nodes, cnt = plan.execute(x.detach(), a, ai, offsets)
            hits, sel = plan.selection(cnt)
            vals = torch.complex(torch.from_numpy(np.ascontiguousarray(nodes["re"])),
                                 torch.from_numpy(np.ascontiguousarray(nodes["im"])))
            keys = torch.from_numpy(nodes["key"].astype(np.int64)).to(x.device)
            ctx.plan = plan
            ctx.draws = (np.array(a, np.int32), np.array(ai, np.int32),
                         np.array(offsets, np.int32))
            ctx.save_for_backward(hits, sel)
            ctx.mark_non_differentiable(keys)
            return vals.to(x.device), keys

        @staticmethod
        def backward(ctx, g_vals, g_keys):
            hits, sel = ctx.saved_tensors
            if g_vals is None:
                return None, None, None, None, None
            gx = ctx.plan.value_adjoint(g_vals, *ctx.draws, hits, sel)
            return gx, None, None, None, None

    return SparseFFTValues


_fn = None


def sparse_fft(x, plan, a, ai, offsets):
    """Differentiable sparse FFT: (values complex128[I_F], keys int64[I_F]) of x with the
    given plan (FP64, filter set) and draws; gradients flow to x through the value map's
    adjoint with the located indices and median choices held fixed."""
    global _fn
    if _fn is None:
        _fn = _make_function()
    return _fn.apply(x, plan, a, ai, offsets)
