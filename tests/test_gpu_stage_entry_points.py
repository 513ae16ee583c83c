"""GPU parity of the reference's per-stage C entry points (innerLoop.h:32-48), called through
ctypes on device buffers exactly as a caller of the reference would, against the oracle's stage
functions on the same inputs.

restr_loc_freq_coeff draws its comb offset with libc drand48() (innerLoop.cu:122-123): the test
seeds libc, predicts the draw, and re-seeds before the call.  reverse_hash appends to d_I in the
order its atomics land, so hit lists are compared as sets; the vote table d_vote is
compared entry for entry.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from helpers import N20_K1000, SMALL_CONFIGS, bits, draws, oracle_filter, params, sparse_signal

pytestmark = pytest.mark.gpu

CONFIGS = SMALL_CONFIGS + [N20_K1000]
IDS = ["q%d_k%d" % (c[0], c[1]) for c in CONFIGS]


@pytest.fixture(scope="module")
def S():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import sfft_amd
    return sfft_amd


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _vp(t):
    return C.c_void_p(t.data_ptr())


def _zeros(n, dtype):
    import torch
    return torch.zeros(n, dtype=dtype, device="cuda")


def _case(cfg, seed):
    prm = params(cfg)
    n, B = prm["n"], prm["B"]
    x, locs, _ = sparse_signal(n, prm["k"], seed)
    ft, ff = oracle_filter(prm)
    a, ai, offs = draws(n, prm["loops"], prm["RLFC_loops"], prm["W_RLFC"], seed + 3)
    return prm, x, locs, ft, ff, a, ai, offs


@pytest.mark.parametrize("cfg", CONFIGS, ids=IDS)
def test_restr_loc_freq_coeff(S, cfg):
    import torch
    libc = C.CDLL("libc.so.6")
    libc.drand48.restype = C.c_double
    prm, x, _, _, _, _, _, _ = _case(cfg, 31)
    n, W, K = prm["n"], prm["W_RLFC"], prm["B_thresh"]
    dx = _dev(x)
    for seed in (5, 6):
        libc.srand48(C.c_long(seed))
        off = int(np.floor(libc.drand48() * (n // W)))
        libc.srand48(C.c_long(seed))
        dJ2 = _zeros(K, torch.int32)
        t = C.c_double(0.0)
        rc = S.lib.restr_loc_freq_coeff(_vp(dx), n, K, W, _vp(dJ2), C.byref(t))
        assert rc == 0
        assert np.array_equal(dJ2.cpu().numpy(), O.rlfc64(x, n, W, off, K)), f"offset {off}"
        assert t.value > 0


@pytest.mark.parametrize("N,K", [(4096, 100), (1 << 16, 5000), (3000, 1)])
def test_samples_indices_random_and_ties(S, N, K):
    import torch
    rng = np.random.default_rng(N + K)
    for X in (rng.standard_normal(N) + 1j * rng.standard_normal(N),
              (rng.integers(-2, 3, N) + 1j * rng.integers(-2, 3, N)).astype(np.complex128)):
        dX = _dev(X)
        out = _zeros(K, torch.int32)
        t = C.c_double(0.0)
        assert S.lib.samples_indices(_vp(dX), K, N, _vp(out), C.byref(t)) == 0
        assert np.array_equal(out.cpu().numpy(), O.cutoff64(O.samples64(X), K))


@pytest.mark.parametrize("cfg", CONFIGS, ids=IDS)
def test_perm_filter_cutoff(S, cfg):
    import torch
    prm, x, _, ft, _, a, ai, _ = _case(cfg, 41)
    n, B, K, w = prm["n"], prm["B"], prm["B_thresh"], prm["w"]
    dx, dft = _dev(x), _dev(ft)
    for l in range(2):
        bins = _zeros(B, torch.complex128)
        J = _zeros(K, torch.int32)
        pf, bc = C.c_double(0.0), C.c_double(0.0)
        rc = S.lib.perm_filter_cutoff(_vp(dx), n, _vp(dft), w, K, B, int(a[l]), int(ai[l]),
                                      _vp(bins), _vp(J), C.byref(pf), C.byref(bc))
        assert rc == 0
        want = O.fft64(O.perm_filter64(x, n, ft, B, int(ai[l]), O.FOLD_REFERENCE))
        assert np.array_equal(bits(bins.cpu().numpy()), bits(want)), f"loop {l} bins_f"
        assert np.array_equal(J.cpu().numpy(), O.cutoff64(O.samples64(want), K)), f"loop {l} J"


@pytest.mark.parametrize("cfg", CONFIGS, ids=IDS)
def test_reverse_hash_votes_and_hits(S, cfg):
    """Every location loop through reverse_hash in turn, one shared vote table and hit list,
    exactly as outer_loop drives it (outerLoop.cu:143-146); the J lists and the merged J2 come
    from the oracle's own transform of the signal."""
    import torch
    prm, x, locs, ft, ff, a, ai, offs = _case(cfg, 51)
    n, B, K, W = prm["n"], prm["B"], prm["B_thresh"], prm["W_RLFC"]
    win = O.filter_window_slice(ff, n, B)
    ref = O.outer_loop64(prm, x, ft, win, a, ai, offs, want_J=True)
    J2 = ref["J2"]
    d_vote = _zeros(n, torch.int32)
    d_I = _zeros(524288, torch.int32)
    d_J2 = _dev(J2.astype(np.int32))
    vote = np.zeros(n, dtype=np.int32)
    hits = np.zeros(524288, dtype=np.int32)
    cnt = 0
    I_F = C.c_int(0)
    for l in range(prm["loops_loc"]):
        dJ = _dev(ref["J"][l].astype(np.int32))
        t = C.c_double(0.0)
        rc = S.lib.reverse_hash(_vp(dJ), n, K, B, int(a[l]), int(ai[l]), prm["loops_thresh"],
                                _vp(d_vote), _vp(d_I), C.byref(I_F), C.byref(t), _vp(d_J2),
                                int(J2.size), W, prm["RLFC_loops"])
        assert rc == 0
        cnt = O.reverse_hash(ref["J"][l], n, B, int(a[l]), int(ai[l]), prm["loops_thresh"],
                             vote, hits, cnt, J2, W)
        assert I_F.value == cnt, f"loop {l}"
        assert np.array_equal(d_vote.cpu().numpy(), vote), f"loop {l} vote table"
        got = np.sort(d_I[:cnt].cpu().numpy())
        assert np.array_equal(got, np.sort(hits[:cnt])), f"loop {l} hit set"
    assert np.array_equal(np.sort(hits[:cnt]), np.sort(ref["hits"]))
    assert set(locs.tolist()) <= set(hits[:cnt].tolist())


@pytest.mark.parametrize("cfg", CONFIGS, ids=IDS)
@pytest.mark.parametrize("freq_on_device", [False, True])
def test_estimate_values(S, cfg, freq_on_device):
    prm, x, locs, ft, ff, a, ai, offs = _case(cfg, 61)
    n, B, L = prm["n"], prm["B"], prm["loops"]
    win = O.filter_window_slice(ff, n, B)
    ref = O.outer_loop64(prm, x, ft, win, a, ai, offs, want_bins=True)
    hits = ref["hits"]
    d_I = _dev(hits.astype(np.int32))
    d_bins = _dev(np.ascontiguousarray(ref["bins_f"]).ravel())
    perm = np.ascontiguousarray(ai, dtype=np.int32)
    if freq_on_device:
        d_ff = _dev(ff)
        ffp = _vp(d_ff)
    else:
        ffp = ff.ctypes.data_as(C.c_void_p)
    I_F = C.c_int(int(hits.size))
    p = S.lib.estimate_values(_vp(d_I), C.byref(I_F), _vp(d_bins), L, n,
                              perm.ctypes.data_as(C.c_void_p), B, ffp, prm["loops_loc"])
    assert p
    got = np.frombuffer(C.string_at(p, I_F.value * 24), dtype=O.NODE_DTYPE).copy()
    C.CDLL("libc.so.6").free(C.c_void_p(C.cast(p, C.c_void_p).value))
    want = O.estimate_values(hits, ref["bins_f"], L, n, ai, B, win)
    assert got.tobytes() == want.tobytes()
    assert got.tobytes() == ref["nodes"].tobytes()


def test_timesmod(S):
    rng = np.random.default_rng(3)
    for _ in range(200):
        n = 1 << int(rng.integers(4, 31))
        x, a = int(rng.integers(0, n)), int(rng.integers(0, n))
        assert S.lib.timesmod(x, a, n) == (x * a) % n == O.timesmod(x, a, n)
