"""GPU checks of the experiment harness (cuExperiment.cu:64-137 synthesis and truth spectrum,
:181-221 error metric) against independent references:

* sfft_metric against the oracle's restatement of run_experiment's metric (orc_metric_dense)
  on the same nodes -- missed counts equal, MSE*k to 1e-12 relative (the device metric sums the
  error as the energy outside the kept support plus the error on it, in another order) -- and
  on hand-built cases with known answers;
* run_experiment's printed ERROR line against the oracle metric of the nodes outer_loop returns
  for the same draws;
* sfft_truth_spectrum bit for bit against the oracle's canonical FFT divided by n, and within
  1e-13 log n of numpy;
* sfft_signal_synthesize against numpy's inverse FFT (and bit for bit against the canonical
  inverse FFT), its complex64 output against the rounded complex128 one, and its noise against
  the SNR formula sigma^2 = k / (2 * 10^(snr/10)) per component (experiment.c:69).
"""
import ctypes as C
import re

import numpy as np
import pytest

import oracle as O
from helpers import N20_K1000, SMALL_CONFIGS, bits, draws, oracle_filter, params, sparse_signal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import sfft_amd
    return sfft_amd


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _nodes(keys, vals):
    nd = np.zeros(len(keys), dtype=O.NODE_DTYPE)
    nd["key"] = keys
    nd["re"] = np.real(vals)
    nd["im"] = np.imag(vals)
    return nd


# ---- metric ----------------------------------------------------------------------------------
def test_metric_hand_cases(S):
    """The hand case of the oracle's own metric test, through the device metric."""
    n, k = 64, 3
    locs = np.array([5, 9, 40], dtype=np.int32)
    X = np.zeros(n, dtype=np.complex128)
    X[locs] = [1.0, 1j, -1.0]
    nodes = _nodes([9, 5, 12, 41], [1j, 1.0, 0.5, 2.0])
    got = S.metric(nodes, k, locs, _dev(X))
    assert (got["missed_est"], got["missed_res"]) == (1, 1)
    assert got["mse_k"] == 5.0
    assert (got["missed_est"], got["missed_res"], got["mse_k"]) == O.metric_dense(
        nodes, k, locs, n, X)


def test_metric_hand_cases_exact_and_empty(S):
    n, k = 128, 4
    locs = np.array([1, 7, 64, 127], dtype=np.int32)
    X = np.zeros(n, dtype=np.complex128)
    X[locs] = [1.0, -1j, 0.5, 2.0]
    exact = _nodes(locs, X[locs])                       # every value exact: no error
    got = S.metric(exact, k, locs, _dev(X))
    assert (got["missed_est"], got["missed_res"], got["mse_k"]) == (0, 0, 0.0)
    # no nodes at all: everything missed, error = the whole energy
    got = S.metric(np.zeros(0, dtype=O.NODE_DTYPE), k, locs, _dev(X))
    assert (got["missed_est"], got["missed_res"]) == (4, 4)
    assert got["mse_k"] == 1.0 + 1.0 + 0.25 + 4.0
    # more nodes than k: only the k largest count, a located but cut location counts as
    # missed in the result and not in the estimation
    nd = _nodes([1, 7, 64, 127, 3, 5], [1.0, -1j, 0.1, 2.0, 3.0, 1.5])
    got = S.metric(nd, k, locs, _dev(X))
    want = O.metric_dense(nd, k, locs, n, X)
    assert (got["missed_est"], got["missed_res"]) == want[:2] == (0, 2)
    # kept: 3 (3.0), 127 (2.0), 5 (1.5), 1 (1.0); error 9 + 0 + 2.25 + 0 + |X[7]|^2 + |X[64]|^2
    assert got["mse_k"] == want[2] == 9.0 + 2.25 + 1.0 + 0.25


@pytest.mark.parametrize("cfg", [SMALL_CONFIGS[2], N20_K1000], ids=["q16_k50", "q20_k1000"])
@pytest.mark.parametrize("snr_db", [None, 20.0])
def test_metric_matches_oracle_on_transform_output(S, cfg, snr_db):
    prm = params(cfg)
    n, B, k = prm["n"], prm["B"], prm["k"]
    x, locs, _ = sparse_signal(n, k, 91, snr_db)
    ft, ff = oracle_filter(prm)
    win = O.filter_window_slice(ff, n, B)
    a, ai, offs = draws(n, prm["loops"], prm["RLFC_loops"], prm["W_RLFC"], 92)
    plan = S.Plan(prm, "fp64")
    plan.set_filter(_dev(ft), win)
    nodes, cnt = plan.execute(_dev(x), a, ai, offs)
    X = np.fft.fft(x) / n
    rng = np.random.default_rng(93)
    fresh = rng.choice(np.setdiff1d(np.arange(n), nodes["key"]), 30, replace=False)
    variants = [nodes,
                nodes[rng.permutation(cnt)[: cnt // 2]],            # half the nodes lost
                np.concatenate([nodes, _nodes(fresh, rng.standard_normal(30))])]  # false hits
    for nd in variants:
        got = S.metric(nd, k, locs, _dev(X))
        me, mr, err = O.metric_dense(nd, k, locs, n, X)
        assert (got["missed_est"], got["missed_res"]) == (me, mr)
        assert abs(got["mse_k"] - err) <= 1e-12 * err
        assert got["mse"] == got["mse_k"] / k


def test_run_experiment_error_line(S, capfd):
    """run_experiment (cuExperiment.cu:159-223) prints MISSED and MSE*k of the nodes that
    outer_loop returns; the same draws replayed give the nodes, the oracle metric gives the
    numbers (printed with %lg, six significant digits)."""
    libc = C.CDLL("libc.so.6")
    prm = params(SMALL_CONFIGS[1])
    n, B, k = prm["n"], prm["B"], prm["k"]
    x, locs, _ = sparse_signal(n, k, 101, 25.0)
    ft, ff = oracle_filter(prm)
    X = np.fft.fft(x) / n
    dx, dft = _dev(x), _dev(ft)
    lf = np.ascontiguousarray(locs, dtype=np.int32)
    C.c_int.in_dll(S.lib, "TIMING").value = 0
    try:
        libc.srand48(C.c_long(11))
        S.lib.sfft_srandom(12)
        I_F = C.c_int(0)
        p = S.lib.outer_loop(C.c_void_p(dx.data_ptr()), n, C.c_void_p(dft.data_ptr()),
                             ff.ctypes.data_as(C.c_void_p), prm["w"], B, prm["B_thresh"],
                             prm["W_RLFC"], prm["RLFC_loops"], prm["loops_thresh"],
                             prm["loops_loc"], prm["loops"], C.byref(I_F), C.c_float(0.0))
        assert p
        nodes = np.frombuffer(C.string_at(p, I_F.value * 24), dtype=O.NODE_DTYPE).copy()
        libc.free(C.c_void_p(C.cast(p, C.c_void_p).value))
        libc.srand48(C.c_long(11))
        S.lib.sfft_srandom(12)
        capfd.readouterr()
        run = S.lib.run_experiment
        run.restype = None
        run(C.c_void_p(dx.data_ptr()), n, C.c_void_p(dft.data_ptr()),
            ff.ctypes.data_as(C.c_void_p), prm["w"], B, prm["B_thresh"], prm["loops_loc"],
            prm["loops_thresh"], prm["loops"] - prm["loops_loc"], prm["W_RLFC"],
            prm["RLFC_loops"], 1, lf.ctypes.data_as(C.c_void_p), k,
            X.ctypes.data_as(C.c_void_p), C.c_float(0.0))
        libc.fflush(None)
    finally:
        C.c_int.in_dll(S.lib, "TIMING").value = 1
    out = capfd.readouterr().out
    m = re.search(r"K=(\d+); MISSED \(estimation, result\) = \((\d+), (\d+)\), MSE\*k = (\S+), "
                  r"Mean Square Error \(MSE\) = (\S+)", out)
    assert m, out
    me, mr, err = O.metric_dense(nodes, k, locs, n, X)
    assert int(m.group(1)) == k
    assert (int(m.group(2)), int(m.group(3))) == (me, mr)
    assert float(m.group(4)) == pytest.approx(err, rel=1e-5)
    assert float(m.group(5)) == pytest.approx(err / k, rel=1e-5)


# ---- synthesis and truth -------------------------------------------------------------------
@pytest.mark.parametrize("log2n", [8, 12, 16, 20, 22])
def test_truth_spectrum_bits(S, log2n):
    import torch
    n = 1 << log2n
    rng = np.random.default_rng(log2n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = S.truth_spectrum(_dev(x)).cpu().numpy()
    assert np.array_equal(bits(X), bits(O.fft64(x) / n))
    ref = np.fft.fft(x) / n
    assert np.max(np.abs(X - ref)) <= 1e-13 * np.max(np.abs(ref)) * (log2n + 1)
    # complex64 input: widened exactly, then the same FFT
    x32 = x.astype(np.complex64)
    X32 = S.truth_spectrum(_dev(x32)).cpu().numpy()
    assert np.array_equal(bits(X32), bits(O.fft64(x32.astype(np.complex128)) / n))
    torch.cuda.synchronize()


@pytest.mark.parametrize("log2n,k", [(10, 5), (16, 50), (20, 1000), (22, 1000)])
def test_synthesize_matches_inverse_fft(S, log2n, k):
    n = 1 << log2n
    locs, vals = S.signal_spectrum(n, k, 1234 + log2n)
    assert np.all(np.diff(locs) > 0) and locs[0] >= 0 and locs[-1] < n
    assert np.max(np.abs(np.abs(vals) - 1)) <= 1e-15
    Xs = np.zeros(n, dtype=np.complex128)
    Xs[locs] = vals
    x = S.synthesize(n, k, locs, vals).cpu().numpy()
    assert np.array_equal(bits(x), bits(O.fft64(Xs, inverse=True)))
    ref = np.fft.ifft(Xs) * n
    assert np.max(np.abs(x - ref)) <= 1e-13 * np.sqrt(k) * (log2n + 1)
    x32 = S.synthesize(n, k, locs, vals, fp32=True).cpu().numpy()
    assert np.array_equal(bits(x32), bits(x.astype(np.complex64)))
    # and the truth spectrum of the synthesised signal gives the spectrum back
    X = S.truth_spectrum(_dev(x)).cpu().numpy()
    assert np.max(np.abs(X[locs] - vals)) <= 1e-14 * (log2n + 1)
    mask = np.ones(n, bool)
    mask[locs] = False
    assert np.max(np.abs(X[mask])) <= 1e-14 * (log2n + 1)


@pytest.mark.parametrize("snr_db", [30.0, 40.0])
def test_synthesize_noise_level(S, snr_db):
    """noise = noisy - clean synthesis (same spectrum): i.i.d. Gaussian components of variance
    k / (2 * 10^(snr/10)) (experiment.c:69), zero mean, real and imaginary uncorrelated."""
    n, k = 1 << 20, 1000
    locs, vals = S.signal_spectrum(n, k, 77)
    clean = S.synthesize(n, k, locs, vals).cpu().numpy()
    noisy = S.synthesize(n, k, locs, vals, snr_db=snr_db, noise_seed=5).cpu().numpy()
    e = noisy - clean
    var = k / (2 * 10 ** (snr_db / 10))
    for comp in (e.real, e.imag):
        assert abs(np.var(comp) / var - 1) < 0.01          # sampling error ~0.14% at n = 2^20
        assert abs(np.mean(comp)) < 5 * np.sqrt(var / n)
        kurt = np.mean(comp ** 4) / np.var(comp) ** 2
        assert abs(kurt - 3) < 0.05
    assert abs(np.mean(e.real * e.imag)) < 5 * var / np.sqrt(n)
    again = S.synthesize(n, k, locs, vals, snr_db=snr_db, noise_seed=5).cpu().numpy()
    assert np.array_equal(bits(again), bits(noisy))     # the seed fixes the noise
    other = S.synthesize(n, k, locs, vals, snr_db=snr_db, noise_seed=6).cpu().numpy()
    assert not np.array_equal(other, noisy)
