"""Filter design (make_multiple_t, cuExperiment.cu:248-447) of the oracle, stage by stage,
against independent computations: closed-form Chebyshev samples (with the reference's
NaN -> 0 rule), scipy's Dolph-Chebyshev window, numpy FFTs, a vectorised boxcar and an exact
phase ramp."""
import numpy as np
import pytest
from scipy.signal import windows

import oracle as O
from helpers import SMALL_CONFIGS, params


def _cheb_exact(m, x):
    """T_m(x) for all real x (m even: T_m(x) = cosh(m acosh|x|) outside [-1, 1])."""
    ax = np.abs(x)
    inside = np.cos(m * np.arccos(np.clip(x, -1, 1)))
    outside = np.cosh(m * np.arccosh(np.maximum(ax, 1.0)))
    return np.where(ax <= 1, inside, outside)


@pytest.mark.parametrize("cfg", SMALL_CONFIGS[:3])
def test_design_stages(cfg):
    prm = params(cfg)
    n, w, b_f, tol = prm["n"], prm["w"], prm["b_f"], prm["tolerance"]
    ft, ff, v, V, G = O.make_multiple_t_taps(prm["lobefrac"], tol, n, b_f, w)
    t0 = np.cosh(np.arccosh(1 / tol) / (w - 1))
    i = np.arange(w)
    x = t0 * np.cos(np.pi * i / w)
    exact = _cheb_exact(w - 1, x) * tol
    quirk = np.where(x < -1, 0.0, exact)               # cuCheb: acosh(x < -1) = NaN -> 0
    assert np.count_nonzero(x < -1) >= 1
    assert np.max(np.abs(v.real - quirk)) <= 1e-9 * np.max(np.abs(quirk))
    assert np.all(v.imag == 0)
    # DFT_w of the samples (Bluestein) vs numpy
    assert np.max(np.abs(V - np.fft.fft(v))) <= 1e-12 * np.max(np.abs(V))
    # shift by w/2 and keep the real part; without the quirk this is scipy's chebwin (scaled)
    h = w // 2
    shifted = np.roll(V.real, h)
    win_exact = np.roll(np.fft.fft(exact).real, h)
    ref = windows.chebwin(w, at=-20 * np.log10(tol), sym=True)
    ratio = win_exact / ref
    assert np.max(np.abs(ratio / ratio[h] - 1)) < 1e-6
    # fold (f[w-1] dropped) and FFT_n
    g = np.zeros(n, dtype=np.complex128)
    g[:h] = shifted[h:2 * h]
    g[n - h:] = shifted[:h]
    assert np.max(np.abs(G - np.fft.fft(g))) <= 1e-12 * np.max(np.abs(G))
    # boxcar of width b_f placed at offset b_f/2, normalised, times the phase ramp
    Gc = np.concatenate([G, G[:b_f]])
    cs = np.concatenate([[0], np.cumsum(Gc)])
    s = cs[np.arange(n) + b_f] - cs[np.arange(n)]          # s_i = sum_{j<b_f} G[i+j]
    box = np.empty(n, dtype=np.complex128)
    box[(np.arange(n) + b_f // 2) % n] = s
    box /= np.max(np.abs(box))
    ramp = np.exp(-2j * np.pi * (w // 2) * np.arange(n) / n)
    want_ff = box * ramp
    assert np.max(np.abs(ff - want_ff)) <= 1e-9
    # time taps = IDFT_n(filter_freq)[0:w] / n
    assert np.max(np.abs(ft - np.fft.ifft(ff)[:w])) <= 1e-15 * n
    # invariants: |filter_freq| peaks at 1, flat passband, stop band at the tolerance
    mag = np.abs(ff)
    assert abs(mag.max() - 1) < 1e-12
    pb = np.concatenate([mag[: b_f // 4], mag[n - b_f // 4:]])   # central half of the flat top
    assert np.min(pb) > 0.97
    far = mag[n // 4: 3 * n // 4]
    # the six zeroed Chebyshev samples raise the leakage at these tiny sizes (4.7e-9 at the
    # n=2^20 configuration); the bound is loose on purpose
    assert np.max(far) < 1e-4


def test_filter_window_slice_layout():
    n, B = 1 << 10, 16
    m = n // B
    ff = np.arange(n).astype(np.complex128)
    win = O.filter_window_slice(ff, n, B)
    assert win.size == m
    assert win[m // 2] == 0 and win[0] == n - m // 2 and win[m - 1] == m // 2 - 1


# ---- the accurate sample form (reference for the parallel design) ---------------------------
@pytest.mark.parametrize("w", [631, 1849, 123349, 2230607])
def test_accurate_chebyshev_samples_against_mpmath(w):
    """The cancellation-free samples against 50-digit evaluations of the definition; the
    reference's direct form cosh(m acosh(x)) is off by up to ~m^2 eps / 19 near x = t0."""
    import mpmath as mp
    from helpers import cheb_samples_accurate
    mp.mp.dps = 50
    tol = 1e-8
    idx = np.unique(np.concatenate([np.arange(40), np.linspace(0, w - 1, 200).astype(np.int64),
                                    np.arange(w - 12, w)]))
    m = w - 1
    a = mp.acosh(1 / mp.mpf(tol)) / m
    t0 = mp.cosh(a)
    exact = []
    for i in idx.tolist():
        x = t0 * mp.cos(mp.pi * i / w)
        v = 0 if x < -1 else (mp.cos(m * mp.acos(x)) if abs(x) <= 1 else mp.cosh(m * mp.acosh(x)))
        exact.append(float(v * mp.mpf(tol)))
    exact = np.array(exact)
    peak = np.max(np.abs(exact))
    acc = cheb_samples_accurate(w, tol)[idx]
    assert np.max(np.abs(acc - exact)) <= 1e-14 * peak
    assert np.array_equal(acc == 0, exact == 0)               # the NaN -> 0 region, 6 samples
    assert np.count_nonzero(exact == 0) == 6
    t0d = np.cosh(np.arccosh(1 / tol) / (w - 1))
    x = t0d * np.cos(np.pi * idx / w)
    direct = np.where(np.abs(x) <= 1, np.cos(m * np.arccos(np.clip(x, -1, 1))),
                      np.cosh(m * np.arccosh(np.maximum(x, 1.0)))) * tol
    direct = np.where(x < -1, 0.0, direct)
    if w > 100000:
        assert np.max(np.abs(direct - exact)) >= 1e3 * np.max(np.abs(acc - exact))


@pytest.mark.parametrize("cfg", SMALL_CONFIGS[:2])
def test_exact_design_agrees_with_oracle(cfg):
    """the accurate design differs from the oracle's restatement of the reference's rounding
    only by the direct sample form's error and the recurrences' drift (small at these n)"""
    from helpers import exact_filter_design
    prm = params(cfg)
    ft, ff = O.make_multiple_t(prm["lobefrac"], prm["tolerance"], prm["n"], prm["b_f"], prm["w"])
    ft_x, ff_x = exact_filter_design(prm["n"], prm["w"], prm["b_f"], prm["tolerance"])
    assert np.max(np.abs(ff - ff_x)) <= 1e-11
    assert np.max(np.abs(ft - ft_x)) <= 1e-11 * np.max(np.abs(ft_x))
