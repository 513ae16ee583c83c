"""The oracle against the committed known-answer fixtures (tests/golden/, made by
make_golden.py with numpy, scipy and a pure-Python restatement of the reverse hash)."""
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = np.load(os.path.join(HERE, "golden.npz"), allow_pickle=False)


@pytest.mark.parametrize("l2", [3, 10, 12])
def test_fft_fixtures(l2):
    x = G[f"fft_in_{l2}"]
    n = x.size
    for got, want in ((O.fft64(x), G[f"fft_out_{l2}"]),
                      (O.fft64(x, inverse=True), G[f"ifft_out_{l2}"])):
        assert np.max(np.abs(got - want)) <= 1e-14 * np.max(np.abs(want)) * (l2 + 1)


@pytest.mark.parametrize("N", [101, 1849])
def test_dft_any_fixtures(N):
    got = O.dft_any64(G[f"dft_in_{N}"])
    want = G[f"dft_out_{N}"]
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))


@pytest.mark.parametrize("N", [101, 1849])
def test_chebwin_fixture_matches_design_window(N):
    """The design window before the quirk (exact Chebyshev samples, DFT, shift by N/2, real
    part) is scipy's chebwin up to scale; the oracle's samples differ only where
    t0 cos(pi i/N) < -1 (the reference's NaN -> 0 rule)."""
    tol = 1e-8
    lobefrac = 1.0 / (np.pi * N / np.arccosh(1 / tol))   # any lobefrac giving w = N is fine
    n = 1 << (int(np.ceil(np.log2(N))) + 1)
    w, t0, ok = O.filter_window(lobefrac * 0.999999, tol, n)
    if w != N:
        pytest.skip(f"lobefrac rounding gives w={w}")
    _, _, v, V, _ = O.make_multiple_t_taps(lobefrac * 0.999999, tol, n, 5, w)
    i = np.arange(N)
    x = t0 * np.cos(np.pi * i / N)
    bad = x < -1
    m = N - 1
    exact = np.where(np.abs(x) <= 1, np.cos(m * np.arccos(np.clip(x, -1, 1))),
                     np.cosh(m * np.arccosh(np.maximum(np.abs(x), 1)))) * tol
    assert np.max(np.abs(v.real[~bad] - exact[~bad])) <= 1e-9 * np.max(np.abs(exact))
    assert np.all(v.real[bad] == 0)
    win = np.roll(np.fft.fft(exact).real, N // 2)
    ref = G[f"chebwin_{N}"]
    r = win / ref
    assert np.max(np.abs(r / r[N // 2] - 1)) < 1e-6


def test_cutoff_fixture():
    s, K = G["cutoff_keys"], int(G["cutoff_K"][0])
    assert np.array_equal(O.cutoff64(s, K), G["cutoff_out"])


def test_reverse_hash_fixture():
    n, B, W, thresh = (int(v) for v in G["rh_dims"])
    vote = np.zeros(n, dtype=np.int32)
    hits = np.zeros(4096, dtype=np.int32)
    cnt = 0
    for a, ai, J in zip(G["rh_a"], G["rh_ai"], G["rh_J"]):
        cnt = O.reverse_hash(J, n, B, int(a), int(ai), thresh, vote, hits, cnt, G["rh_J2"], W)
    assert np.array_equal(hits[:cnt], G["rh_hits"])
    assert np.array_equal(vote, G["rh_vote"])


def test_readme_sample_fixture():
    with open(os.path.join(HERE, "readme_sample.json")) as fh:
        r = json.load(fh)
    p = O.derive_params(*r["args"])
    for key in ("B_thresh", "W_RLFC", "b_f", "B", "w"):
        assert p[key] == r[key]
    assert round(0.5 / p["lobefrac"], 1) == r["numlobes"]
