"""The canonical FFT of the oracle against numpy.fft (an independent implementation),
closed forms and identities; Bluestein for the odd design lengths."""
import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("log2n", range(0, 21))
def test_fft64_matches_numpy(log2n):
    n = 1 << log2n
    rng = np.random.default_rng(log2n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for inverse in (False, True):
        got = O.fft64(x, inverse)
        ref = np.fft.ifft(x) * n if inverse else np.fft.fft(x)
        scale = max(1.0, np.max(np.abs(ref)))
        assert np.max(np.abs(got - ref)) <= 2e-16 * scale * (log2n + 2) * 4


@pytest.mark.parametrize("log2n", [1, 4, 9, 12, 17])
def test_fft32_matches_numpy(log2n):
    n = 1 << log2n
    rng = np.random.default_rng(50 + log2n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    got = O.fft32(x)
    ref = np.fft.fft(x.astype(np.complex128))
    assert np.max(np.abs(got - ref)) <= 1.2e-7 * np.max(np.abs(ref)) * (log2n + 2) * 4


def test_impulse_and_tone_closed_forms():
    n = 1 << 12
    x = np.zeros(n, dtype=np.complex128)
    x[0] = 1.0
    assert np.array_equal(O.fft64(x), np.ones(n, dtype=np.complex128))   # exact: no rounding
    x[:] = 0
    x[5] = 1.0
    X = O.fft64(x)
    want = np.exp(-2j * np.pi * 5 * np.arange(n) / n)
    assert np.max(np.abs(X - want)) < 1e-14
    t = np.arange(n)
    tone = np.exp(2j * np.pi * 37 * t / n)
    X = O.fft64(tone)
    assert abs(X[37] - n) < 1e-10 and np.max(np.abs(np.delete(X, 37))) < 1e-10


def test_round_trip_and_parseval():
    n = 1 << 15
    rng = np.random.default_rng(7)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = O.fft64(x)
    back = O.fft64(X, inverse=True) / n
    assert np.max(np.abs(back - x)) < 1e-13
    assert abs(np.sum(np.abs(X) ** 2) / n - np.sum(np.abs(x) ** 2)) < 1e-9 * np.sum(np.abs(x) ** 2)


def test_pass_plan():
    for l2 in range(0, 31):
        p = O.fft_plan(l2)
        assert sum(p) == l2
        if l2:
            assert len(p) == -(-l2 // 11) and max(p) - min(p) <= 1 and max(p) <= 11
            assert p == sorted(p, reverse=True)


def test_twiddle_values():
    """Within one ulp of exp(-2 pi i e/T) evaluated in x87 extended precision."""
    pi_ld = np.longdouble("3.14159265358979323846264338327950288")
    for T in (1, 2, 4, 8, 16, 1 << 11, 1 << 20, 1 << 27):
        for e in {0, 1, 3, 11, T // 8, T // 4, T // 2, (3 * T) // 4, T - 1, T // 3, (5 * T) // 7}:
            if e >= T:
                continue
            ang = 2 * pi_ld * np.longdouble(e) / np.longdouble(T)
            want = complex(float(np.cos(ang)), -float(np.sin(ang)))
            got = O.twiddle(e, T)
            assert abs(got.real - want.real) <= 1.2e-16 and abs(got.imag - want.imag) <= 1.2e-16, \
                (e, T)
            assert O.twiddle(e, T, inverse=True) == np.conj(got)
    # exact quarter turns
    assert O.twiddle(1, 4) == complex(0.0, -1.0) and O.twiddle(2, 8) == complex(0.0, -1.0)


@pytest.mark.parametrize("N", [1, 2, 3, 5, 7, 15, 123, 631, 1001, 4095, 5499])
def test_bluestein_matches_numpy(N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    for inverse in (False, True):
        got = O.dft_any64(x, inverse)
        ref = np.fft.ifft(x) * N if inverse else np.fft.fft(x)
        assert np.max(np.abs(got - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))
