"""Shared test helpers: small valid configurations, numpy signals, comparisons."""
import numpy as np

import oracle as O

# (q, k, Bcst, loops_loc, loops_thresh, loops_est, RLFC_cst, RLFC_loops) -- every tuple
# satisfies the reference's asserts (checked in test_oracle_params)
SMALL_CONFIGS = [
    (12, 8, 1.0, 2, 2, 3, 8, 2),
    (14, 20, 1.0, 2, 2, 5, 4, 2),
    (16, 50, 1.0, 3, 2, 4, 16, 2),
    (17, 100, 0.8, 2, 1, 3, 32, 1),
]
N20_K1000 = (20, 1000, 1.40, 2, 2, 5, 128, 2)
N27_K16000 = (27, 16000, 0.65, 2, 2, 7, 4096, 2)   # the reference README's sample run


def params(cfg):
    return O.derive_params(*cfg)


def sparse_signal(n, k, seed, snr_db=None):
    """k distinct unit-magnitude random-phase coefficients; x = unnormalised IDFT."""
    rng = np.random.default_rng(seed)
    locs = np.sort(rng.choice(n, size=k, replace=False)).astype(np.int32)
    X = np.zeros(n, dtype=np.complex128)
    X[locs] = np.exp(2j * np.pi * rng.random(k))
    x = np.fft.ifft(X) * n
    if snr_db is not None:
        sigma = np.sqrt(k / (2 * 10 ** (snr_db / 10)))
        x = x + sigma * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    return x.astype(np.complex128), locs, X[locs]


def draws(n, loops, RLFC_loops, W, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, n // 2, size=loops) * 2 + 1
    ai = np.array([pow(int(v), -1, n) for v in a])
    offs = rng.integers(0, n // W, size=RLFC_loops)
    return a.astype(np.int32), ai.astype(np.int32), offs.astype(np.int32)


def oracle_filter(prm):
    ft, ff = O.make_multiple_t(prm["lobefrac"], prm["tolerance"], prm["n"], prm["b_f"], prm["w"])
    return ft, ff


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64) if a.dtype.itemsize % 8 == 0 else a.view(np.uint32)


# ---- the filter of make_multiple_t computed to full FP64 accuracy (reference for the parallel
# design; the oracle restates the reference's own rounding instead) ------------------------------
def cheb_samples_accurate(w, tol):
    """T_{w-1}(t0 cos(pi i / w)) * tol, zero where the argument is below -1, without the direct
    form's cancellation: x - 1 = 2 sinh^2(a/2) - 2 t0 sin^2(phi/2) (t0 = cosh a), then log1p /
    asin forms of acosh and acos.  Pinned against mpmath in test_oracle_filter."""
    m = w - 1
    a = np.arccosh(1 / tol) / m
    t0, t0m1 = np.cosh(a), 2 * np.sinh(a / 2) ** 2
    half = np.pi * np.arange(w) / (2 * w)
    d = t0m1 - 2 * t0 * np.sin(half) ** 2                  # x - 1
    xp1 = 2 * t0 * np.cos(half) ** 2 - t0m1                # x + 1
    out = np.zeros(w)
    up, mid = d > 0, (d <= 0) & (d >= -1)
    low = (d < -1) & (xp1 >= 0)
    du = d[up]
    out[up] = np.cosh(m * np.log1p(du + np.sqrt(du * (2 + du))))
    out[mid] = np.cos(2 * m * np.arcsin(np.sqrt(-d[mid] / 2)))
    out[low] = np.cos(m * (np.pi - 2 * np.arcsin(np.sqrt(xp1[low] / 2))))
    return out * tol


def exact_filter_design(n, w, b_f, tol):
    """(taps[w], response[n]) of make_multiple_t from accurate samples: numpy FFTs, the boxcar
    from an extended-precision prefix sum, the phase ramp exp(-2 pi i ((w/2) i mod n) / n) with
    the angle reduced in integers."""
    V = np.fft.fft(cheb_samples_accurate(w, tol).astype(np.complex128))
    h = w // 2
    shifted = np.roll(V.real, h)
    g = np.zeros(n, dtype=np.complex128)
    g[:h] = shifted[h:2 * h]
    g[n - h:] = shifted[:h]
    G = np.fft.fft(g)
    Gc = np.concatenate([G, G[:b_f]]).astype(np.clongdouble)
    cs = np.concatenate([np.zeros(1, dtype=np.clongdouble), np.cumsum(Gc)])
    idx = np.arange(n)
    box = np.empty(n, dtype=np.clongdouble)
    box[(idx + b_f // 2) % n] = cs[idx + b_f] - cs[idx]
    box /= np.max(np.abs(box))
    r = (np.int64(h) * idx.astype(np.int64)) % n
    ang = (-2 * np.pi) * (r.astype(np.longdouble) / n)
    ff = (box * (np.cos(ang) + 1j * np.sin(ang))).astype(np.complex128)
    return np.fft.ifft(ff)[:w], ff
