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
CFG1 = (20, 1000, 1.40, 2, 2, 5, 128, 2)
CFG2 = (27, 16000, 0.65, 2, 2, 7, 4096, 2)    # BASELINE configs[1] (README sample)


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
