"""The oracle's FP32 path (complex64 storage, FP32 FFTs, Kahan fold, FP64 estimation) against
its FP64 path on the same signal, filter and draws: same recovered locations, values within a
few FP32 ulps of the signal scale."""
import numpy as np

import oracle as O
from helpers import N20_K1000, SMALL_CONFIGS, draws, oracle_filter, params, sparse_signal


def _as_dict(nodes):
    return dict(zip(nodes["key"].tolist(), (nodes["re"] + 1j * nodes["im"]).tolist()))


def test_fp32_path_matches_fp64_path():
    for cfg, seed in ((N20_K1000, 2024), (SMALL_CONFIGS[3], 5)):
        prm = params(cfg)
        n, k = prm["n"], prm["k"]
        x, locs, vals = sparse_signal(n, k, seed)
        ft, ff = oracle_filter(prm)
        win = O.filter_window_slice(ff, n, prm["B"])
        a, ai, offs = draws(n, prm["loops"], prm["RLFC_loops"], prm["W_RLFC"], seed + 7)
        r64 = O.outer_loop64(prm, x, ft, win, a, ai, offs)
        r32 = O.outer_loop32(prm, x.astype(np.complex64), ft.astype(np.complex64), win, a, ai,
                             offs)
        d64, d32 = _as_dict(r64["nodes"]), _as_dict(r32["nodes"])
        common = [int(l) for l in locs if int(l) in d64]
        assert set(common) <= set(d32)
        diff = max(abs(d32[l] - d64[l]) for l in common)
        assert diff < 2e-5, diff                   # unit-magnitude coefficients
