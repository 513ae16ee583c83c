"""The oracle's whole transform (outerLoop.cu:65-190) on the reference's tuned n=2^20 set and
the harness metric (cuExperiment.cu:181-221) on hand-built cases."""
import numpy as np

import oracle as O
from helpers import N20_K1000, draws, oracle_filter, params, sparse_signal


def test_n20_k1000_recovers_every_location():
    prm = params(N20_K1000)
    n, k = prm["n"], prm["k"]
    x, locs, vals = sparse_signal(n, k, 2024)
    ft, ff = oracle_filter(prm)
    win = O.filter_window_slice(ff, n, prm["B"])
    a, ai, offs = draws(n, prm["loops"], prm["RLFC_loops"], prm["W_RLFC"], 2031)
    r = O.outer_loop64(prm, x, ft, win, a, ai, offs)
    nodes = r["nodes"]
    assert r["I_F"] == nodes.size >= k
    X = np.zeros(n, dtype=np.complex128)
    X[locs] = vals
    me, mr, err = O.metric_dense(nodes, k, locs, n, X)
    assert (me, mr) == (0, 0)
    assert err / k < 1e-3                         # MSE of the reference's filter at this set
    # deterministic
    r2 = O.outer_loop64(prm, x, ft, win, a, ai, offs)
    assert r2["nodes"].tobytes() == nodes.tobytes()


def test_metric_dense_hand_cases():
    n, k = 64, 3
    locs = np.array([5, 9, 40], dtype=np.int32)
    X = np.zeros(n, dtype=np.complex128)
    X[locs] = [1.0, 1j, -1.0]
    nodes = np.zeros(4, dtype=O.NODE_DTYPE)
    nodes["key"] = [9, 5, 12, 41]
    nodes["re"] = [0.0, 1.0, 0.5, 2.0]
    nodes["im"] = [1.0, 0.0, 0.0, 0.0]
    # top-3 by |value|: 41 (2), then 5 and 9 (1, ties by key) -> 12 dropped
    me, mr, err = O.metric_dense(nodes, k, locs, n, X)
    assert me == 1                  # 40 never located
    assert mr == 1
    # error: |2 - 0|^2 at 41, 0 at 5 and 9, |0 - (-1)|^2 at 40
    assert err == 5.0
