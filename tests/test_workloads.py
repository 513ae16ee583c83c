"""Host-side workload generators: the batched workload's parameter draws (q uniform, k
log-uniform) and its redraw rule for draws that violate the reference's asserts
(experiment.c:77-84): at q = 18 a large k gives B = 2048 or 4096 with B_thresh = 2k >= B."""
import math

import numpy as np

import sfft_amd as S


def _first_k(rng, lo, hi):
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))


def test_batch_workload_ranges_and_validity():
    for seed in range(20):
        for p in S.batch_workload(64, seed):
            q = p["n"].bit_length() - 1
            assert 18 <= q <= 24 and 50 <= p["k"] <= 4000
            assert p["assert_ok"] and S.validate(p) == 0
            assert p["B_thresh"] == 2 * p["k"] < p["B"] and p["B_thresh"] < p["W_RLFC"]
            assert (p["Bcst"], p["loops_loc"], p["loops_thresh"], p["loops"]) == (1.0, 2, 2, 7)
            assert p["RLFC_loops"] == 2


def test_batch_workload_redraws_k_only():
    """Replays the generator's stream: where the first k of a signal is invalid for its q, the
    returned k is the first valid later draw and q is kept."""
    redraws = 0
    for seed in range(40):
        got = S.batch_workload(64, seed)
        rng = np.random.default_rng(seed)
        for p in got:
            q = int(rng.integers(18, 25))
            assert p["n"] == 1 << q
            while True:
                k = _first_k(rng, 50, 4000)
                cand = S.derive_params(q, k, 1.0, 2, 2, 5, 256, 2)
                if cand["assert_ok"] and S.validate(cand) == 0:
                    break
                # the corner of this range that breaks B_thresh < B (experiment.c:83)
                assert not cand["assert_ok"] and cand["B_thresh"] >= cand["B"] and q <= 19
                redraws += 1
            assert p["k"] == k
    assert redraws > 0                                  # the rule was exercised
