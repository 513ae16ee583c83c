"""Parameter derivation (experiment.c:54-84, cuExperiment.cu:279-284) and number theory.

Pinned by the README sample's printed values (README.md:84-90) and closed forms."""
import math

import numpy as np
import pytest

import oracle as O
from helpers import N20_K1000, N27_K16000, SMALL_CONFIGS


def test_readme_sample_parameters():
    # ./exec 27 16000 0.65 2 2 7 4096 2  ->  README.md:84-90
    p = O.derive_params(*N27_K16000)
    assert p["n"] == 134217728
    assert p["B_thresh"] == 32000
    assert p["W_RLFC"] == 4194304
    assert 0.5 / p["lobefrac"] == pytest.approx(183314.0, abs=0.05)   # "numlobes=183314.0"
    assert p["b_f"] == 966
    assert p["B"] == 131072
    assert p["w"] == 2230607                                           # "Window size: 2230607"
    assert p["assert_ok"] == 1


def test_n20_k1000_closed_forms():
    p = O.derive_params(*N20_K1000)
    n, k = 1 << 20, 1000
    BB = float(int(1.40 * math.sqrt(n * k / 20)))
    assert p["BB"] == BB == 10137.0
    assert p["B"] == 8192 and p["b_f"] == int(1.32 * n / BB) == 136
    assert p["W_RLFC"] == 16384 and p["w"] == 123349 and p["loops"] == 7


def test_small_configs_satisfy_reference_asserts():
    for cfg in SMALL_CONFIGS + [N20_K1000, N27_K16000]:
        p = O.derive_params(*cfg)
        assert p["assert_ok"] == 1, cfg
        assert p["B_thresh"] < p["B"] and p["B_thresh"] < p["W_RLFC"]
        assert p["w"] <= p["n"] and p["w"] >= p["B"] and p["w"] % 2 == 1


def test_invalid_config4_draw_is_flagged():
    # q=18, k=2048 -> B=4096 and B_thresh=4096: assert(B_thresh < B) fails in the reference
    p = O.derive_params(18, 2048, 1.0, 2, 2, 5, 256, 2)
    assert p["B"] == 4096 and p["B_thresh"] == 4096 and p["assert_ok"] == 0


def test_floor_to_pow2_brute_force():
    for x in [0.5, 1, 1.5, 2, 3, 7.9, 8, 1023.99, 1024, 183314.0, 2 ** 30 - 1]:
        want = 1
        while want * 2 <= x:
            want *= 2
        assert O.floor_to_pow2(x) == (want if x >= 1 else 0)


def test_mod_inverse_and_gcd():
    rng = np.random.default_rng(1)
    for q in (4, 10, 20, 27, 30):
        n = 1 << q
        for a in rng.integers(0, n // 2, size=50) * 2 + 1:
            a = int(a)
            ai = O.mod_inverse(a, n)
            assert 0 < ai < n and (a * ai) % n == 1
            assert O.gcd(a, n) == 1
        assert O.gcd(6, n) == 2


def test_timesmod_64bit_product():
    rng = np.random.default_rng(2)
    for _ in range(200):
        n = 1 << int(rng.integers(4, 31))
        x, a = int(rng.integers(0, n)), int(rng.integers(0, n))
        assert O.timesmod(x, a, n) == (x * a) % n
