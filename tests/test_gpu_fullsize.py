"""Parity at the headline sizes: one whole transform at n=2^27 on the GPU against the CPU
oracle on the same signal, filter and draws, bit for bit (hit list in order, node keys and
values) -- FP64 at n=2^27, k=16000 and FP32 at the same shapes, plus the n=2^27, k=40000,
30 dB workload in both precisions (its comb FFT of 2^24 points takes three passes, the first
reading comb rows that interleave two sub-FFTs each).  The signal and the filter are made on the device (both are
bit-exact with the oracle's own restatements, see test_filter_design_bitexact) and copied to
the host for the oracle.  Each case takes seconds on the CPU side."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CASES = {
    # name: (q, k, Bcst, loops_loc, loops_thresh, loops_est, RLFC_cst, RLFC_loops, snr_db, prec)
    "n27_k16000_fp64": (27, 16000, 0.65, 2, 2, 7, 4096, 2, 0.0, "fp64"),
    "n27_k16000_fp32": (27, 16000, 0.65, 2, 2, 7, 4096, 2, 0.0, "fp32"),
    "n27_k40000_30dB_fp32": (27, 40000, 0.7, 3, 2, 8, 32768, 2, 30.0, "fp32"),
    "n27_k40000_30dB_fp64": (27, 40000, 0.7, 3, 2, 8, 32768, 2, 30.0, "fp64"),
}


@pytest.fixture(scope="module")
def S():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import sfft_amd
    return sfft_amd


@pytest.mark.parametrize("case", sorted(CASES))
def test_full_size_transform_matches_oracle(S, case):
    q, k, Bcst, ll, lt, le, rc, rl, snr, prec = CASES[case]
    print(f"full-size parity case {case}", flush=True)
    prm = S.derive_params(q, k, Bcst, ll, lt, le, rc, rl)
    n, B = prm["n"], prm["B"]
    seed = 0x5FF70000 + q * 1000 + k % 997
    locs, vals = S.signal_spectrum(n, k, seed)
    fp32 = prec == "fp32"
    x = S.synthesize(n, k, locs, vals, snr_db=snr, noise_seed=seed + 1, fp32=fp32)
    ft, ff = S.make_filter(prm["lobefrac"], prm["tolerance"], n, prm["b_f"], prm["w"])
    win = S.filter_window(ff, n, B)
    del ff
    if fp32:
        ft = S.to_complex64(ft)
    a, ai, offs = S.draw_permutations(n, prm["loops"], prm["RLFC_loops"], prm["W_RLFC"],
                                      seed + 2)
    plan = S.Plan(prm, prec)
    plan.set_filter(ft, win)
    nodes, cnt = plan.execute(x, a, ai, offs)
    hits = plan.debug_array(3, cnt, np.int32)
    xh, fth = x.cpu().numpy(), ft.cpu().numpy()
    del plan
    ref = (O.outer_loop32 if fp32 else O.outer_loop64)(prm, xh, fth, win, a, ai, offs)
    assert cnt == ref["I_F"]
    assert np.array_equal(hits, ref["hits"])
    assert nodes.tobytes() == ref["nodes"].tobytes()
    # every true location is found
    assert set(locs.tolist()) <= set(nodes["key"].tolist())
