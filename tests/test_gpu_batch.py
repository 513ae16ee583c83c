"""Batches of uneven signals (sfft_batch_*): every signal's result equals a single-signal
transform with the same plan parameters and draws, bit for bit -- eagerly, replayed from a
hipGraph (several replays with other device work in between, new signal contents behind the
same pointers), under every launch shape the autotuner may pick, and after tuning."""
import numpy as np
import pytest

import oracle as O
from helpers import bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import sfft_amd
    return sfft_amd


def _setup(S, precision, count=6, seed=3):
    import torch
    # small uneven signals; a comb constant of 8 keeps W_RLFC <= n at these sizes
    prms = S.batch_workload(count, seed, q_range=(12, 17), k_range=(6, 80), RLFC_cst=8)
    cdt = np.complex64 if precision == "fp32" else np.complex128
    xs, fts, wins, draws = [], [], [], []
    rng = np.random.default_rng(seed)
    for i, p in enumerate(prms):
        n, k = p["n"], p["k"]
        locs = np.sort(rng.choice(n, size=k, replace=False))
        X = np.zeros(n, dtype=np.complex128)
        X[locs] = np.exp(2j * np.pi * rng.random(k))
        x = np.fft.ifft(X) * n
        x = x + 0.01 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
        xs.append(torch.from_numpy(x.astype(cdt)).to("cuda"))
        ft, ff = O.make_multiple_t(p["lobefrac"], p["tolerance"], n, p["b_f"], p["w"])
        fts.append(torch.from_numpy(ft.astype(cdt)).to("cuda"))
        wins.append(O.filter_window_slice(ff, n, p["B"]))
        draws.append(S.draw_permutations(n, p["loops"], p["RLFC_loops"], p["W_RLFC"], 100 + i))
    return prms, xs, fts, wins, draws


def _singles(S, prms, xs, fts, wins, draws, precision):
    out = []
    for p, x, ft, win, d in zip(prms, xs, fts, wins, draws):
        plan = S.Plan(p, precision)
        plan.set_filter(ft, win)
        out.append(plan.execute(x, *d))
    return out


def _same(got, want):
    assert len(got) == len(want)
    for i, ((gn, gc), (wn, wc)) in enumerate(zip(got, want)):
        assert gc == wc, f"signal {i}: {gc} vs {wc} hits"
        assert gn.tobytes() == wn.tobytes(), f"signal {i}"


def test_batch_front_end_param(S):
    """The batch gathers: front_end AUTO is accepted (and computes the same bits), COLUMNS and
    STREAM are refused by sfft_batch_create."""
    prms, xs, fts, wins, draws = _setup(S, "fp32", count=3)
    b = S.Batch(prms, "fp32", front_end="auto")
    for i in range(len(prms)):
        b.set_filter(i, fts[i], wins[i])
    _same(b.execute(xs, draws), _singles(S, prms, xs, fts, wins, draws, "fp32"))
    b.close()
    for fe in ("columns", "stream"):
        with pytest.raises(S.SfftError):
            S.Batch(prms, "fp32", front_end=fe)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_batch_equals_single_transforms(S, precision):
    import torch
    prms, xs, fts, wins, draws = _setup(S, precision)
    assert len({p["n"] for p in prms}) > 1 and len({p["k"] for p in prms}) > 1   # uneven
    want = _singles(S, prms, xs, fts, wins, draws, precision)
    b = S.Batch(prms, precision)
    for i in range(len(prms)):
        b.set_filter(i, fts[i], wins[i])
    _same(b.execute(xs, draws), want)
    # every launch shape the autotuner chooses among computes the same bits
    for shape in [(1, 128, 7, 1, 1), (2, 256, 64, 4, 16), (4, 128, 4096, 16, 2),
                  (3, 256, 2048, 0, 0), (7, 256, 1024, 0, 0), (9, 128, 2048, 4, 4),
                  (5, 256, 2048, 32, 32)]:
        b.set_shape(*shape)
        _same(b.execute(xs, draws), want)
    # hipGraph: capture once, replay several times with other device work in between
    b.set_graph(True)
    for r in range(3):
        _same(b.execute(xs, draws), want)
        y = torch.randn(1 << 20, device="cuda")
        (y * 2).sum().item()
        torch.cuda.synchronize()
    # new contents behind the same pointers (the graph is replayed, not re-captured)
    for x in xs:
        x.mul_(0.5).add_(torch.roll(x, 3) * 0.25)
    want2 = _singles(S, prms, xs, fts, wins, draws, precision)
    _same(b.execute(xs, draws), want2)
    b.set_graph(False)
    _same(b.execute(xs, draws), want2)


def test_batch_autotuner_keeps_results(S):
    prms, xs, fts, wins, draws = _setup(S, "fp32", count=5, seed=9)
    want = _singles(S, prms, xs, fts, wins, draws, "fp32")
    b = S.Batch(prms, "fp32")
    for i in range(len(prms)):
        b.set_filter(i, fts[i], wins[i])
    choice = b.tune(xs, draws, reps=2)
    assert 1 <= choice["loops_per_thread"] <= 9 and choice["block"] in (128, 256)
    assert choice["grid"] in (1024, 2048, 4096)
    assert choice["bucket_fft_columns"] in (1, 2, 4, 8, 16, 32)
    _same(b.execute(xs, draws), want)
    b.set_graph(True)
    _same(b.execute(xs, draws), want)
    _same(b.execute(xs, draws), want)


def test_batch_fp64_matches_oracle(S):
    """one batch signal against the CPU oracle directly (not only against a single plan)"""
    prms, xs, fts, wins, draws = _setup(S, "fp64", count=3, seed=21)
    b = S.Batch(prms, "fp64")
    for i in range(len(prms)):
        b.set_filter(i, fts[i], wins[i])
    res = b.execute(xs, draws)
    for i, p in enumerate(prms):
        a, ai, off = draws[i]
        ref = O.outer_loop64(p, xs[i].cpu().numpy(), fts[i].cpu().numpy(), wins[i], a, ai, off)
        nodes, cnt = res[i]
        assert cnt == ref["I_F"]
        assert np.array_equal(nodes["key"], ref["nodes"]["key"])
        assert np.array_equal(bits(nodes["re"]), bits(ref["nodes"]["re"]))
        assert np.array_equal(bits(nodes["im"]), bits(ref["nodes"]["im"]))


def _workload64(S, seed):
    """The batched workload at its real size: 64 complex64 signals, n = 2^q with q uniform on
    18..24, k log-uniform on [50, 4000], unit-magnitude random-phase spectra, 40 dB noise,
    Bcst 1.0, loops 2+5, thresh 2, RLFC 256 x 2; filters from the parallel device design."""
    import torch
    prms = S.batch_workload(64, seed)
    xs, fts, wins, draws = [], [], [], []
    for i, p in enumerate(prms):
        n, k = p["n"], p["k"]
        locs, vals = S.signal_spectrum(n, k, 1000 * seed + i)
        xs.append(S.synthesize(n, k, locs, vals, snr_db=40.0, noise_seed=i + 1, fp32=True))
        ft, ff = S.make_filter(p["lobefrac"], p["tolerance"], n, p["b_f"], p["w"],
                               mode="parallel", freq="device")
        fts.append(S.to_complex64(ft))
        wins.append(S.filter_window(ff, n, p["B"]))
        del ff
        draws.append(S.draw_permutations(n, p["loops"], p["RLFC_loops"], p["W_RLFC"], 500 + i))
    torch.cuda.synchronize()
    return prms, xs, fts, wins, draws


def test_batch64_workload_equals_singles_and_oracle(S):
    """The 64-signal uneven workload: eager batch and graph replays equal the single-signal
    transforms bit for bit; one signal of every drawn size equals the CPU oracle's FP32 path."""
    import torch
    prms, xs, fts, wins, draws = _workload64(S, seed=4)
    qs = [p["n"].bit_length() - 1 for p in prms]
    assert min(qs) >= 18 and max(qs) <= 24 and len(set(qs)) >= 5
    assert all(50 <= p["k"] <= 4000 for p in prms)
    assert all(p["B_thresh"] < p["B"] and p["B_thresh"] < p["W_RLFC"] for p in prms)
    want = _singles(S, prms, xs, fts, wins, draws, "fp32")
    b = S.Batch(prms, "fp32")
    for i in range(len(prms)):
        b.set_filter(i, fts[i], wins[i])
    _same(b.execute(xs, draws), want)
    b.set_graph(True)
    for _ in range(2):
        _same(b.execute(xs, draws), want)
        (torch.randn(1 << 20, device="cuda") * 2).sum().item()
        torch.cuda.synchronize()
    b.set_graph(False)
    res = b.execute(xs, draws)
    _same(res, want)
    seen = set()
    for i, p in enumerate(prms):
        q = qs[i]
        if q in seen:
            continue
        seen.add(q)
        a, ai, off = draws[i]
        ref = O.outer_loop32(p, xs[i].cpu().numpy(), fts[i].cpu().numpy(), wins[i], a, ai, off)
        nodes, cnt = res[i]
        assert cnt == ref["I_F"], f"signal {i} (q={q})"
        assert nodes.tobytes() == ref["nodes"].tobytes(), f"signal {i} (q={q})"
    b.close()
