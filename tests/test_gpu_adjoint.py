"""The value map and its adjoint on the GPU (sfft_value_apply / sfft_value_adjoint through the C
ABI) against the oracle, bit for bit, and the differentiable transform (sparse_fft, a
torch.autograd.Function) against FP64 central finite differences and the dense DFT gradient."""
import numpy as np
import pytest

import oracle as O
from helpers import N20_K1000, SMALL_CONFIGS, bits, draws, oracle_filter, params, sparse_signal

pytestmark = pytest.mark.gpu

# n=2^22, k=1000, Bcst=0.99, loops 2+5, thresh 2, RLFC 256 x 2 (the differentiable-path
# workload)
N22_K1000 = (22, 1000, 0.99, 2, 2, 5, 256, 2)


@pytest.fixture(scope="module")
def S():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import sfft_amd
    return sfft_amd


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _setup(S, cfg, seed, fold_mode=O.FOLD_REFERENCE):
    prm = params(cfg)
    n, B = prm["n"], prm["B"]
    x, locs, _ = sparse_signal(n, prm["k"], seed)
    ft, ff = oracle_filter(prm)
    win = O.filter_window_slice(ff, n, B)
    a, ai, offs = draws(n, prm["loops"], prm["RLFC_loops"], prm["W_RLFC"], seed + 3)
    plan = S.Plan(prm, "fp64", fold_mode=fold_mode)
    dft = _dev(ft)
    plan.set_filter(dft, win)
    return prm, x, ft, win, a, ai, offs, plan, locs


@pytest.mark.parametrize("fold_mode", [O.FOLD_REFERENCE, O.FOLD_ALL_TAPS])
@pytest.mark.parametrize("cfg", SMALL_CONFIGS[:3] + [N20_K1000])
def test_value_map_and_adjoint_bitexact(S, cfg, fold_mode):
    if cfg == N20_K1000 and fold_mode == O.FOLD_ALL_TAPS:
        pytest.skip("covered by the small configurations")
    prm, x, ft, win, a, ai, offs, plan, _ = _setup(S, cfg, 61, fold_mode)
    dx = _dev(x)
    nodes, cnt = plan.execute(dx, a, ai, offs)
    hits, sel = plan.selection(cnt)
    ref = O.outer_loop64(prm, x, ft, win, a, ai, offs, fold_mode=fold_mode, want_bins=True)
    sel_ref = O.estimate_select(ref["hits"], ref["bins_f"], prm["loops"], prm["n"], ai,
                                prm["B"], win)
    assert np.array_equal(sel.cpu().numpy(), sel_ref)
    # applied to the transform's own signal the map returns the node values
    v = plan.value_apply(dx, a, ai, offs, hits, sel).cpu().numpy()
    assert np.array_equal(bits(v.real), bits(nodes["re"]))
    assert np.array_equal(bits(v.imag), bits(nodes["im"]))
    rng = np.random.default_rng(7)
    xr = rng.standard_normal(prm["n"]) + 1j * rng.standard_normal(prm["n"])
    g = rng.standard_normal(cnt) + 1j * rng.standard_normal(cnt)
    hits_np = ref["hits"]
    v2 = plan.value_apply(_dev(xr), a, ai, offs, hits, sel).cpu().numpy()
    assert np.array_equal(bits(v2), bits(O.value_apply64(prm, xr, ft, win, ai, hits_np, sel_ref,
                                                         fold_mode)))
    gx = plan.value_adjoint(_dev(g), a, ai, offs, hits, sel).cpu().numpy()
    gx_ref = O.value_adjoint64(prm, ft, win, ai, hits_np, sel_ref, g, fold_mode)
    assert np.array_equal(bits(gx), bits(gx_ref))
    # adjoint identity on the GPU results
    lhs, rhs = np.vdot(v2, g).real, np.vdot(xr, gx).real
    assert abs(lhs - rhs) <= 1e-12 * (np.linalg.norm(v2) * np.linalg.norm(g) +
                                      np.linalg.norm(xr) * np.linalg.norm(gx))


def test_adjoint_is_deterministic(S):
    prm, x, ft, win, a, ai, offs, plan, _ = _setup(S, SMALL_CONFIGS[2], 71)
    nodes, cnt = plan.execute(_dev(x), a, ai, offs)
    hits, sel = plan.selection(cnt)
    g = _dev(np.random.default_rng(3).standard_normal(cnt) + 0j)
    r1 = plan.value_adjoint(g, a, ai, offs, hits, sel).cpu().numpy()
    r2 = plan.value_adjoint(g, a, ai, offs, hits, sel).cpu().numpy()
    assert np.array_equal(bits(r1), bits(r2))


def test_fp32_plan_rejects_the_adjoint(S):
    import torch
    prm = params(SMALL_CONFIGS[1])
    plan = S.Plan(prm, "fp32")
    hits = torch.zeros(1, dtype=torch.int32, device="cuda")
    sel = torch.zeros((1, 2), dtype=torch.uint8, device="cuda")
    a, ai, offs = draws(prm["n"], prm["loops"], prm["RLFC_loops"], prm["W_RLFC"], 1)
    with pytest.raises(S.SfftError) as e:
        plan.value_adjoint(torch.zeros(1, dtype=torch.complex128, device="cuda"), a, ai, offs,
                           hits, sel)
    assert e.value.status in (1, 5)


def _loss(vals, w):
    return float(np.sum(w * np.abs(vals) ** 2))


def test_autograd_n22_k1000_finite_differences_and_dense_gradient(S):
    """n=2^22, k=1000 (complex64 signal, differentiated in FP64): dLoss/dx of
    Loss = sum_j w_j |Xhat_j|^2, w_j ~ U(0,1), by the adjoint kernels, against FP64 central
    finite differences (step 1e-6) at 256 random signal indices with the located indices and
    median choices held fixed, and against the exact gradient of the dense DFT coefficients
    at the located indices."""
    import torch
    prm = params(N22_K1000)
    n, k, B = prm["n"], prm["k"], prm["B"]
    x, locs, _ = sparse_signal(n, k, 2022)
    x = x.astype(np.complex64).astype(np.complex128)
    ft, ff = oracle_filter(prm)
    win = O.filter_window_slice(ff, n, B)
    a, ai, offs = draws(n, prm["loops"], prm["RLFC_loops"], prm["W_RLFC"], 2023)
    plan = S.Plan(prm, "fp64")
    plan.set_filter(_dev(ft), win)
    xt = _dev(x).requires_grad_(True)
    vals, keys = S.sparse_fft(xt, plan, a, ai, offs)
    keys_np = keys.cpu().numpy()
    assert set(locs.tolist()) <= set(keys_np.tolist())
    w = np.random.default_rng(5).random(keys_np.size)
    wt = torch.from_numpy(w).to("cuda")
    loss = (wt * vals.abs() ** 2).sum()
    loss.backward()
    grad = xt.grad.cpu().numpy()

    # central differences of the loss with the located indices and median choices held fixed
    # (they are integer decisions), i.e. through the forward value map (permute+filter, bucket
    # FFT, filter division: sfft_value_apply), at 256 random indices, real and imaginary
    # parts.  The loss is then a quadratic form in x, so a central difference has no
    # truncation error at any step; what limits it is the rounding of the loss (|Loss| ~ 5e2
    # against changes of |grad| * h).  At the 1e-6 step the rounding dominates; the 1e-2
    # step resolves the gradient to ~1e-8.
    hits0, sel0 = plan.selection(keys_np.size)
    v0 = plan.value_apply(_dev(x), a, ai, offs, hits0, sel0)
    assert torch.equal(v0, vals.detach())
    rng = np.random.default_rng(11)
    idx = rng.choice(n, size=256, replace=False)
    xd = _dev(x)
    fd = {1e-6: np.zeros(256, dtype=np.complex128), 1e-2: np.zeros(256, dtype=np.complex128)}
    for j, t in enumerate(idx):
        for part, unit in ((0, 1.0), (1, 1j)):
            for h in fd:
                lv = []
                for sgn in (1.0, -1.0):
                    xd[t] = complex(x[t] + sgn * h * unit)
                    v = plan.value_apply(xd, a, ai, offs, hits0, sel0).cpu().numpy()
                    lv.append(_loss(v, w))
                d = (lv[0] - lv[1]) / (2 * h)
                fd[h][j] += d if part == 0 else 1j * d
            xd[t] = complex(x[t])
    g = grad[idx]
    scale = np.max(np.abs(g))
    L0 = float(loss.item())
    rel = {h: float(np.max(np.abs(f - g)) / scale) for h, f in fd.items()}
    print(f"n=2^22 k=1000 |Loss| = {L0:.3f}, max |dLoss/dx| = {scale:.3e}; central differences: "
          f"max |fd - grad| / max |grad| = {rel[1e-6]:.3e} (step 1e-6), "
          f"{rel[1e-2]:.3e} (step 1e-2)", flush=True)
    # rounding bound of a 1e-6 difference: a few ulps of the loss over 2h
    assert np.max(np.abs(fd[1e-6] - g)) <= 64 * np.finfo(float).eps * L0 / 2e-6
    assert rel[1e-2] <= 1e-6

    # dense DFT: Xhat = FFT(x)/n, and d/dXhat_j of sum_j w_j |Xhat_j|^2 is 2 w_j Xhat_j; the
    # spectrum of a signal gradient (FFT of dLoss/dx) at the located indices is the gradient
    # with respect to those coefficients.  Compared over the true locations among the hits
    # (loss restricted to them): what remains is the estimator's own cross-sensitivity (each
    # value also responds, through the filter's transition band, to coefficients in nearby
    # buckets of its median-selected loop).  Over all hits the false positives, whose values
    # are leakage of true coefficients, add gradient at the true frequencies: that figure is
    # printed for the record.
    Xh = np.fft.fft(x) / n
    true = np.isin(keys_np, locs)
    rel_dense = {}
    for name, mask in (("true locations", true), ("all located indices", np.ones_like(true))):
        xt2 = _dev(x).requires_grad_(True)
        v2, _ = S.sparse_fft(xt2, plan, a, ai, offs)
        ((torch.from_numpy(np.where(mask, w, 0.0)).to("cuda") * v2.abs() ** 2).sum()).backward()
        spec = np.fft.fft(xt2.grad.cpu().numpy())[keys_np[mask]]
        dense = 2 * w[mask] * Xh[keys_np[mask]]
        rel_dense[name] = float(np.linalg.norm(spec - dense) / np.linalg.norm(dense))
    print("n=2^22 k=1000 dense-DFT gradient at the located indices, relative error: "
          + ", ".join(f"{k_}: {v_:.3e}" for k_, v_ in rel_dense.items()), flush=True)
    assert rel_dense["true locations"] <= 0.1
