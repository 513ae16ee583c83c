"""The oracle's value map (estimated values of fixed hits as a function of the signal) and its
adjoint: the map reproduces the transform's node values bit for bit, the adjoint is the
transpose of the map (adjoint identity, and entry by entry against the dense matrix of the
map on a tiny signal)."""
import numpy as np
import pytest

import oracle as O
from helpers import SMALL_CONFIGS, draws, oracle_filter, params, sparse_signal


def _transform(cfg, seed, fold_mode):
    prm = params(cfg)
    n, B = prm["n"], prm["B"]
    x, locs, _ = sparse_signal(n, prm["k"], seed)
    ft, ff = oracle_filter(prm)
    win = O.filter_window_slice(ff, n, B)
    a, ai, offs = draws(n, prm["loops"], prm["RLFC_loops"], prm["W_RLFC"], seed + 1)
    ref = O.outer_loop64(prm, x, ft, win, a, ai, offs, fold_mode=fold_mode, want_bins=True)
    sel = O.estimate_select(ref["hits"], ref["bins_f"], prm["loops"], n, ai, B, win)
    return prm, x, ft, win, ai, ref, sel


@pytest.mark.parametrize("fold_mode", [O.FOLD_REFERENCE, O.FOLD_ALL_TAPS])
@pytest.mark.parametrize("cfg", SMALL_CONFIGS[:3])
def test_value_apply_reproduces_the_transform(cfg, fold_mode):
    prm, x, ft, win, ai, ref, sel = _transform(cfg, 41, fold_mode)
    vals = O.value_apply64(prm, x, ft, win, ai, ref["hits"], sel, fold_mode)
    assert np.array_equal(vals.real.view(np.uint64), ref["nodes"]["re"].view(np.uint64))
    assert np.array_equal(vals.imag.view(np.uint64), ref["nodes"]["im"].view(np.uint64))
    # the selected loops are the ones whose per-loop estimates sit at the lower median
    assert sel.max() < prm["loops"]


@pytest.mark.parametrize("fold_mode", [O.FOLD_REFERENCE, O.FOLD_ALL_TAPS])
@pytest.mark.parametrize("cfg", SMALL_CONFIGS[:3])
def test_adjoint_identity(cfg, fold_mode):
    prm, x, ft, win, ai, ref, sel = _transform(cfg, 43, fold_mode)
    rng = np.random.default_rng(5)
    n, nh = prm["n"], ref["hits"].size
    for _ in range(3):
        xr = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        g = rng.standard_normal(nh) + 1j * rng.standard_normal(nh)
        Ax = O.value_apply64(prm, xr, ft, win, ai, ref["hits"], sel, fold_mode)
        ATg = O.value_adjoint64(prm, ft, win, ai, ref["hits"], sel, g, fold_mode)
        lhs = np.vdot(Ax, g).real          # <A x, g> over R^2
        rhs = np.vdot(xr, ATg).real        # <x, A^T g>
        scale = np.linalg.norm(Ax) * np.linalg.norm(g) + np.linalg.norm(xr) * np.linalg.norm(ATg)
        assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs)


def test_adjoint_is_the_dense_transpose():
    """Build the real 2nh x 2n matrix of the map column by column (unit real and imaginary
    impulses) and compare its transpose with the adjoint applied to unit gradients."""
    cfg = SMALL_CONFIGS[0]
    prm, x, ft, win, ai, ref, sel = _transform(cfg, 47, O.FOLD_REFERENCE)
    n, nh = prm["n"], ref["hits"].size
    hits = ref["hits"]
    M = np.zeros((2 * nh, 2 * n))
    e = np.zeros(n, dtype=np.complex128)
    for m in range(n):
        for part, unit in ((0, 1.0), (1, 1j)):
            e[m] = unit
            col = O.value_apply64(prm, e, ft, win, ai, hits, sel)
            M[:nh, 2 * m + part] = col.real
            M[nh:, 2 * m + part] = col.imag
            e[m] = 0
    rng = np.random.default_rng(9)
    for _ in range(4):
        g = rng.standard_normal(nh) + 1j * rng.standard_normal(nh)
        gx = O.value_adjoint64(prm, ft, win, ai, hits, sel, g)
        want = M.T @ np.concatenate([g.real, g.imag])
        got = np.empty(2 * n)
        got[0::2], got[1::2] = gx.real, gx.imag
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    # and the map itself is the matrix (linearity on the transform's own signal)
    v = O.value_apply64(prm, x, ft, win, ai, hits, sel)
    xv = np.empty(2 * n)
    xv[0::2], xv[1::2] = x.real, x.imag
    mv = M @ xv
    assert np.max(np.abs(mv[:nh] - v.real)) <= 1e-9 * np.max(np.abs(v))
    assert np.max(np.abs(mv[nh:] - v.imag)) <= 1e-9 * np.max(np.abs(v))
