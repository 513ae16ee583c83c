"""samples_indices_noalloc (no counterpart in the reference: samples_indices for callers that
cannot allocate) through the C ABI against samples_indices (innerLoop.h:35-36) and the oracle's
cutoff, on two calls in a row -- the second smaller, so it runs on the contents the first left
in the shared device scratch -- with ties, and the size limit rejected without touching the
output."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import sfft_amd
    return sfft_amd


def _vp(t):
    return C.c_void_p(t.data_ptr())


def _both(S, X, K):
    import torch
    dX = torch.from_numpy(np.ascontiguousarray(X)).to("cuda")
    N = X.size
    a = torch.full((K,), -7, dtype=torch.int32, device="cuda")
    b = torch.full((K,), -7, dtype=torch.int32, device="cuda")
    ta, tb = C.c_double(0.0), C.c_double(0.0)
    assert S.lib.samples_indices_noalloc(_vp(dX), K, N, _vp(a), C.byref(ta)) == 0
    assert S.lib.samples_indices(_vp(dX), K, N, _vp(b), C.byref(tb)) == 0
    assert ta.value > 0
    return a.cpu().numpy(), b.cpu().numpy()


def test_noalloc_cutoff_matches_allocating_entry_point(S):
    rng = np.random.default_rng(2024)
    N1 = 262144                                   # SFFT_NOALLOC_MAX_SAMPLES
    X1 = rng.standard_normal(N1) + 1j * rng.standard_normal(N1)
    X2 = (rng.integers(-2, 3, 4096) + 1j * rng.integers(-2, 3, 4096)).astype(np.complex128)
    for X, K in ((X1, 20000), (X2, 300), (X1[:100000], 1), (X2[:1000], 999)):
        got, want = _both(S, X, K)
        assert np.array_equal(got, want)
        assert np.array_equal(got, O.cutoff64(O.samples64(X), K))


def test_noalloc_rejects_sizes_past_the_maximum(S):
    import torch
    N = 262144 + 1
    dX = torch.zeros(N, dtype=torch.complex128, device="cuda")
    out = torch.full((10,), -7, dtype=torch.int32, device="cuda")
    t = C.c_double(0.0)
    assert S.lib.samples_indices_noalloc(_vp(dX), 10, N, _vp(out), C.byref(t)) == -2
    assert S.lib.samples_indices_noalloc(_vp(dX), 0, 100, _vp(out), C.byref(t)) == -1
    torch.cuda.synchronize()
    assert np.all(out.cpu().numpy() == -7)
