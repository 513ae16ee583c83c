"""Stage functions of the oracle against brute force written independently here in numpy /
pure Python: permute+filter (incl. the reference's tiled tail, innerLoop.cu:725-757),
samples, cutoff (ties), comb restriction (aliasing identity), J2 merge, reverse hash
(set semantics with the findUpperBound rule), estimation (exact medians)."""
import numpy as np
import pytest

import oracle as O
from helpers import bits


def _int_complex(rng, n, lo=-8, hi=9):
    return (rng.integers(lo, hi, n) + 1j * rng.integers(lo, hi, n)).astype(np.complex128)


def _dyadic_complex(rng, n):
    return ((rng.integers(-64, 65, n) + 1j * rng.integers(-64, 65, n)) / 64.0).astype(np.complex128)


def _pf_brute(x, n, f, B, ai, tiled):
    """Exactly-representable inputs make every summation order give the same bits."""
    w = f.size
    bins = np.zeros(B, dtype=np.complex128)
    if tiled:
        tiles, residue = w // B, w % B
        for r in range(B):
            for t in range(tiles):
                i = t * B + r
                bins[r] += x[(i * ai) % n] * f[i]
        for j in range(residue):                      # the thread with index B
            i = j * tiles + B
            bins[j] += x[(i * ai) % n] * f[i]
    else:
        for i in range(w):
            bins[i % B] += x[(i * ai) % n] * f[i]
    return bins


@pytest.mark.parametrize("n,B,w", [(1 << 10, 16, 101), (1 << 12, 64, 1001), (1 << 14, 128, 5000)])
@pytest.mark.parametrize("mode", [O.FOLD_REFERENCE, O.FOLD_ALL_TAPS])
def test_perm_filter64_brute_force(n, B, w, mode):
    rng = np.random.default_rng(n + w + mode)
    x = _int_complex(rng, n)
    f = _dyadic_complex(rng, w)
    ai = int(rng.integers(0, n // 2)) * 2 + 1
    got = O.perm_filter64(x, n, f, B, ai, mode)
    want = _pf_brute(x, n, f, B, ai, tiled=(mode == O.FOLD_REFERENCE))
    assert np.array_equal(got, want)


def test_perm_filter_reference_mode_is_flat_from_2_27():
    """n >= 2^27: the reference switches to the every-tap kernel (innerLoop.cu:230-237)."""
    n, B, w = 1 << 27, 32, 200
    x = np.zeros(n, dtype=np.complex128)
    rng = np.random.default_rng(3)
    ai = 12345679
    idx = (np.arange(w, dtype=np.int64) * ai) % n
    x[idx] = _int_complex(rng, w)
    f = _dyadic_complex(rng, w)
    got = O.perm_filter64(x, n, f, B, ai, O.FOLD_REFERENCE)
    want = O.perm_filter64(x, n, f, B, ai, O.FOLD_ALL_TAPS)
    assert np.array_equal(got, want)
    assert np.array_equal(got, _pf_brute(x, n, f, B, ai, tiled=False))


def test_perm_filter32_kahan_bound():
    rng = np.random.default_rng(9)
    n, B, w = 1 << 14, 64, 4001
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    f = (rng.standard_normal(w) + 1j * rng.standard_normal(w)).astype(np.complex64)
    ai = 4097
    got = O.perm_filter32(x, n, f, B, ai, O.FOLD_ALL_TAPS).astype(np.complex128)
    xx, ff = x.astype(np.complex128), f.astype(np.complex128)
    terms = np.zeros((B, -(-w // B)), dtype=np.complex128)
    for i in range(w):
        terms[i % B, i // B] = xx[(i * ai) % n] * ff[i]
    exact = terms.sum(axis=1)
    mag = np.abs(terms.real).sum(axis=1) + np.abs(terms.imag).sum(axis=1)
    eps = np.finfo(np.float32).eps
    # one rounding per product term plus the compensated sum: |err| <= 4 eps sum|terms|
    assert np.all(np.abs(got - exact) <= 4 * eps * mag + 1e-30)


def test_samples_rounding():
    """|X|^2 = re*re + im*im with each square rounded, then added (the reference's
    pow(re,2) + pow(im,2), innerLoop.cu:715): bit for bit numpy's elementwise form, which does
    exactly that.  A fused form (one rounding) differs from it on random inputs, so this pins
    the choice."""
    rng = np.random.default_rng(4)
    X = _int_complex(rng, 1000, -1000, 1000)
    assert np.array_equal(O.samples64(X), X.real ** 2 + X.imag ** 2)   # exact for integers
    Y = rng.standard_normal(4000) + 1j * rng.standard_normal(4000)
    s = O.samples64(Y)
    assert np.array_equal(s, Y.real * Y.real + Y.imag * Y.imag)
    re, im = Y.real.astype(np.longdouble), Y.imag.astype(np.longdouble)
    fused = (re * re + np.longdouble(1) * (Y.imag * Y.imag)).astype(np.float64)
    assert np.count_nonzero(fused != s) > 100          # the test tells the two forms apart
    Z = Y.astype(np.complex64)
    assert np.array_equal(O.samples32(Z), Z.real * Z.real + Z.imag * Z.imag)


def _pf_sequential(x, n, f, B, ai, tiled, fp32=False):
    """The fold as the reference writes it (innerLoop.cu:740-741, 770-771), restated in numpy:
    per tap the term x.x*f.x - x.y*f.y (products rounded, then combined), added to the bucket
    in ascending tap order; FP32 sums are Kahan-compensated.  Vectorised over buckets, so it
    is exact for any inputs, not only exactly representable ones."""
    w = f.size
    dt = np.float32 if fp32 else np.float64
    xr, xi = x.real.astype(dt), x.imag.astype(dt)
    fr, fi = f.real.astype(dt), f.imag.astype(dt)
    r = np.arange(B, dtype=np.int64)
    sx, sy = np.zeros(B, dt), np.zeros(B, dt)
    cx, cy = np.zeros(B, dt), np.zeros(B, dt)

    def add(ids, rows):
        j = (ids * ai) % n
        tx = xr[j] * fr[ids] - xi[j] * fi[ids]
        ty = xr[j] * fi[ids] + xi[j] * fr[ids]
        if not fp32:
            sx[rows] = sx[rows] + tx
            sy[rows] = sy[rows] + ty
            return
        for s_, c_, t_ in ((sx, cx, tx), (sy, cy, ty)):
            yv = t_ - c_[rows]
            tv = s_[rows] + yv
            c_[rows] = (tv - s_[rows]) - yv
            s_[rows] = tv

    if tiled:
        tiles, residue = w // B, w % B
        for t in range(tiles):
            add(t * B + r, r)
        rr = r[:residue]
        add(rr * tiles + B, rr)
    else:
        for t in range(-(-w // B)):
            ids = t * B + r
            ok = ids < w
            add(ids[ok], r[ok])
    return sx + 1j * sy if not fp32 else (sx + 1j * sy).astype(np.complex64)


@pytest.mark.parametrize("mode", [O.FOLD_REFERENCE, O.FOLD_ALL_TAPS])
def test_perm_filter_term_then_add_random_inputs(mode):
    rng = np.random.default_rng(21 + mode)
    n, B, w = 1 << 14, 128, 5000
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = rng.standard_normal(w) + 1j * rng.standard_normal(w)
    ai = int(rng.integers(0, n // 2)) * 2 + 1
    tiled = mode == O.FOLD_REFERENCE
    got = O.perm_filter64(x, n, f, B, ai, mode)
    assert np.array_equal(bits(got), bits(_pf_sequential(x, n, f, B, ai, tiled)))
    x32, f32 = x.astype(np.complex64), f.astype(np.complex64)
    got32 = O.perm_filter32(x32, n, f32, B, ai, mode)
    assert np.array_equal(bits(got32), bits(_pf_sequential(x32, n, f32, B, ai, tiled, True)))


def test_estimate_values_rounding_random_bins():
    """Estimation divides as innerLoop.cu:848-852 writes it: |F|^2 = F.x*F.x + F.y*F.y and
    bins*conj(F) with products rounded and combined, then the lower median; restated in
    numpy on random (non-exact) bins and filter values, bit for bit."""
    rng = np.random.default_rng(12)
    n, B, loops = 1 << 12, 64, 7
    m = n // B
    win = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    ai = (rng.integers(0, n // 2, loops) * 2 + 1).astype(np.int32)
    b = rng.standard_normal((loops, B)) + 1j * rng.standard_normal((loops, B))
    I = rng.choice(n, 50, replace=False).astype(np.int32)
    got = O.estimate_values(I, b, loops, n, ai, B, win)
    med = (loops - 1) // 2
    for h, Iv in enumerate(I):
        vr, vi = np.empty(loops), np.empty(loops)
        for l in range(loops):
            p = (int(ai[l]) * int(Iv)) % n
            hh, d = p // m, p % m
            if d > m // 2:
                hh, d = (hh + 1) % B, d - m
            F = win[((n - d) % n + m // 2) % n]
            bb = b[l, hh]
            a2 = F.real * F.real + F.imag * F.imag
            vr[l] = (bb.real * F.real + bb.imag * F.imag) / a2
            vi[l] = (bb.imag * F.real - bb.real * F.imag) / a2
        assert got[h]["key"] == Iv
        assert got[h]["re"] == np.sort(vr)[med] and got[h]["im"] == np.sort(vi)[med]


def _cutoff_brute(s, K):
    order = np.lexsort((np.arange(s.size), -s))      # value descending, then index ascending
    return np.sort(order[:K]).astype(np.int32)


@pytest.mark.parametrize("N,K", [(64, 10), (1024, 200), (8192, 2000), (1 << 16, 1000), (1000, 100),
                                 (5000, 1234), (70001, 30000)])
def test_cutoff_brute_force_with_ties(N, K):
    rng = np.random.default_rng(N + K)
    for s in (rng.random(N),                                   # distinct
              rng.integers(0, 5, N).astype(np.float64),       # massive ties
              np.zeros(N),                                      # all equal
              np.concatenate([np.ones(K // 2), rng.random(N - K // 2) * 0.5])):
        assert np.array_equal(O.cutoff64(s, K), _cutoff_brute(s, K))
        s32 = s.astype(np.float32)
        assert np.array_equal(O.cutoff32(s32, K), _cutoff_brute(s32.astype(np.float64), K))


def test_rlfc_aliasing_identity():
    """FFT_W(x[off + i sigma])[g] = W sum_{f = g mod W} X[f] exp(2 pi i f off / n)."""
    n, W, K = 1 << 14, 1 << 9, 40
    rng = np.random.default_rng(5)
    X = np.zeros(n, dtype=np.complex128)
    locs = rng.choice(n, 60, replace=False)
    X[locs] = np.exp(2j * np.pi * rng.random(60)) * (1 + rng.random(60))
    x = np.fft.ifft(X) * n
    for off in (0, 7, n // W - 1):
        f = np.arange(n)
        contrib = X * np.exp(2j * np.pi * f * off / n)
        Y = W * (np.bincount(f % W, weights=contrib.real, minlength=W) +
                 1j * np.bincount(f % W, weights=contrib.imag, minlength=W))
        want = _cutoff_brute(np.abs(Y) ** 2, K)
        got = O.rlfc64(x, n, W, off, K)
        assert np.array_equal(got, want)


def test_merge_J2():
    rng = np.random.default_rng(6)
    a = np.sort(rng.choice(1000, 50, replace=False))
    b = np.sort(rng.choice(1000, 50, replace=False))
    assert np.array_equal(O.merge_J2(np.concatenate([a, b])), np.union1d(a, b))


def _reverse_hash_brute(J, n, B, a, ai, thresh, vote, hits, J2, W):
    """Set semantics: candidates of bucket J are the permuted positions p of the window
    [J m - m/2, J m + m/2) (mod n, increasing from low) with p mod W in S = {J2 ai mod W};
    none when (low mod W) > max(S) (findUpperBound's last-index return, no wrap)."""
    m = n // B
    S = {(int(j) * ai) % W for j in J2}
    mx = max(S)
    for Jv in J:
        low = (int(Jv) * m - m // 2) % n
        if (low % W) > mx:
            continue
        for d in range(m):
            p = (low + d) % n
            if (p % W) in S:
                loc = (p * a) % n
                vote[loc] += 1
                if vote[loc] == thresh:
                    hits.append(loc)


@pytest.mark.parametrize("n,B,W,K,nJ2", [(1 << 12, 32, 1 << 9, 16, 60), (1 << 14, 128, 1 << 9, 40, 120),
                                         (1 << 16, 256, 1 << 12, 100, 300)])
def test_reverse_hash_brute_force(n, B, W, K, nJ2):
    rng = np.random.default_rng(n + K)
    J2 = np.sort(rng.choice(W, nJ2, replace=False)).astype(np.int32)
    vote_o = np.zeros(n, dtype=np.int32)
    hits_o = np.zeros(4096, dtype=np.int32)
    vote_b = np.zeros(n, dtype=np.int64)
    hits_b = []
    cnt = 0
    for loop in range(3):
        a = int(rng.integers(0, n // 2)) * 2 + 1
        ai = pow(a, -1, n)
        J = np.sort(rng.choice(B, K, replace=False)).astype(np.int32)
        cnt = O.reverse_hash(J, n, B, a, ai, 2, vote_o, hits_o, cnt, J2, W)
        _reverse_hash_brute(J, n, B, a, ai, 2, vote_b, hits_b, J2, W)
        assert cnt == len(hits_b)
        assert np.array_equal(hits_o[:cnt], np.array(hits_b, dtype=np.int32))
        assert np.array_equal(vote_o, vote_b.astype(np.int32))


def test_reverse_hash_upper_bound_rule():
    """A window whose start residue exceeds every approved residue yields no candidates even
    when the window runs into the next W-block (where approved residues exist)."""
    n, B, W = 1 << 12, 64, 128          # m = 64
    J2 = np.array([0, 1, 2, 3], dtype=np.int32)
    a = ai = 1
    # J = 2: low = 2*64 - 32 = 96 -> key 96 > max(S) = 3, window [96, 160) reaches p=128..131
    vote = np.zeros(n, dtype=np.int32)
    hits = np.zeros(16, dtype=np.int32)
    c = O.reverse_hash(np.array([2], dtype=np.int32), n, B, a, ai, 1, vote, hits, 0, J2, W)
    assert c == 0 and vote.sum() == 0
    # J = 0: low = n - 32 (key 96 > 3 as well) -> nothing; J = 1: low = 32 -> key 32 > 3 too
    # J = 2 with a residue set reaching above 96 finds 128..131 through the wrap
    J2b = np.array([0, 1, 2, 3, 100], dtype=np.int32)
    c = O.reverse_hash(np.array([2], dtype=np.int32), n, B, a, ai, 1, vote, hits, 0, J2b, W)
    assert c == 5 and sorted(hits[:c].tolist()) == [100, 128, 129, 130, 131]


def test_reverse_hash_overflow_reported():
    n, B, W = 1 << 12, 32, 1 << 9
    J2 = np.arange(0, W, 2, dtype=np.int32)
    vote = np.zeros(n, dtype=np.int32)
    hits = np.zeros(8, dtype=np.int32)
    c = O.reverse_hash(np.arange(16, dtype=np.int32), n, B, 3, pow(3, -1, n), 1, vote, hits, 0,
                       J2, W, hit_capacity=8)
    assert c == -1


def test_estimate_values_exact_medians():
    """bins chosen as v_l * F_l with |F| = 1 (exact quarter turns) give v_l back exactly; the
    node value is the lower median of real and imaginary parts separately."""
    rng = np.random.default_rng(8)
    n, B, loops = 1 << 12, 64, 7
    m = n // B
    win = np.array([[1, 1j, -1, -1j][j % 4] for j in range(m)], dtype=np.complex128)
    ai = (rng.integers(0, n // 2, loops) * 2 + 1).astype(np.int32)
    med = (loops - 1) // 2
    for I in rng.choice(n, 20, replace=False):
        vals = (rng.integers(-50, 50, loops) + 1j * rng.integers(-50, 50, loops)) / 4.0
        b = np.zeros((loops, B), dtype=np.complex128)
        for l in range(loops):
            p = (int(ai[l]) * int(I)) % n
            h, d = p // m, p % m
            if d > m // 2:
                h, d = (h + 1) % B, d - m
            b[l, h] = vals[l] * win[((n - d) % n + m // 2) % n]
        node = O.estimate_values(np.array([I], dtype=np.int32), b, loops, n, ai, B, win)[0]
        assert node["key"] == I
        assert node["re"] == np.sort(vals.real)[med] and node["im"] == np.sort(vals.imag)[med]


@pytest.mark.parametrize("log2n,log2B", [(12, 5), (20, 13), (27, 17)])
def test_bucket_taps_share_one_column(log2n, log2B):
    """The fact the column-ordered fold rests on: every tap id = t*B + r of bucket r gathers
    x[(id*ai) mod n], whose column (id*ai) mod B is (r*ai) mod B for every t; column c holds
    bucket (c*a) mod B of the loop, a = ai^-1 mod n."""
    rng = np.random.default_rng(log2n)
    n, B = 1 << log2n, 1 << log2B
    for _ in range(4):
        a = int(rng.integers(0, n // 2)) * 2 + 1
        ai = pow(a, -1, n)
        r = rng.integers(0, B, 64).astype(np.int64)
        t = rng.integers(0, 40, 64).astype(np.int64)
        j = ((t * B + r) * ai) % n
        assert np.array_equal(j % B, (r * ai) % B)
        assert np.array_equal(((j % B) * a) % B, r)
