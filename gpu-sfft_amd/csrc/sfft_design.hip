// sfft_design.hip -- FP64 filter design (the make_multiple_t computation), seeded synthetic
// k-sparse signals, the truth spectrum and the run_experiment error metric.  None of this is
// inside the timed transform; the FFTs it needs run on the GPU with the canonical kernels.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unordered_set>
#include <vector>

#include "sfft_aux.h"

using namespace sfft;

#define HCK(x)                               \
  do {                                       \
    if ((x) != hipSuccess) {                 \
      status = SFFT_ERR_HIP;                 \
      goto done;                             \
    }                                        \
  } while (0)

static int ilog2_pow2(long long v) {
  if (v <= 0) return -1;
  int l = 0;
  while ((1LL << l) < v) ++l;
  return ((1LL << l) == v) ? l : -1;
}

// ---------------------------------------------------------------------------------------------
// filter design: cuExperiment.cu:248-447
// ---------------------------------------------------------------------------------------------
extern "C" int sfft_filter_taps(double lobefrac, double tolerance, double* t0_out) {
  // cuExperiment.cu:279-284
  int w = (int)((1 / M_PI) * (1 / lobefrac) * acosh(1. / tolerance));
  if (!(w % 2)) w--;
  if (t0_out) *t0_out = cosh(acosh(1 / tolerance) / (w - 1));
  return w;
}

static double cheb(double m, double x) {  // cuCheb, cuExperiment.cu:593-601
  if (fabs(x) <= 1) return cos(m * acos(x));
  return cosh(m * acosh(x));               // NaN below -1, zeroed by the caller
}

// DFT of arbitrary length N on the device: canonical FFT when N is a power of two, else
// Bluestein with the chirp c[m] = exp(-+ i pi (m^2 mod 2N) / N), built on the host with libm's
// sincos (faithful design) or on the device with sincospi (device_chirp).
static sfft_status dft_any_device(double2* d_data, int N, int inverse, hipStream_t st,
                                  bool device_chirp = false) {
  sfft_status status = SFFT_OK;
  const int l2 = ilog2_pow2(N);
  void *scr = nullptr, *A = nullptr, *Bv = nullptr, *c = nullptr, *tmp = nullptr;
  if (N <= 1) return SFFT_OK;
  if (l2 >= 0) {
    HCK(hipMalloc(&scr, (size_t)N * 16));
    HCK(hipMalloc(&tmp, (size_t)N * 16));
    HCK(hipMemcpyAsync(tmp, d_data, (size_t)N * 16, hipMemcpyDeviceToDevice, st));
    HCK(fft_plain(tmp, d_data, scr, l2, 1, inverse, 0, st));
    HCK(hipStreamSynchronize(st));
    goto done;
  }
  {
    int lm = 0;
    while ((1LL << lm) < 2LL * N - 1) ++lm;
    const long long M = 1LL << lm;
    std::vector<double2> hc, hb;
    const double sgn = inverse ? 1.0 : -1.0;
    HCK(hipMalloc(&c, (size_t)N * 16));
    HCK(hipMalloc(&A, (size_t)M * 16));
    HCK(hipMalloc(&Bv, (size_t)M * 16));
    HCK(hipMalloc(&tmp, (size_t)M * 16));
    HCK(hipMalloc(&scr, (size_t)M * 16));
    if (device_chirp) {
      HCK(aux_chirp((double2*)c, (double2*)tmp, N, M, inverse, st));
      goto chirp_ready;
    }
    hc.resize((size_t)N);
    hb.assign((size_t)M, make_double2(0.0, 0.0));
    for (long long m = 0; m < N; ++m) {
      const unsigned long long r =
          ((unsigned long long)m * (unsigned long long)m) % (2ULL * (unsigned long long)N);
      const double ang = M_PI * (double)r / (double)N;
      double sa, ca;
      sincos(ang, &sa, &ca);   // one libm call: sin()/cos() may round differently from it
      hc[m].x = ca;
      hc[m].y = sgn * sa;
    }
    hb[0] = make_double2(hc[0].x, -hc[0].y);
    for (long long m = 1; m < N; ++m) {
      hb[m] = make_double2(hc[m].x, -hc[m].y);
      hb[M - m] = hb[m];
    }
    HCK(hipMemcpy(c, hc.data(), (size_t)N * 16, hipMemcpyHostToDevice));
    HCK(hipMemcpy(tmp, hb.data(), (size_t)M * 16, hipMemcpyHostToDevice));
  chirp_ready:
    HCK(fft_plain(tmp, Bv, scr, lm, 1, 0, 0, st));
    HCK(aux_bluestein_pre(d_data, (const double2*)c, (double2*)tmp, N, M, st));
    HCK(fft_plain(tmp, A, scr, lm, 1, 0, 0, st));
    HCK(aux_cmul_pointwise((double2*)A, (const double2*)Bv, M, st));
    HCK(fft_plain(A, tmp, scr, lm, 1, 1, 0, st));
    HCK(aux_bluestein_post((const double2*)tmp, (const double2*)c, d_data, N, (double)M, st));
    HCK(hipStreamSynchronize(st));
  }
done:
  if (scr) (void)hipFree(scr);
  if (A) (void)hipFree(A);
  if (Bv) (void)hipFree(Bv);
  if (c) (void)hipFree(c);
  if (tmp) (void)hipFree(tmp);
  return status;
}

extern "C" sfft_status sfft_dft_any_device(complex_t* d_data, int N, int inverse, void* stream) {
  if (!d_data || N < 1) return SFFT_ERR_INVALID;
  return dft_any_device((double2*)d_data, N, inverse, (hipStream_t)stream);
}

// The parallel design: the same filter with every n-length step on the device and no
// sequential recurrence (device Chebyshev samples and chirp, panel-sum boxcar, device max,
// exactly reduced phase ramp).  Not bit-equal to the faithful design: it evaluates the
// Chebyshev samples without the direct form's cancellation (relative error ~1e-15 instead of
// up to m^2 eps / 19) and its rounding does not drift with n.
static sfft_status make_filter_parallel(double tolerance, int n, int ln, int b_f, int w,
                                        complex_t* d_filter_time, complex_t* h_filter_freq,
                                        complex_t* d_filter_freq, hipStream_t st) {
  sfft_status status = SFFT_OK;
  void *dv = nullptr, *g = nullptr, *G = nullptr, *scr = nullptr, *pan = nullptr, *mx = nullptr;
  const long long nb = aux_boxcar_blocks(n);
  double2* F = nullptr;
  HCK(hipMalloc(&dv, (size_t)w * 16));
  HCK(aux_cheb_samples((double2*)dv, w, tolerance, st));
  status = dft_any_device((double2*)dv, w, 0, st, true);
  if (status != SFFT_OK) goto done;
  HCK(aux_shift_real((const double2*)dv, (double2*)d_filter_time, w, st));
  HCK(hipMalloc(&g, (size_t)n * 16));
  HCK(hipMalloc(&G, (size_t)n * 16));
  HCK(hipMalloc(&scr, (size_t)n * 16));
  HCK(hipMalloc(&pan, (size_t)(n / 64 + n / 4096) * 16));
  HCK(hipMalloc(&mx, (size_t)(nb + 1) * 8));
  HCK(aux_fold((const double2*)d_filter_time, (double2*)g, w, n, st));
  HCK(fft_plain(g, G, scr, ln, 1, 0, 0, st));
  HCK(aux_boxcar((const double2*)G, (double2*)g, n, b_f, (double2*)pan, (double*)mx + 1,
                 (double*)mx, st));
  F = d_filter_freq ? (double2*)d_filter_freq : (double2*)g;
  HCK(aux_ramp((const double2*)g, F, n, w / 2, (const double*)mx, st));
  if (h_filter_freq)
    HCK(hipMemcpyAsync(h_filter_freq, F, (size_t)n * 16, hipMemcpyDeviceToHost, st));
  HCK(fft_plain(F, G, scr, ln, 1, 1, 0, st));
  HCK(aux_scale_head((const double2*)G, (double2*)d_filter_time, w, (double)n, st));
  HCK(hipStreamSynchronize(st));
done:
  if (dv) (void)hipFree(dv);
  if (g) (void)hipFree(g);
  if (G) (void)hipFree(G);
  if (scr) (void)hipFree(scr);
  if (pan) (void)hipFree(pan);
  if (mx) (void)hipFree(mx);
  return status;
}

// make_multiple_t's filter computation.  d_filter_time: device complex128[w] (output).
// h_filter_freq: host complex128[n], d_filter_freq: device complex128[n] (outputs, may be NULL).
extern "C" sfft_status sfft_make_filter_ex(double lobefrac, double tolerance, int n, int b_f,
                                           int w, int mode, complex_t* d_filter_time,
                                           complex_t* h_filter_freq, complex_t* d_filter_freq,
                                           void* stream) {
  (void)lobefrac;
  const int ln = ilog2_pow2(n);
  if (ln < 1 || w < 3 || w > n || !d_filter_time || b_f < 1 || b_f > n)
    return SFFT_ERR_INVALID;
  if (mode != SFFT_DESIGN_FAITHFUL && mode != SFFT_DESIGN_PARALLEL) return SFFT_ERR_INVALID;
  if (!(tolerance > 0.0 && tolerance < 1.0)) return SFFT_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  if (mode == SFFT_DESIGN_PARALLEL) {
    if (n < 4096) return SFFT_ERR_INVALID;
    return make_filter_parallel(tolerance, n, ln, b_f, w, d_filter_time, h_filter_freq,
                                d_filter_freq, st);
  }
  sfft_status status = SFFT_OK;
  void *dv = nullptr, *g = nullptr, *G = nullptr, *scr = nullptr;
  std::vector<double2> hv((size_t)w);
  std::vector<double2> hG;
  complex_t* ff = h_filter_freq;
  std::vector<complex_t> ff_own;
  const double t0 = cosh(acosh(1 / tolerance) / (w - 1));
  // Dolph-Chebyshev samples, NaN -> 0 (makeDolphchebyshevKernel1, cuExperiment.cu:538-552)
  for (int i = 0; i < w; ++i) {
    const double val = cheb(w - 1, t0 * cos(M_PI * i / (w))) * tolerance;
    hv[i].x = std::isnan(val) ? 0.0 : val;
    hv[i].y = 0.0;
  }
  HCK(hipMalloc(&dv, (size_t)w * 16));
  HCK(hipMemcpy(dv, hv.data(), (size_t)w * 16, hipMemcpyHostToDevice));
  status = dft_any_device((double2*)dv, w, 0, st);   // FFT_w, cuExperiment.cu:305-308
  if (status != SFFT_OK) goto done;
  HCK(aux_shift_real((const double2*)dv, (double2*)d_filter_time, w, st));   // cuShift + Kernel2
  HCK(hipMalloc(&g, (size_t)n * 16));
  HCK(hipMalloc(&G, (size_t)n * 16));
  HCK(hipMalloc(&scr, (size_t)n * 16));
  HCK(aux_fold((const double2*)d_filter_time, (double2*)g, w, n, st));       // :336-352 (drops f[w-1])
  HCK(fft_plain(g, G, scr, ln, 1, 0, 0, st));                               // FFT_n, :362-365
  hG.resize((size_t)n);
  HCK(hipMemcpyAsync(hG.data(), G, (size_t)n * 16, hipMemcpyDeviceToHost, st));
  HCK(hipStreamSynchronize(st));
  if (!ff) {
    ff_own.resize((size_t)n);
    ff = ff_own.data();
  }
  {
    // boxcar running sum (cuExperiment.cu:373-396), a sequential recurrence
    double sx = 0.0, sy = 0.0;
    for (int i = 0; i < b_f; ++i) {
      sx += hG[i].x;
      sy += hG[i].y;
    }
    double mx = 0.0;
    const long long offset = b_f / 2, mask = (long long)n - 1;
    for (long long i = 0; i < n; ++i) {
      ff[(i + n + offset) & mask].x = sx;
      ff[(i + n + offset) & mask].y = sy;
      const double ab = sqrt(sx * sx + sy * sy);
      mx = mx > ab ? mx : ab;
      sx += hG[(i + b_f) & mask].x - hG[i].x;
      sy += hG[(i + b_f) & mask].y - hG[i].y;
    }
    for (long long i = 0; i < n; ++i) {   // makeMultipleKernel1
      ff[i].x /= mx;
      ff[i].y /= mx;
    }
    // phase ramp by a sequential complex recurrence (cuExperiment.cu:407-421)
    double ox = 1.0, oy = 0.0;
    double stx, sty;
    sincos(-2 * M_PI * (w / 2) / n, &sty, &stx);
    for (long long i = 0; i < n; ++i) {
      const double ax = ff[i].x, ay = ff[i].y;
      ff[i].x = ax * ox - ay * oy;
      ff[i].y = ax * oy + ay * ox;
      const double nx = ox * stx - oy * sty;
      const double ny = ox * sty + oy * stx;
      ox = nx;
      oy = ny;
    }
  }
  HCK(hipMemcpy(G, ff, (size_t)n * 16, hipMemcpyHostToDevice));
  if (d_filter_freq)
    HCK(hipMemcpyAsync(d_filter_freq, G, (size_t)n * 16, hipMemcpyDeviceToDevice, st));
  HCK(fft_plain(G, g, scr, ln, 1, 1, 0, st));                               // IFFT_n, :424-426
  HCK(aux_scale_head((const double2*)g, (double2*)d_filter_time, w, (double)n, st));  // Kernel2
  HCK(hipStreamSynchronize(st));
done:
  if (dv) (void)hipFree(dv);
  if (g) (void)hipFree(g);
  if (G) (void)hipFree(G);
  if (scr) (void)hipFree(scr);
  return status;
}

extern "C" sfft_status sfft_make_filter(double lobefrac, double tolerance, int n, int b_f, int w,
                                        complex_t* d_filter_time, complex_t* h_filter_freq,
                                        void* stream) {
  return sfft_make_filter_ex(lobefrac, tolerance, n, b_f, w, SFFT_DESIGN_FAITHFUL,
                             d_filter_time, h_filter_freq, nullptr, stream);
}

// win[j] = filter_freq[(j - m/2) mod n] from a device response: two contiguous copies
extern "C" sfft_status sfft_filter_window_device(const complex_t* d_filter_freq, int n, int B,
                                                 complex_t* h_win, void* stream) {
  if (!d_filter_freq || !h_win || ilog2_pow2(n) < 0 || B < 1 || B > n || n % B)
    return SFFT_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  const long long m = n / B, h = m / 2;
  if (h > 0 && hipMemcpyAsync(h_win, d_filter_freq + (n - h), (size_t)h * 16,
                              hipMemcpyDeviceToHost, st) != hipSuccess)
    return SFFT_ERR_HIP;
  if (hipMemcpyAsync(h_win + h, d_filter_freq, (size_t)(m - h) * 16, hipMemcpyDeviceToHost,
                     st) != hipSuccess)
    return SFFT_ERR_HIP;
  return hipStreamSynchronize(st) == hipSuccess ? SFFT_OK : SFFT_ERR_HIP;
}

extern "C" void sfft_filter_window(const complex_t* filter_freq, int n, int B, complex_t* win) {
  const long long m = n / B;
  for (long long j = 0; j < m; ++j) win[j] = filter_freq[(j - m / 2 + n) & ((long long)n - 1)];
}

extern "C" sfft_status sfft_convert_c128_to_c64(const void* d_in, void* d_out, size_t count,
                                                void* stream) {
  return aux_c128_to_c64((const double2*)d_in, (float2*)d_out, (long long)count,
                         (hipStream_t)stream) == hipSuccess
             ? SFFT_OK
             : SFFT_ERR_HIP;
}

extern "C" sfft_status sfft_convert_c64_to_c128(const void* d_in, void* d_out, size_t count,
                                                void* stream) {
  return aux_c64_to_c128((const float2*)d_in, (double2*)d_out, (long long)count,
                         (hipStream_t)stream) == hipSuccess
             ? SFFT_OK
             : SFFT_ERR_HIP;
}

// ---------------------------------------------------------------------------------------------
// seeded synthetic signals: xoshiro256** seeded by SplitMix64
// ---------------------------------------------------------------------------------------------
namespace {
struct Xoshiro {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Xoshiro(uint64_t seed) {
    for (auto& v : s) v = splitmix(&seed);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  uint64_t below(uint64_t bound) {  // unbiased (Lemire)
    uint64_t x = next();
    __uint128_t m = (__uint128_t)x * bound;
    uint64_t l = (uint64_t)m;
    if (l < bound) {
      const uint64_t t = (0 - bound) % bound;
      while (l < t) {
        x = next();
        m = (__uint128_t)x * bound;
        l = (uint64_t)m;
      }
    }
    return (uint64_t)(m >> 64);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};
}  // namespace

// k distinct locations (Floyd), ascending; unit-magnitude values with uniform phase.
extern "C" sfft_status sfft_signal_spectrum(int n, int k, uint64_t seed, int* h_locs,
                                            complex_t* h_vals) {
  if (n < 2 || k < 1 || k > n || !h_locs || !h_vals) return SFFT_ERR_INVALID;
  Xoshiro rng(seed);
  std::unordered_set<long long> S;
  S.reserve((size_t)k * 2);
  for (long long j = (long long)n - k; j < n; ++j) {
    const long long t = (long long)rng.below((uint64_t)j + 1);
    if (!S.insert(t).second) S.insert(j);
  }
  std::vector<long long> v(S.begin(), S.end());
  std::sort(v.begin(), v.end());
  for (int i = 0; i < k; ++i) h_locs[i] = (int)v[i];
  Xoshiro rph(seed ^ 0x5DEECE66Dull);
  for (int i = 0; i < k; ++i) {
    const double phi = 2.0 * M_PI * rph.uniform();
    h_vals[i].x = cos(phi);
    h_vals[i].y = sin(phi);
  }
  return SFFT_OK;
}

// x = unnormalised inverse DFT of the sparse spectrum (locs, vals) (cuExperiment.cu:91-93),
// plus complex Gaussian noise of per-component std sqrt(k / (2 * 10^(snr_db/10)))
// (experiment.c:69) when snr_db > 0.  d_x: complex128[n] or complex64[n] (x_fp32).
extern "C" sfft_status sfft_signal_synthesize(int n, int k, const int* h_locs,
                                              const complex_t* h_vals, double snr_db,
                                              uint64_t noise_seed, int x_fp32, void* d_x,
                                              void* stream) {
  const int ln = ilog2_pow2(n);
  if (ln < 1 || k < 1 || !h_locs || !h_vals || !d_x) return SFFT_ERR_INVALID;
  for (int i = 0; i < k; ++i)
    if (h_locs[i] < 0 || h_locs[i] >= n) return SFFT_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  sfft_status status = SFFT_OK;
  void *X = nullptr, *y = nullptr, *scr = nullptr, *dl = nullptr, *dv = nullptr;
  HCK(hipMalloc(&X, (size_t)n * 16));
  HCK(hipMalloc(&scr, (size_t)n * 16));
  HCK(hipMalloc(&dl, (size_t)k * 4));
  HCK(hipMalloc(&dv, (size_t)k * 16));
  HCK(hipMemsetAsync(X, 0, (size_t)n * 16, st));
  HCK(hipMemcpyAsync(dl, h_locs, (size_t)k * 4, hipMemcpyHostToDevice, st));
  HCK(hipMemcpyAsync(dv, h_vals, (size_t)k * 16, hipMemcpyHostToDevice, st));
  HCK(aux_scatter_c128((double2*)X, (const int*)dl, (const double2*)dv, k, st));
  if (x_fp32) {
    HCK(hipMalloc(&y, (size_t)n * 16));
  } else {
    y = d_x;
  }
  HCK(fft_plain(X, y, scr, ln, 1, 1, 0, st));
  if (snr_db > 0) {
    const double sigma = sqrt((double)k / (2.0 * pow(10.0, snr_db / 10.0)));
    HCK(aux_add_noise((double2*)y, n, sigma, noise_seed, st));
  }
  if (x_fp32) HCK(aux_c128_to_c64((const double2*)y, (float2*)d_x, n, st));
  HCK(hipStreamSynchronize(st));
done:
  if (X) (void)hipFree(X);
  if (scr) (void)hipFree(scr);
  if (dl) (void)hipFree(dl);
  if (dv) (void)hipFree(dv);
  if (x_fp32 && y) (void)hipFree(y);
  return status;
}

// truth spectrum X = FFT_n(x) / n in FP64 (cuExperiment.cu:103-111), device complex128[n]
extern "C" sfft_status sfft_truth_spectrum(const void* d_x, int n, int x_fp32, complex_t* d_X,
                                           void* stream) {
  const int ln = ilog2_pow2(n);
  if (ln < 1 || !d_x || !d_X) return SFFT_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  sfft_status status = SFFT_OK;
  void *in = nullptr, *scr = nullptr;
  HCK(hipMalloc(&scr, (size_t)n * 16));
  HCK(hipMalloc(&in, (size_t)n * 16));
  if (x_fp32)
    HCK(aux_c64_to_c128((const float2*)d_x, (double2*)in, n, st));
  else
    HCK(hipMemcpyAsync(in, d_x, (size_t)n * 16, hipMemcpyDeviceToDevice, st));
  HCK(fft_plain(in, d_X, scr, ln, 1, 0, 0, st));
  HCK(aux_scale_c128((double2*)d_X, n, (double)n, st));
  HCK(hipStreamSynchronize(st));
done:
  if (in) (void)hipFree(in);
  if (scr) (void)hipFree(scr);
  return status;
}

static double device_energy(const double2* a, long long n, hipStream_t st, hipError_t* err) {
  const int nb = 2048;
  double* part = nullptr;
  *err = hipMalloc(&part, nb * sizeof(double));
  if (*err != hipSuccess) return 0.0;
  std::vector<double> h(nb);
  *err = aux_energy_partials(a, n, part, nb, st);
  if (*err == hipSuccess) *err = hipMemcpyAsync(h.data(), part, nb * 8, hipMemcpyDeviceToHost, st);
  if (*err == hipSuccess) *err = hipStreamSynchronize(st);
  (void)hipFree(part);
  double s = 0.0;
  for (double v : h) s += v;
  return s;
}

// run_experiment's error metric (cuExperiment.cu:181-221) against the truth spectrum d_X:
// nodes sorted by |value| descending (ties by key), the first k kept; MISSED counts as the
// reference; err_total = sum over all n of |yhat - X|^2 ("MSE*k"), computed as the energy of X
// outside the kept support plus the error on it (d_X is restored before returning).
// support_err = error on the kept support + energy of X at missed true locations.
extern "C" sfft_status sfft_metric(const Node* nodes, int I_F, int k, const int* locs,
                                   complex_t* d_X, int n, int* missed_est, int* missed_res,
                                   double* err_total, double* err_support, double* max_abs,
                                   void* stream) {
  if ((!nodes && I_F > 0) || I_F < 0 || k < 1 || !locs || !d_X) return SFFT_ERR_INVALID;
  for (int i = 0; i < k; ++i)
    if (locs[i] < 0 || locs[i] >= n) return SFFT_ERR_INVALID;
  for (int i = 0; i < I_F; ++i)
    if (nodes[i].key >= (unsigned)n) return SFFT_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  sfft_status status = SFFT_OK;
  std::vector<Node> v(nodes, nodes + I_F);
  std::sort(v.begin(), v.end(), [](const Node& a, const Node& b) {
    const double aa = sqrt(a.value.x * a.value.x + a.value.y * a.value.y);
    const double bb = sqrt(b.value.x * b.value.x + b.value.y * b.value.y);
    if (aa != bb) return aa > bb;
    return a.key < b.key;
  });
  const int top = std::min(k, I_F);
  std::vector<int> supp(top);
  for (int i = 0; i < top; ++i) supp[i] = (int)v[i].key;
  std::unordered_set<long long> all, kept;
  for (int i = 0; i < I_F; ++i) all.insert(nodes[i].key);
  for (int i = 0; i < top; ++i)
    if (v[i].value.x != 0.0 || v[i].value.y != 0.0) kept.insert(v[i].key);
  int found = 0, large = 0;
  for (int i = 0; i < k; ++i) {
    found += all.count(locs[i]) ? 1 : 0;
    large += kept.count(locs[i]) ? 1 : 0;
  }
  *missed_est = k - found;
  *missed_res = k - large;
  // truth at the support and at the true locations
  std::vector<int> idx(supp);
  idx.insert(idx.end(), locs, locs + k);
  const int ni = (int)idx.size();
  std::vector<double2> tv(ni);
  void *di = nullptr, *dt = nullptr, *dz = nullptr;
  double e_out = 0.0, e_sup = 0.0, e_miss = 0.0, mabs = 0.0;
  hipError_t herr = hipSuccess;
  HCK(hipMalloc(&di, (size_t)ni * 4));
  HCK(hipMalloc(&dt, (size_t)ni * 16));
  HCK(hipMemcpy(di, idx.data(), (size_t)ni * 4, hipMemcpyHostToDevice));
  HCK(aux_gather_c128((const double2*)d_X, (const int*)di, (double2*)dt, ni, st));
  HCK(hipMemcpyAsync(tv.data(), dt, (size_t)ni * 16, hipMemcpyDeviceToHost, st));
  HCK(hipStreamSynchronize(st));
  // energy outside the support: zero the support entries, sum, restore
  if (top > 0) {
    std::vector<double2> zeros(top, make_double2(0.0, 0.0));
    HCK(hipMalloc(&dz, (size_t)top * 16));
    HCK(hipMemcpy(dz, zeros.data(), (size_t)top * 16, hipMemcpyHostToDevice));
    HCK(aux_scatter_c128((double2*)d_X, (const int*)di, (const double2*)dz, top, st));
  }
  e_out = device_energy((const double2*)d_X, n, st, &herr);
  HCK(herr);
  if (top > 0) HCK(aux_scatter_c128((double2*)d_X, (const int*)di, (const double2*)dt, top, st));
  HCK(hipStreamSynchronize(st));
  {
    std::unordered_set<long long> sset(supp.begin(), supp.end());
    for (int i = 0; i < top; ++i) {
      const double dx = v[i].value.x - tv[i].x, dy = v[i].value.y - tv[i].y;
      e_sup += dx * dx + dy * dy;
      mabs = std::max(mabs, sqrt(dx * dx + dy * dy));
    }
    for (int i = 0; i < k; ++i)
      if (!sset.count(locs[i])) {
        const double2 t = tv[top + i];
        e_miss += t.x * t.x + t.y * t.y;
        mabs = std::max(mabs, sqrt(t.x * t.x + t.y * t.y));
      }
  }
  if (err_total) *err_total = e_out + e_sup;
  if (err_support) *err_support = e_sup + e_miss;
  if (max_abs) *max_abs = mabs;
done:
  if (di) (void)hipFree(di);
  if (dt) (void)hipFree(dt);
  if (dz) (void)hipFree(dz);
  return status;
}
