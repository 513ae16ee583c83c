/*
 * sfft_oracle.c -- CPU oracle for the sparse-FFT hot path.  TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
 * product.  A sequential restatement of the reference semantics (pcdslab/gpu-sfft), written
 * from the reference's behaviour, with citations to the reference file:line it follows.
 * Build: gcc -O2 -ffp-contract=off (oracle/Makefile).
 */
#define _GNU_SOURCE
#include "sfft_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------------------------
 * number theory (utils.c:81-105, innerLoop.cu:647-649)
 * ---------------------------------------------------------------------------------------- */
int orc_gcd(int a, int b) {           /* Euclid, utils.c:81-84 (a, b > 0) */
  while (b != 0) {
    int r = a % b;
    a = b;
    b = r;
  }
  return a;
}

int orc_mod_inverse(int a, int n) {   /* extended Euclid, utils.c:92-105 */
  long long r0 = n, r1 = a, t0 = 0, t1 = 1;
  while (r1 != 0) {
    long long qq = r0 / r1, tmp;
    tmp = r0 - qq * r1; r0 = r1; r1 = tmp;
    tmp = t0 - qq * t1; t0 = t1; t1 = tmp;
  }
  long long v = t0 % n;
  if (v < 0) v += n;
  return (int)v;
}

int orc_timesmod(int x, int a, int n) { return (int)(((long long)x * a) % n); }

int orc_floor_to_pow2(double x) {     /* largest power of two <= x, utils.c:180-185 */
  unsigned int ans = 1;
  while (ans <= x) ans <<= 1;
  return (int)(ans / 2);
}

/* ------------------------------------------------------------------------------------------
 * parameters (experiment.c:54-84; w and t0 from cuExperiment.cu:279-284)
 * ---------------------------------------------------------------------------------------- */
int orc_filter_window(double lobefrac, double tolerance, int n, int *w_out, double *t0_out) {
  int w = (int)((1 / M_PI) * (1 / lobefrac) * acosh(1. / tolerance));
  if (!(w % 2)) w--;
  *w_out = w;
  *t0_out = cosh(acosh(1 / tolerance) / (w - 1));
  return w <= n;
}

void orc_derive_params(int q, int k, double Bcst, int loops_loc, int loops_thresh, int loops_est,
                       int RLFC_cst, int RLFC_loops, double tolerance, orc_params *o) {
  memset(o, 0, sizeof(*o));
  o->q = q;
  o->k = k;
  o->Bcst = Bcst;
  o->loops_loc = loops_loc;
  o->loops_thresh = loops_thresh;
  o->loops_est = loops_est;
  o->loops = loops_loc + loops_est;
  o->RLFC_cst = (double)RLFC_cst;        /* parsed with atoi (experiment.c:62) */
  o->RLFC_loops = RLFC_loops;
  o->tolerance = tolerance;
  int n = (int)pow(2, q);
  n = orc_floor_to_pow2(n);
  o->n = n;
  o->BB = (unsigned)(Bcst * sqrt((double)n * k / (log2(n))));
  o->lobefrac = 0.5 / (o->BB);
  o->b_f = (int)(1.2 * 1.1 * ((double)n / o->BB));
  o->B = orc_floor_to_pow2(o->BB);
  o->B_thresh = 2 * k;
  o->W_RLFC = orc_floor_to_pow2(o->RLFC_cst * n / o->B);
  int ok = (o->b_f <= n) && (o->B_thresh < o->W_RLFC) && (o->B_thresh < o->B) &&
           (loops_thresh <= loops_loc);
  ok = orc_filter_window(o->lobefrac, tolerance, n, &o->w, &o->t0) && ok;
  o->assert_ok = ok;
}

/* ------------------------------------------------------------------------------------------
 * canonical twiddles and FFT plan (definition shared with the HIP kernels)
 * tw(e, T) = exp(-+ 2 pi i e / T), octant-reduced so that cos/sin only see [0, pi/4].
 * ---------------------------------------------------------------------------------------- */
void orc_twiddle64(uint64_t e, uint64_t T, int inverse, orc_c64 *w) {
  double c, s;
  e %= T;
  if (T < 8) {
    /* exact values for T = 1, 2, 4 */
    const uint64_t e4 = e * (4 / T);  /* quarter turns */
    static const double cq[4] = {1.0, 0.0, -1.0, 0.0};
    static const double sq[4] = {0.0, 1.0, 0.0, -1.0};
    c = cq[e4 & 3];
    s = sq[e4 & 3];
  } else {
    const uint64_t quarter = T / 4;
    const uint64_t qd = e / quarter;
    const uint64_t r = e % quarter;
    const uint64_t rr = (2 * r > quarter) ? (quarter - r) : r;
    const double ang = (2.0 * M_PI * (double)rr) / (double)T;
    double c0, s0;
    sincos(ang, &s0, &c0);   /* one glibc call on both sides (sin/cos may differ from it) */
    if (2 * r > quarter) { double t = c0; c0 = s0; s0 = t; }
    switch (qd) {
      case 0: c = c0; s = s0; break;
      case 1: c = -s0; s = c0; break;
      case 2: c = -c0; s = -s0; break;
      default: c = s0; s = -c0; break;
    }
  }
  w->x = c;
  w->y = inverse ? s : -s;
}

int orc_fft_plan(int log2N, int *pass_logs) {
  if (log2N <= 0) return 0;
  const int p = (log2N + 10) / 11;
  const int base = log2N / p, extra = log2N % p;
  for (int i = 0; i < p; ++i) pass_logs[i] = base + (i < extra ? 1 : 0);
  return p;
}

#define CF_REAL double
#define CF_CPLX orc_c64
#define CF_FMA fma
#define CF_SFX 64
#include "canon_fft_impl.h"
#undef CF_REAL
#undef CF_CPLX
#undef CF_FMA
#undef CF_SFX

#define CF_REAL float
#define CF_CPLX orc_c32
#define CF_FMA fmaf
#define CF_SFX 32
#include "canon_fft_impl.h"
#undef CF_REAL
#undef CF_CPLX
#undef CF_FMA
#undef CF_SFX

static int ilog2_exact(uint64_t N) {
  int l = 0;
  while (((uint64_t)1 << l) < N) ++l;
  return (((uint64_t)1 << l) == N) ? l : -1;
}

/* Bluestein: X[k] = c[k] sum_j (x[j] c[j]) conj(c[k-j]), c[m] = exp(-+ i pi m^2 / N) */
void orc_dft_any64(orc_c64 *data, int N, int inverse) {
  if (N <= 1) return;
  int l2 = ilog2_exact((uint64_t)N);
  if (l2 >= 0) {
    orc_c64 *scr = (orc_c64 *)malloc(sizeof(orc_c64) * (size_t)N);
    orc_fft64(data, l2, inverse, scr);
    free(scr);
    return;
  }
  int lm = 0;
  while (((size_t)1 << lm) < (size_t)(2 * N - 1)) ++lm;
  const size_t M = (size_t)1 << lm;
  orc_c64 *c = (orc_c64 *)malloc(sizeof(orc_c64) * (size_t)N);
  orc_c64 *A = (orc_c64 *)calloc(M, sizeof(orc_c64));
  orc_c64 *Bv = (orc_c64 *)calloc(M, sizeof(orc_c64));
  orc_c64 *scr = (orc_c64 *)malloc(sizeof(orc_c64) * M);
  const double sgn = inverse ? 1.0 : -1.0;
  for (int m = 0; m < N; ++m) {
    const unsigned long long r = ((unsigned long long)m * (unsigned long long)m) %
                                 (2ULL * (unsigned long long)N);
    const double ang = M_PI * (double)r / (double)N;
    double sa, ca;
    sincos(ang, &sa, &ca);
    c[m].x = ca;
    c[m].y = sgn * sa;
  }
  for (int j = 0; j < N; ++j) {
    A[j].x = data[j].x * c[j].x - data[j].y * c[j].y;
    A[j].y = data[j].x * c[j].y + data[j].y * c[j].x;
  }
  Bv[0].x = c[0].x; Bv[0].y = -c[0].y;
  for (int m = 1; m < N; ++m) {
    Bv[m].x = c[m].x; Bv[m].y = -c[m].y;
    Bv[M - m] = Bv[m];
  }
  orc_fft64(A, lm, 0, scr);
  orc_fft64(Bv, lm, 0, scr);
  for (size_t i = 0; i < M; ++i) {
    const double xr = A[i].x * Bv[i].x - A[i].y * Bv[i].y;
    const double xi = A[i].x * Bv[i].y + A[i].y * Bv[i].x;
    A[i].x = xr;
    A[i].y = xi;
  }
  orc_fft64(A, lm, 1, scr);
  for (int k = 0; k < N; ++k) {
    const double yr = A[k].x / (double)M, yi = A[k].y / (double)M;
    data[k].x = yr * c[k].x - yi * c[k].y;
    data[k].y = yr * c[k].y + yi * c[k].x;
  }
  free(c); free(A); free(Bv); free(scr);
}

/* ------------------------------------------------------------------------------------------
 * filter design (cuExperiment.cu:248-447, 507-522, 538-601)
 * ---------------------------------------------------------------------------------------- */
static double orc_cheb(double m, double x) {   /* cuCheb, cuExperiment.cu:593-601 */
  if (fabs(x) <= 1) return cos(m * acos(x));
  return cosh(m * acosh(x));                  /* NaN for x < -1: zeroed by the caller */
}

void orc_make_multiple_t(double lobefrac, double tolerance, int n, int b_f, int w,
                         orc_c64 *filter_time, orc_c64 *filter_freq) {
  orc_make_multiple_t_taps(lobefrac, tolerance, n, b_f, w, filter_time, filter_freq, NULL, NULL,
                           NULL);
}

void orc_make_multiple_t_taps(double lobefrac, double tolerance, int n, int b_f, int w,
                              orc_c64 *filter_time, orc_c64 *filter_freq, orc_c64 *tap_v,
                              orc_c64 *tap_V, orc_c64 *tap_G) {
  (void)lobefrac;
  const double t0 = cosh(acosh(1 / tolerance) / (w - 1));
  /* Dolph-Chebyshev samples (makeDolphchebyshevKernel1, cuExperiment.cu:538-552) */
  orc_c64 *v = (orc_c64 *)malloc(sizeof(orc_c64) * (size_t)w);
  for (int i = 0; i < w; ++i) {
    double val = orc_cheb(w - 1, t0 * cos(M_PI * i / (w))) * tolerance;
    v[i].x = isnan(val) ? 0.0 : val;
    v[i].y = 0.0;
  }
  if (tap_v) memcpy(tap_v, v, sizeof(orc_c64) * (size_t)w);
  orc_dft_any64(v, w, 0);                     /* FFT_w forward, cuExperiment.cu:305-308 */
  if (tap_V) memcpy(tap_V, v, sizeof(orc_c64) * (size_t)w);
  /* cuShift(v, w, w/2): new[j] = old[(j - w/2) mod w] (cuExperiment.cu:507-522) */
  const int h = w / 2;
  for (int j = 0; j < w; ++j) {
    filter_time[j].x = v[(j - h + w) % w].x;
    filter_time[j].y = 0.0;                   /* makeDolphchebyshevKernel2 */
  }
  free(v);
  /* fold around 0: g[0..h) = f[h..2h), g[n-h..n) = f[0..h)  (f[w-1] dropped) */
  orc_c64 *g = (orc_c64 *)calloc((size_t)n, sizeof(orc_c64));
  memcpy(g, filter_time + h, sizeof(orc_c64) * (size_t)h);
  memcpy(g + n - h, filter_time, sizeof(orc_c64) * (size_t)h);
  int ln = ilog2_exact((uint64_t)n);
  orc_c64 *scr = (orc_c64 *)malloc(sizeof(orc_c64) * (size_t)n);
  orc_fft64(g, ln, 0, scr);                   /* FFT_n forward, cuExperiment.cu:362-365 */
  if (tap_G) memcpy(tap_G, g, sizeof(orc_c64) * (size_t)n);
  /* boxcar running sum (cuExperiment.cu:373-396), a sequential recurrence */
  orc_c64 s = {0.0, 0.0};
  for (int i = 0; i < b_f; ++i) {
    s.x += g[i].x;
    s.y += g[i].y;
  }
  double mx = 0.0;
  const long long offset = b_f / 2;
  const long long mask = (long long)n - 1;
  for (long long i = 0; i < n; ++i) {
    filter_freq[(i + n + offset) & mask] = s;
    const double ab = sqrt(s.x * s.x + s.y * s.y);
    mx = mx > ab ? mx : ab;
    s.x += g[(i + b_f) & mask].x - g[i].x;
    s.y += g[(i + b_f) & mask].y - g[i].y;
  }
  for (long long i = 0; i < n; ++i) {         /* makeMultipleKernel1 */
    filter_freq[i].x /= mx;
    filter_freq[i].y /= mx;
  }
  /* phase ramp by a sequential complex recurrence (cuExperiment.cu:407-421) */
  orc_c64 off = {1.0, 0.0}, step;
  sincos(-2 * M_PI * (w / 2) / n, &step.y, &step.x);
  for (long long i = 0; i < n; ++i) {
    orc_c64 a = filter_freq[i], r;
    r.x = a.x * off.x - a.y * off.y;
    r.y = a.x * off.y + a.y * off.x;
    filter_freq[i] = r;
    orc_c64 o2;
    o2.x = off.x * step.x - off.y * step.y;
    o2.y = off.x * step.y + off.y * step.x;
    off = o2;
  }
  memcpy(g, filter_freq, sizeof(orc_c64) * (size_t)n);
  orc_fft64(g, ln, 1, scr);                   /* inverse FFT_n, cuExperiment.cu:424-426 */
  for (int i = 0; i < w; ++i) {               /* makeMultipleKernel2 */
    filter_time[i].x = g[i].x / n;
    filter_time[i].y = g[i].y / n;
  }
  free(g);
  free(scr);
}

/* ------------------------------------------------------------------------------------------
 * permutation + filtering (innerLoop.cu:194-246, 725-773)
 * ---------------------------------------------------------------------------------------- */
/* the term as the reference writes it (innerLoop.cu:740-741, 770-771): both products
 * rounded, combined, then added to the bucket (no fused multiply-add) */
static inline void cmac64(orc_c64 *acc, orc_c64 x, orc_c64 f) {
  const double tx = x.x * f.x - x.y * f.y;
  const double ty = x.x * f.y + x.y * f.x;
  acc->x = acc->x + tx;
  acc->y = acc->y + ty;
}

void orc_perm_filter64(const orc_c64 *x, int n, const orc_c64 *f, int w, int B, int ai,
                       int fold_mode, orc_c64 *bins_t) {
  const uint32_t mask = (uint32_t)n - 1u, uai = (uint32_t)ai;
  const int tiled = (fold_mode == ORC_FOLD_REFERENCE) && (n < (1 << 27));
  const int tiles = w / B, residue = w % B;
  for (int r = 0; r < B; ++r) {
    orc_c64 acc = {0.0, 0.0};
    if (tiled) {
      for (int t = 0; t < tiles; ++t) {               /* permFilterTillKernel :735-742 */
        const uint32_t id = (uint32_t)(t * B + r);
        cmac64(&acc, x[(id * uai) & mask], f[id]);
      }
      if (r < residue) {                              /* faulty tail thread :744-755 (n < 2^27) */
        const uint32_t id = (uint32_t)(r * tiles + B);
        cmac64(&acc, x[(id * uai) & mask], f[id]);
      }
    } else {
      for (int i = r; i < w; i += B) {                /* permFilterKernel :766-772, without its racy += */
        const uint32_t id = (uint32_t)i;
        cmac64(&acc, x[(id * uai) & mask], f[id]);
      }
    }
    bins_t[r] = acc;
  }
}

/* FP32: the same term in FP32 (products rounded, then combined), Kahan-compensated
 * accumulation */
static inline void kahan_add(float *sum, float *comp, float term) {
  const float y = term - *comp;
  const float t = *sum + y;
  *comp = (t - *sum) - y;
  *sum = t;
}

static inline void cmac32k(float *sx, float *cx, float *sy, float *cy, orc_c32 x, orc_c32 f) {
  const float tx = x.x * f.x - x.y * f.y;
  const float ty = x.x * f.y + x.y * f.x;
  kahan_add(sx, cx, tx);
  kahan_add(sy, cy, ty);
}

void orc_perm_filter32(const orc_c32 *x, int n, const orc_c32 *f, int w, int B, int ai,
                       int fold_mode, orc_c32 *bins_t) {
  const uint32_t mask = (uint32_t)n - 1u, uai = (uint32_t)ai;
  const int tiled = (fold_mode == ORC_FOLD_REFERENCE) && (n < (1 << 27));
  const int tiles = w / B, residue = w % B;
  for (int r = 0; r < B; ++r) {
    float sx = 0.f, cx = 0.f, sy = 0.f, cy = 0.f;
    if (tiled) {
      for (int t = 0; t < tiles; ++t) {
        const uint32_t id = (uint32_t)(t * B + r);
        cmac32k(&sx, &cx, &sy, &cy, x[(id * uai) & mask], f[id]);
      }
      if (r < residue) {
        const uint32_t id = (uint32_t)(r * tiles + B);
        cmac32k(&sx, &cx, &sy, &cy, x[(id * uai) & mask], f[id]);
      }
    } else {
      for (int i = r; i < w; i += B) {
        const uint32_t id = (uint32_t)i;
        cmac32k(&sx, &cx, &sy, &cy, x[(id * uai) & mask], f[id]);
      }
    }
    bins_t[r].x = sx;
    bins_t[r].y = sy;
  }
}

/* ------------------------------------------------------------------------------------------
 * samples and cutoff (innerLoop.cu:484-642, 708-718, 875-890)
 * ---------------------------------------------------------------------------------------- */
/* |X|^2 = re*re + im*im, the reference's pow(re,2) + pow(im,2) (innerLoop.cu:715) with
 * each square rounded and then added */
void orc_samples64(const orc_c64 *X, int N, double *s) {
  for (int i = 0; i < N; ++i) s[i] = X[i].x * X[i].x + X[i].y * X[i].y;
}

void orc_samples32(const orc_c32 *X, int N, float *s) {
  for (int i = 0; i < N; ++i) s[i] = X[i].x * X[i].x + X[i].y * X[i].y;
}

/* K-th largest by quickselect on a copy (the reference sorts; the order statistic is the
 * same value), then one ascending pass: s > T, plus the first `need` ties (lowest indices first). */
#define DEFINE_CUTOFF(SFX, T)                                                              \
  static T kth_largest_##SFX(T *a, int N, int K) {                                         \
    int lo = 0, hi = N - 1, target = K - 1; /* index in descending order */               \
    while (lo < hi) {                                                                      \
      T pivot = a[lo + (hi - lo) / 2];                                                     \
      int i = lo, j = hi;                                                                  \
      while (i <= j) {                                                                     \
        while (a[i] > pivot) ++i;                                                          \
        while (a[j] < pivot) --j;                                                          \
        if (i <= j) { T t = a[i]; a[i] = a[j]; a[j] = t; ++i; --j; }                       \
      }                                                                                    \
      if (target <= j) hi = j;                                                             \
      else if (target >= i) lo = i;                                                        \
      else break;                                                                          \
    }                                                                                      \
    return a[target];                                                                      \
  }                                                                                        \
  void orc_cutoff##SFX(const T *s, int N, int K, int *out) {                               \
    T *tmp = (T *)malloc(sizeof(T) * (size_t)N);                                          \
    memcpy(tmp, s, sizeof(T) * (size_t)N);                                                \
    const T thr = kth_largest_##SFX(tmp, N, K);                                            \
    free(tmp);                                                                             \
    int gt = 0;                                                                            \
    for (int i = 0; i < N; ++i) gt += (s[i] > thr);                                        \
    int need = K - gt, c = 0;                                                              \
    for (int i = 0; i < N; ++i) {                                                          \
      if (s[i] > thr) out[c++] = i;                                                        \
      else if (s[i] == thr && need > 0) { out[c++] = i; --need; }                          \
    }                                                                                      \
  }
DEFINE_CUTOFF(64, double)
DEFINE_CUTOFF(32, float)

/* ------------------------------------------------------------------------------------------
 * RLFC comb restriction (innerLoop.cu:112-177, 693-703) and J2 merge (outerLoop.cu:104-121)
 * ---------------------------------------------------------------------------------------- */
void orc_rlfc64(const orc_c64 *x, int n, int W, int offset, int B_thresh, int *J2_out) {
  const int sigma = n / W;
  orc_c64 *y = (orc_c64 *)malloc(sizeof(orc_c64) * (size_t)W);
  orc_c64 *scr = (orc_c64 *)malloc(sizeof(orc_c64) * (size_t)W);
  double *s = (double *)malloc(sizeof(double) * (size_t)W);
  for (int i = 0; i < W; ++i) y[i] = x[(size_t)offset + (size_t)i * sigma];
  orc_fft64(y, ilog2_exact((uint64_t)W), 0, scr);
  orc_samples64(y, W, s);
  orc_cutoff64(s, W, B_thresh, J2_out);
  free(y); free(scr); free(s);
}

static int cmp_int(const void *a, const void *b) {
  const int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

int orc_merge_J2(int *J2, int count) {
  if (count <= 0) return 0;
  qsort(J2, (size_t)count, sizeof(int), cmp_int);
  int last = 0;
  for (int i = 1; i < count; ++i)
    if (J2[i] != J2[last]) J2[++last] = J2[i];
  return last + 1;
}

/* ------------------------------------------------------------------------------------------
 * reverse hash + voting (innerLoop.cu:300-384, 783-817, 901-920)
 * ---------------------------------------------------------------------------------------- */
typedef struct { int first, second; } orc_pair;

static int cmp_pair_first(const void *a, const void *b) {   /* comp_struct3, utils.c:246-253 */
  const orc_pair *x = (const orc_pair *)a, *y = (const orc_pair *)b;
  return (x->first > y->first) - (x->first < y->first);
}

/* findUpperBound (innerLoop.cu:901-920): lower bound, but the last index when
 * val >= last.first (findUpperBound's last-index rule) */
static int find_upper_bound(const orc_pair *P, int n, int val) {
  int low = 0, high = n - 1;
  if (val >= P[high].first) return high;
  int mid = (low + high) / 2;
  while (high > low) {
    if (P[mid].first >= val) high = mid;
    else low = mid + 1;
    mid = (low + high) / 2;
  }
  return mid;
}

int orc_reverse_hash(const int *J, int B_thresh, int n, int B, int a, int ai, int loop_thresh,
                     int *vote, int *hits, int hit_count, int hit_capacity,
                     const int *J2, int num_RLFC, int W) {
  orc_pair *P = (orc_pair *)malloc(sizeof(orc_pair) * (size_t)num_RLFC);
  for (int m = 0; m < num_RLFC; ++m) {                        /* innerLoop.cu:319-324 */
    const int prev = orc_timesmod(J2[m], ai, W);
    P[m].first = prev;
    P[m].second = orc_timesmod(prev, a, n);
  }
  qsort(P, (size_t)num_RLFC, sizeof(orc_pair), cmp_pair_first);
  const long long nmask = (long long)n - 1;
  const long long max_steps = (long long)num_RLFC * ((long long)(n / B) / W + 2) + 2;
  int overflow = 0;
  for (int i = 0; i < B_thresh && !overflow; ++i) {             /* reverseHashKernel */
    const long long low = ((long long)(ceil((J[i] - 0.5) * n / B)) + n) & nmask;
    const long long high = ((long long)(ceil((J[i] + 0.5) * n / B)) + n) & nmask;
    const int key = (int)(low & (W - 1));
    int j = find_upper_bound(P, num_RLFC, key);
    long long location = low - (low & (W - 1));
    long long locinv = ((long long)location * a) % n;
    for (long long steps = 0; steps < max_steps; ++steps, ++j) {
      if (j == num_RLFC) {
        j -= num_RLFC;
        location = (location + W) & nmask;
        locinv = ((long long)location * a) % n;
      }
      const long long approved = location + P[j].first;
      if ((low < high && (approved >= high || approved < low)) ||
          (low > high && (approved >= high && approved < low)))
        break;
      const long long loc = (locinv + P[j].second) & nmask;
      vote[loc] += 1;
      if (vote[loc] == loop_thresh) {
        if (hit_count >= hit_capacity) { overflow = 1; break; }
        hits[hit_count++] = (int)loc;
      }
    }
  }
  free(P);
  return overflow ? -1 : hit_count;
}

/* ------------------------------------------------------------------------------------------
 * estimation (innerLoop.cu:396-466, 825-870, 924-937)
 * ---------------------------------------------------------------------------------------- */
static void insertion_sort(double *v, int n) {                /* insertionSortDev */
  for (int i = 1; i < n; ++i) {
    const double key = v[i];
    int j = i - 1;
    while (j >= 0 && v[j] > key) { v[j + 1] = v[j]; --j; }
    v[j + 1] = key;
  }
}

/* bins are complex128 (bins_f) or complex64 (bins32, widened exactly to double) */
static void estimate_values_any(const int *hits, int I_F, const orc_c64 *bins_f,
                                const orc_c32 *bins32, int loops, int n, const int *ai, int B,
                                const orc_c64 *filter_win, orc_node *out);

void orc_estimate_values(const int *hits, int I_F, const orc_c64 *bins_f, int loops, int n,
                         const int *ai, int B, const orc_c64 *filter_win, orc_node *out) {
  estimate_values_any(hits, I_F, bins_f, NULL, loops, n, ai, B, filter_win, out);
}

void orc_estimate_values32(const int *hits, int I_F, const orc_c32 *bins_f, int loops, int n,
                           const int *ai, int B, const orc_c64 *filter_win, orc_node *out) {
  estimate_values_any(hits, I_F, NULL, bins_f, loops, n, ai, B, filter_win, out);
}

static void estimate_values_any(const int *hits, int I_F, const orc_c64 *bins_f,
                                const orc_c32 *bins32, int loops, int n, const int *ai, int B,
                                const orc_c64 *filter_win, orc_node *out) {
  const int m = n / B;
  const long long nmask = (long long)n - 1;
  double *vr = (double *)malloc(sizeof(double) * (size_t)loops);
  double *vi = (double *)malloc(sizeof(double) * (size_t)loops);
  for (int h = 0; h < I_F; ++h) {
    const int I = hits[h];
    for (int l = 0; l < loops; ++l) {
      const long long p = ((long long)ai[l] * I) % n;
      long long hashed = p / m;
      long long dist = p & (m - 1);
      if (dist > m / 2) {
        hashed = (hashed + 1) & (B - 1);
        dist -= m;
      }
      dist = ((long long)n - dist) & nmask;
      const long long widx = (dist + m / 2) & nmask;
      const orc_c64 F = filter_win[widx];
      orc_c64 b;
      if (bins_f) {
        b = bins_f[(size_t)l * B + hashed];
      } else {
        b.x = (double)bins32[(size_t)l * B + hashed].x;
        b.y = (double)bins32[(size_t)l * B + hashed].y;
      }
      /* innerLoop.cu:848-852: |F|^2, then bins * conj(F), products rounded and combined */
      const double abs2 = F.x * F.x + F.y * F.y;
      const double mr = b.x * F.x + b.y * F.y;
      const double mi = b.y * F.x - b.x * F.y;
      vr[l] = mr / abs2;
      vi[l] = mi / abs2;
    }
    insertion_sort(vr, loops);
    insertion_sort(vi, loops);
    const int med = (loops - 1) / 2;
    out[h].key = (unsigned)I;
    out[h].value.x = vr[med];
    out[h].value.y = vi[med];
  }
  free(vr);
  free(vi);
}

/* ------------------------------------------------------------------------------------------
 * outer loop (outerLoop.cu:65-190)
 * ---------------------------------------------------------------------------------------- */
static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int orc_outer_loop64(const orc_plan *p, const orc_c64 *x, const orc_c64 *filter_time,
                     const orc_c64 *filter_win, const int *a, const int *ai, const int *offsets,
                     int *hits_out, orc_node *nodes_out, orc_c64 *bins_f_out, int *J_out,
                     int *J2_out, int *num_RLFC_out, double *st) {
  const int n = p->n, B = p->B, K = p->B_thresh, W = p->W_RLFC;
  double t0 = now_s(), t;
  double T_vote = 0, T_rlfc = 0, T_pf = 0, T_rh = 0, T_est = 0, T_fc = 0;

  t = now_s();
  int *vote = (int *)calloc((size_t)n, sizeof(int));
  T_vote = now_s() - t;

  t = now_s();
  int *J2 = (int *)malloc(sizeof(int) * (size_t)p->RLFC_loops * K);
  for (int r = 0; r < p->RLFC_loops; ++r) orc_rlfc64(x, n, W, offsets[r], K, J2 + (size_t)r * K);
  int num_RLFC = K;
  if (p->RLFC_loops > 1) num_RLFC = orc_merge_J2(J2, p->RLFC_loops * K);
  T_rlfc = now_s() - t;

  orc_c64 *bins_f = (orc_c64 *)malloc(sizeof(orc_c64) * (size_t)p->loops * B);
  orc_c64 *scr = (orc_c64 *)malloc(sizeof(orc_c64) * (size_t)B);
  double *s = (double *)malloc(sizeof(double) * (size_t)B);
  int *J = (int *)malloc(sizeof(int) * (size_t)K);
  int *hits = hits_out;
  int I_F = 0;
  const int lB = ilog2_exact((uint64_t)B);
  for (int l = 0; l < p->loops && I_F >= 0; ++l) {
    orc_c64 *row = bins_f + (size_t)l * B;
    t = now_s();
    orc_perm_filter64(x, n, filter_time, p->w, B, ai[l], p->fold_mode, row);
    T_pf += now_s() - t;
    t = now_s();
    orc_fft64(row, lB, 0, scr);
    if (l < p->loops_loc) {   /* estimation loops skip the cutoff, whose result is unused */
      orc_samples64(row, B, s);
      orc_cutoff64(s, B, K, J);
      if (J_out) memcpy(J_out + (size_t)l * K, J, sizeof(int) * (size_t)K);
    }
    T_fc += now_s() - t;
    if (l < p->loops_loc) {
      t = now_s();
      I_F = orc_reverse_hash(J, K, n, B, a[l], ai[l], p->loops_thresh, vote, hits, I_F,
                             p->hit_capacity, J2, num_RLFC, W);
      T_rh += now_s() - t;
    }
  }
  if (I_F >= 0) {
    t = now_s();
    orc_estimate_values(hits, I_F, bins_f, p->loops, n, ai, B, filter_win, nodes_out);
    T_est = now_s() - t;
  }
  if (bins_f_out) memcpy(bins_f_out, bins_f, sizeof(orc_c64) * (size_t)p->loops * B);
  if (J2_out) memcpy(J2_out, J2, sizeof(int) * (size_t)num_RLFC);
  if (num_RLFC_out) *num_RLFC_out = num_RLFC;
  free(vote); free(J2); free(bins_f); free(scr); free(s); free(J);
  if (st) {
    st[0] = T_vote; st[1] = T_rlfc; st[2] = T_pf; st[3] = T_rh; st[4] = T_est; st[5] = T_fc;
    st[6] = now_s() - t0;
  }
  return I_F;
}

/* FP32 path: complex64 signal and taps, FP32 comb and bucket FFTs and |X|^2 keys, Kahan
 * folding; the filter division and medians in FP64 on the widened bins. */
void orc_rlfc32(const orc_c32 *x, int n, int W, int offset, int B_thresh, int *J2_out) {
  const int sigma = n / W;
  orc_c32 *y = (orc_c32 *)malloc(sizeof(orc_c32) * (size_t)W);
  orc_c32 *scr = (orc_c32 *)malloc(sizeof(orc_c32) * (size_t)W);
  float *s = (float *)malloc(sizeof(float) * (size_t)W);
  for (int i = 0; i < W; ++i) y[i] = x[(size_t)offset + (size_t)i * sigma];
  orc_fft32(y, ilog2_exact((uint64_t)W), 0, scr);
  orc_samples32(y, W, s);
  orc_cutoff32(s, W, B_thresh, J2_out);
  free(y); free(scr); free(s);
}

int orc_outer_loop32(const orc_plan *p, const orc_c32 *x, const orc_c32 *filter_time,
                     const orc_c64 *filter_win, const int *a, const int *ai, const int *offsets,
                     int *hits_out, orc_node *nodes_out, orc_c32 *bins_f_out, int *J_out,
                     int *J2_out, int *num_RLFC_out) {
  const int n = p->n, B = p->B, K = p->B_thresh, W = p->W_RLFC;
  int *vote = (int *)calloc((size_t)n, sizeof(int));
  int *J2 = (int *)malloc(sizeof(int) * (size_t)p->RLFC_loops * K);
  for (int r = 0; r < p->RLFC_loops; ++r) orc_rlfc32(x, n, W, offsets[r], K, J2 + (size_t)r * K);
  int num_RLFC = K;
  if (p->RLFC_loops > 1) num_RLFC = orc_merge_J2(J2, p->RLFC_loops * K);
  orc_c32 *bins_f = (orc_c32 *)malloc(sizeof(orc_c32) * (size_t)p->loops * B);
  orc_c32 *scr = (orc_c32 *)malloc(sizeof(orc_c32) * (size_t)B);
  float *s = (float *)malloc(sizeof(float) * (size_t)B);
  int *J = (int *)malloc(sizeof(int) * (size_t)K);
  int I_F = 0;
  const int lB = ilog2_exact((uint64_t)B);
  for (int l = 0; l < p->loops && I_F >= 0; ++l) {
    orc_c32 *row = bins_f + (size_t)l * B;
    orc_perm_filter32(x, n, filter_time, p->w, B, ai[l], p->fold_mode, row);
    orc_fft32(row, lB, 0, scr);
    if (l < p->loops_loc) {
      orc_samples32(row, B, s);
      orc_cutoff32(s, B, K, J);
      if (J_out) memcpy(J_out + (size_t)l * K, J, sizeof(int) * (size_t)K);
      I_F = orc_reverse_hash(J, K, n, B, a[l], ai[l], p->loops_thresh, vote, hits_out, I_F,
                             p->hit_capacity, J2, num_RLFC, W);
    }
  }
  if (I_F >= 0)
    orc_estimate_values32(hits_out, I_F, bins_f, p->loops, n, ai, B, filter_win, nodes_out);
  if (bins_f_out) memcpy(bins_f_out, bins_f, sizeof(orc_c32) * (size_t)p->loops * B);
  if (J2_out) memcpy(J2_out, J2, sizeof(int) * (size_t)num_RLFC);
  if (num_RLFC_out) *num_RLFC_out = num_RLFC;
  free(vote); free(J2); free(bins_f); free(scr); free(s); free(J);
  return I_F;
}

/* ------------------------------------------------------------------------------------------
 * harness metric (cuExperiment.cu:181-221; comparators utils.c:231-244)
 * ---------------------------------------------------------------------------------------- */
static int cmp_node_abs_desc(const void *a, const void *b) {
  const orc_node *x = (const orc_node *)a, *y = (const orc_node *)b;
  const double ax = sqrt(x->value.x * x->value.x + x->value.y * x->value.y);
  const double ay = sqrt(y->value.x * y->value.x + y->value.y * y->value.y);
  if (ax != ay) return ax < ay ? 1 : -1;
  return (x->key > y->key) - (x->key < y->key);   /* deterministic tie order */
}

static int cmp_node_key(const void *a, const void *b) {
  const orc_node *x = (const orc_node *)a, *y = (const orc_node *)b;
  return (x->key > y->key) - (x->key < y->key);
}

void orc_metric_dense(orc_node *nodes, int I_F, int k, const int *large_freq, int n,
                      const orc_c64 *x_f, int *missed_est, int *missed_res, double *err_total) {
  orc_c64 *hat = (orc_c64 *)calloc((size_t)n, sizeof(orc_c64));
  qsort(nodes, (size_t)I_F, sizeof(orc_node), cmp_node_abs_desc);
  const int top = k < I_F ? k : I_F;   /* the reference reads k nodes even if I_F < k */
  for (int i = 0; i < top; ++i) hat[nodes[i].key] = nodes[i].value;
  qsort(nodes, (size_t)I_F, sizeof(orc_node), cmp_node_key);
  int found = 0, large_found = 0;
  for (int i = 0; i < k; ++i) {
    orc_node probe;
    probe.key = (unsigned)large_freq[i];
    if (I_F > 0 && bsearch(&probe, nodes, (size_t)I_F, sizeof(orc_node), cmp_node_key)) ++found;
    large_found += (hat[large_freq[i]].x != 0.0 || hat[large_freq[i]].y != 0.0);
  }
  double err = 0.0;
  for (int i = 0; i < n; ++i) {
    const double dx = hat[i].x - x_f[i].x, dy = hat[i].y - x_f[i].y;
    err += dx * dx + dy * dy;
  }
  free(hat);
  *missed_est = k - found;
  *missed_res = k - large_found;
  *err_total = err;
}

/* ------------------------------------------------------------------------------------------
 * value map and its adjoint (a project addition, not in the reference)
 *
 * With the located indices I_h and, per hit, the loops whose values the real and the
 * imaginary lower medians returned (sel[2h], sel[2h+1]) held fixed, the estimated values are
 * a real-linear function of the signal:
 *   yhat_h = Re(v_{sel[2h]}(I_h)) + i Im(v_{sel[2h+1]}(I_h)),
 *   v_l(I) = bins_f[l][hashed_l(I)] conj(F) / |F|^2,  bins_f[l] = FFT_B(bins_t[l]),
 *   bins_t[l][r] = sum of x[(i ai_l) mod n] f[i] over the taps i of bucket r,
 * every step exactly as innerLoop.cu:396-466 / :725-773 compute it.  The adjoint maps a
 * gradient g (g_h = dL/dRe yhat_h + i dL/dIm yhat_h) back to the signal:
 *   u[l][h'] = sum over the (hit, part) entries landing in bucket h' of loop l, in hit order
 *              (real part before imaginary part): (gr F.x/d, gr F.y/d) for the real part,
 *              (-gi F.y/d, gi F.x/d) for the imaginary part  (d = |F|^2),
 *   g_bins_t[l] = unnormalised inverse FFT_B(u[l])  (the conjugate transpose of FFT_B),
 *   g_x[(i ai_l) mod n] += conj(f[i]) g_bins_t[l][r(i)], loop by loop, taps of the tiled
 *              region (or all taps) first, then the reference fold's tail taps (n < 2^27).
 * pinned by: value_apply == the transform's node values (bit for bit), the adjoint identity
 * <A x, y> = <x, A^T y> on random vectors, and the dense matrix of A on tiny n.
 * ---------------------------------------------------------------------------------------- */
static void est_term(int I, int l, int n, const int *ai, int B, const orc_c64 *win,
                     long long *hashed_out, orc_c64 *F_out) {
  const int m = n / B;
  const long long nmask = (long long)n - 1;
  const long long p = ((long long)ai[l] * I) % n;
  long long hashed = p / m;
  long long dist = p & (m - 1);
  if (dist > m / 2) {
    hashed = (hashed + 1) & (B - 1);
    dist -= m;
  }
  dist = ((long long)n - dist) & nmask;
  *hashed_out = hashed;
  *F_out = win[(dist + m / 2) & nmask];
}

void orc_estimate_select(const int *hits, int I_F, const orc_c64 *bins_f, int loops, int n,
                         const int *ai, int B, const orc_c64 *filter_win, unsigned char *sel) {
  double *vr = (double *)malloc(sizeof(double) * (size_t)loops);
  double *vi = (double *)malloc(sizeof(double) * (size_t)loops);
  const int med = (loops - 1) / 2;
  for (int h = 0; h < I_F; ++h) {
    for (int l = 0; l < loops; ++l) {
      long long hashed;
      orc_c64 F;
      est_term(hits[h], l, n, ai, B, filter_win, &hashed, &F);
      const orc_c64 b = bins_f[(size_t)l * B + hashed];
      const double abs2 = F.x * F.x + F.y * F.y;
      vr[l] = (b.x * F.x + b.y * F.y) / abs2;
      vi[l] = (b.y * F.x - b.x * F.y) / abs2;
    }
    /* the loop a stable insertion sort puts at position med */
    for (int j = 0; j < loops; ++j) {
      int pr = 0, pi = 0;
      for (int i = 0; i < loops; ++i) {
        pr += (vr[i] < vr[j]) || (i < j && vr[i] == vr[j]);
        pi += (vi[i] < vi[j]) || (i < j && vi[i] == vi[j]);
      }
      if (pr == med) sel[2 * h] = (unsigned char)j;
      if (pi == med) sel[2 * h + 1] = (unsigned char)j;
    }
  }
  free(vr);
  free(vi);
}

static void bins_of_signal(const orc_plan *p, const orc_c64 *x, const orc_c64 *f, const int *ai,
                           orc_c64 *bins_f) {
  const int B = p->B;
  orc_c64 *scr = (orc_c64 *)malloc(sizeof(orc_c64) * (size_t)B);
  for (int l = 0; l < p->loops; ++l) {
    orc_c64 *row = bins_f + (size_t)l * B;
    orc_perm_filter64(x, p->n, f, p->w, B, ai[l], p->fold_mode, row);
    orc_fft64(row, ilog2_exact((uint64_t)B), 0, scr);
  }
  free(scr);
}

void orc_value_apply64(const orc_plan *p, const orc_c64 *x, const orc_c64 *f,
                       const orc_c64 *win, const int *ai, const int *hits, int nh,
                       const unsigned char *sel, orc_c64 *out) {
  const int B = p->B;
  orc_c64 *bins_f = (orc_c64 *)malloc(sizeof(orc_c64) * (size_t)p->loops * B);
  bins_of_signal(p, x, f, ai, bins_f);
  for (int h = 0; h < nh; ++h) {
    long long hr, hi;
    orc_c64 Fr, Fi;
    est_term(hits[h], sel[2 * h], p->n, ai, B, win, &hr, &Fr);
    est_term(hits[h], sel[2 * h + 1], p->n, ai, B, win, &hi, &Fi);
    const orc_c64 br = bins_f[(size_t)sel[2 * h] * B + hr];
    const orc_c64 bi = bins_f[(size_t)sel[2 * h + 1] * B + hi];
    out[h].x = (br.x * Fr.x + br.y * Fr.y) / (Fr.x * Fr.x + Fr.y * Fr.y);
    out[h].y = (bi.y * Fi.x - bi.x * Fi.y) / (Fi.x * Fi.x + Fi.y * Fi.y);
  }
  free(bins_f);
}

void orc_value_adjoint64(const orc_plan *p, const orc_c64 *f, const orc_c64 *win,
                         const int *ai, const int *hits, int nh, const unsigned char *sel,
                         const orc_c64 *g, orc_c64 *gx) {
  const int n = p->n, B = p->B, L = p->loops, w = p->w;
  orc_c64 *u = (orc_c64 *)calloc((size_t)L * B, sizeof(orc_c64));
  orc_c64 *scr = (orc_c64 *)malloc(sizeof(orc_c64) * (size_t)B);
  for (int h = 0; h < nh; ++h) {
    for (int part = 0; part < 2; ++part) {
      const int l = sel[2 * h + part];
      long long hb;
      orc_c64 F;
      est_term(hits[h], l, n, ai, B, win, &hb, &F);
      const double d = F.x * F.x + F.y * F.y;
      orc_c64 *ub = &u[(size_t)l * B + hb];
      if (part == 0) {
        const double q = g[h].x / d;
        ub->x = ub->x + q * F.x;
        ub->y = ub->y + q * F.y;
      } else {
        const double q = g[h].y / d;
        ub->x = ub->x - q * F.y;
        ub->y = ub->y + q * F.x;
      }
    }
  }
  for (size_t i = 0; i < (size_t)n; ++i) gx[i].x = gx[i].y = 0.0;
  const uint32_t mask = (uint32_t)n - 1u;
  const int tiled = (p->fold_mode == ORC_FOLD_REFERENCE) && (n < (1 << 27));
  const int tiles = w / B, residue = w % B;
  for (int l = 0; l < L; ++l) {
    orc_c64 *gb = u + (size_t)l * B;
    orc_fft64(gb, ilog2_exact((uint64_t)B), 1, scr);
    const uint32_t uai = (uint32_t)ai[l];
    const int head = tiled ? tiles * B : w;
    for (int i = 0; i < head; ++i) {
      const orc_c64 fv = f[i], gv = gb[i & (B - 1)];
      orc_c64 *dst = &gx[((uint32_t)i * uai) & mask];
      dst->x = dst->x + fma(fv.x, gv.x, fv.y * gv.y);
      dst->y = dst->y + fma(fv.x, gv.y, -(fv.y * gv.x));
    }
    if (tiled) {
      for (int j = 0; j < residue; ++j) {
        const int i = j * tiles + B;
        const orc_c64 fv = f[i], gv = gb[j];
        orc_c64 *dst = &gx[((uint32_t)i * uai) & mask];
        dst->x = dst->x + fma(fv.x, gv.x, fv.y * gv.y);
        dst->y = dst->y + fma(fv.x, gv.y, -(fv.y * gv.x));
      }
    }
  }
  free(u);
  free(scr);
}
