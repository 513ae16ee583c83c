/* Host-side logic of libsfft exercised without a GPU, built with AddressSanitizer and
 * UndefinedBehaviorSanitizer (tests/sanitize/Makefile): parameter validation against the
 * reference's asserts (experiment.c:77-84, cuExperiment.cu:283), the filter-width and window
 * helpers, the canonical twiddles and pass split, timesmod, the library-private random()
 * generator against glibc's own, the seeded spectrum generator, status strings, and the
 * failure paths of plan / batch creation and of the no-allocation cutoff without a device.
 * Exit status 0 when every check holds; a sanitizer report aborts the run. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/sfft.h"

static int failures = 0;
#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      fprintf(stderr, "host_checks: %s:%d: %s\n", __FILE__, __LINE__, #c); \
      ++failures;                                                         \
    }                                                                     \
  } while (0)

static sfft_params base_params(void) {
  /* n = 2^20, k = 1000, Bcst 1.4, loops 2+5, RLFC 128 x 2 */
  sfft_params p;
  memset(&p, 0, sizeof p);
  p.n = 1 << 20;
  p.B = 8192;
  p.B_thresh = 2000;
  p.W_RLFC = 16384;
  p.RLFC_loops = 2;
  p.loops_thresh = 2;
  p.loops_loc = 2;
  p.loops = 7;
  p.w = 123349;
  p.fold_mode = SFFT_FOLD_REFERENCE;
  p.hit_capacity = 0;
  p.front_end = SFFT_FRONT_AUTO;
  return p;
}

static void check_validate(void) {
  sfft_params p = base_params();
  CHECK(sfft_params_validate(&p) == SFFT_OK);
  CHECK(sfft_params_validate(NULL) == SFFT_ERR_INVALID);
  p = base_params(); p.n = 3 << 18;            CHECK(sfft_params_validate(&p) == SFFT_ERR_INVALID);
  p = base_params(); p.B = 6000;               CHECK(sfft_params_validate(&p) == SFFT_ERR_INVALID);
  p = base_params(); p.B_thresh = p.B;         CHECK(sfft_params_validate(&p) == SFFT_ERR_INVALID);
  p = base_params(); p.W_RLFC = 1024;          CHECK(sfft_params_validate(&p) == SFFT_ERR_INVALID);
  p = base_params(); p.loops_thresh = 3;       CHECK(sfft_params_validate(&p) == SFFT_ERR_INVALID);
  p = base_params(); p.loops = 33;             CHECK(sfft_params_validate(&p) == SFFT_ERR_INVALID);
  p = base_params(); p.w = p.n + 1;            CHECK(sfft_params_validate(&p) == SFFT_ERR_INVALID);
  p = base_params(); p.w = p.B - 1;            CHECK(sfft_params_validate(&p) == SFFT_ERR_INVALID);
  p.fold_mode = SFFT_FOLD_ALL_TAPS;            CHECK(sfft_params_validate(&p) == SFFT_OK);
  p = base_params(); p.fold_mode = 7;          CHECK(sfft_params_validate(&p) == SFFT_ERR_INVALID);
  p = base_params(); p.front_end = 4;          CHECK(sfft_params_validate(&p) == SFFT_ERR_INVALID);
  p = base_params(); p.hit_capacity = -1;      CHECK(sfft_params_validate(&p) == SFFT_ERR_INVALID);
  p = base_params(); p.RLFC_loops = 0;         CHECK(sfft_params_validate(&p) == SFFT_ERR_INVALID);
  /* every power-of-two size the plan accepts, and one past each end */
  for (int q = 4; q <= 31; ++q) {
    p = base_params();
    p.n = (q <= 30) ? (1 << q) : (int)0x7fffffff;
    p.B = 16;
    p.B_thresh = 4;
    p.W_RLFC = (q > 9) ? (1 << (q - 4)) : 32;   /* W >= n/B, the comb spans a window */
    p.w = 16;
    p.fold_mode = SFFT_FOLD_ALL_TAPS;
    const sfft_status s = sfft_params_validate(&p);
    CHECK((q >= 5 && q <= 30) ? s == SFFT_OK : s == SFFT_ERR_INVALID);
  }
}

static void check_helpers(void) {
  double t0 = 0;
  const int w = sfft_filter_taps(0.5 / 10137.0, 1e-8, &t0);   /* n = 2^20, k = 1000 */
  CHECK(w == 123349 && t0 > 1.0);
  for (int log2n = 0; log2n <= 30; ++log2n) {
    int logs[8] = {0};
    const int np = sfft_fft_plan_passes(log2n, logs);
    int sum = 0;
    for (int i = 0; i < np; ++i) {
      CHECK(logs[i] >= 1 && logs[i] <= 11);
      sum += logs[i];
    }
    CHECK(sum == log2n);
  }
  for (uint64_t T = 1; T <= (1ull << 30); T <<= 3) {
    for (uint64_t e = 0; e < T && e < 64; ++e) {
      double re, im;
      sfft_twiddle(e, T, 0, &re, &im);
      const double ang = -2.0 * M_PI * (double)e / (double)T;
      CHECK(fabs(re - cos(ang)) < 1e-15 && fabs(im - sin(ang)) < 1e-15);
    }
  }
  for (int i = 0; i < 1000; ++i) {
    const int n = 1 << (4 + i % 27), x = (i * 7919) & (n - 1), a = (i * 104729 + 1) & (n - 1);
    CHECK(timesmod(x, a, n) == (int)(((long long)x * a) % n));
  }
  const int n = 4096, B = 64, m = n / B;
  complex_t* ff = (complex_t*)calloc((size_t)n, sizeof(complex_t));
  complex_t* win = (complex_t*)calloc((size_t)m, sizeof(complex_t));
  for (int i = 0; i < n; ++i) ff[i].x = i;
  sfft_filter_window(ff, n, B, win);
  for (int j = 0; j < m; ++j) CHECK(win[j].x == (double)((j - m / 2 + n) % n));
  free(ff);
  free(win);
  for (int s = 0; s <= 6; ++s) CHECK(sfft_status_string((sfft_status)s) != NULL);
  CHECK(strcmp(sfft_status_string(SFFT_OK), "ok") == 0);
}

static void check_random(void) {
  /* the library-private generator is glibc's random(): same seed, same sequence */
  for (unsigned seed = 1; seed < 5; ++seed) {
    sfft_srandom(seed);
    srandom(seed);
    for (int i = 0; i < 1000; ++i) CHECK(sfft_random() == random());
  }
}

static void check_spectrum(void) {
  const int n = 1 << 16, k = 500;
  int* locs = (int*)malloc(sizeof(int) * k);
  complex_t* vals = (complex_t*)malloc(sizeof(complex_t) * k);
  CHECK(sfft_signal_spectrum(n, k, 12345u, locs, vals) == SFFT_OK);
  for (int i = 0; i < k; ++i) {
    CHECK(locs[i] >= 0 && locs[i] < n);
    if (i) CHECK(locs[i] > locs[i - 1]);
    CHECK(fabs(hypot(vals[i].x, vals[i].y) - 1.0) < 1e-15);
  }
  CHECK(sfft_signal_spectrum(n, n + 1, 1u, locs, vals) == SFFT_ERR_INVALID);
  free(locs);
  free(vals);
}

static void check_no_device_paths(void) {
  sfft_params p = base_params();
  sfft_plan* plan = (sfft_plan*)0x1;
  const sfft_status s = sfft_plan_create(&p, SFFT_FP64, NULL, &plan);
  if (sfft_device_count() == 0) CHECK(s == SFFT_ERR_NO_DEVICE && plan == NULL);
  if (s == SFFT_OK) sfft_plan_destroy(plan);
  p.B_thresh = p.B;
  CHECK(sfft_plan_create(&p, SFFT_FP64, NULL, &plan) == SFFT_ERR_INVALID && plan == NULL);
  sfft_batch* b = (sfft_batch*)0x1;
  CHECK(sfft_batch_create(&p, 1, SFFT_FP32, NULL, &b) == SFFT_ERR_INVALID && b == NULL);
  CHECK(sfft_batch_create(NULL, 1, SFFT_FP32, NULL, &b) == SFFT_ERR_INVALID);
  double t = 0;
  /* size limit and argument errors are decided before any device call */
  CHECK(samples_indices_noalloc(NULL, 10, SFFT_NOALLOC_MAX_SAMPLES + 1, NULL, &t) == -2);
  CHECK(samples_indices_noalloc(NULL, 0, 100, NULL, &t) == -1);
}

int main(void) {
  check_validate();
  check_helpers();
  check_random();
  check_spectrum();
  check_no_device_paths();
  if (failures) {
    fprintf(stderr, "host_checks: %d failure(s)\n", failures);
    return 1;
  }
  printf("host_checks: ok\n");
  return 0;
}
