// sfft_compat.hip -- the reference's C entry points (innerLoop.h, outerLoop.h, filters.h,
// experiment.h) on top of the MI355X kernels, with the reference's argument meaning,
// printed output and libc random draws.  The native plan API does the work.
#include <stdlib.h>
#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "sfft_aux.h"

using namespace sfft;

extern "C" {
bool_t ALGORITHM1 = 1;   // outerLoop.cu:49 (unused by the reference too)
bool_t VERBOSE = 0;      // outerLoop.cu:50
bool_t TIMING = 1;       // outerLoop.cu:51
}

#define vprintf_(...)          \
  do {                         \
    if (VERBOSE) printf(__VA_ARGS__); \
  } while (0)

static struct timeval g_start;   // timer.c:31
static void reset_timer_() { gettimeofday(&g_start, NULL); }
static double get_time_() {
  struct timeval end;
  gettimeofday(&end, NULL);
  return (double)(end.tv_sec - g_start.tv_sec) +
         ((double)end.tv_usec / 1e6 - (double)g_start.tv_usec / 1e6);
}

static double wall_now() {
  struct timeval t;
  gettimeofday(&t, NULL);
  return (double)t.tv_sec + 1e-6 * (double)t.tv_usec;
}

// The permutation draws of outer_loop (outerLoop.cu:129-134) use glibc's random() generator
// (the TYPE_3 additive-feedback generator with a 128-byte state), but on a library-private
// state: the HIP runtime calls random() from its own threads, so once a device is
// initialised the process-wide state advances by a timing-dependent amount and the draws
// would differ from run to run.  The private state starts at seed 1, which is the state of
// a program that never calls srandom(); sfft_srandom reseeds it.  drand48() (comb offsets,
// signal locations) is not touched by the runtime and stays the process-wide generator.
static std::mutex g_rand_mu;
static struct random_data g_rand;
static char g_rand_state[128];
static bool g_rand_ready = false;

static void rand_ready_locked() {
  if (g_rand_ready) return;
  memset(&g_rand, 0, sizeof g_rand);
  initstate_r(1u, g_rand_state, sizeof g_rand_state, &g_rand);
  g_rand_ready = true;
}

extern "C" void sfft_srandom(unsigned int seed) {
  std::lock_guard<std::mutex> lk(g_rand_mu);
  rand_ready_locked();
  srandom_r(seed, &g_rand);
}

extern "C" long sfft_random(void) {
  std::lock_guard<std::mutex> lk(g_rand_mu);
  rand_ready_locked();
  int32_t r = 0;
  random_r(&g_rand, &r);
  return (long)r;
}

// outer_loop's plan and pinned node buffer, kept between calls (released with the process)
namespace {
struct OuterCache {
  sfft_plan* plan = nullptr;
  sfft_params p{};
  int device = -1;
  Node* nodes = nullptr;
};
OuterCache g_outer;
std::mutex g_outer_mu;
}  // namespace

static int gcd_(int a, int b) {   // utils.c:81-84
  while (b) {
    const int r = a % b;
    a = b;
    b = r;
  }
  return a;
}

static int mod_inverse_(int a, int n) {   // utils.c:92-105 (extended Euclid)
  long long r0 = n, r1 = a, t0 = 0, t1 = 1;
  while (r1) {
    const long long q = r0 / r1;
    long long t = r0 - q * r1;
    r0 = r1;
    r1 = t;
    t = t0 - q * t1;
    t0 = t1;
    t1 = t;
  }
  long long v = t0 % n;
  if (v < 0) v += n;
  return (int)v;
}

extern "C" int timesmod(const int x, const int a, const int n) {   // innerLoop.cu:647-649
  return (int)((((long long)x) * a) % n);
}

static int ilog2_(long long v) {
  int l = 0;
  while ((1LL << l) < v) ++l;
  return ((1LL << l) == v) ? l : -1;
}

// ---------------------------------------------------------------------------------------------
// outer_loop (outerLoop.cu:65-190)
// ---------------------------------------------------------------------------------------------
extern "C" Node* outer_loop(complex_t* d_x, int n, complex_t* d_filter_time,
                            complex_t* d_filter_freq, int d_filter_size, int B, int B_thresh,
                            int W_RLFC, int RLFC_loops, int loop_threshold, int loops_loc,
                            int loops, int* I_F, float t_H_to_D) {
  const double t_begin = wall_now();
  sfft_params p;
  p.n = n;
  p.B = B;
  p.B_thresh = B_thresh;
  p.W_RLFC = W_RLFC;
  p.RLFC_loops = RLFC_loops;
  p.loops_thresh = loop_threshold;
  p.loops_loc = loops_loc;
  p.loops = loops;
  p.w = d_filter_size;
  p.fold_mode = SFFT_FOLD_REFERENCE;
  p.hit_capacity = MAXSIZE_HITS;
  p.front_end = SFFT_FRONT_AUTO;
  // draws in the reference's order: one drand48 per comb loop (innerLoop.cu:122-123), then
  // random() until gcd(a, n) == 1 per loop (outerLoop.cu:129-134)
  std::vector<int> a(loops), ai(loops), offs(RLFC_loops);
  const int sigma = n / W_RLFC;
  for (int r = 0; r < RLFC_loops; ++r) offs[r] = (int)(unsigned)floor(drand48() * sigma);
  for (int l = 0; l < loops; ++l) {
    int av = 0;
    while (gcd_(av, n) != 1) av = (int)(sfft_random() % n);
    a[l] = av;
    ai[l] = mod_inverse_(av, n);
  }
  // the plan (workspace, FFT tables) and a pinned node buffer are kept from one call to the
  // next while the parameters and the device stay the same (callers transform repeatedly)
  std::lock_guard<std::mutex> lk(g_outer_mu);
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (g_outer.plan && (g_outer.device != dev || memcmp(&g_outer.p, &p, sizeof p) != 0)) {
    sfft_plan_destroy(g_outer.plan);
    g_outer.plan = nullptr;
  }
  sfft_status s = SFFT_OK;
  if (!g_outer.plan) {
    s = sfft_plan_create(&p, SFFT_FP64, nullptr, &g_outer.plan);
    if (s != SFFT_OK) {
      g_outer.plan = nullptr;
      fprintf(stderr, "outer_loop: %s\n", sfft_status_string(s));
      return nullptr;
    }
    g_outer.p = p;
    g_outer.device = dev;
  }
  if (!g_outer.nodes &&
      hipHostMalloc((void**)&g_outer.nodes, sizeof(Node) * MAXSIZE_HITS, hipHostMallocDefault) !=
          hipSuccess) {
    g_outer.nodes = nullptr;
    fprintf(stderr, "outer_loop: cannot allocate the pinned node buffer\n");
    return nullptr;
  }
  sfft_plan* plan = g_outer.plan;
  const int m = n / B;
  std::vector<complex_t> win(m), seg(m);
  // the m entries of filter_freq around 0 (host or device pointer)
  hipError_t e1 = hipMemcpy(win.data(), d_filter_freq + (n - m / 2),
                            (size_t)(m / 2) * sizeof(complex_t), hipMemcpyDefault);
  hipError_t e2 = hipMemcpy(win.data() + m / 2, d_filter_freq,
                            (size_t)(m - m / 2) * sizeof(complex_t), hipMemcpyDefault);
  if (e1 != hipSuccess || e2 != hipSuccess) {
    fprintf(stderr, "outer_loop: cannot read filter_freq\n");
    return nullptr;
  }
  sfft_plan_set_filter(plan, d_filter_time, win.data());
  sfft_plan_set_profiling(plan, TIMING ? 1 : 0);
  sfft_draws dr{a.data(), ai.data(), offs.data()};
  Node* nodes = g_outer.nodes;
  int count = 0;
  s = sfft_execute(plan, d_x, &dr, nodes, &count);
  double st7[7] = {0, 0, 0, 0, 0, 0, 0};
  if (TIMING) sfft_plan_stage_times(plan, st7);
  if (s != SFFT_OK) {
    fprintf(stderr, "outer_loop: %s (%d hits)\n", sfft_status_string(s), count);
    return nullptr;
  }
  *I_F = count;
  Node* out = (Node*)malloc(sizeof(Node) * (size_t)(count > 0 ? count : 1));
  memcpy(out, nodes, sizeof(Node) * (size_t)count);
  const double DDD = wall_now() - t_begin;
  if (TIMING) {   // outerLoop.cu:174-187
    const double oth = DDD - st7[0] - st7[1] - st7[2] - st7[3] - st7[4] - st7[5];
    printf("Total GPU-SFFT time: %lf\n", DDD);
    printf("Time distribution: votetable  restr-loc-index  perm+filter reverse-hash  estimation "
           "FFT+Cutoff    other    total\n");
    printf("                    %2.4lf       %2.4lf          %2.4lf      %2.4lf        %2.4lf     "
           "%2.4lf      %2.4lf   %2.4lf\n",
           st7[0], st7[1], st7[2], st7[3], st7[4], st7[5], oth, DDD);
    const double tott = DDD / 100;
    printf("                    %4.1f%%         %4.1f%%           %4.1f%%      %4.1f%%         "
           "%4.1f%%      %4.1f%%       %4.1f%%   %5.1f%%\n",
           st7[0] / tott, st7[1] / tott, st7[2] / tott, st7[3] / tott, st7[4] / tott,
           st7[5] / tott, oth / tott, DDD / tott);
    printf("\n");
    const double thd = t_H_to_D / 1.0e3;
    printf("HtoD time = %lfs, HtoD time/total time  = %lf%%, total time + HtoD time = %lfs \n",
           thd, thd / tott, DDD + thd);
    printf("GPU device time (events, first kernel to last) = %lfs\n", st7[6]);
  }
  return out;
}

// ---------------------------------------------------------------------------------------------
// stage entry points (innerLoop.cu)
// ---------------------------------------------------------------------------------------------
static int cutoff_device(int fp32, const void* d_keys, int N, int K, int* d_out, hipStream_t st) {
  void* w = nullptr;
  if (hipMalloc(&w, select_work_bytes(1, N)) != hipSuccess) return -1;
  (void)hipMemsetAsync(w, 0, select_work_bytes(1, N), st);
  SelectWork sw;
  select_work_bind(&sw, w, 1, N);
  const hipError_t e = select_topk(fp32, d_keys, N, N, K, 1, d_out, 1, sw, st);
  (void)hipStreamSynchronize(st);
  (void)hipFree(w);
  return e == hipSuccess ? 0 : -1;
}

// |X|^2 keys of a complex128 vector: a one-pass identity "FFT" is not available, so the keys
// come from a tiny kernel expressed through the FFT's key epilogue on a copy.
__global__ void keys_c128_kernel(const double2* __restrict__ X, unsigned long long* __restrict__ k,
                                 int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) k[i] = (unsigned long long)__double_as_longlong(X[i].x * X[i].x + X[i].y * X[i].y);
}

extern "C" int samples_indices(complex_t* d_samp, int B, int samp_size, int* d_output,
                               double* CTime) {   // innerLoop.cu:484-539
  const double t0 = wall_now();
  if (B < 1 || B >= samp_size) return -1;   // cutoff asserts samp_size >= B + 1
  unsigned long long* keys = nullptr;
  if (hipMalloc(&keys, (size_t)samp_size * 8) != hipSuccess) return -1;
  hipLaunchKernelGGL(keys_c128_kernel, dim3((samp_size + 255) / 256), dim3(256), 0, nullptr,
                     (const double2*)d_samp, keys, samp_size);
  const int rc = cutoff_device(0, keys, samp_size, B, d_output, nullptr);
  (void)hipFree(keys);
  if (CTime) *CTime += wall_now() - t0;
  return rc;
}

// samples_indices without allocation: the |X|^2 keys and the select workspace live in one
// file-scope device array (each device holds its own copy for the life of the process).  Host
// code reaches it only through the address hipGetSymbolAddress returns on that device, looked
// up once per device and kept.
static constexpr int kNoAllocMaxN = SFFT_NOALLOC_MAX_SAMPLES;
static constexpr size_t kNoAllocBytes = (size_t)12 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];

static void* scratch_on_current_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

extern "C" int samples_indices_noalloc(complex_t* d_samp, int B, int samp_size, int* d_output,
                                       double* CTime) {
  const double t0 = wall_now();
  if (B < 1 || B >= samp_size) return -1;   // as samples_indices
  if (samp_size > kNoAllocMaxN) {
    fprintf(stderr, "samples_indices_noalloc: samp_size %d exceeds the scratch maximum %d\n",
            samp_size, kNoAllocMaxN);
    return -2;
  }
  const size_t keys_bytes = ((size_t)samp_size * 8 + 255) / 256 * 256;
  const size_t work_bytes = select_work_bytes(1, samp_size);
  if (keys_bytes + work_bytes > kNoAllocBytes) return -2;   // cannot happen below the maximum
  char* base = (char*)scratch_on_current_device();
  if (!base) return -1;
  unsigned long long* keys = (unsigned long long*)base;
  void* work = base + keys_bytes;
  // the select state must start at zero; the previous call left its own contents behind
  if (fill_u32(work, 0u, (work_bytes + 3) / 4, nullptr) != hipSuccess) return -1;
  hipLaunchKernelGGL(keys_c128_kernel, dim3((samp_size + 255) / 256), dim3(256), 0, nullptr,
                     (const double2*)d_samp, keys, samp_size);
  if (hipGetLastError() != hipSuccess) return -1;
  SelectWork sw;
  select_work_bind(&sw, work, 1, samp_size);
  const hipError_t e = select_topk(0, keys, samp_size, samp_size, B, 1, d_output, 1, sw, nullptr);
  const hipError_t s = hipStreamSynchronize(nullptr);
  if (CTime) *CTime += wall_now() - t0;
  return (e == hipSuccess && s == hipSuccess) ? 0 : -1;
}

extern "C" int restr_loc_freq_coeff(complex_t* d_x, int n, int B_thresh, int W_RLFC, int* d_J2,
                                    double* RLFC_Time) {   // innerLoop.cu:112-177
  const double t0 = wall_now();
  const int lW = ilog2_(W_RLFC);
  if (lW < 1 || W_RLFC > n || B_thresh >= W_RLFC) return -1;
  const int sigma = n / W_RLFC;
  const int offset = (int)(unsigned)floor(drand48() * sigma);
  FftTables tab;
  if (fft_tables_create(lW, 0, 0, &tab) != hipSuccess) return -1;
  void *b0 = nullptr, *b1 = nullptr, *keys = nullptr;
  int* doff = nullptr;
  int rc = -1;
  if (hipMalloc(&b0, (size_t)W_RLFC * 16) == hipSuccess &&
      hipMalloc(&b1, (size_t)W_RLFC * 16) == hipSuccess &&
      hipMalloc(&keys, (size_t)W_RLFC * 8) == hipSuccess &&
      hipMalloc(&doff, 4) == hipSuccess &&
      hipMemcpy(doff, &offset, 4, hipMemcpyHostToDevice) == hipSuccess) {
    FftLaunch c;
    c.tab = &tab;
    c.out = b0;
    c.out_bstride = W_RLFC;
    c.scratch = b1;
    c.keys = keys;
    c.keys_bstride = W_RLFC;
    c.key_batches = 1;
    c.batch = 1;
    c.load = kLoadComb;
    c.store = kStoreKeys;
    c.comb_x = d_x;
    c.comb_offs = doff;
    c.comb_sigma = sigma;
    if (fft_run(c, nullptr) == hipSuccess)
      rc = cutoff_device(0, keys, W_RLFC, B_thresh, d_J2, nullptr);
  }
  (void)hipDeviceSynchronize();
  if (b0) (void)hipFree(b0);
  if (b1) (void)hipFree(b1);
  if (keys) (void)hipFree(keys);
  if (doff) (void)hipFree(doff);
  fft_tables_destroy(&tab);
  if (RLFC_Time) *RLFC_Time += wall_now() - t0;
  return rc;
}

extern "C" int perm_filter_cutoff(complex_t* d_x, int n, complex_t* d_filter_time,
                                  int d_filter_size, int B_thresh, int B, int a, int ai,
                                  complex_t* d_bins_f, int* d_J, double* PF_T,
                                  double* BC_T) {   // innerLoop.cu:194-286
  (void)a;
  const int lB = ilog2_(B);
  if (lB < 1 || B > n || B_thresh >= B) return -1;
  if (n % B) fprintf(stderr, "Warning: n is not divisible by B, which algorithm expects.\n");
  const int tiled = n < (1 << 27);
  if (tiled && d_filter_size < B) return -1;
  double t0 = wall_now();
  void *bins_t = nullptr, *scr = nullptr, *keys = nullptr;
  int* dai = nullptr;
  int rc = -1;
  FftTables tab;
  if (fft_tables_create(lB, 0, 0, &tab) != hipSuccess) return -1;
  if (hipMalloc(&bins_t, (size_t)B * 16) == hipSuccess &&
      hipMalloc(&scr, (size_t)B * 16) == hipSuccess &&
      hipMalloc(&keys, (size_t)B * 8) == hipSuccess && hipMalloc(&dai, 4) == hipSuccess &&
      hipMemcpy(dai, &ai, 4, hipMemcpyHostToDevice) == hipSuccess &&
      perm_filter_launch(0, d_x, d_filter_time, bins_t, dai, n, B, d_filter_size, 1, tiled,
                         nullptr) == hipSuccess) {
    (void)hipDeviceSynchronize();
    if (PF_T) *PF_T = wall_now() - t0;
    t0 = wall_now();
    FftLaunch f;
    f.tab = &tab;
    f.in = bins_t;
    f.in_bstride = B;
    f.out = d_bins_f;
    f.out_bstride = B;
    f.scratch = scr;
    f.keys = keys;
    f.keys_bstride = B;
    f.key_batches = 1;
    f.batch = 1;
    f.store = kStoreBoth;
    if (fft_run(f, nullptr) == hipSuccess) rc = cutoff_device(0, keys, B, B_thresh, d_J, nullptr);
    if (BC_T) *BC_T = wall_now() - t0;
  }
  (void)hipDeviceSynchronize();
  if (bins_t) (void)hipFree(bins_t);
  if (scr) (void)hipFree(scr);
  if (keys) (void)hipFree(keys);
  if (dai) (void)hipFree(dai);
  fft_tables_destroy(&tab);
  return rc;
}

extern "C" int reverse_hash(int* d_J, int n, int B_thresh, int B, int a, int ai, int loop_thresh,
                            int* d_vote, int* d_I, int* I_F, double* G_T, int* d_J2,
                            int num_RLFC, int W_RLFC, int RLFC_loops) {   // innerLoop.cu:300-384
  (void)RLFC_loops;
  const double t0 = wall_now();
  if (ilog2_(W_RLFC) < 5 || num_RLFC < 1 || !I_F) return -1;
  unsigned* bm = nullptr;
  int *maxfirst = nullptr, *cnt = nullptr, *j2 = nullptr;
  const size_t words = ((size_t)W_RLFC + 31) / 32;
  int rc = -1;
  if (hipMalloc(&bm, words * 4) == hipSuccess && hipMalloc(&maxfirst, 4) == hipSuccess &&
      hipMalloc(&cnt, 4) == hipSuccess && hipMalloc(&j2, (size_t)num_RLFC * 4) == hipSuccess &&
      hipMemcpy(j2, d_J2, (size_t)num_RLFC * 4, hipMemcpyDefault) == hipSuccess &&
      hipMemcpy(cnt, I_F, 4, hipMemcpyHostToDevice) == hipSuccess &&
      fill_u32(bm, 0u, words, nullptr) == hipSuccess &&
      fill_u32(maxfirst, 0xFFFFFFFFu, 1, nullptr) == hipSuccess &&
      aux_bitmap_from_list(j2, num_RLFC, ai, W_RLFC, bm, maxfirst, nullptr) == hipSuccess &&
      aux_vote_counts(d_J, B_thresh, n, B, W_RLFC, a, bm, maxfirst, loop_thresh, d_vote, d_I, cnt,
                      MAXSIZE_HITS, nullptr) == hipSuccess &&
      hipMemcpy(I_F, cnt, 4, hipMemcpyDeviceToHost) == hipSuccess) {
    rc = (*I_F > MAXSIZE_HITS) ? -1 : 0;
  }
  if (bm) (void)hipFree(bm);
  if (maxfirst) (void)hipFree(maxfirst);
  if (cnt) (void)hipFree(cnt);
  if (j2) (void)hipFree(j2);
  if (G_T) *G_T = wall_now() - t0;
  return rc;
}

extern "C" Node* estimate_values(int* d_I, int* I_F, complex_t* d_bins_f, int loops, int n,
                                 int* permute, int B, complex_t* d_filter_freq,
                                 int loops_loc) {   // innerLoop.cu:396-466
  (void)loops_loc;
  if (!I_F || *I_F < 0 || loops < 1 || loops > 32) return nullptr;
  const int cnt = *I_F, m = n / B;
  std::vector<complex_t> win(m);
  if (hipMemcpy(win.data(), d_filter_freq + (n - m / 2), (size_t)(m / 2) * 16,
                hipMemcpyDefault) != hipSuccess ||
      hipMemcpy(win.data() + m / 2, d_filter_freq, (size_t)(m - m / 2) * 16,
                hipMemcpyDefault) != hipSuccess)
    return nullptr;
  Node* h = (Node*)malloc(sizeof(Node) * (size_t)(cnt > 0 ? cnt : 1));
  void *dwin = nullptr, *dnodes = nullptr;
  int *dperm = nullptr, *dcnt = nullptr;
  bool ok = hipMalloc(&dwin, (size_t)m * 16) == hipSuccess &&
            hipMalloc(&dnodes, sizeof(Node) * (size_t)(cnt > 0 ? cnt : 1)) == hipSuccess &&
            hipMalloc(&dperm, (size_t)loops * 4) == hipSuccess &&
            hipMalloc(&dcnt, 4) == hipSuccess &&
            hipMemcpy(dwin, win.data(), (size_t)m * 16, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dperm, permute, (size_t)loops * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dcnt, &cnt, 4, hipMemcpyHostToDevice) == hipSuccess &&
            estimate_launch(0, d_I, dcnt, cnt, d_bins_f, loops, n, dperm, B,
                            (const double2*)dwin, (Node*)dnodes, nullptr, nullptr) ==
                hipSuccess &&
            hipMemcpy(h, dnodes, sizeof(Node) * (size_t)cnt, hipMemcpyDeviceToHost) == hipSuccess;
  if (dwin) (void)hipFree(dwin);
  if (dnodes) (void)hipFree(dnodes);
  if (dperm) (void)hipFree(dperm);
  if (dcnt) (void)hipFree(dcnt);
  if (!ok) {
    free(h);
    return nullptr;
  }
  return h;
}

// ---------------------------------------------------------------------------------------------
// harness (cuExperiment.cu)
// ---------------------------------------------------------------------------------------------
extern "C" sfft_status sfft_make_filter(double lobefrac, double tolerance, int n, int b_f, int w,
                                        complex_t* d_filter_time, complex_t* h_filter_freq,
                                        void* stream);
extern "C" int sfft_filter_taps(double lobefrac, double tolerance, double* t0_out);

extern "C" void cuShift(complex_t* d_x, int n, int r) {   // cuExperiment.cu:507-522
  r = (n + r) % n;
  complex_t* tmp = nullptr;
  if (hipMalloc(&tmp, (size_t)n * sizeof(complex_t)) != hipSuccess) return;
  (void)hipMemcpy(tmp, d_x + (n - r), (size_t)r * sizeof(complex_t), hipMemcpyDeviceToDevice);
  (void)hipMemcpy(tmp + r, d_x, (size_t)(n - r) * sizeof(complex_t), hipMemcpyDeviceToDevice);
  (void)hipMemcpy(d_x, tmp, (size_t)n * sizeof(complex_t), hipMemcpyDeviceToDevice);
  (void)hipFree(tmp);
}

static double cabs2_(complex_t v) { return v.x * v.x + v.y * v.y; }

extern "C" void run_experiment(complex_t* d_x, int n, complex_t* d_filter_time,
                               complex_t* d_filter_freq, int d_filter_size, int B, int B_thresh,
                               int loops_loc, int loops_thresh, int loops_est, int W_RLFC,
                               int RLFC_loops, int repetitions, int* LARGE_FREQ, int k,
                               complex_t* x_f, float t_H_to_D) {   // cuExperiment.cu:159-223
  printf("GPU-SFFT Results\n");
  printf("******************************************************************************\n");
  Node* hat_x = nullptr;
  int I_F = 0;
  const int loops = loops_loc + loops_est;
  for (int i = 0; i < repetitions; ++i) {
    reset_timer_();
    if (hat_x) free(hat_x);
    I_F = 0;
    hat_x = outer_loop(d_x, n, d_filter_time, d_filter_freq, d_filter_size, B, B_thresh, W_RLFC,
                       RLFC_loops, loops_thresh, loops_loc, loops, &I_F, t_H_to_D);
  }
  if (!hat_x) {
    printf("ERROR: transform failed\n");
    return;
  }
  const double t0 = get_time_();
  std::vector<Node> v(hat_x, hat_x + I_F);
  std::sort(v.begin(), v.end(), [](const Node& a, const Node& b) {
    const double aa = sqrt(cabs2_(a.value)), bb = sqrt(cabs2_(b.value));
    if (aa != bb) return aa > bb;
    return a.key < b.key;
  });
  std::vector<complex_t> large((size_t)n, complex_t{0.0, 0.0});
  const int top = std::min(k, I_F);   // the reference reads k entries even when I_F < k
  for (int i = 0; i < top; ++i) large[v[i].key] = v[i].value;
  const double t_large = get_time_() - t0;
  std::vector<unsigned> keys(I_F);
  for (int i = 0; i < I_F; ++i) keys[i] = hat_x[i].key;
  std::sort(keys.begin(), keys.end());
  int found = 0, large_found = 0;
  for (int i = 0; i < k; ++i) {
    found += std::binary_search(keys.begin(), keys.end(), (unsigned)LARGE_FREQ[i]) ? 1 : 0;
    large_found += (large[LARGE_FREQ[i]].x != 0.0 || large[LARGE_FREQ[i]].y != 0.0);
  }
  double ERROR = 0.0;
  for (int i = 0; i < n; ++i) {
    const double dx = large[i].x - x_f[i].x, dy = large[i].y - x_f[i].y;
    ERROR += dx * dx + dy * dy;
  }
  const double MSE = ERROR / k;
  printf("ERROR:\n");
  printf("K=%d; MISSED (estimation, result) = (%d, %d), MSE*k = %lg, Mean Square Error (MSE) = "
         "%lg\n",
         k, k - found, k - large_found, ERROR, MSE);
  printf("time to extract the values of the largest frequencies: %lfs \n", t_large);
  printf("******************************************************************************\n\n");
  free(hat_x);
}

static void fft_results(complex_t* d_x, int n) {   // cudafft_results + print_timing
  void *out = nullptr, *scr = nullptr;
  const int ln = ilog2_(n);
  if (ln < 1 || hipMalloc(&out, (size_t)n * 16) != hipSuccess ||
      hipMalloc(&scr, (size_t)n * 16) != hipSuccess) {
    if (out) (void)hipFree(out);
    return;
  }
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  (void)fft_plain(d_x, out, scr, ln, 1, 0, 0, nullptr);   // warm-up (tables)
  (void)hipEventRecord(a, nullptr);
  (void)fft_plain(d_x, out, scr, ln, 1, 0, 0, nullptr);
  (void)hipEventRecord(b, nullptr);
  (void)hipEventSynchronize(b);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  printf("canonical FFT (gfx950 kernels) n=%d execution time    %4.2fms\n", n, ms);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(out);
  (void)hipFree(scr);
}

extern "C" void make_multiple_t(double lobefrac, double tolerance, int n, int b_f, complex_t* d_x,
                                int B, int B_thresh, int loops_loc, int loops_thresh,
                                int loops_est, int W_Comb, int Comb_loops, int repetitions,
                                int* LARGE_FREQ, int k, complex_t* x_f,
                                float t_H_to_D) {   // cuExperiment.cu:248-503
  const double t0 = wall_now();
  const int w = sfft_filter_taps(lobefrac, tolerance, nullptr);
  if (w > n) {
    fprintf(stderr, "make_multiple_t: filter wider than the signal (w=%d > n=%d)\n", w, n);
    return;
  }
  complex_t* d_ft = nullptr;
  if (hipMalloc(&d_ft, (size_t)w * sizeof(complex_t)) != hipSuccess) return;
  std::vector<complex_t> ff((size_t)n);
  const sfft_status s = sfft_make_filter(lobefrac, tolerance, n, b_f, w, d_ft, ff.data(), nullptr);
  if (s != SFFT_OK) {
    fprintf(stderr, "make_multiple_t: %s\n", sfft_status_string(s));
    (void)hipFree(d_ft);
    return;
  }
  printf(" Window size: Filter : %d\n", w);
  printf(" Time for computing  filter:  %lfs\n", wall_now() - t0);
  printf("******************************************************************************\n\n");
  printf("Simulation starting:\n");
  run_experiment(d_x, n, d_ft, ff.data(), w, B, B_thresh, loops_loc, loops_thresh, loops_est,
                 W_Comb, Comb_loops, repetitions, LARGE_FREQ, k, x_f, t_H_to_D);
  (void)hipFree(d_ft);
  printf("canonical FFT (no vendor library) results\n");
  printf("******************************************************************************\n");
  fft_results(d_x, n);
  printf("******************************************************************************\n\n");
}

extern "C" int cudafft_experiment(int n, double lobefrac, double tolerance, int b_f, int B,
                                  int B_thresh, int loops_loc, int loops_thresh, int loops_est,
                                  int W_RLFC, int RLFC_loops, int repetitions, int k,
                                  double std_noise) {   // cuExperiment.cu:64-137
  const int ln = ilog2_(n);
  if (ln < 1) return -1;
  std::vector<complex_t> h_x((size_t)n), x_f((size_t)n, complex_t{0.0, 0.0});
  std::vector<int> LARGE_FREQ(k);
  for (int i = 0; i < k; ++i) {   // unit real coefficients at drand48 locations
    LARGE_FREQ[i] = (int)(unsigned)floor(drand48() * n);
    x_f[LARGE_FREQ[i]].x = 1.0;
    x_f[LARGE_FREQ[i]].y = 0.0;
  }
  void *a = nullptr, *b = nullptr, *scr = nullptr;
  if (hipMalloc(&a, (size_t)n * 16) != hipSuccess || hipMalloc(&b, (size_t)n * 16) != hipSuccess ||
      hipMalloc(&scr, (size_t)n * 16) != hipSuccess)
    return -1;
  // inverse FFT -> time signal (cuExperiment.cu:91-93)
  (void)hipMemcpy(a, x_f.data(), (size_t)n * 16, hipMemcpyHostToDevice);
  (void)fft_plain(a, b, scr, ln, 1, 1, 0, nullptr);
  (void)hipMemcpy(h_x.data(), b, (size_t)n * 16, hipMemcpyDeviceToHost);
  if (std_noise != 0) {   // AWGN (utils.c:187-219)
    double sig = 0, noise = 0;
    for (int h = 0; h < n; ++h) {
      sig += cabs2_(h_x[h]);
      const double u = drand48(), v = drand48();
      h_x[h].x += std_noise * sqrt(-2 * log(u)) * cos(2 * M_PI * v);
      h_x[h].y += std_noise * sqrt(-2 * log(u)) * sin(2 * M_PI * v);
      noise += -2 * log(u);
    }
    const double snr = sig / (noise * std_noise * std_noise);
    printf("SNR = %g / %.2f dB \n\n", snr, 10 * log10(snr));
  }
  // truth spectrum (cuExperiment.cu:103-111)
  (void)hipMemcpy(a, h_x.data(), (size_t)n * 16, hipMemcpyHostToDevice);
  (void)fft_plain(a, b, scr, ln, 1, 0, 0, nullptr);
  (void)hipMemcpy(x_f.data(), b, (size_t)n * 16, hipMemcpyDeviceToHost);
  for (int i = 0; i < n; ++i) {
    x_f[i].x /= n;
    x_f[i].y /= n;
  }
  (void)hipFree(b);
  (void)hipFree(scr);
  complex_t* d_x = (complex_t*)a;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, nullptr);
  (void)hipMemcpy(d_x, h_x.data(), (size_t)n * 16, hipMemcpyHostToDevice);
  (void)hipEventRecord(e1, nullptr);
  (void)hipEventSynchronize(e1);
  float t_H_to_D = 0;
  (void)hipEventElapsedTime(&t_H_to_D, e0, e1);
  printf("\n computing filter starts\n");
  make_multiple_t(lobefrac, tolerance, n, b_f, d_x, B, B_thresh, loops_loc, loops_thresh,
                  loops_est, W_RLFC, RLFC_loops, repetitions, LARGE_FREQ.data(), k, x_f.data(),
                  t_H_to_D);
  (void)hipFree(a);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}
