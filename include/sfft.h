/*
 * sfft.h -- C ABI of libsfft, an MI355X-native sparse FFT (gfx950, hand-written HIP kernels).
 *
 * Two layers live behind this header.
 *
 * 1. The reference's own entry points, with the reference's names, argument meaning and POD
 *    layouts, so a caller of pcdslab/gpu-sfft links against libsfft.so unchanged.  Each one
 *    cites the declaration it replaces.  Device pointers are HIP device memory.
 *
 * 2. A native plan API (sfft_plan_*): explicit stream, plan-owned workspace, explicit
 *    permutation/offset draws, status codes instead of exit()/assert().
 *
 * No function here takes a torch type; every pointer is a plain host or device address.
 */
#ifndef SFFT_H
#define SFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- POD types, byte-identical to the reference (utils.h:32-51) -------------------------- */
typedef struct double2C {
  double x;
  double y;
} complex_t;                                   /* utils.h:32-35, 16 B, = hipDoubleComplex */

typedef int bool_t;                            /* utils.h:37 */
#ifndef MAXSIZE_HITS
#define MAXSIZE_HITS 524288                    /* utils.h:40 */
#endif

typedef struct nodes {
  unsigned int key;
  complex_t value;
} Node;                                        /* utils.h:42-45, 24 B (4 B padding) */

typedef struct pairs {
  int first;
  int second;
} Pair;                                        /* utils.h:48-51 */

/* globals of outerLoop.cu:49-51 (declared in outerLoop.h:29-31) */
extern bool_t ALGORITHM1;
extern bool_t VERBOSE;
extern bool_t TIMING;

/* ---- reference entry points ------------------------------------------------------------- */

/* outerLoop.h:35-38.  d_x: device complex_t[n].  d_filter_time: device complex_t[w].
 * d_filter_freq: complex_t[n], host-readable or device (only the m = n/B entries around 0
 * are read).  Returns a malloc'd host Node[*I_F] (caller frees), NULL on failure.
 * Permutations and RLFC offsets are drawn with glibc's random() generator (library-private
 * state, see sfft_srandom) and libc drand48() in the reference's order (outerLoop.cu:128-134,
 * innerLoop.cu:122-123). */
Node *outer_loop(complex_t *d_x, int n, complex_t *d_filter_time, complex_t *d_filter_freq,
                 int d_filter_size, int B, int B_thresh, int W_RLFC, int RLFC_loops,
                 int loop_threshold, int loops_loc, int loops, int *I_F, float t_H_to_D);

/* innerLoop.h:32-33.  Writes B_thresh ascending residues to device int d_J2[B_thresh];
 * adds the elapsed seconds to *RLFC_Time. */
int restr_loc_freq_coeff(complex_t *d_x, int n, int B_thresh, int W_RLFC, int *d_J2,
                         double *RLFC_Time);

/* innerLoop.h:35-36.  Top-B of |d_samp|^2 over samp_size entries, ascending indices into
 * device int d_output[B]; adds seconds to *CTime. */
int samples_indices(complex_t *d_samp, int B, int samp_size, int *d_output, double *CTime);

/* innerLoop.h:38-41.  One loop: permute+filter into B buckets, FFT_B into device
 * d_bins_f[B], top-B_thresh bucket indices into device d_J[B_thresh]. */
int perm_filter_cutoff(complex_t *d_x, int n, complex_t *d_filter_time, int d_filter_size,
                       int B_thresh, int B, int a, int ai, complex_t *d_bins_f, int *d_J,
                       double *PF_T, double *BC_T);

/* innerLoop.h:43-45.  Votes for one location loop into device int d_vote[n]; appends
 * locations that reach loop_thresh to device d_I; *I_F is the host in/out hit count.
 * d_J2 (device or host-readable, num_RLFC sorted distinct residues). */
int reverse_hash(int *d_J, int n, int B_thresh, int B, int a, int ai, int loop_thresh,
                 int *d_vote, int *d_I, int *I_F, double *G_T, int *d_J2, int num_RLFC,
                 int W_RLFC, int RLFC_loops);

/* innerLoop.h:47-48.  Medians over `loops` bucket rows; returns malloc'd host Node[*I_F]. */
Node *estimate_values(int *d_I, int *I_F, complex_t *d_bins_f, int loops, int n, int *permute,
                      int B, complex_t *d_filter_freq, int loops_loc);

/* innerLoop.h:50 */
int timesmod(const int x, const int a, const int n);

/* No counterpart in the reference: samples_indices (above) for callers that cannot allocate
 * device memory (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, result and synchronisation as samples_indices, but the |X|^2
 * keys and the select workspace live in a file-scope device array of 12 MiB that each device
 * holds until the process ends; no hipMalloc/hipFree.  samp_size may be at most
 * SFFT_NOALLOC_MAX_SAMPLES: larger sizes return -2 (and print why to stderr), other argument
 * errors -1 as samples_indices.  Every call on one device shares that device's array, so
 * calls that overlap on one GPU (from several host threads) race: serialise them. */
#define SFFT_NOALLOC_MAX_SAMPLES 262144
int samples_indices_noalloc(complex_t *d_samp, int B, int samp_size, int *d_output,
                            double *CTime);

/* outer_loop draws its permutations a[l] with glibc's random() generator in the reference's
 * order (outerLoop.cu:129-134) -- on a library-private state, because the HIP runtime calls
 * random() from its own threads and would make the process-wide sequence timing-dependent.
 * The private state starts as a program that never called srandom() (seed 1), so the
 * reference's command line gets the reference's draws; these reseed it and draw from it
 * (the counterparts of srandom()/random() for callers that set or replay the draws). */
void sfft_srandom(unsigned int seed);
long sfft_random(void);

/* filters.h:44-48.  Designs the filter (FP64) for d_x's size and runs the experiment
 * (run_experiment) exactly as the reference's harness does. */
void make_multiple_t(double lobefrac, double tolerance, int n, int b_f, complex_t *d_x, int B,
                     int B_thresh, int loops_loc, int loops_thresh, int loops_est, int W_Comb,
                     int Comb_loops, int repetitions, int *LARGE_FREQ, int k, complex_t *x_f,
                     float t_H_to_D);

/* experiment.h:32-46 */
int cudafft_experiment(int n, double lobefrac, double tolerance, int b_f, int B, int B_thresh,
                       int loops_loc, int loops_thresh, int loops_est, int W_RLFC,
                       int RLFC_loops, int repetitions, int k, double std_noise);
void run_experiment(complex_t *d_x, int n, complex_t *d_filter_time, complex_t *d_filter_freq,
                    int d_filter_size, int B, int B_thresh, int loops_loc, int loops_thresh,
                    int loops_est, int W_RLFC, int RLFC_loops, int repetitions,
                    int *LARGE_FREQ, int k, complex_t *x_f, float t_H_to_D);
void cuShift(complex_t *d_x, int n, int r);

/* ---- native plan API --------------------------------------------------------------------- */
typedef enum {
  SFFT_OK = 0,
  SFFT_ERR_INVALID = 1,        /* parameters violate the reference's asserts or ABI limits */
  SFFT_ERR_ALLOC = 2,
  SFFT_ERR_HIP = 3,
  SFFT_ERR_HIT_OVERFLOW = 4,   /* more located indices than hit_capacity (reference: UB) */
  SFFT_ERR_NO_FILTER = 5,
  SFFT_ERR_NO_DEVICE = 6
} sfft_status;

typedef enum { SFFT_FP64 = 0, SFFT_FP32 = 1 } sfft_precision;

typedef enum {
  SFFT_FOLD_REFERENCE = 0,     /* n < 2^27: tiled fold + the reference's tail taps */
  SFFT_FOLD_ALL_TAPS = 1       /* every tap i < w into bucket i mod B */
} sfft_fold_mode;

typedef struct {
  int n;              /* signal length, power of two, 2^4 .. 2^30 */
  int B;              /* buckets, power of two */
  int B_thresh;       /* cutoff size (2k) */
  int W_RLFC;         /* comb length, power of two */
  int RLFC_loops;     /* comb repetitions */
  int loops_thresh;   /* votes needed */
  int loops_loc;      /* location loops */
  int loops;          /* loops_loc + loops_est */
  int w;              /* filter taps */
  int fold_mode;      /* sfft_fold_mode */
  int hit_capacity;   /* 0 -> MAXSIZE_HITS */
  int front_end;      /* sfft_front_end: how the taps and comb samples are read from x */
} sfft_params;

/* Every front end produces bit-identical results.  GATHER reads x[(i*ai) mod n] per tap in
 * bucket order and x[off + j*sigma] per comb sample (random HBM reads).  STREAM reads x once
 * in address order and scatters each used element to a tap-ordered staging array (loops*w
 * entries) and the comb arrays, which the fold then reads contiguously; the scatter is as
 * random as the gathers it replaces (one partial cache line per used element).  COLUMNS folds
 * in column order: all taps of bucket r lie in column (r*ai) mod B of x viewed as n/B rows of
 * B columns, so the fold walks the columns in blocks of one 128-byte line width for every loop
 * at once and the taps of different loops that fall in the same line share its fetch; the
 * filter taps are read from a bucket-major copy built when the filter is set (B*(w/B+1)
 * entries in the plan workspace).  AUTO picks COLUMNS wherever the fold can stream the
 * 256-byte-wide column slabs through LDS (both precisions at n = 2^27, k = 16000), else GATHER;
 * sfft_plan_front_end reports the choice. */
typedef enum {
  SFFT_FRONT_AUTO = 0,
  SFFT_FRONT_GATHER = 1,
  SFFT_FRONT_STREAM = 2,
  SFFT_FRONT_COLUMNS = 3
} sfft_front_end;

typedef struct sfft_plan sfft_plan;

/* draws for one transform: a[l] odd, ai[l] = a[l]^-1 mod n (l < loops), comb offsets
 * offsets[r] in [0, n/W_RLFC) (r < RLFC_loops).  Host arrays. */
typedef struct {
  const int *a;
  const int *ai;
  const int *offsets;
} sfft_draws;

const char *sfft_status_string(sfft_status s);
/* validates against the reference's asserts (experiment.c:77,82-84; cuExperiment.cu:283) */
sfft_status sfft_params_validate(const sfft_params *p);
/* stream: a hipStream_t (NULL = default stream) */
sfft_status sfft_plan_create(const sfft_params *p, sfft_precision prec, void *stream,
                             sfft_plan **out);
void sfft_plan_destroy(sfft_plan *plan);
size_t sfft_plan_workspace_bytes(const sfft_plan *plan);

/* filter: filter_time device pointer of w taps (complex128 for SFFT_FP64, complex64 for
 * SFFT_FP32; the plan keeps the pointer, caller keeps ownership), and the m = n/B FP64
 * frequency-response entries win[j] = filter_freq[(j - m/2) mod n] (host pointer, copied). */
sfft_status sfft_plan_set_filter(sfft_plan *plan, const void *d_filter_time,
                                 const complex_t *h_filter_win);

/* One transform of device signal d_x (complex128[n] or complex64[n]).  Results stay on the
 * device: nodes in d_nodes (Node[hit_capacity], device) and the count in *d_count (device
 * int).  Asynchronous on the plan's stream. */
sfft_status sfft_execute_async(sfft_plan *plan, const void *d_x, const sfft_draws *draws,
                               Node *d_nodes, int *d_count);
/* Same, with the count and nodes in host memory (h_nodes has hit_capacity entries), then
 * synchronises the stream.  When h_nodes is pinned host memory (hipHostMalloc, or memory the
 * HIP runtime reports as pinned) the estimation kernel stores the nodes and the count into it
 * directly; otherwise they are copied after the transform.  Returns SFFT_ERR_HIT_OVERFLOW
 * when the hits exceed capacity. */
sfft_status sfft_execute(sfft_plan *plan, const void *d_x, const sfft_draws *draws,
                         Node *h_nodes, int *h_count);

/* Graph mode (off by default): the transform's kernel sequence is captured into a hipGraph
 * on first use and replayed by later sfft_execute[_async] calls while the signal and result
 * pointers stay the same (re-captured when they change); the draws are uploaded before each
 * replay.  A plan created with stream NULL gets a stream of its own for this.  Profiling
 * (sfft_plan_set_profiling) runs the transform eagerly. */
sfft_status sfft_plan_set_graph(sfft_plan *plan, int on);

/* Stage timing with hipEvents on the plan stream (off by default).  After a completed
 * transform, sfft_plan_stage_times writes seconds in the reference's table order
 * (outerLoop.cu:174-187): votetable, restr-loc-index, perm+filter, reverse-hash,
 * estimation, FFT+Cutoff, total. */
sfft_status sfft_plan_set_profiling(sfft_plan *plan, int on);
sfft_status sfft_plan_stage_times(sfft_plan *plan, double *seconds7);
/* raw intervals: clear, streaming pass (0 with gathers), comb FFT + cutoff + bitmaps,
 * permute+filter (or the tap fold), bucket FFT + cutoff, reverse hash, estimation.  With the
 * column front end the slab fold does the clear and the comb copy: clear and permute+filter
 * read 0 and the streaming pass interval is the fold. */
sfft_status sfft_plan_interval_times(sfft_plan *plan, double *seconds7);

/* Debug/parity taps into the last transform (device pointers owned by the plan):
 * 0 bins_f (loops x B complex), 1 J (loops_loc x B_thresh int), 2 J2 lists (RLFC_loops x
 * B_thresh int), 3 hits (int, count from the last execute), 4 bins_t (loops x B complex),
 * 5 nodes (Node[hit_capacity]), 6 count (int), 7 sel (2 unsigned char per hit: the loops
 * whose values the real and the imaginary lower medians returned), 8 comb keys (RLFC_loops x
 * W_RLFC bit patterns of |Y|^2: uint64 for SFFT_FP64, uint32 for SFFT_FP32), in the storage
 * order sfft_plan_comb_key_order reports. */
const void *sfft_plan_debug_buffer(const sfft_plan *plan, int which);
/* No counterpart in the reference (a debugging aid for the comb keys above): position s of a
 * comb loop's keys holds bin (s mod 2^lo) | ((s >> lo) mod 2^mid) << (log2 W_RLFC - mid) |
 * (s >> (lo + mid)) << lo; mid = 0 means natural order. */
sfft_status sfft_plan_comb_key_order(const sfft_plan *plan, int *lo, int *mid);
/* No counterpart in the reference (an inspection aid): the front end the plan runs, with
 * SFFT_FRONT_AUTO resolved (SFFT_FRONT_GATHER, _STREAM or _COLUMNS; -1 for a NULL plan). */
int sfft_plan_front_end(const sfft_plan *plan);

/* ---- the value map and its adjoint (FP64 plans) ----------------------------------------
 * No counterpart in the reference: the gradient of the estimated values (innerLoop.cu:
 * 396-466) with respect to the signal.  With nh located indices d_hits and their
 * median-selected loops d_sel (2 per hit, as debug buffer 7 holds them after a transform)
 * held fixed, each estimated value is a real-linear function of the signal:
 *   values[h] = Re v_{sel[2h]}(hits[h]) + i Im v_{sel[2h+1]}(hits[h]),
 *   v_l(I) = FFT_B(permute+filter of x, loop l)[bucket of I] * conj(F) / |F|^2.
 * sfft_value_apply evaluates it for any device signal d_x (complex128[n]); applied to the
 * transform's own signal it returns the transform's node values bit for bit.
 * sfft_value_adjoint applies its transpose: d_grad[h] = dL/dRe values[h] + i dL/dIm values[h]
 * (complex128[nh]) -> d_grad_x[m] = dL/dRe x[m] + i dL/dIm x[m] (complex128[n], overwritten).
 * draws: the transform's draws (a, ai used).  Deterministic (no floating-point atomics).
 * Both reuse the plan's bucket workspace (debug buffers 0 and 4 are overwritten).  They need
 * n/B <= 8192 and return SFFT_ERR_INVALID for FP32 plans.  Asynchronous on the plan stream. */
sfft_status sfft_value_apply(sfft_plan *plan, const void *d_x, const sfft_draws *draws, int nh,
                             const int *d_hits, const unsigned char *d_sel, complex_t *d_values);
sfft_status sfft_value_adjoint(sfft_plan *plan, const complex_t *d_grad, const sfft_draws *draws,
                               int nh, const int *d_hits, const unsigned char *d_sel,
                               complex_t *d_grad_x);

/* ---- batches of uneven signals -----------------------------------------------------------
 * No counterpart in the reference (which transforms one signal per outer_loop call): count
 * signals, each with its own parameters (n, k, B, ... may all differ), transformed together.
 * The permute+filter of all signals runs as one persistent launch (the gathering front end:
 * params[i].front_end must be SFFT_FRONT_AUTO or SFFT_FRONT_GATHER); every signal's other
 * stages run on worker streams; with sfft_batch_set_graph(b, 1) each worker stream's launch
 * sequence is captured into hipGraphs once and replayed (re-captured when the signal
 * pointers, filters or shapes change).  Each signal's results are bit-identical to
 * sfft_execute of the same signal with the same plan parameters and draws.  stream NULL:
 * the batch makes its own. */
typedef struct sfft_batch sfft_batch;
sfft_status sfft_batch_create(const sfft_params *params, int count, sfft_precision prec,
                              void *stream, sfft_batch **out);
void sfft_batch_destroy(sfft_batch *b);
/* the plan of signal i (its debug buffers hold signal i's last transform) */
sfft_plan *sfft_batch_plan(sfft_batch *b, int i);
sfft_status sfft_batch_set_filter(sfft_batch *b, int i, const void *d_filter_time,
                                  const complex_t *h_filter_win);
sfft_status sfft_batch_set_graph(sfft_batch *b, int on);
/* d_x[count] device signals, draws[count]; results stay on the device (per-signal plan
 * buffers; counts in sfft_batch_counts_device) */
sfft_status sfft_batch_execute_async(sfft_batch *b, const void *const *d_x,
                                     const sfft_draws *draws);
/* same, then h_counts[count] and, where h_nodes[i] is not NULL, the nodes of signal i
 * (h_nodes[i] holds hit_capacity entries); synchronous */
sfft_status sfft_batch_execute(sfft_batch *b, const void *const *d_x, const sfft_draws *draws,
                               Node *const *h_nodes, int *h_counts);
const int *sfft_batch_counts_device(const sfft_batch *b);
/* Autotuner: times bit-equivalent launch shapes on this batch (reps runs each) and keeps the
 * fastest: permute+filter loops per thread {1..9} x block {128, 256} x persistent grid
 * {1024, 2048, 4096}, and FFT columns per tile {1..32} of the bucket and comb FFTs.
 * choice4 (optional): lpt, block, grid, bucket columns | comb columns << 8. */
sfft_status sfft_batch_tune(sfft_batch *b, const void *const *d_x, const sfft_draws *draws,
                            int reps, int *choice4);
/* sets a shape explicitly (lpt 1..9, block 128/256, grid >= 1, columns 0..32, 0 = default) */
sfft_status sfft_batch_set_shape(sfft_batch *b, int lpt, int block, int grid, int qmax_bucket,
                                 int qmax_comb);

/* ---- canonical FFT on the device (the same kernels the transform uses) ------------------ */
/* In-place-or-not batched power-of-two FFT: count transforms of 2^log2n points, contiguous,
 * complex128 (SFFT_FP64) or complex64 (SFFT_FP32).  Unnormalised, forward = exp(-2 pi i). */
sfft_status sfft_fft_device(const void *d_in, void *d_out, int log2n, int count, int inverse,
                            sfft_precision prec, void *stream);

/* canonical twiddle exp(-+2 pi i e / T) (T power of two), as used by every FFT table */
void sfft_twiddle(uint64_t e, uint64_t T, int inverse, double *re, double *im);
/* pass split of a 2^log2n FFT; returns the number of passes */
int sfft_fft_plan_passes(int log2n, int *pass_logs);

/* arbitrary-length FP64 DFT in place (Bluestein over the canonical FFT; cuFFT's role for the
 * odd window lengths of cuExperiment.cu:305-308).  Synchronous. */
sfft_status sfft_dft_any_device(complex_t *d_data, int N, int inverse, void *stream);

/* ---- FP64 filter design (make_multiple_t without the experiment) ------------------------- */
/* taps w of the Dolph-Chebyshev window (cuExperiment.cu:279-284); *t0 optional */
int sfft_filter_taps(double lobefrac, double tolerance, double *t0);
/* filter time taps into device d_filter_time[w] and the n-entry frequency response into host
 * h_filter_freq[n] (may be NULL), exactly the computation of cuExperiment.cu:279-447
 * (including the reference's NaN->0 Chebyshev rule, the dropped fold tap and the sequential
 * boxcar and phase-ramp recurrences).  Synchronous. */
sfft_status sfft_make_filter(double lobefrac, double tolerance, int n, int b_f, int w,
                             complex_t *d_filter_time, complex_t *h_filter_freq, void *stream);
/* design modes: FAITHFUL is the computation above (host boxcar and phase-ramp recurrences, bit
 * for bit the CPU restatement); PARALLEL computes the same filter with every n-length step on
 * the device (Chebyshev samples, Bluestein chirp, boxcar from panel sums, max, phase ramp with
 * the angle (w/2) i mod n reduced exactly) -- its rounding differs from FAITHFUL's and does not
 * drift with n.  PARALLEL needs n >= 4096. */
typedef enum { SFFT_DESIGN_FAITHFUL = 0, SFFT_DESIGN_PARALLEL = 1 } sfft_design_mode;
/* sfft_make_filter with a mode and an optional device copy of the response d_filter_freq[n]
 * (h_filter_freq and d_filter_freq may each be NULL; with both NULL the PARALLEL mode moves no
 * n-length array to the host).  Replaces make_multiple_t's design (cuExperiment.cu:279-447).
 * Synchronous. */
sfft_status sfft_make_filter_ex(double lobefrac, double tolerance, int n, int b_f, int w,
                                int mode, complex_t *d_filter_time, complex_t *h_filter_freq,
                                complex_t *d_filter_freq, void *stream);
/* win[j] = filter_freq[(j - m/2) mod n], j < m = n/B: the entries estimation reads */
void sfft_filter_window(const complex_t *filter_freq, int n, int B, complex_t *win);
/* the same window read from a device response (two contiguous copies).  Synchronous. */
sfft_status sfft_filter_window_device(const complex_t *d_filter_freq, int n, int B,
                                      complex_t *h_win, void *stream);
sfft_status sfft_convert_c128_to_c64(const void *d_in, void *d_out, size_t count, void *stream);
sfft_status sfft_convert_c64_to_c128(const void *d_in, void *d_out, size_t count, void *stream);

/* ---- synthetic signals and the experiment metric (cuExperiment.cu:64-137, 181-221) ------ */
/* k distinct locations (ascending) and unit-magnitude uniform-phase values from a seed */
sfft_status sfft_signal_spectrum(int n, int k, uint64_t seed, int *h_locs, complex_t *h_vals);
/* x = unnormalised inverse DFT of the sparse spectrum, plus complex Gaussian noise of
 * per-component std sqrt(k / (2 * 10^(snr_db/10))) when snr_db > 0; x is complex128[n] or,
 * with x_fp32, complex64[n].  Synchronous. */
sfft_status sfft_signal_synthesize(int n, int k, const int *h_locs, const complex_t *h_vals,
                                   double snr_db, uint64_t noise_seed, int x_fp32, void *d_x,
                                   void *stream);
/* X = FFT_n(x) / n in FP64 (device complex128[n]).  Synchronous. */
sfft_status sfft_truth_spectrum(const void *d_x, int n, int x_fp32, complex_t *d_X, void *stream);
/* run_experiment's metric against the truth spectrum d_X: MISSED (estimation, result) and
 * err_total = sum over all n of |yhat - X|^2 ("MSE*k"); err_support and max_abs restricted
 * to the kept support and the true locations.  d_X is restored before returning. */
sfft_status sfft_metric(const Node *nodes, int I_F, int k, const int *locs, complex_t *d_X,
                        int n, int *missed_est, int *missed_res, double *err_total,
                        double *err_support, double *max_abs, void *stream);

/* ---- library-wide ------------------------------------------------------------------------ */
const char *sfft_version(void);
int sfft_device_count(void);

#ifdef __cplusplus
}
#endif
#endif /* SFFT_H */
