/*
 * sfft_oracle.h -- CPU oracle for the sparse-FFT hot path (outer_loop and its six stages).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a sequential, deterministic restatement in plain C of
 * the reference semantics of pcdslab/gpu-sfft (outerLoop.cu / innerLoop.cu /
 * cuExperiment.cu / experiment.c / utils.c).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or as the timed CPU
 * baseline.  The product library never links or calls it.
 *
 * Floating-point rules: built with -ffp-contract=off; every fused multiply-add is an explicit
 * fma()/fmaf() call so the GPU kernels can reproduce the exact rounding sequence.
 *
 * Parity status per function is stated next to each declaration; "pinned by" names the
 * independent check in tests/ that would fail on a plausible mistake.
 */
#ifndef SFFT_ORACLE_H
#define SFFT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y; } orc_c64;   /* layout of complex_t, utils.h:32-35 */
typedef struct { float x, y; } orc_c32;
typedef struct { unsigned int key; orc_c64 value; } orc_node; /* Node, utils.h:43-46 (24 B) */

/* ---- parameters (experiment.c:54-84, cuExperiment.cu:279-284) -------------------------
 * pinned by: README sample values (README.md:84-90) and closed forms in test_oracle_params. */
typedef struct {
  int q, n, k;
  double Bcst;
  int loops_loc, loops_thresh, loops_est, loops;
  double RLFC_cst;
  int RLFC_loops;
  double tolerance;
  double BB, lobefrac;
  int b_f, B, B_thresh, W_RLFC, w;
  double t0;
  int assert_ok;   /* 1 if every assert of experiment.c:77,82-84 and cuExperiment.cu:283 holds */
} orc_params;

void orc_derive_params(int q, int k, double Bcst, int loops_loc, int loops_thresh, int loops_est,
                       int RLFC_cst, int RLFC_loops, double tolerance, orc_params *out);
int orc_floor_to_pow2(double x);             /* utils.c:180-185 */
int orc_gcd(int a, int b);                   /* utils.c:81-84 */
int orc_mod_inverse(int a, int n);           /* utils.c:92-105 */
int orc_timesmod(int x, int a, int n);       /* innerLoop.cu:647-649 */

/* ---- canonical power-of-two FFT -------------------------------------------------------
 * The exact operation sequence the HIP FFT kernels execute (multi-pass big-radix Stockham,
 * radix-4/2 sub-stages, two-level twiddle tables).  Replaces cuFFT (innerLoop.cu:144-152,
 * 252-256); parity with cuFFT bits is unpinnable (cuFFT absent).
 * pinned by: numpy.fft at every size 2^0..2^20 (rel err < 1e-13), impulses/tones closed
 * forms, inverse(forward(x)) = N x. */
int orc_fft_plan(int log2N, int *pass_logs);            /* returns number of passes */
void orc_twiddle64(uint64_t e, uint64_t T, int inverse, orc_c64 *w);  /* exp(-+2 pi i e/T) */
void orc_fft64(orc_c64 *data, int log2N, int inverse, orc_c64 *scratch);
void orc_fft32(orc_c32 *data, int log2N, int inverse, orc_c32 *scratch);
/* arbitrary-length DFT (Bluestein over orc_fft64) for filter design; pinned by numpy.fft. */
void orc_dft_any64(orc_c64 *data, int N, int inverse);

/* ---- filter design: make_multiple_t (cuExperiment.cu:248-447) ---------------------------
 * Dolph-Chebyshev window (cuExperiment.cu:279-328, 538-552, 593-601, NaN->0 quirk kept),
 * circular shift (cuShift :507-522), fold with the dropped last tap f[w-1] (:336-352),
 * sequential boxcar running sum (:373-396), normalisation (:400-403), sequential
 * phase-ramp recurrence (:407-421), inverse FFT and /n (:424-447).
 * filter_time[w], filter_freq[n].  pinned by: closed-form Chebyshev samples, brute-force
 * O(w^2) DFT of the window at small w, max|filter_freq| = 1, passband flatness, and the
 * time/frequency pair identity IDFT(filter_freq)[0..w)/n = filter_time. */
int orc_filter_window(double lobefrac, double tolerance, int n, int *w_out, double *t0_out);
void orc_make_multiple_t(double lobefrac, double tolerance, int n, int b_f, int w,
                         orc_c64 *filter_time, orc_c64 *filter_freq);
/* same, with optional intermediate taps: the Chebyshev samples (w), their DFT (w) and the
 * n-point spectrum of the folded window (n); NULL skips a tap. */
void orc_make_multiple_t_taps(double lobefrac, double tolerance, int n, int b_f, int w,
                              orc_c64 *filter_time, orc_c64 *filter_freq, orc_c64 *tap_v,
                              orc_c64 *tap_V, orc_c64 *tap_G);

/* ---- stage functions -------------------------------------------------------------------- */
/* fold modes: 0 = reference (n < 2^27: tiled fold + faulty tail, innerLoop.cu:725-757;
 * n >= 2^27: every tap, innerLoop.cu:760-773, without the race), 1 = every tap. */
enum { ORC_FOLD_REFERENCE = 0, ORC_FOLD_ALL_TAPS = 1 };

/* bins_t[r] = sum over taps i = r (mod B) of x[(i*ai) & (n-1)] * f[i]  (innerLoop.cu:194-246).
 * pinned by: brute force over exactly-representable integer inputs (order-free sums) in
 * test_oracle_stages, including the tiled fold's tail taps at n < 2^27. */
void orc_perm_filter64(const orc_c64 *x, int n, const orc_c64 *f, int w, int B, int ai,
                       int fold_mode, orc_c64 *bins_t);
/* FP32 path: complex64 input/taps, Kahan-compensated per-bucket fold. */
void orc_perm_filter32(const orc_c32 *x, int n, const orc_c32 *f, int w, int B, int ai,
                       int fold_mode, orc_c32 *bins_t);

/* |X|^2 = re*re + im*im, each product rounded (samplesKernel, innerLoop.cu:708-718). */
void orc_samples64(const orc_c64 *X, int N, double *s);
void orc_samples32(const orc_c32 *X, int N, float *s);
/* cutoff (innerLoop.cu:551-642): the K indices with largest s, ties by lower index,
 * written ascending.  pinned by: brute-force argsort with (-s, index) keys incl. forced ties. */
void orc_cutoff64(const double *s, int N, int K, int *out);
void orc_cutoff32(const float *s, int N, int K, int *out);

/* RLFC comb restriction (innerLoop.cu:112-177 + outerLoop.cu:104-121): y[i] = x[offset+i*sigma],
 * Y = FFT_W(y), top-B_thresh residues; the merge sorts and de-duplicates across loops.
 * pinned by: aliasing identity FFT_W(x[off+i sigma])[g] = sum_{f=g mod W} X[f] e^{2 pi i f off/n}
 * against numpy on small n, plus the cutoff pins. */
void orc_rlfc64(const orc_c64 *x, int n, int W, int offset, int B_thresh, int *J2_out);
int orc_merge_J2(int *J2, int count);   /* radix_sort_t + dedup; returns num_RLFC */

/* reverse hash + voting (innerLoop.cu:300-384, 783-817, 901-920).
 * J: B_thresh sorted bucket indices.  vote[n] (int32) accumulates; hits appended in the
 * sequential order (J ascending, walk order).  Returns the new hit count, or -1 when the hit
 * list would exceed capacity.  pinned by: brute-force enumeration of approved
 * positions in each window with findUpperBound's last-index rule, votes by dictionary
 * (test_oracle_stages). */
int orc_reverse_hash(const int *J, int B_thresh, int n, int B, int a, int ai, int loop_thresh,
                     int *vote, int *hits, int hit_count, int hit_capacity,
                     const int *J2, int num_RLFC, int W);

/* estimation (innerLoop.cu:396-466, 825-870, 924-937): lower median over all loops,
 * real and imaginary independently.  filter_freq is read at index (n-d) mod n only; the
 * caller passes a window of m = n/B entries laid out as win[(idx + m/2) mod n] for
 * idx in [n-m/2, n) U [0, m/2), i.e. win[j] = filter_freq[(j - m/2) mod n].
 * pinned by: 1-sparse exact signals (closed form) and median by full sort. */
void orc_estimate_values(const int *hits, int I_F, const orc_c64 *bins_f, int loops, int n,
                         const int *ai, int B, const orc_c64 *filter_win, orc_node *out);

/* the whole transform with explicit draws (outerLoop.cu:65-190).  a[l], ai[l] per loop,
 * offsets[r] per RLFC loop.  Returns I_F (>= 0) or -1 on hit overflow.  bins_f_out (optional,
 * loops*B) receives the bucket spectra; J_out (optional, loops_loc*B_thresh) the cutoffs;
 * J2_out (optional, RLFC_loops*B_thresh) the merged restriction list (first num_RLFC valid).
 * stage_seconds (optional, 7 entries) receives the reference's timing table columns:
 * votetable, restr-loc-index, perm+filter, reverse-hash, estimation, FFT+Cutoff, total. */
typedef struct {
  int n, B, B_thresh, W_RLFC, RLFC_loops, loops_thresh, loops_loc, loops, w;
  int fold_mode;
  int hit_capacity;
} orc_plan;

int orc_outer_loop64(const orc_plan *p, const orc_c64 *x, const orc_c64 *filter_time,
                     const orc_c64 *filter_win, const int *a, const int *ai, const int *offsets,
                     int *hits_out, orc_node *nodes_out, orc_c64 *bins_f_out, int *J_out,
                     int *J2_out, int *num_RLFC_out, double *stage_seconds);

/* FP32 path (the project's mixed-precision addition, not in the reference): complex64 x and
 * taps, FP32 comb/bucket FFTs and keys, Kahan-compensated fold, FP64 filter division and
 * medians on the widened bins.  pinned by: the FP32 stage pins above plus error bounds against
 * the FP64 path in test_oracle_fp32. */
void orc_rlfc32(const orc_c32 *x, int n, int W, int offset, int B_thresh, int *J2_out);
void orc_estimate_values32(const int *hits, int I_F, const orc_c32 *bins_f, int loops, int n,
                           const int *ai, int B, const orc_c64 *filter_win, orc_node *out);
int orc_outer_loop32(const orc_plan *p, const orc_c32 *x, const orc_c32 *filter_time,
                     const orc_c64 *filter_win, const int *a, const int *ai, const int *offsets,
                     int *hits_out, orc_node *nodes_out, orc_c32 *bins_f_out, int *J_out,
                     int *J2_out, int *num_RLFC_out);

/* ---- value map and its adjoint (project addition; see sfft_oracle.c) ---------------------
 * sel[2h], sel[2h+1]: the loops whose values the real / imaginary lower medians of hit h
 * returned (stable-rank order, as the reference's insertion sort orders ties).
 * value_apply: the estimated values of the fixed hits as a function of x.
 * value_adjoint: its transpose; g[h] = dL/dRe yhat_h + i dL/dIm yhat_h, gx[n] is overwritten.
 * pinned by: apply == transform node values; <A x, y> = <x, A^T y>; dense matrix on tiny n. */
void orc_estimate_select(const int *hits, int I_F, const orc_c64 *bins_f, int loops, int n,
                         const int *ai, int B, const orc_c64 *filter_win, unsigned char *sel);
void orc_value_apply64(const orc_plan *p, const orc_c64 *x, const orc_c64 *f,
                       const orc_c64 *win, const int *ai, const int *hits, int nh,
                       const unsigned char *sel, orc_c64 *out);
void orc_value_adjoint64(const orc_plan *p, const orc_c64 *f, const orc_c64 *win,
                         const int *ai, const int *hits, int nh, const unsigned char *sel,
                         const orc_c64 *g, orc_c64 *gx);

/* ---- harness metric (run_experiment, cuExperiment.cu:181-221) ---------------------------
 * qsort by |value| descending (comp_struct, utils.c:231-235), keep the first k in a dense
 * n-vector, MISSED = (k - FOUND, k - large_found) with FOUND by key search over all I_F
 * nodes, ERROR = sum over all n of |yhat - x_f|^2 ("MSE*k").  Dense version for small n.
 * nodes is sorted in place (as the reference does).
 * pinned by: hand-built cases with known answers in test_oracle_metric. */
void orc_metric_dense(orc_node *nodes, int I_F, int k, const int *large_freq, int n,
                      const orc_c64 *x_f, int *missed_est, int *missed_res, double *err_total);

#ifdef __cplusplus
}
#endif
#endif
