// sfft_internal.h -- shared declarations between the HIP kernels (sfft_fft.hip,
// sfft_select.hip, sfft_hash.hip, sfft_adjoint.hip) and the host orchestration
// (sfft_plan.hip, sfft_batch.hip, sfft_compat.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "../../include/sfft.h"

namespace sfft {

// ---- canonical FFT description -----------------------------------------------------------
// Pass split: log2N bits over p = ceil(log2N / 11) passes, the first (log2N mod p) passes
// one bit larger.  Every table value comes from canon_twiddle().
constexpr int kMaxPassLog = 11;
constexpr int kMaxPasses = 4;

int fft_plan_passes(int log2n, int* pass_logs);
void canon_twiddle(uint64_t e, uint64_t T, int inverse, double* c, double* s);

// Device twiddle tables of one FFT size/direction/precision.
struct FftTables {
  int log2n = -1;
  int inverse = 0;
  int is_fp32 = 0;
  int npass = 0;
  int logs[kMaxPasses] = {0, 0, 0, 0};
  const void* twR[kMaxPasses] = {nullptr, nullptr, nullptr, nullptr};   // R entries
  const void* tdir[kMaxPasses] = {nullptr, nullptr, nullptr, nullptr};  // Lcur entries
  const void* thi[kMaxPasses] = {nullptr, nullptr, nullptr, nullptr};   // Lcur >> s
  const void* tlo[kMaxPasses] = {nullptr, nullptr, nullptr, nullptr};   // 1 << s
  const void* tw4[kMaxPasses] = {nullptr, nullptr, nullptr, nullptr};   // radix-4 quads
  int s[kMaxPasses] = {0, 0, 0, 0};
  void* storage = nullptr;  // one device allocation holding every table
  size_t bytes = 0;
};

hipError_t fft_tables_create(int log2n, int inverse, int is_fp32, FftTables* t);
void fft_tables_destroy(FftTables* t);

// The select's first digit of a |X|^2 key (12 bits, a non-decreasing function of the key, so
// that digit > d0 means key > every key of digit d0): FP32 keys, their top 12 bits below the
// sign; FP64 keys, the top 12 bits below the sign of the value rounded to FP32 (8 exponent and
// 4 mantissa bits -- the FP64 pattern's own top 12 bits are its 11 exponent bits and one
// mantissa bit, which leaves a whole binade of candidates below the threshold digit).
__device__ __forceinline__ unsigned sel_first_digit(unsigned key) { return (key >> 19) & 0xFFFu; }
__device__ __forceinline__ unsigned sel_first_digit(unsigned long long key) {
  const float f = __double2float_rn(__longlong_as_double((long long)key));
  return (__float_as_uint(f) >> 19) & 0xFFFu;
}

// kLoadCombRowsG: kLoadCombRows with sibling sub-FFTs grouped per tile (chosen per pass by
// fft_run; callers set kLoadCombRows)
enum FftLoad { kLoadPlain = 0, kLoadComb = 1, kLoadCombRows = 2, kLoadPerm = 3, kLoadCombRowsG = 4 };
enum FftStore { kStoreCplx = 0, kStoreKeys = 1, kStoreBoth = 2 };

// One canonical FFT of `batch` transforms.  Passes ping-pong between out and scratch so that
// the last pass lands in `out` (keys for kStoreKeys/kStoreBoth).  in may equal out only for
// single-pass sizes.  The comb loader reads element j of batch b from
// x[offs[b] + j * sigma] (pass 0); in is then ignored.
struct FftLaunch {
  const FftTables* tab = nullptr;
  const void* in = nullptr;
  long long in_bstride = 0;      // elements
  void* out = nullptr;
  long long out_bstride = 0;
  void* scratch = nullptr;       // batch * N elements, stride N
  void* keys = nullptr;          // u64 (fp64) / u32 (fp32) keys |X|^2, stride keys_bstride
  long long keys_bstride = 0;
  int key_batches = 0;
  int batch = 1;
  int load = kLoadPlain;
  int store = kStoreCplx;
  const void* comb_x = nullptr;
  const int* comb_offs = nullptr;
  int comb_sigma = 1;
  int comb_log2nlo = 0;          // kLoadCombRows: element g of batch b at
  int comb_log2m = 0;            // in[b*in_bstride + (g mod 2^log2nlo) << log2m | g >> log2nlo]
  const int* perm = nullptr;     // kLoadPerm: element g of batch b at
                                 // in[b*in_bstride + (g * perm[b] mod N)] (perm[b] odd)
  const int* perm_inv = nullptr; // kLoadPerm: perm_inv[b] * perm[b] = 1 mod N (the first
                                 // pass reads whole lines of in and regroups by it)
  int qmax = 0;                  // most columns per tile (power of two; 0 = the default)
  int keys_tiled = 0;            // keys in the last pass's tile order (see KeyMap)
  uint16_t* digits = nullptr;    // optional: each key's top 12 bits below the sign bit (the
                                 // select's first digit), same positions and stride as keys
};
hipError_t fft_run(const FftLaunch& L, hipStream_t st);

// Where the keys of an FFT run with keys_tiled land: the last pass of a multi-pass FFT stores
// its tile's keys contiguously (whole-line stores), so storage position s holds the key of
// natural index  (s mod 2^lo) | ((s >> lo) mod 2^mid) << (log2n - mid) | (s >> (lo + mid)) << lo
// (lo: the last pass's columns per tile, mid: its radix bits).  mid == 0: natural order.
struct KeyMap {
  int lo = 0, mid = 0, log2n = 0;
};
__host__ __device__ inline long long key_natural(const KeyMap& m, long long s) {
  if (m.mid == 0) return s;
  const long long c = s & ((1LL << m.lo) - 1);
  const long long kk = (s >> m.lo) & ((1LL << m.mid) - 1);
  const long long t = s >> (m.lo + m.mid);
  return c | (kk << (m.log2n - m.mid)) | (t << m.lo);
}
KeyMap fft_keys_map(const FftLaunch& L);

// ---- permutation + filtering -------------------------------------------------------------
// bins[l*B + r] = canonical fold of loop l (ai_dev[l]) for r < B, l < loops.
hipError_t perm_filter_launch(int is_fp32, const void* x, const void* f, void* bins,
                              const int* ai_dev, int n, int B, int w, int loops, int tiled,
                              hipStream_t st);
// column-ordered fold (same sums as perm_filter_launch): fT is the bucket-major filter copy
// filter_rows_launch writes (B rows of fold_rows_pitch(B, w) taps); a_dev and ai_dev per loop.
// by_column = 1: the slab fold (LDS-staged, the whole signal streamed once; only where
// fold_slab_applies, else hipErrorInvalidValue) writing the bucket sums in column order,
// bins[l*B + c] = sum of bucket (c * a_l mod B), for a bucket FFT that loads with kLoadPerm
// (perm = ai).  by_column = 0: each thread gathers its taps from global memory (remap:
// XCD-aware workgroup order) and writes bucket order.  With yc, the comb samples of the Rl
// offsets offs[] are copied too, column-major: yc[r*W + lo*m + row] =
// x[off_r + (row*(B/sigma) + lo)*sigma] (m = n/B, sigma = n/W; needs sigma <= B).  clr: the
// slab fold (by_column = 1) clears it first; the per-thread fold takes only an empty one.
int fold_rows_pitch(int B, int w);
// whether the LDS-staged slab fold takes this shape (at most 64 comb loops)
int fold_slab_applies(int is_fp32, int n, int B, int w, int loops);
hipError_t filter_rows_launch(int is_fp32, const void* f, void* fT, int B, int w, int tiled,
                              hipStream_t st);
// A clear the slab fold does before its first loads: bm[0, nbm) = 0 and maxf[0, nmax) = all
// ones (a transform's approved-residue bitmaps and largest approved residues, which the comb
// cutoff's bitmap kernel fills later on the same stream).  Empty: nothing is cleared.
struct FoldClear {
  uint32_t* bm = nullptr;
  size_t nbm = 0;
  uint32_t* maxf = nullptr;
  int nmax = 0;
};
hipError_t fold_columns_launch(int is_fp32, const void* x, const void* fT, void* bins,
                               const int* a_dev, const int* ai_dev, int n, int B, int w,
                               int loops, int tiled, int by_column, int remap, void* yc,
                               const int* offs, int Rl, int W, const FoldClear& clr,
                               hipStream_t st);
// same fold, reading the tap values from the staging array T[l*w + i] of the streaming pass
// batch: one persistent launch over the permute+filter of several signals
struct PfJob {
  const void* x;       // signal (complex128 / complex64)
  const void* f;       // taps
  void* bins;          // loops x B folded buckets
  const int* ai;       // device ai[loops]
  uint32_t mask;       // n - 1
  int B, w, loops, tiled;
  int nbx;             // bucket blocks of nt threads
  int item0;           // first work item of this signal
};
// items of one signal for a (loops per thread, threads per block) shape; *nbx = bucket blocks
int pf_batch_items(int B, int loops, int lpt, int nt, int* nbx);
hipError_t perm_filter_batch_launch(int is_fp32, const PfJob* d_jobs, int njobs, int total,
                                    int lpt, int nt, int grid, hipStream_t st);
hipError_t fold_taps_launch(int is_fp32, const void* T, const void* f, void* bins,
                            const int* ai_dev, int n, int B, int w, int loops, int tiled,
                            hipStream_t st);
// streaming pass over x: T[l*w + i] = x[m] for i = m * a_l mod n < w, and the comb samples
// y[r*W + j] = x[off_r + j * n/W]
hipError_t stream_scatter_launch(int is_fp32, const void* x, void* T, void* y, const int* a_dev,
                                 const int* offs, int n, int w, int loops, int Rl, int W,
                                 hipStream_t st);

// ---- top-K select over `segs` key segments of N keys (stride seg_stride) ----------------
struct SelectWork {
  void* state = nullptr;      // per segment: threshold, top digit, counters
  unsigned* hist = nullptr;   // segs * 4096 column sums, then segs * nchunks * 4096 rows
  unsigned* done = nullptr;   // segs (zero between calls)
  int* cnt = nullptr;         // segs * nchunks: selected before each block (ordered mode)
  void* cand_key = nullptr;   // segs * N keys
  int* cand_idx = nullptr;    // segs * N
  int* emit_idx = nullptr;    // segs * N (per-block regions)
  void* blk_key = nullptr;    // segs * N candidate keys in per-block regions
  int* blk_idx = nullptr;     // segs * N
  int* ccnt = nullptr;        // segs * nchunks candidate counts
  int* ecnt = nullptr;        // segs * nchunks counts of digit > d0
  int max_segs = 0;
  int max_chunks = 0;
};
// any N >= 2 (the composite key's index field takes ceil(log2 N) bits); the workspace must
// start zeroed
size_t select_work_bytes(int segs, int N);
void select_work_bind(SelectWork* w, void* base, int segs, int N);
// out[s*K + j]: the K largest keys of segment s (ties by lower index); ascending indices when
// `ordered`, otherwise the same set in arbitrary order.  bits (optional, ordered mode): the
// selection as a bitmap, ceil(N/32) words per segment, every word written.  map (optional,
// unordered mode only): the keys are in that storage order; indices (ties, output) are the
// natural ones
// digits (optional, unordered mode only): the keys' first digits (as the FFT's `digits`
// output), read instead of the keys wherever the first digit is all a pass needs
hipError_t select_topk(int is_fp32, const void* keys, long long seg_stride, int N, int K,
                       int segs, int* out, int ordered, const SelectWork& w, hipStream_t st,
                       unsigned* bits = nullptr, const KeyMap* map = nullptr,
                       const uint16_t* digits = nullptr);

// ---- reverse hash (bitmap of approved residues, deterministic hit list) -----------------
struct HashArgs {
  int n, B, K, W, loops_loc, loops_thresh;
  const int* J;          // loops_loc x K
  const int* J2;         // nJ2 entries (all comb lists, duplicates allowed)
  int nJ2;
  const int* a_dev;      // per loop
  const int* ai_dev;
  unsigned* bitmap;      // loops_loc x words
  int bm_words;
  int* maxfirst;         // loops_loc
  unsigned* jsel;        // loops_loc x jsel_words: the buckets each location loop selected
  int jsel_words;
  int* hitoff;           // loops_loc x K (local prefix within block)
  int* blocksum;         // hit total of each count/write block
  unsigned* hmask;       // hash_mask_bytes(): one hit mask per walk slice
  int* hits;             // capacity
  int capacity;
  int* count;            // device: total hits (may exceed capacity)
};
hipError_t hash_clear_launch(const HashArgs& h, hipStream_t st);
hipError_t hash_bitmap_launch(const HashArgs& h, hipStream_t st);
hipError_t hash_hits_launch(const HashArgs& h, hipStream_t st);
int hash_blocks(int loops_loc, int K, long long m);   // m = n / B
size_t hash_mask_bytes(int loops_loc, int K, long long m);   // hit masks of the slices

// ---- estimation ---------------------------------------------------------------------------
// sel (optional, 2 bytes per hit): the loops whose values the real / imaginary lower medians
// returned (stable-rank order).  out may be device memory or the device view of pinned host
// memory (stored in coalesced 8-byte words either way); count_copy (optional, the same
// kinds of memory) receives *count.
hipError_t estimate_launch(int is_fp32, const int* hits, const int* count, int capacity,
                           const void* bins_f, int loops, int n, const int* ai_dev, int B,
                           const double2* filter_win, Node* out, uint8_t* sel,
                           hipStream_t st, int* count_copy = nullptr);

// ---- value map and its adjoint (FP64) -----------------------------------------------------
// values[h] = Re v_{sel[2h]}(hits[h]) + i Im v_{sel[2h+1]}(hits[h]) from the bucket spectra
hipError_t value_pick_launch(const int* hits, int nh, const uint8_t* sel,
                             const double2* bins_f, int n, const int* ai_dev, int B,
                             const double2* win, double2* values, hipStream_t st);
// u[l][b] = the adjoint seeds of the (hit, part) entries landing in bucket b of loop l,
// summed in entry order; head (loops*B ints) and next (2*nh ints) are scratch
hipError_t adjoint_seed_launch(const int* hits, int nh, const uint8_t* sel, const double2* g,
                               int n, const int* ai_dev, int B, int loops, const double2* win,
                               int* head, int* next, double2* u, hipStream_t st);
// gx[(i ai_l) mod n] += conj(f[i]) gb[l][bucket(i)], loop by loop (gx cleared first)
hipError_t adjoint_taps_launch(const double2* f, const double2* gb, const int* ai_dev, int n,
                               int B, int w, int loops, int tiled, double2* gx, hipStream_t st);

// ---- small utilities ---------------------------------------------------------------------
hipError_t fill_u32(void* p, uint32_t v, size_t words, hipStream_t st);

}  // namespace sfft
