// sfft_aux.h -- launchers of the kernels outside the timed transform (filter design, signal
// synthesis, metric helpers, reference-compatible vote counting).  Not part of the public ABI.
#pragma once

#include "sfft_internal.h"

namespace sfft {

hipError_t aux_bluestein_pre(const double2* x, const double2* c, double2* A, int N, long long M,
                             hipStream_t st);
hipError_t aux_cmul_pointwise(double2* A, const double2* B, long long M, hipStream_t st);
hipError_t aux_bluestein_post(const double2* A, const double2* c, double2* out, int N, double M,
                              hipStream_t st);
hipError_t aux_shift_real(const double2* V, double2* ft, int w, hipStream_t st);
hipError_t aux_fold(const double2* ft, double2* g, int w, long long n, hipStream_t st);
hipError_t aux_scale_head(const double2* g, double2* ft, int w, double n, hipStream_t st);
// parallel filter design: Chebyshev samples (cancellation-free), Bluestein chirp (c[N], b[M]),
// boxcar over G into out with *mx2 = max |out|^2 (panels: n/64 + n/4096 entries, blkmax:
// aux_boxcar_blocks(n) entries; n >= 4096), normalisation and phase ramp from in into ff
// (in == ff allowed)
hipError_t aux_cheb_samples(double2* v, int w, double tolerance, hipStream_t st);
hipError_t aux_chirp(double2* c, double2* b, int N, long long M, int inverse, hipStream_t st);
long long aux_boxcar_blocks(long long n);
hipError_t aux_boxcar(const double2* G, double2* out, long long n, long long b_f,
                      double2* panels, double* blkmax, double* mx2, hipStream_t st);
hipError_t aux_ramp(const double2* in, double2* ff, long long n, long long hw, const double* mx2,
                    hipStream_t st);
hipError_t aux_scatter_c128(double2* dst, const int* idx, const double2* val, int k,
                            hipStream_t st);
hipError_t aux_add_noise(double2* x, long long n, double sigma, uint64_t seed, hipStream_t st);
hipError_t aux_c128_to_c64(const double2* a, float2* b, long long n, hipStream_t st);
hipError_t aux_c64_to_c128(const float2* a, double2* b, long long n, hipStream_t st);
hipError_t aux_scale_c128(double2* a, long long n, double s, hipStream_t st);
hipError_t aux_energy_partials(const double2* a, long long n, double* partial, int nblocks,
                               hipStream_t st);
hipError_t aux_gather_c128(const double2* a, const int* idx, double2* out, int k,
                           hipStream_t st);
hipError_t aux_vote_counts(const int* J, int nJ, long long n, int B, int W, int a,
                           const unsigned* bm, const int* maxfirst, int thresh, int* vote,
                           int* d_I, int* cnt, int cap, hipStream_t st);
hipError_t aux_bitmap_from_list(const int* J2, int nJ2, int ai, int W, unsigned* bm,
                                int* maxfirst, hipStream_t st);

// canonical device FFT with caller-provided scratch (in != out for multi-pass sizes)
hipError_t fft_plain(const void* in, void* out, void* scratch, int log2n, int count, int inverse,
                     int is_fp32, hipStream_t st);

}  // namespace sfft
