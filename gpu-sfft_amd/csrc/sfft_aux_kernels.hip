// sfft_aux_kernels.hip -- kernels outside the timed transform: FP64 filter design
// (Bluestein chirp products, shift/fold, scaling), synthetic-signal helpers (scatter, noise,
// precision conversion), metric helpers (energy, gather) and the reference-compatible
// reverse_hash that counts votes in an int32[n] table.
//
// The filter-design element-wise products are written without fusion (the library is built
// with -ffp-contract=off) so that, together with the canonical FFT, the design reproduces the
// CPU restatement bit for bit.
#include "sfft_aux.h"

namespace sfft {

// ---- filter design -------------------------------------------------------------------------
// A[j] = x[j] * c[j] for j < N, 0 for N <= j < M
__global__ void bluestein_pre_kernel(const double2* __restrict__ x, const double2* __restrict__ c,
                                     double2* __restrict__ A, int N, long long M) {
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < M;
       j += (long long)gridDim.x * blockDim.x) {
    double2 r = make_double2(0.0, 0.0);
    if (j < N) {
      const double2 a = x[j], w = c[j];
      r.x = a.x * w.x - a.y * w.y;
      r.y = a.x * w.y + a.y * w.x;
    }
    A[j] = r;
  }
}

__global__ void cmul_pointwise_kernel(double2* __restrict__ A, const double2* __restrict__ B,
                                      long long M) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < M;
       i += (long long)gridDim.x * blockDim.x) {
    const double2 a = A[i], b = B[i];
    double2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    A[i] = r;
  }
}

// out[k] = (A[k] / M) * c[k]
__global__ void bluestein_post_kernel(const double2* __restrict__ A, const double2* __restrict__ c,
                                      double2* __restrict__ out, int N, double M) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < N;
       k += (long long)gridDim.x * blockDim.x) {
    const double yr = A[k].x / M, yi = A[k].y / M;
    const double2 w = c[k];
    double2 r;
    r.x = yr * w.x - yi * w.y;
    r.y = yr * w.y + yi * w.x;
    out[k] = r;
  }
}

// ft[j] = (V[(j - h + w) mod w].x, 0)   (cuShift by w/2, then the imaginary part cleared)
__global__ void shift_real_kernel(const double2* __restrict__ V, double2* __restrict__ ft, int w) {
  const int h = w / 2;
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < w;
       j += (long long)gridDim.x * blockDim.x) {
    const long long src = (j - h + w) % w;
    ft[j] = make_double2(V[src].x, 0.0);
  }
}

// g = 0; g[0..h) = ft[h..2h); g[n-h..n) = ft[0..h)
__global__ void fold_kernel(const double2* __restrict__ ft, double2* __restrict__ g, int w,
                            long long n) {
  const long long h = w / 2;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    double2 v = make_double2(0.0, 0.0);
    if (i < h)
      v = ft[h + i];
    else if (i >= n - h)
      v = ft[i - (n - h)];
    g[i] = v;
  }
}

__global__ void scale_head_kernel(const double2* __restrict__ g, double2* __restrict__ ft, int w,
                                  double n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < w;
       i += (long long)gridDim.x * blockDim.x) {
    ft[i] = make_double2(g[i].x / n, g[i].y / n);
  }
}

// ---- filter design, parallel mode ------------------------------------------------------------
// The same mathematical filter as the faithful design, computed without sequential recurrences:
// device Chebyshev samples and Bluestein chirp, a boxcar from two levels of panel sums, a
// device max, and a phase ramp whose angle is reduced exactly in integers.

// v[i] = (T_m(x_i) * tol, 0) with m = w - 1, x_i = t0 cos(phi), phi = pi i / w, and 0 where
// x_i < -1 (the faithful design's NaN -> 0 rule).  Evaluated without cancellation: the direct
// form cosh(m acosh(x)) loses about m^2 eps / 19 relative near x = t0 (4e-5 at w = 2.2M),
// because x^2 - 1 cancels.  Here x - 1 = (t0 - 1) - 2 t0 sin^2(phi/2) with
// t0 - 1 = 2 sinh^2(a/2) (t0 = cosh a) comes without cancellation, and then
//   x > 1:      T_m = cosh(m log1p(d + sqrt(d (2 + d)))),   d = x - 1
//   0 <= x <= 1: T_m = cos(2 m asin(sqrt((1 - x) / 2)))
//   -1 <= x < 0: T_m = cos(m (pi - 2 asin(sqrt((1 + x) / 2)))),  1 + x = 2 t0 cos^2(phi/2) - (t0 - 1)
__global__ void cheb_samples_kernel(double2* __restrict__ v, int w, double t0, double t0m1,
                                    double tol) {
  const double m = (double)(w - 1);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < w;
       i += (long long)gridDim.x * blockDim.x) {
    const double hq = (double)i / (2.0 * (double)w);   // phi / (2 pi)
    const double sh = sinpi(hq), ch = cospi(hq);
    const double d = t0m1 - 2.0 * t0 * (sh * sh);       // x - 1
    double val;
    if (d > 0.0) {
      val = cosh(m * log1p(d + sqrt(d * (2.0 + d))));
    } else if (d >= -1.0) {
      val = cos(2.0 * m * asin(sqrt(-d * 0.5)));
    } else {
      const double xp1 = 2.0 * t0 * (ch * ch) - t0m1;   // x + 1
      val = (xp1 >= 0.0) ? cos(m * (M_PI - 2.0 * asin(sqrt(xp1 * 0.5)))) : 0.0;
    }
    v[i] = make_double2(val * tol, 0.0);
  }
}

// Bluestein chirp c[j] = exp(sgn i pi (j^2 mod 2N) / N), j < N, and its conjugate mirrored into
// b[0..M): b[j] = b[M-j] = conj(c[j]) for j < N, zero elsewhere (M >= 2N - 1)
__global__ void chirp_kernel(double2* __restrict__ c, double2* __restrict__ b, int N, long long M,
                             double sgn) {
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < M;
       j += (long long)gridDim.x * blockDim.x) {
    const long long m = (j < N) ? j : ((j > M - N) ? M - j : -1);
    double2 r = make_double2(0.0, 0.0);
    if (m >= 0) {
      const unsigned long long q =
          ((unsigned long long)m * (unsigned long long)m) % (2ULL * (unsigned long long)N);
      double s, co;
      sincospi((double)q / (double)N, &s, &co);
      if (j < N) c[j] = make_double2(co, sgn * s);
      r = make_double2(co, -sgn * s);
    }
    b[j] = r;
  }
}

// ps[p] = sum of in[P p .. P p + P), one thread per panel
__global__ void panel_sum_kernel(const double2* __restrict__ in, double2* __restrict__ ps,
                                 long long npan, int P) {
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < npan;
       p += (long long)gridDim.x * blockDim.x) {
    double sx = 0.0, sy = 0.0;
    const double2* src = in + p * P;
    for (int j = 0; j < P; ++j) {
      sx += src[j].x;
      sy += src[j].y;
    }
    ps[p] = make_double2(sx, sy);
  }
}

constexpr int kPan1 = 64;     // level-1 panel (entries of G)
constexpr int kPan2 = 4096;   // level-2 panel (64 level-1 panels)

// boxcar s_i = sum of G[i .. i + b_f) (indices mod n) stored at (i + b_f/2) mod n, i.e. the
// faithful running sum without its n-step rounding drift: each thread owns kPan1 consecutive
// outputs, starts from a window sum over panel sums (at most 3*63 + b_f/4096 terms) and slides
// kPan1 - 1 steps.  blkmax[block] = max |s|^2 over the block's outputs.  n >= 4096.
__global__ __launch_bounds__(256) void boxcar_kernel(const double2* __restrict__ G,
                                                     const double2* __restrict__ p1,
                                                     const double2* __restrict__ p2,
                                                     double2* __restrict__ out, long long n,
                                                     long long b_f, double* __restrict__ blkmax) {
  __shared__ double sh[256];
  const long long mask = n - 1, np1 = n / kPan1, np2 = n / kPan2;
  const long long c = blockIdx.x * 256LL + threadIdx.x;
  double mx2 = 0.0;
  if (c < np1) {
    const long long i0 = c * kPan1;
    double sx = 0.0, sy = 0.0;
    long long pos = i0, len = b_f;
    while (len > 0) {
      if ((pos & (kPan2 - 1)) == 0 && len >= kPan2) {
        const double2 v = p2[(pos / kPan2) % np2];
        sx += v.x;
        sy += v.y;
        pos += kPan2;
        len -= kPan2;
      } else if ((pos & (kPan1 - 1)) == 0 && len >= kPan1) {
        const double2 v = p1[(pos / kPan1) % np1];
        sx += v.x;
        sy += v.y;
        pos += kPan1;
        len -= kPan1;
      } else {
        const double2 v = G[pos & mask];
        sx += v.x;
        sy += v.y;
        ++pos;
        --len;
      }
    }
    const long long offset = b_f / 2;
    for (int t = 0; t < kPan1; ++t) {
      const long long i = i0 + t;
      out[(i + offset) & mask] = make_double2(sx, sy);
      const double a2 = sx * sx + sy * sy;
      mx2 = mx2 > a2 ? mx2 : a2;
      const double2 ahead = G[(i + b_f) & mask], behind = G[i];
      sx += ahead.x - behind.x;
      sy += ahead.y - behind.y;
    }
  }
  sh[threadIdx.x] = mx2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) blkmax[blockIdx.x] = sh[0];
}

// out[0] = max of part[0..np) (one block)
__global__ __launch_bounds__(256) void max_reduce_kernel(const double* __restrict__ part, int np,
                                                         double* __restrict__ out) {
  __shared__ double sh[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) m = fmax(m, part[i]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// ff[i] = (ff[i] / sqrt(*mx2)) * exp(-2 pi i (hw i mod n) / n): the normalisation and the phase
// ramp, with the angle reduced exactly in integers (n a power of two, hw i < 2^63)
__global__ void ramp_kernel(const double2* in, double2* ff, long long n,
                            long long hw, const double* __restrict__ mx2) {
  const double mx = sqrt(*mx2);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const unsigned long long r =
        ((unsigned long long)hw * (unsigned long long)i) & (unsigned long long)(n - 1);
    double s, co;
    sincospi(-2.0 * (double)r / (double)n, &s, &co);
    const double ax = in[i].x / mx, ay = in[i].y / mx;
    ff[i] = make_double2(ax * co - ay * s, ax * s + ay * co);
  }
}

// ---- signals ---------------------------------------------------------------------------------
__global__ void scatter_c128_kernel(double2* __restrict__ dst, const int* __restrict__ idx,
                                    const double2* __restrict__ val, int k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k) dst[idx[i]] = val[i];
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// x[i] += sigma * (N(0,1), N(0,1)), Box-Muller on two counter-hashed uniforms
__global__ void add_noise_kernel(double2* __restrict__ x, long long n, double sigma,
                                 uint64_t seed) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const uint64_t h1 = mix64(seed ^ (2ull * (uint64_t)i));
    const uint64_t h2 = mix64(seed ^ (2ull * (uint64_t)i + 1ull));
    const double u = ((double)(h1 >> 11) + 1.0) * (1.0 / 9007199254740993.0);  // (0, 1]
    const double v = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);          // [0, 1)
    const double rad = sigma * sqrt(-2.0 * log(u));
    double sv, cv;
    sincospi(2.0 * v, &sv, &cv);
    x[i].x += rad * cv;
    x[i].y += rad * sv;
  }
}

__global__ void c128_to_c64_kernel(const double2* __restrict__ a, float2* __restrict__ b,
                                   long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = make_float2((float)a[i].x, (float)a[i].y);
}

__global__ void c64_to_c128_kernel(const float2* __restrict__ a, double2* __restrict__ b,
                                   long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = make_double2((double)a[i].x, (double)a[i].y);
}

__global__ void scale_c128_kernel(double2* __restrict__ a, long long n, double s) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    a[i].x /= s;
    a[i].y /= s;
  }
}

// per-block partial sums of |a|^2 (fixed order inside a block; the host adds the partials in
// block order, so the total is deterministic)
__global__ __launch_bounds__(256) void energy_kernel(const double2* __restrict__ a, long long n,
                                                     double* __restrict__ partial) {
  __shared__ double sh[256];
  double s = 0.0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    s += a[i].x * a[i].x + a[i].y * a[i].y;
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void gather_c128_kernel(const double2* __restrict__ a, const int* __restrict__ idx,
                                   double2* __restrict__ out, int k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k) out[i] = a[idx[i]];
}

// ---- reference-compatible reverse hash: int32 vote counts, atomic append -------------------
__global__ void vote_counts_kernel(const int* __restrict__ J, int nJ, long long n, int B, int W,
                                   int a, const unsigned* __restrict__ bm,
                                   const int* __restrict__ maxfirst, int thresh,
                                   int* __restrict__ vote, int* __restrict__ d_I,
                                   int* __restrict__ cnt, int cap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nJ) return;
  const long long m = n / B;
  long long low = (long long)J[i] * m - (m >> 1);
  low = (low + n) & (n - 1);
  if ((int)(low & (W - 1)) > *maxfirst) return;
  const uint64_t nmask = (uint64_t)n - 1;
  for (long long d = 0; d < m;) {
    const uint64_t p = (uint64_t)((low + d) & (n - 1));
    const uint32_t bpos = (uint32_t)(p & (uint64_t)(W - 1));
    const int off = (int)(bpos & 31u);
    const long long left = m - d;
    const int take = (left < 32 - off) ? (int)left : (32 - off);
    uint32_t word = bm[bpos >> 5] >> off;
    if (take < 32) word &= (1u << take) - 1u;
    while (word) {
      const int bit = __ffs(word) - 1;
      word &= word - 1u;
      const uint32_t loc = (uint32_t)(((p + (uint64_t)bit) * (uint64_t)(uint32_t)a) & nmask);
      const int old = atomicAdd(&vote[loc], 1);
      if (old + 1 == thresh) {
        const int pos = atomicAdd(cnt, 1);
        if (pos < cap) d_I[pos] = (int)loc;
      }
    }
    d += take;
  }
}

__global__ void bitmap_from_list_kernel(const int* __restrict__ J2, int nJ2, int ai, int W,
                                        unsigned* __restrict__ bm, int* __restrict__ maxfirst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nJ2) return;
  const uint32_t first =
      (uint32_t)(((uint64_t)(uint32_t)J2[i] * (uint64_t)(uint32_t)ai) & (uint64_t)(W - 1));
  atomicOr(&bm[first >> 5], 1u << (first & 31));
  atomicMax(maxfirst, (int)first);
}

// ---- launchers ---------------------------------------------------------------------------------
static dim3 grid_for(long long n) {
  long long b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

hipError_t aux_bluestein_pre(const double2* x, const double2* c, double2* A, int N, long long M,
                             hipStream_t st) {
  hipLaunchKernelGGL(bluestein_pre_kernel, grid_for(M), dim3(256), 0, st, x, c, A, N, M);
  return hipGetLastError();
}
hipError_t aux_cmul_pointwise(double2* A, const double2* B, long long M, hipStream_t st) {
  hipLaunchKernelGGL(cmul_pointwise_kernel, grid_for(M), dim3(256), 0, st, A, B, M);
  return hipGetLastError();
}
hipError_t aux_bluestein_post(const double2* A, const double2* c, double2* out, int N, double M,
                              hipStream_t st) {
  hipLaunchKernelGGL(bluestein_post_kernel, grid_for(N), dim3(256), 0, st, A, c, out, N, M);
  return hipGetLastError();
}
hipError_t aux_shift_real(const double2* V, double2* ft, int w, hipStream_t st) {
  hipLaunchKernelGGL(shift_real_kernel, grid_for(w), dim3(256), 0, st, V, ft, w);
  return hipGetLastError();
}
hipError_t aux_fold(const double2* ft, double2* g, int w, long long n, hipStream_t st) {
  hipLaunchKernelGGL(fold_kernel, grid_for(n), dim3(256), 0, st, ft, g, w, n);
  return hipGetLastError();
}
hipError_t aux_scale_head(const double2* g, double2* ft, int w, double n, hipStream_t st) {
  hipLaunchKernelGGL(scale_head_kernel, grid_for(w), dim3(256), 0, st, g, ft, w, n);
  return hipGetLastError();
}
hipError_t aux_cheb_samples(double2* v, int w, double tolerance, hipStream_t st) {
  const double a = acosh(1.0 / tolerance) / (double)(w - 1);
  const double sh = sinh(0.5 * a);
  hipLaunchKernelGGL(cheb_samples_kernel, grid_for(w), dim3(256), 0, st, v, w, cosh(a),
                     2.0 * sh * sh, tolerance);
  return hipGetLastError();
}
hipError_t aux_chirp(double2* c, double2* b, int N, long long M, int inverse, hipStream_t st) {
  hipLaunchKernelGGL(chirp_kernel, grid_for(M), dim3(256), 0, st, c, b, N, M,
                     inverse ? 1.0 : -1.0);
  return hipGetLastError();
}
long long aux_boxcar_blocks(long long n) { return (n / kPan1 + 255) / 256; }
hipError_t aux_boxcar(const double2* G, double2* out, long long n, long long b_f,
                      double2* panels, double* blkmax, double* mx2, hipStream_t st) {
  if (n < kPan2 || b_f < 1 || b_f > n) return hipErrorInvalidValue;
  const long long np1 = n / kPan1, np2 = n / kPan2;
  double2* p1 = panels;
  double2* p2 = panels + np1;
  hipLaunchKernelGGL(panel_sum_kernel, grid_for(np1), dim3(256), 0, st, G, p1, np1, kPan1);
  hipLaunchKernelGGL(panel_sum_kernel, grid_for(np2), dim3(256), 0, st, (const double2*)p1, p2,
                     np2, kPan2 / kPan1);
  const long long nb = aux_boxcar_blocks(n);
  hipLaunchKernelGGL(boxcar_kernel, dim3((unsigned)nb), dim3(256), 0, st, G,
                     (const double2*)p1, (const double2*)p2, out, n, b_f, blkmax);
  hipLaunchKernelGGL(max_reduce_kernel, dim3(1), dim3(256), 0, st, (const double*)blkmax,
                     (int)nb, mx2);
  return hipGetLastError();
}
hipError_t aux_ramp(const double2* in, double2* ff, long long n, long long hw, const double* mx2,
                    hipStream_t st) {
  hipLaunchKernelGGL(ramp_kernel, grid_for(n), dim3(256), 0, st, in, ff, n, hw, mx2);
  return hipGetLastError();
}
hipError_t aux_scatter_c128(double2* dst, const int* idx, const double2* val, int k,
                            hipStream_t st) {
  hipLaunchKernelGGL(scatter_c128_kernel, dim3((k + 255) / 256), dim3(256), 0, st, dst, idx, val,
                     k);
  return hipGetLastError();
}
hipError_t aux_add_noise(double2* x, long long n, double sigma, uint64_t seed, hipStream_t st) {
  hipLaunchKernelGGL(add_noise_kernel, grid_for(n), dim3(256), 0, st, x, n, sigma, seed);
  return hipGetLastError();
}
hipError_t aux_c128_to_c64(const double2* a, float2* b, long long n, hipStream_t st) {
  hipLaunchKernelGGL(c128_to_c64_kernel, grid_for(n), dim3(256), 0, st, a, b, n);
  return hipGetLastError();
}
hipError_t aux_c64_to_c128(const float2* a, double2* b, long long n, hipStream_t st) {
  hipLaunchKernelGGL(c64_to_c128_kernel, grid_for(n), dim3(256), 0, st, a, b, n);
  return hipGetLastError();
}
hipError_t aux_scale_c128(double2* a, long long n, double s, hipStream_t st) {
  hipLaunchKernelGGL(scale_c128_kernel, grid_for(n), dim3(256), 0, st, a, n, s);
  return hipGetLastError();
}
hipError_t aux_energy_partials(const double2* a, long long n, double* partial, int nblocks,
                               hipStream_t st) {
  hipLaunchKernelGGL(energy_kernel, dim3(nblocks), dim3(256), 0, st, a, n, partial);
  return hipGetLastError();
}
hipError_t aux_gather_c128(const double2* a, const int* idx, double2* out, int k,
                           hipStream_t st) {
  hipLaunchKernelGGL(gather_c128_kernel, dim3((k + 255) / 256), dim3(256), 0, st, a, idx, out, k);
  return hipGetLastError();
}
hipError_t aux_vote_counts(const int* J, int nJ, long long n, int B, int W, int a,
                           const unsigned* bm, const int* maxfirst, int thresh, int* vote,
                           int* d_I, int* cnt, int cap, hipStream_t st) {
  hipLaunchKernelGGL(vote_counts_kernel, dim3((nJ + 255) / 256), dim3(256), 0, st, J, nJ, n, B,
                     W, a, bm, maxfirst, thresh, vote, d_I, cnt, cap);
  return hipGetLastError();
}
hipError_t aux_bitmap_from_list(const int* J2, int nJ2, int ai, int W, unsigned* bm,
                                int* maxfirst, hipStream_t st) {
  hipLaunchKernelGGL(bitmap_from_list_kernel, dim3((nJ2 + 255) / 256), dim3(256), 0, st, J2, nJ2,
                     ai, W, bm, maxfirst);
  return hipGetLastError();
}

}  // namespace sfft
