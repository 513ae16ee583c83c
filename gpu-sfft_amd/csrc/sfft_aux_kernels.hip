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
