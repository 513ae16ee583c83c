// sfft_adjoint.hip -- the estimated values as a function of the signal, and its adjoint.
//
// With the located indices and the median-selected loops of a transform held fixed, each
// estimated value is a real-linear function of the signal (the permute+filter fold, the
// bucket FFT, the filter division of innerLoop.cu:396-466 / :725-773):
//   yhat_h = Re v_{sel[2h]}(I_h) + i Im v_{sel[2h+1]}(I_h)
// value_pick evaluates it from bucket spectra; the adjoint maps g_h = dL/dRe yhat_h +
// i dL/dIm yhat_h back to dL/dRe x + i dL/dIm x in three steps:
//   1. seeds: every (hit, part) entry adds (gr F.x/d, gr F.y/d) (real part) or
//      (-gi F.y/d, gi F.x/d) (imaginary part) to its bucket; entries of one bucket are summed
//      in entry order, so the result does not depend on thread timing;
//   2. the unnormalised inverse FFT of each loop's seed row (the transpose of FFT_B), run by
//      the caller with the canonical FFT kernels;
//   3. taps: gx[(i ai_l) mod n] += conj(f[i]) gb[l][bucket(i)], one loop after the other, the
//      tiled taps (or all taps) first, then the reference fold's tail taps.  Within one launch
//      every tap writes a different signal index, so no atomics are needed.
// The library is built with -ffp-contract=off: the arithmetic below is exactly the oracle's.
#include "sfft_internal.h"

namespace sfft {

// bucket and filter response of location I in loop l (innerLoop.cu:837-848)
__device__ __forceinline__ void est_term(uint32_t I, int l, uint32_t n, const int* ai_dev,
                                         int B, const double2* win, long long* hashed_out,
                                         double2* F_out) {
  const long long m = (long long)(n / (uint32_t)B);
  const long long nn = n;
  const long long p = (long long)(((uint64_t)(uint32_t)ai_dev[l] * I) & (uint64_t)(n - 1));
  long long hashed = p / m;
  long long dist = p & (m - 1);
  if (dist > m / 2) {
    hashed = (hashed + 1) & (B - 1);
    dist -= m;
  }
  dist = (nn - dist) & (nn - 1);
  *hashed_out = hashed;
  *F_out = win[(dist + m / 2) & (nn - 1)];
}

__global__ __launch_bounds__(256) void value_pick_kernel(
    const int* __restrict__ hits, int nh, const uint8_t* __restrict__ sel,
    const double2* __restrict__ bins_f, uint32_t n, const int* __restrict__ ai_dev, int B,
    const double2* __restrict__ win, double2* __restrict__ values) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  const uint32_t I = (uint32_t)hits[h];
  const int lr = sel[2 * h], li = sel[2 * h + 1];
  long long hr, hi;
  double2 Fr, Fi;
  est_term(I, lr, n, ai_dev, B, win, &hr, &Fr);
  est_term(I, li, n, ai_dev, B, win, &hi, &Fi);
  const double2 br = bins_f[(long long)lr * B + hr];
  const double2 bi = bins_f[(long long)li * B + hi];
  double2 v;
  v.x = (br.x * Fr.x + br.y * Fr.y) / (Fr.x * Fr.x + Fr.y * Fr.y);
  v.y = (bi.y * Fi.x - bi.x * Fi.y) / (Fi.x * Fi.x + Fi.y * Fi.y);
  values[h] = v;
}

hipError_t value_pick_launch(const int* hits, int nh, const uint8_t* sel,
                             const double2* bins_f, int n, const int* ai_dev, int B,
                             const double2* win, double2* values, hipStream_t st) {
  if (nh <= 0) return hipSuccess;
  hipLaunchKernelGGL(value_pick_kernel, dim3((nh + 255) / 256), dim3(256), 0, st, hits, nh, sel,
                     bins_f, (uint32_t)n, ai_dev, B, win, values);
  return hipGetLastError();
}

// entry e = 2h + part -> bucket key l*B + hashed; per-key singly linked lists (any order)
__global__ __launch_bounds__(256) void adj_link_kernel(
    const int* __restrict__ hits, int nh, const uint8_t* __restrict__ sel, uint32_t n,
    const int* __restrict__ ai_dev, int B, const double2* __restrict__ win, int* head,
    int* __restrict__ next) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * nh) return;
  const int l = sel[e];
  long long hb;
  double2 F;
  est_term((uint32_t)hits[e >> 1], l, n, ai_dev, B, win, &hb, &F);
  next[e] = atomicExch(&head[(long long)l * B + hb], e);
}

// one thread per bucket key: sum its entries in ascending entry order (repeated minimum
// search over the short list)
__global__ __launch_bounds__(256) void adj_sum_kernel(
    const int* __restrict__ hits, const uint8_t* __restrict__ sel,
    const double2* __restrict__ g, uint32_t n, const int* __restrict__ ai_dev, int B,
    int loops, const double2* __restrict__ win, const int* __restrict__ head,
    const int* __restrict__ next, double2* __restrict__ u) {
  const long long key = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (key >= (long long)loops * B) return;
  const int l = (int)(key / B);
  const int first = head[key];
  double2 acc = make_double2(0.0, 0.0);
  int cur = -1;
  while (first >= 0) {
    int best = 0x7fffffff;
    for (int e = first; e >= 0; e = next[e])
      if (e > cur && e < best) best = e;
    if (best == 0x7fffffff) break;
    long long hb;
    double2 F;
    est_term((uint32_t)hits[best >> 1], l, n, ai_dev, B, win, &hb, &F);
    const double d = F.x * F.x + F.y * F.y;
    const double2 gv = g[best >> 1];
    if ((best & 1) == 0) {
      const double q = gv.x / d;
      acc.x = acc.x + q * F.x;
      acc.y = acc.y + q * F.y;
    } else {
      const double q = gv.y / d;
      acc.x = acc.x - q * F.y;
      acc.y = acc.y + q * F.x;
    }
    cur = best;
  }
  u[key] = acc;
}

hipError_t adjoint_seed_launch(const int* hits, int nh, const uint8_t* sel, const double2* g,
                               int n, const int* ai_dev, int B, int loops, const double2* win,
                               int* head, int* next, double2* u, hipStream_t st) {
  const long long keys = (long long)loops * B;
  hipError_t e = fill_u32(head, 0xffffffffu, (size_t)keys, st);
  if (e != hipSuccess) return e;
  if (nh > 0) {
    hipLaunchKernelGGL(adj_link_kernel, dim3((2 * nh + 255) / 256), dim3(256), 0, st, hits, nh,
                       sel, (uint32_t)n, ai_dev, B, win, head, next);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(adj_sum_kernel, dim3((unsigned)((keys + 255) / 256)), dim3(256), 0, st, hits,
                     sel, g, (uint32_t)n, ai_dev, B, loops, win, head, next, u);
  return hipGetLastError();
}

// one loop, one phase: tap i of [0, count) is id = i (head phase, bucket i mod B) or
// id = i*tiles + B (tail phase, bucket i)
__global__ __launch_bounds__(256) void adj_taps_kernel(
    const double2* __restrict__ f, const double2* __restrict__ gb, const int* __restrict__ ai_dev,
    int l, uint32_t mask, int B, int count, int tail, int tiles, double2* __restrict__ gx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t id = tail ? (uint32_t)(i * tiles + B) : (uint32_t)i;
  const int r = tail ? i : (i & (B - 1));
  const double2 fv = f[id], gv = gb[r];
  double2* dst = gx + ((id * (uint32_t)ai_dev[l]) & mask);
  double2 acc = *dst;
  acc.x = acc.x + fma(fv.x, gv.x, fv.y * gv.y);
  acc.y = acc.y + fma(fv.x, gv.y, -(fv.y * gv.x));
  *dst = acc;
}

hipError_t adjoint_taps_launch(const double2* f, const double2* gb, const int* ai_dev, int n,
                               int B, int w, int loops, int tiled, double2* gx, hipStream_t st) {
  hipError_t e = fill_u32(gx, 0u, (size_t)n * 4, st);
  if (e != hipSuccess) return e;
  const int tiles = w / B, residue = w % B;
  const int head = tiled ? tiles * B : w;
  const uint32_t mask = (uint32_t)n - 1u;
  for (int l = 0; l < loops; ++l) {
    const double2* row = gb + (long long)l * B;
    hipLaunchKernelGGL(adj_taps_kernel, dim3((head + 255) / 256), dim3(256), 0, st, f, row,
                       ai_dev, l, mask, B, head, 0, tiles, gx);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (tiled && residue > 0) {
      hipLaunchKernelGGL(adj_taps_kernel, dim3((residue + 255) / 256), dim3(256), 0, st, f, row,
                         ai_dev, l, mask, B, residue, 1, tiles, gx);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

}  // namespace sfft
