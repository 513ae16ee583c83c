// sfft_kernels.hip -- hand-written CDNA4 (gfx950) kernels of the sparse-FFT hot path.
//
//  * fft_pass_kernel      canonical multi-pass big-radix Stockham FFT (LDS sub-FFT, radix-4/2
//                         stages), optional fused comb gather (RLFC) and |X|^2 key epilogue
//  * perm_filter_kernel   permutation + filtering: bucket-owner gather fold, all loops of a
//                         transform in one launch, fixed (deterministic) tap order
//  * select_*_kernel      top-K: MSD radix select on the IEEE bits of |X|^2 + order-preserving
//                         compaction (ascending indices, ties by lower index)
//  * hash_*_kernel        reverse hash: approved-residue bitmaps, packed per-loop vote masks,
//                         deterministic hit list (loop, bucket, walk order)
//  * estimate_kernel      per-location FP64 filter division and stable lower median
//
// Every floating-point operation that must match the CPU oracle bit for bit is written out
// explicitly (fma where fused, plain ops elsewhere); the library is built with
// -ffp-contract=off so the compiler adds no fusion of its own.
#include "sfft_internal.h"

namespace sfft {

template <typename R>
struct CT;
template <>
struct CT<double> {
  using T2 = double2;
  using K = unsigned long long;
};
template <>
struct CT<float> {
  using T2 = float2;
  using K = unsigned int;
};

__device__ __forceinline__ double2 cmul(double2 a, double2 w) {
  return make_double2(fma(a.x, w.x, -(a.y * w.y)), fma(a.x, w.y, a.y * w.x));
}
__device__ __forceinline__ float2 cmul(float2 a, float2 w) {
  return make_float2(fmaf(a.x, w.x, -(a.y * w.y)), fmaf(a.x, w.y, a.y * w.x));
}
__device__ __forceinline__ unsigned long long sample_key(double2 v) {
  return (unsigned long long)__double_as_longlong(fma(v.x, v.x, v.y * v.y));
}
__device__ __forceinline__ unsigned int sample_key(float2 v) {
  return __float_as_uint(fmaf(v.x, v.x, v.y * v.y));
}

template <typename T2>
__device__ __forceinline__ T2 cadd(T2 a, T2 b) {
  T2 r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}
template <typename T2>
__device__ __forceinline__ T2 csub(T2 a, T2 b) {
  T2 r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}

// radix-4 butterfly: t0=a0+a2, t1=a0-a2, t2=a1+a3, t3=(a1-a3)*(-+i); y=(t0+t2, t1+t3, t0-t2, t1-t3)
template <typename T2>
__device__ __forceinline__ void bfly4(T2& a0, T2& a1, T2& a2, T2& a3, int inverse) {
  const T2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), d = csub(a1, a3);
  T2 t3;
  if (inverse) {
    t3.x = -d.y;
    t3.y = d.x;
  } else {
    t3.x = d.y;
    t3.y = -d.x;
  }
  a0 = cadd(t0, t2);
  a1 = cadd(t1, t3);
  a2 = csub(t0, t2);
  a3 = csub(t1, t3);
}

// ============================================================================================
// FFT pass
// ============================================================================================
constexpr int FFT_NT = 512;
constexpr int FFT_CAP = 8192;                // complex elements per tile (LDS)
constexpr int FFT_EPT = FFT_CAP / FFT_NT;    // elements per thread
constexpr int FFT_QMAX = 16;                 // columns per tile

struct PassArgs {
  const void* in;
  long long in_bstride;
  void* out;
  long long out_bstride;
  void* keys;
  long long keys_bstride;
  int key_batches;
  int log2N, r, log2Lprev, log2Q;
  int inverse;
  const void* twR;
  const void* tdir;
  const void* thi;
  const void* tlo;
  int s;
  const void* comb_x;
  const int* comb_offs;
  int comb_sigma;
};

template <typename T2>
__device__ __forceinline__ T2 outer_tw(const PassArgs& A, uint32_t e, int log2Lcur) {
  if (log2Lcur <= kMaxPassLog) return ((const T2*)A.tdir)[e];
  const uint32_t lo = e & ((1u << A.s) - 1u), hi = e >> A.s;
  return cmul(((const T2*)A.thi)[hi], ((const T2*)A.tlo)[lo]);
}

template <typename T2>
__device__ __forceinline__ void fft_stage4(T2* lds, int RP, int r, int log2ncol, int log2ns,
                                           const T2* __restrict__ twR, int inverse) {
  const int tid = threadIdx.x;
  const int Rq = 1 << (r - 2);
  const int nbf = (1 << log2ncol) * Rq;
  const int ns = 1 << log2ns;
  const int tstep = 1 << (r - log2ns - 2);
  T2 v[FFT_EPT];
#pragma unroll
  for (int kk = 0; kk < FFT_EPT / 4; ++kk) {
    const int bi = tid + kk * FFT_NT;
    if (bi < nbf) {
      const int c = bi >> (r - 2);
      const int j = bi & (Rq - 1);
      const T2* col = lds + c * RP;
      T2 a0 = col[j], a1 = col[j + Rq], a2 = col[j + 2 * Rq], a3 = col[j + 3 * Rq];
      if (ns > 1) {
        const int jm = j & (ns - 1);
        a1 = cmul(a1, twR[jm * tstep]);
        a2 = cmul(a2, twR[2 * jm * tstep]);
        a3 = cmul(a3, twR[3 * jm * tstep]);
      }
      bfly4(a0, a1, a2, a3, inverse);
      v[4 * kk + 0] = a0;
      v[4 * kk + 1] = a1;
      v[4 * kk + 2] = a2;
      v[4 * kk + 3] = a3;
    }
  }
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < FFT_EPT / 4; ++kk) {
    const int bi = tid + kk * FFT_NT;
    if (bi < nbf) {
      const int c = bi >> (r - 2);
      const int j = bi & (Rq - 1);
      T2* col = lds + c * RP;
      const int base = ((j >> log2ns) << (log2ns + 2)) + (j & (ns - 1));
      col[base] = v[4 * kk + 0];
      col[base + ns] = v[4 * kk + 1];
      col[base + 2 * ns] = v[4 * kk + 2];
      col[base + 3 * ns] = v[4 * kk + 3];
    }
  }
  __syncthreads();
}

template <typename T2>
__device__ __forceinline__ void fft_stage2(T2* lds, int RP, int r, int log2ncol, int log2ns,
                                           const T2* __restrict__ twR) {
  const int tid = threadIdx.x;
  const int Rh = 1 << (r - 1);
  const int nbf = (1 << log2ncol) * Rh;
  const int ns = 1 << log2ns;
  const int tstep = 1 << (r - log2ns - 1);
  T2 v[FFT_EPT];
#pragma unroll
  for (int kk = 0; kk < FFT_EPT / 2; ++kk) {
    const int bi = tid + kk * FFT_NT;
    if (bi < nbf) {
      const int c = bi >> (r - 1);
      const int j = bi & (Rh - 1);
      const T2* col = lds + c * RP;
      T2 a0 = col[j], a1 = col[j + Rh];
      if (ns > 1) a1 = cmul(a1, twR[(j & (ns - 1)) * tstep]);
      v[2 * kk + 0] = cadd(a0, a1);
      v[2 * kk + 1] = csub(a0, a1);
    }
  }
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < FFT_EPT / 2; ++kk) {
    const int bi = tid + kk * FFT_NT;
    if (bi < nbf) {
      const int c = bi >> (r - 1);
      const int j = bi & (Rh - 1);
      T2* col = lds + c * RP;
      const int base = ((j >> log2ns) << (log2ns + 1)) + (j & (ns - 1));
      col[base] = v[2 * kk + 0];
      col[base + ns] = v[2 * kk + 1];
    }
  }
  __syncthreads();
}

template <typename Real, int LOAD, int STORE>
__global__ __launch_bounds__(FFT_NT) void fft_pass_kernel(PassArgs A) {
  using T2 = typename CT<Real>::T2;
  using KT = typename CT<Real>::K;
  __shared__ T2 lds[FFT_CAP + FFT_QMAX];
  const int tid = threadIdx.x;
  const long long b = blockIdx.y;
  const int r = A.r;
  const int R = 1 << r;
  const int RP = R + 1;
  const int log2Lprev = A.log2Lprev;
  const int log2Lcur = log2Lprev + r;
  const long long Lprev = 1LL << log2Lprev;
  const long long nq = 1LL << (A.log2N - log2Lcur);
  const bool first = (log2Lprev == 0);
  int log2ncol;
  long long q0 = 0, q = 0, t0 = 0;
  if (first) {
    log2ncol = min(A.log2Q, A.log2N - log2Lcur);
    q0 = (long long)blockIdx.x << log2ncol;
  } else {
    log2ncol = A.log2Q;
    const int log2tpq = log2Lprev - A.log2Q;
    q = (long long)blockIdx.x >> log2tpq;
    t0 = ((long long)blockIdx.x & ((1LL << log2tpq) - 1)) << A.log2Q;
  }
  const int ncol = 1 << log2ncol;
  const int TE = ncol << r;

  // ---- load (+ outer twiddle) into LDS, column c at lds[c*RP + v]
  const T2* __restrict__ in = (const T2*)A.in + b * A.in_bstride;
#pragma unroll
  for (int k = 0; k < FFT_EPT; ++k) {
    const int e = tid + k * FFT_NT;
    if (e < TE) {
      const int c = e & (ncol - 1);
      const int v = e >> log2ncol;
      long long g;
      if (first)
        g = q0 + c + nq * v;
      else
        g = (q + nq * v) * Lprev + t0 + c;
      T2 val;
      if (LOAD == kLoadComb)
        val = ((const T2*)A.comb_x)[(long long)A.comb_offs[b] + g * (long long)A.comb_sigma];
      else
        val = in[g];
      if (!first && v > 0) val = cmul(val, outer_tw<T2>(A, (uint32_t)((t0 + c) * v), log2Lcur));
      lds[c * RP + v] = val;
    }
  }
  __syncthreads();

  // ---- sub-FFT of every column: radix-4 stages, then one radix-2 stage if r is odd
  const T2* __restrict__ twR = (const T2*)A.twR;
  int log2ns = 0;
  int rem = r;
  while (rem >= 2) {
    fft_stage4<T2>(lds, RP, r, log2ncol, log2ns, twR, A.inverse);
    log2ns += 2;
    rem -= 2;
  }
  if (rem == 1) fft_stage2<T2>(lds, RP, r, log2ncol, log2ns, twR);

  // ---- store: out[q*Lcur + t + Lprev*k]
  T2* __restrict__ out = (T2*)A.out + b * A.out_bstride;
  KT* __restrict__ keys = (KT*)A.keys + b * A.keys_bstride;
  const bool wkeys = (STORE != kStoreCplx) && (b < A.key_batches);
#pragma unroll
  for (int k = 0; k < FFT_EPT; ++k) {
    const int e = tid + k * FFT_NT;
    if (e < TE) {
      int c, kk;
      long long g;
      if (first) {
        c = e >> r;
        kk = e & (R - 1);
        g = (q0 + c) * R + kk;
      } else {
        c = e & (ncol - 1);
        kk = e >> log2ncol;
        g = (q << log2Lcur) + t0 + c + Lprev * kk;
      }
      const T2 val = lds[c * RP + kk];
      if (STORE != kStoreKeys) out[g] = val;
      if (wkeys) keys[g] = sample_key(val);
    }
  }
}

template <typename Real, int LOAD, int STORE>
static hipError_t launch_pass(const PassArgs& A, int tiles, int batch, hipStream_t st) {
  hipLaunchKernelGGL((fft_pass_kernel<Real, LOAD, STORE>), dim3(tiles, batch), dim3(FFT_NT), 0,
                     st, A);
  return hipGetLastError();
}

template <typename Real>
static hipError_t dispatch_pass(const PassArgs& A, int load, int store, int tiles, int batch,
                                hipStream_t st) {
  if (load == kLoadComb) {
    if (store == kStoreCplx) return launch_pass<Real, kLoadComb, kStoreCplx>(A, tiles, batch, st);
    if (store == kStoreKeys) return launch_pass<Real, kLoadComb, kStoreKeys>(A, tiles, batch, st);
    return launch_pass<Real, kLoadComb, kStoreBoth>(A, tiles, batch, st);
  }
  if (store == kStoreCplx) return launch_pass<Real, kLoadPlain, kStoreCplx>(A, tiles, batch, st);
  if (store == kStoreKeys) return launch_pass<Real, kLoadPlain, kStoreKeys>(A, tiles, batch, st);
  return launch_pass<Real, kLoadPlain, kStoreBoth>(A, tiles, batch, st);
}

hipError_t fft_run(const FftLaunch& L, hipStream_t st) {
  const FftTables& T = *L.tab;
  const int log2N = T.log2n;
  const long long N = 1LL << log2N;
  const int np = T.npass;
  if (np == 0) {
    // N == 1: the DFT is the identity (copy / keys)
    return hipErrorInvalidValue;
  }
  int log2Lprev = 0;
  const void* cur_in = L.in;
  long long cur_in_stride = L.in_bstride;
  for (int p = 0; p < np; ++p) {
    const int r = T.logs[p];
    const int log2Lcur = log2Lprev + r;
    const bool last = (p == np - 1);
    // pass p writes to out when (np - 1 - p) is even, else to scratch
    void* dst = (((np - 1 - p) & 1) == 0) ? L.out : L.scratch;
    const long long dst_stride = (dst == L.out) ? L.out_bstride : N;
    PassArgs A;
    A.in = cur_in;
    A.in_bstride = cur_in_stride;
    A.out = dst;
    A.out_bstride = dst_stride;
    A.keys = L.keys;
    A.keys_bstride = L.keys_bstride;
    A.key_batches = L.key_batches;
    A.log2N = log2N;
    A.r = r;
    A.log2Lprev = log2Lprev;
    int log2Q = 0;
    while ((1 << (log2Q + 1)) <= FFT_QMAX && ((1 << (log2Q + 1)) << r) <= FFT_CAP) ++log2Q;
    int tiles;
    if (p == 0) {
      const int lnq = log2N - log2Lcur;
      const int lq = log2Q < lnq ? log2Q : lnq;
      tiles = 1 << (lnq - lq);
    } else {
      if (log2Q > log2Lprev) log2Q = log2Lprev;
      tiles = (int)((1LL << (log2N - log2Lcur)) << (log2Lprev - log2Q));
    }
    A.log2Q = log2Q;
    A.inverse = T.inverse;
    A.twR = T.twR[p];
    A.tdir = T.tdir[p];
    A.thi = T.thi[p];
    A.tlo = T.tlo[p];
    A.s = T.s[p];
    A.comb_x = L.comb_x;
    A.comb_offs = L.comb_offs;
    A.comb_sigma = L.comb_sigma;
    const int load = (p == 0) ? L.load : kLoadPlain;
    const int store = last ? L.store : kStoreCplx;
    hipError_t e = T.is_fp32 ? dispatch_pass<float>(A, load, store, tiles, L.batch, st)
                             : dispatch_pass<double>(A, load, store, tiles, L.batch, st);
    if (e != hipSuccess) return e;
    cur_in = dst;
    cur_in_stride = dst_stride;
    log2Lprev = log2Lcur;
  }
  return hipSuccess;
}

// ============================================================================================
// permutation + filtering (bucket-owner fold)
// ============================================================================================
struct Acc64 {
  double2 s;
  __device__ __forceinline__ void init() { s = make_double2(0.0, 0.0); }
  __device__ __forceinline__ void add(double2 x, double2 f) {
    s.x = fma(x.x, f.x, s.x);
    s.x = fma(-x.y, f.y, s.x);
    s.y = fma(x.x, f.y, s.y);
    s.y = fma(x.y, f.x, s.y);
  }
  __device__ __forceinline__ double2 get() const { return s; }
};

struct Acc32 {  // Kahan-compensated FP32 fold
  float sx, cx, sy, cy;
  __device__ __forceinline__ void init() { sx = cx = sy = cy = 0.f; }
  __device__ __forceinline__ static void kahan(float& sum, float& comp, float term) {
    const float y = term - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
  }
  __device__ __forceinline__ void add(float2 x, float2 f) {
    const float tx = fmaf(x.x, f.x, -(x.y * f.y));
    const float ty = fmaf(x.x, f.y, x.y * f.x);
    kahan(sx, cx, tx);
    kahan(sy, cy, ty);
  }
  __device__ __forceinline__ float2 get() const { return make_float2(sx, sy); }
};

template <typename Real>
struct AccOf;
template <>
struct AccOf<double> {
  using A = Acc64;
};
template <>
struct AccOf<float> {
  using A = Acc32;
};

constexpr int PF_NT = 256;

template <typename Real, int LPT>
__global__ __launch_bounds__(PF_NT) void perm_filter_kernel(
    const typename CT<Real>::T2* __restrict__ x, const typename CT<Real>::T2* __restrict__ f,
    typename CT<Real>::T2* __restrict__ bins, const int* __restrict__ ai_dev, uint32_t mask,
    int B, int w, int loops, int tiled) {
  using T2 = typename CT<Real>::T2;
  using AC = typename AccOf<Real>::A;
  const int r = blockIdx.x * PF_NT + threadIdx.x;
  if (r >= B) return;
  const int l0 = blockIdx.y * LPT;
  uint32_t ai[LPT];
  AC acc[LPT];
#pragma unroll
  for (int li = 0; li < LPT; ++li) {
    ai[li] = (l0 + li < loops) ? (uint32_t)ai_dev[l0 + li] : 0u;
    acc[li].init();
  }
  auto tap = [&](uint32_t id) {
    const T2 fv = f[id];
    T2 xv[LPT];
#pragma unroll
    for (int li = 0; li < LPT; ++li) xv[li] = x[(id * ai[li]) & mask];
#pragma unroll
    for (int li = 0; li < LPT; ++li) acc[li].add(xv[li], fv);
  };
  if (tiled) {
    const int T = w / B;
#pragma unroll 2
    for (int t = 0; t < T; ++t) tap((uint32_t)(t * B + r));
    if (r < w - T * B) tap((uint32_t)(r * T + B));
  } else {
#pragma unroll 2
    for (uint32_t id = (uint32_t)r; id < (uint32_t)w; id += (uint32_t)B) tap(id);
  }
#pragma unroll
  for (int li = 0; li < LPT; ++li)
    if (l0 + li < loops) bins[(long long)(l0 + li) * B + r] = acc[li].get();
}

hipError_t perm_filter_launch(int is_fp32, const void* x, const void* f, void* bins,
                              const int* ai_dev, int n, int B, int w, int loops, int tiled,
                              hipStream_t st) {
  constexpr int LPT = 3;
  dim3 grid((B + PF_NT - 1) / PF_NT, (loops + LPT - 1) / LPT);
  const uint32_t mask = (uint32_t)n - 1u;
  if (is_fp32)
    hipLaunchKernelGGL((perm_filter_kernel<float, LPT>), grid, dim3(PF_NT), 0, st,
                       (const float2*)x, (const float2*)f, (float2*)bins, ai_dev, mask, B, w,
                       loops, tiled);
  else
    hipLaunchKernelGGL((perm_filter_kernel<double, LPT>), grid, dim3(PF_NT), 0, st,
                       (const double2*)x, (const double2*)f, (double2*)bins, ai_dev, mask, B, w,
                       loops, tiled);
  return hipGetLastError();
}

// ============================================================================================
// block helpers
// ============================================================================================
// exclusive scan over a block of up to 1024 threads (wave64)
__device__ __forceinline__ int block_excl_scan(int v, int* sh) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[wid] = x;
  __syncthreads();
  int off = 0;
  for (int i = 0; i < wid; ++i) off += sh[i];
  __syncthreads();
  return off + x - v;
}

__device__ __forceinline__ int block_sum(int v, int* sh) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  int x = v;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  if (lane == 0) sh[wid] = x;
  __syncthreads();
  int tot = 0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nw; ++i) tot += sh[i];
  __syncthreads();
  return tot;
}

// ============================================================================================
// top-K select
// ============================================================================================
struct SelState {
  unsigned long long prefix;
  int remaining;
  int pad;
};

constexpr int SEL_NT = 512;
constexpr int SEL_BINS = 2048;

static int sel_chunk(int N) {
  int c = 8192;
  while ((long long)c * 256 < N) c <<= 1;
  return c;
}

size_t select_work_bytes(int segs, int N) {
  const int nchunks = (N + sel_chunk(N) - 1) / sel_chunk(N);
  return (size_t)segs * sizeof(SelState) + (size_t)segs * SEL_BINS * 4 + (size_t)segs * 4 +
         (size_t)segs * nchunks * 2 * 4 + 256;
}

void select_work_bind(SelectWork* w, void* base, int segs, int N) {
  const int nchunks = (N + sel_chunk(N) - 1) / sel_chunk(N);
  char* p = (char*)base;
  w->state = p;
  p += (size_t)segs * sizeof(SelState);
  w->hist = (unsigned*)p;
  p += (size_t)segs * SEL_BINS * 4;
  w->done = (unsigned*)p;
  p += (size_t)segs * 4;
  w->cnt = (int*)p;
  w->max_segs = segs;
  w->max_chunks = nchunks;
}

template <typename K>
__global__ __launch_bounds__(SEL_NT) void select_hist_kernel(
    const K* __restrict__ keys, long long seg_stride, int N, int chunk, int nchunks, int shift,
    int D, int first, int Ksel, SelState* __restrict__ st, unsigned* __restrict__ ghist,
    unsigned* __restrict__ done) {
  __shared__ unsigned h[SEL_BINS];
  __shared__ int s_last;
  __shared__ int sh[16];
  const int tid = threadIdx.x;
  const int seg = blockIdx.y;
  const int nb = 1 << D;
  for (int i = tid; i < nb; i += SEL_NT) h[i] = 0u;
  __syncthreads();
  const K* kp = keys + (long long)seg * seg_stride;
  const unsigned long long prefix = first ? 0ull : st[seg].prefix;
  const int hs = shift + D;
  const unsigned dmask = (unsigned)nb - 1u;
  const int begin = blockIdx.x * chunk;
  const int end = min(N, begin + chunk);
  for (int i = begin + tid; i < end; i += SEL_NT) {
    const K k = kp[i];
    const bool match = first || ((unsigned long long)(k >> hs) == prefix);
    if (match) atomicAdd(&h[(unsigned)(k >> shift) & dmask], 1u);
  }
  __syncthreads();
  unsigned* gh = ghist + (size_t)seg * SEL_BINS;
  for (int i = tid; i < nb; i += SEL_NT)
    if (h[i]) atomicAdd(&gh[i], h[i]);
  __threadfence();
  __syncthreads();
  if (tid == 0) s_last = (atomicAdd(&done[seg], 1u) == (unsigned)(nchunks - 1));
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  // last block of this segment: find the digit holding the remaining-th largest key
  const int rem = first ? Ksel : st[seg].remaining;
  unsigned vals[4];
  int S = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int idx = 4 * tid + j;  // rank from the top
    vals[j] = (idx < nb) ? __hip_atomic_load(&gh[nb - 1 - idx], __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT)
                         : 0u;
    S += (int)vals[j];
  }
  const int E = block_excl_scan(S, sh);
  if (E < rem && rem <= E + S) {
    int cum = E;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (cum + (int)vals[j] >= rem) {
        const unsigned digit = (unsigned)(nb - 1 - (4 * tid + j));
        st[seg].prefix = (prefix << D) | digit;
        st[seg].remaining = rem - cum;
        break;
      }
      cum += (int)vals[j];
    }
  }
  for (int i = tid; i < nb; i += SEL_NT) gh[i] = 0u;
  if (tid == 0) done[seg] = 0u;
}

template <typename K>
__global__ __launch_bounds__(SEL_NT) void select_count_kernel(const K* __restrict__ keys,
                                                              long long seg_stride, int N,
                                                              int chunk, int nchunks,
                                                              const SelState* __restrict__ st,
                                                              int* __restrict__ cnt) {
  __shared__ int sh[16];
  const int seg = blockIdx.y;
  const K T = (K)st[seg].prefix;
  const K* kp = keys + (long long)seg * seg_stride;
  const int begin = blockIdx.x * chunk;
  const int end = min(N, begin + chunk);
  int gt = 0, eq = 0;
  for (int i = begin + threadIdx.x; i < end; i += SEL_NT) {
    const K k = kp[i];
    gt += (k > T);
    eq += (k == T);
  }
  gt = block_sum(gt, sh);
  eq = block_sum(eq, sh);
  if (threadIdx.x == 0) {
    cnt[((long long)seg * nchunks + blockIdx.x) * 2 + 0] = gt;
    cnt[((long long)seg * nchunks + blockIdx.x) * 2 + 1] = eq;
  }
}

template <typename K>
__global__ __launch_bounds__(SEL_NT) void select_write_kernel(const K* __restrict__ keys,
                                                              long long seg_stride, int N,
                                                              int chunk, int nchunks, int Ksel,
                                                              const SelState* __restrict__ st,
                                                              const int* __restrict__ cnt,
                                                              int* __restrict__ out) {
  __shared__ int sh[16];
  __shared__ int s_base[2];
  const int tid = threadIdx.x;
  const int seg = blockIdx.y;
  const K T = (K)st[seg].prefix;
  const int need = st[seg].remaining;
  const K* kp = keys + (long long)seg * seg_stride;
  // chunk base counts
  int bg = 0, be = 0;
  for (int c = tid; c < (int)blockIdx.x; c += SEL_NT) {
    bg += cnt[((long long)seg * nchunks + c) * 2 + 0];
    be += cnt[((long long)seg * nchunks + c) * 2 + 1];
  }
  bg = block_sum(bg, sh);
  be = block_sum(be, sh);
  if (tid == 0) {
    s_base[0] = bg;
    s_base[1] = be;
  }
  __syncthreads();
  const int begin = blockIdx.x * chunk;
  const int per = chunk / SEL_NT;
  const int lo = begin + tid * per;
  const int hi = min(N, lo + per);
  int gt = 0, eq = 0;
  for (int i = lo; i < hi; ++i) {
    const K k = kp[i];
    gt += (k > T);
    eq += (k == T);
  }
  int G = s_base[0] + block_excl_scan(gt, sh);
  int Eq = s_base[1] + block_excl_scan(eq, sh);
  int* o = out + (long long)seg * Ksel;
  for (int i = lo; i < hi; ++i) {
    const K k = kp[i];
    if (k > T) {
      const int pos = G + min(Eq, need);
      if (pos < Ksel) o[pos] = i;
      ++G;
    } else if (k == T) {
      if (Eq < need && G + Eq < Ksel) o[G + Eq] = i;
      ++Eq;
    }
  }
}

template <typename K>
static hipError_t select_impl(const K* keys, long long seg_stride, int N, int Ksel, int segs,
                              int* out, const SelectWork& w, hipStream_t st) {
  const int chunk = sel_chunk(N);
  const int nchunks = (N + chunk - 1) / chunk;
  const int bits = (int)sizeof(K) * 8;
  dim3 grid(nchunks, segs);
  int hi = bits;
  bool first = true;
  while (hi > 0) {
    const int D = hi >= 11 ? 11 : hi;
    const int shift = hi - D;
    hipLaunchKernelGGL((select_hist_kernel<K>), grid, dim3(SEL_NT), 0, st, keys, seg_stride, N,
                       chunk, nchunks, shift, D, first ? 1 : 0, Ksel, (SelState*)w.state,
                       w.hist, w.done);
    first = false;
    hi = shift;
  }
  hipLaunchKernelGGL((select_count_kernel<K>), grid, dim3(SEL_NT), 0, st, keys, seg_stride, N,
                     chunk, nchunks, (const SelState*)w.state, w.cnt);
  hipLaunchKernelGGL((select_write_kernel<K>), grid, dim3(SEL_NT), 0, st, keys, seg_stride, N,
                     chunk, nchunks, Ksel, (const SelState*)w.state, w.cnt, out);
  return hipGetLastError();
}

hipError_t select_topk(int is_fp32, const void* keys, long long seg_stride, int N, int K,
                       int segs, int* out, const SelectWork& w, hipStream_t st) {
  if (is_fp32)
    return select_impl<unsigned int>((const unsigned int*)keys, seg_stride, N, K, segs, out, w,
                                     st);
  return select_impl<unsigned long long>((const unsigned long long*)keys, seg_stride, N, K,
                                         segs, out, w, st);
}

// ============================================================================================
// reverse hash
// ============================================================================================
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v, size_t words) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t s = (size_t)gridDim.x * blockDim.x;
  for (; i < words; i += s) p[i] = v;
}

hipError_t fill_u32(void* p, uint32_t v, size_t words, hipStream_t st) {
  if (words == 0) return hipSuccess;
  size_t blocks = (words + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (uint32_t*)p, v,
                     words);
  return hipGetLastError();
}

hipError_t hash_clear_launch(const HashArgs& h, hipStream_t st) {
  const size_t vote_words = ((size_t)h.n * h.bpe + 31) / 32;
  hipError_t e = fill_u32(h.vote, 0u, vote_words, st);
  if (e != hipSuccess) return e;
  e = fill_u32(h.bitmap, 0u, (size_t)h.loops_loc * h.bm_words, st);
  if (e != hipSuccess) return e;
  return fill_u32(h.maxfirst, 0xFFFFFFFFu, (size_t)h.loops_loc, st);
}

__global__ void hash_bitmap_kernel(HashArgs h) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int l = blockIdx.y;
  if (i >= h.nJ2) return;
  const uint64_t j2 = (uint64_t)h.J2[i];
  const uint32_t first = (uint32_t)((j2 * (uint64_t)(uint32_t)h.ai_dev[l]) & (uint64_t)(h.W - 1));
  atomicOr(&h.bitmap[(size_t)l * h.bm_words + (first >> 5)], 1u << (first & 31));
  atomicMax(&h.maxfirst[l], (int)first);
}

hipError_t hash_bitmap_launch(const HashArgs& h, hipStream_t st) {
  dim3 grid((h.nJ2 + 255) / 256, h.loops_loc);
  hipLaunchKernelGGL(hash_bitmap_kernel, grid, dim3(256), 0, st, h);
  return hipGetLastError();
}

// Enumerate the approved permuted positions p of bucket J's window [J m - m/2, J m + m/2)
// (mod n) in increasing circular order; the window is empty when (low mod W) exceeds the
// largest approved residue (the reference's findUpperBound never wraps then).  Calls
// fn(loc) with loc = p * a mod n.
template <typename F>
__device__ __forceinline__ void hash_walk(const HashArgs& h, int l, int i, F&& fn) {
  const long long n = h.n;
  const long long m = n / h.B;
  const long long J = h.J[(long long)l * h.K + i];
  long long low = J * m - (m >> 1);
  low = (low + n) & (n - 1);
  const int key = (int)(low & (h.W - 1));
  if (key > h.maxfirst[l]) return;
  const unsigned* bm = h.bitmap + (size_t)l * h.bm_words;
  const uint64_t a = (uint64_t)(uint32_t)h.a_dev[l];
  const uint64_t nmask = (uint64_t)n - 1;
  for (long long d = 0; d < m;) {
    const uint64_t p = (uint64_t)((low + d) & (n - 1));
    const uint32_t bpos = (uint32_t)(p & (uint64_t)(h.W - 1));
    const int off = (int)(bpos & 31u);
    const long long left = m - d;
    const int take = (left < 32 - off) ? (int)left : (32 - off);
    uint32_t word = bm[bpos >> 5] >> off;
    if (take < 32) word &= (1u << take) - 1u;
    while (word) {
      const int bit = __ffs(word) - 1;
      word &= word - 1u;
      const uint64_t pp = p + (uint64_t)bit;
      fn((uint32_t)((pp * a) & nmask));
    }
    d += take;
  }
}

__global__ void hash_vote_kernel(HashArgs h) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= h.loops_loc * h.K) return;
  const int l = idx / h.K, i = idx - l * h.K;
  const int lgepw = 5 - __ffs(h.bpe) + 1;  // log2(32 / bpe)
  const uint32_t epw_mask = (1u << lgepw) - 1u;
  hash_walk(h, l, i, [&](uint32_t loc) {
    atomicOr(&h.vote[loc >> lgepw], (1u << l) << ((loc & epw_mask) * (uint32_t)h.bpe));
  });
}

hipError_t hash_vote_launch(const HashArgs& h, hipStream_t st) {
  const int total = h.loops_loc * h.K;
  hipLaunchKernelGGL(hash_vote_kernel, dim3((total + 255) / 256), dim3(256), 0, st, h);
  return hipGetLastError();
}

__device__ __forceinline__ bool hash_is_hit(const HashArgs& h, int l, uint32_t loc, int lgepw,
                                            uint32_t epw_mask) {
  const uint32_t word = h.vote[loc >> lgepw];
  const uint32_t mask = (word >> ((loc & epw_mask) * (uint32_t)h.bpe)) &
                        (uint32_t)((1ull << h.bpe) - 1ull);
  return __popc(mask & ((2u << l) - 1u)) == h.loops_thresh;
}

constexpr int HASH_NT = 256;
int hash_blocks(int loops_loc, int K) { return (loops_loc * K + HASH_NT - 1) / HASH_NT; }

__global__ __launch_bounds__(HASH_NT) void hash_count_kernel(HashArgs h) {
  __shared__ int sh[16];
  const int idx = blockIdx.x * HASH_NT + threadIdx.x;
  const int lgepw = 5 - __ffs(h.bpe) + 1;
  const uint32_t epw_mask = (1u << lgepw) - 1u;
  int c = 0;
  if (idx < h.loops_loc * h.K) {
    const int l = idx / h.K, i = idx - l * h.K;
    hash_walk(h, l, i, [&](uint32_t loc) { c += hash_is_hit(h, l, loc, lgepw, epw_mask); });
  }
  const int ex = block_excl_scan(c, sh);
  if (idx < h.loops_loc * h.K) h.hitoff[idx] = ex;
  if (threadIdx.x == HASH_NT - 1) h.blocksum[blockIdx.x] = ex + c;
}

__global__ __launch_bounds__(1024) void hash_scan_kernel(HashArgs h, int nblocks) {
  __shared__ int sh[16];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nblocks; base += 1024) {
    const int i = base + threadIdx.x;
    const int v = (i < nblocks) ? h.blocksum[i] : 0;
    const int ex = block_excl_scan(v, sh);
    const int c0 = carry;
    if (i < nblocks) h.blockoff[i] = c0 + ex;
    __syncthreads();
    if (threadIdx.x == 1023) carry = c0 + ex + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) *h.count = carry;
}

__global__ __launch_bounds__(HASH_NT) void hash_write_kernel(HashArgs h) {
  const int idx = blockIdx.x * HASH_NT + threadIdx.x;
  if (idx >= h.loops_loc * h.K) return;
  const int lgepw = 5 - __ffs(h.bpe) + 1;
  const uint32_t epw_mask = (1u << lgepw) - 1u;
  const int l = idx / h.K, i = idx - l * h.K;
  int pos = h.blockoff[blockIdx.x] + h.hitoff[idx];
  hash_walk(h, l, i, [&](uint32_t loc) {
    if (hash_is_hit(h, l, loc, lgepw, epw_mask)) {
      if (pos < h.capacity) h.hits[pos] = (int)loc;
      ++pos;
    }
  });
}

hipError_t hash_hits_launch(const HashArgs& h, hipStream_t st) {
  const int nb = hash_blocks(h.loops_loc, h.K);
  hipLaunchKernelGGL(hash_count_kernel, dim3(nb), dim3(HASH_NT), 0, st, h);
  hipLaunchKernelGGL(hash_scan_kernel, dim3(1), dim3(1024), 0, st, h, nb);
  hipLaunchKernelGGL(hash_write_kernel, dim3(nb), dim3(HASH_NT), 0, st, h);
  return hipGetLastError();
}

// ============================================================================================
// estimation
// ============================================================================================
__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
__device__ __forceinline__ double2 to_d2(float2 v) {
  return make_double2((double)v.x, (double)v.y);
}

template <typename Real, int MAXL>
__global__ __launch_bounds__(256) void estimate_kernel(
    const int* __restrict__ hits, const int* __restrict__ count, int capacity,
    const typename CT<Real>::T2* __restrict__ bins, int loops, uint32_t n,
    const int* __restrict__ ai_dev, int B, const double2* __restrict__ win,
    Node* __restrict__ out) {
  const int total = min(*count, capacity);
  const long long m = (long long)(n / (uint32_t)B);
  const long long nn = n;
  const int med = (loops - 1) / 2;
  for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < total; h += gridDim.x * blockDim.x) {
    const uint32_t I = (uint32_t)hits[h];
    double vr[MAXL], vi[MAXL];
#pragma unroll
    for (int l = 0; l < MAXL; ++l) {
      vr[l] = 0.0;
      vi[l] = 0.0;
      if (l < loops) {
        const long long p = (long long)(((uint64_t)(uint32_t)ai_dev[l] * I) & (uint64_t)(n - 1));
        long long hashed = p / m;
        long long dist = p & (m - 1);
        if (dist > m / 2) {
          hashed = (hashed + 1) & (B - 1);
          dist -= m;
        }
        dist = (nn - dist) & (nn - 1);
        const double2 F = win[(dist + m / 2) & (nn - 1)];
        const double2 bb = to_d2(bins[(long long)l * B + hashed]);
        const double abs2 = fma(F.x, F.x, F.y * F.y);
        const double mr = fma(bb.x, F.x, bb.y * F.y);
        const double mi = fma(bb.y, F.x, -(bb.x * F.y));
        vr[l] = mr / abs2;
        vi[l] = mi / abs2;
      }
    }
    // stable rank selection: position of j in an insertion sort = #(v_i < v_j) + #(i<j, v_i==v_j)
    double rr = 0.0, ri = 0.0;
#pragma unroll
    for (int j = 0; j < MAXL; ++j) {
      if (j < loops) {
        int pr = 0, pi = 0;
#pragma unroll
        for (int i = 0; i < MAXL; ++i) {
          if (i < loops) {
            pr += (vr[i] < vr[j]) || (i < j && vr[i] == vr[j]);
            pi += (vi[i] < vi[j]) || (i < j && vi[i] == vi[j]);
          }
        }
        if (pr == med) rr = vr[j];
        if (pi == med) ri = vi[j];
      }
    }
    Node nd;
    nd.key = I;
    nd.value.x = rr;
    nd.value.y = ri;
    out[h] = nd;
  }
}

template <typename Real>
static hipError_t estimate_dispatch(const int* hits, const int* count, int capacity,
                                    const void* bins, int loops, int n, const int* ai_dev, int B,
                                    const double2* win, Node* out, hipStream_t st) {
  using T2 = typename CT<Real>::T2;
  const dim3 grid(512), block(256);
  if (loops <= 8)
    hipLaunchKernelGGL((estimate_kernel<Real, 8>), grid, block, 0, st, hits, count, capacity,
                       (const T2*)bins, loops, (uint32_t)n, ai_dev, B, win, out);
  else if (loops <= 16)
    hipLaunchKernelGGL((estimate_kernel<Real, 16>), grid, block, 0, st, hits, count, capacity,
                       (const T2*)bins, loops, (uint32_t)n, ai_dev, B, win, out);
  else
    hipLaunchKernelGGL((estimate_kernel<Real, 32>), grid, block, 0, st, hits, count, capacity,
                       (const T2*)bins, loops, (uint32_t)n, ai_dev, B, win, out);
  return hipGetLastError();
}

hipError_t estimate_launch(int is_fp32, const int* hits, const int* count, int capacity,
                           const void* bins_f, int loops, int n, const int* ai_dev, int B,
                           const double2* filter_win, Node* out, hipStream_t st) {
  if (is_fp32)
    return estimate_dispatch<float>(hits, count, capacity, bins_f, loops, n, ai_dev, B,
                                    filter_win, out, st);
  return estimate_dispatch<double>(hits, count, capacity, bins_f, loops, n, ai_dev, B,
                                   filter_win, out, st);
}

}  // namespace sfft
