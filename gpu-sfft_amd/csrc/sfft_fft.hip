// sfft_fft.hip -- hand-written CDNA4 (gfx950) kernels of the sparse-FFT hot path, part 1.
//
//  * fft_pass_kernel      canonical multi-pass big-radix Stockham FFT (LDS sub-FFT, radix-4/2
//                         stages), optional fused comb gather (RLFC) and |X|^2 key epilogue
//  * perm_filter_kernel   permutation + filtering: bucket-owner gather fold, all loops of a
//                         transform in one launch, fixed (deterministic) tap order
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
// |X|^2 as the reference's pow(re,2) + pow(im,2) (innerLoop.cu:715): each square rounded,
// then added
__device__ __forceinline__ unsigned long long sample_key(double2 v) {
  return (unsigned long long)__double_as_longlong(v.x * v.x + v.y * v.y);
}
__device__ __forceinline__ unsigned int sample_key(float2 v) {
  return __float_as_uint(v.x * v.x + v.y * v.y);
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
constexpr int FFT_QMAX = 16;                 // columns per tile
// complex elements per tile (LDS): 8192 complex64 (66 KB) or 4096 complex128 (66 KB) with
// 1024 threads, so that two workgroups (32 waves) fill a CU and one loads while the other
// computes (1024 threads rather than 512: the n = 2^27 comb and bucket FFT passes run 3-10%
// faster).
// The pass that gathers the comb samples from the signal (random lines) uses one 2048-point
// column per tile (32 KB FP64 / 16 KB FP32) with 512 threads (FP64: four workgroups per CU)
// or 256 (FP32: eight), so more gathers are in flight while other workgroups compute
// (measured at n = 2^27: 256 threads cost FP64 10 us, 1024 cost it 65 us; FP32 gains 6 us).
template <typename Real, int LOAD = 0>
struct Tile {
  static constexpr int cap = LOAD == 1 ? 2048 : (sizeof(Real) == 8 ? 4096 : 8192);
  // threads per workgroup
  static constexpr int nt = LOAD == 1 ? (sizeof(Real) == 8 ? 512 : 256) : 1024;
  static constexpr int ept = cap / nt;                 // elements per thread
};
static int tile_cap(int is_fp32, int load) {
  if (load == 1) return Tile<float, 1>::cap;
  return is_fp32 ? Tile<float>::cap : Tile<double>::cap;
}

// Workgroups are dealt to the 8 XCDs round-robin by launch order.  Tiles t, t+1, ... read
// neighbouring columns of the same 128-byte lines, so give the workgroups of one XCD (ids
// x, x+8, x+16, ...) consecutive tiles: the lines are then shared in that XCD's L2 instead
// of being fetched once per XCD.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
}

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
  int nbatch;
  int keys_tiled;   // store keys at tile * TE + e (the tile's own contiguous range)
  uint16_t* digits; // optional first digits of the keys (same positions)
};

template <typename T2>
__device__ __forceinline__ T2 outer_tw(const PassArgs& A, uint32_t e, int log2Lcur) {
  if (log2Lcur <= kMaxPassLog) return ((const T2*)A.tdir)[e];
  const uint32_t lo = e & ((1u << A.s) - 1u), hi = e >> A.s;
  return cmul(((const T2*)A.thi)[hi], ((const T2*)A.tlo)[lo]);
}

// LDS layout of a tile: column c at lds[c * RP + LPAD(v)] with one padding slot after every 16
// elements and a one-slot skew between columns, so the radix-4 stage's stride-4 and stride-16
// stores and the column-interleaved loads spread over the LDS banks.
#define LPAD(x) ((x) + ((x) >> 4))

template <typename T2, int EPT, int NT>
__device__ __forceinline__ void fft_stage4(T2* lds, int RP, int r, int log2ncol, int log2ns,
                                           const T2* __restrict__ twR, int inverse) {
  const int tid = threadIdx.x;
  const int Rq = 1 << (r - 2);
  const int nbf = (1 << log2ncol) * Rq;
  const int ns = 1 << log2ns;
  const int tstep = 1 << (r - log2ns - 2);
  T2 v[EPT];
#pragma unroll
  for (int kk = 0; kk < EPT / 4; ++kk) {
    const int bi = tid + kk * NT;
    if (bi < nbf) {
      const int c = bi >> (r - 2);
      const int j = bi & (Rq - 1);
      const T2* col = lds + c * RP;
      T2 a0 = col[LPAD(j)], a1 = col[LPAD(j + Rq)], a2 = col[LPAD(j + 2 * Rq)],
         a3 = col[LPAD(j + 3 * Rq)];
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
  for (int kk = 0; kk < EPT / 4; ++kk) {
    const int bi = tid + kk * NT;
    if (bi < nbf) {
      const int c = bi >> (r - 2);
      const int j = bi & (Rq - 1);
      T2* col = lds + c * RP;
      const int base = ((j >> log2ns) << (log2ns + 2)) + (j & (ns - 1));
      col[LPAD(base)] = v[4 * kk + 0];
      col[LPAD(base + ns)] = v[4 * kk + 1];
      col[LPAD(base + 2 * ns)] = v[4 * kk + 2];
      col[LPAD(base + 3 * ns)] = v[4 * kk + 3];
    }
  }
  __syncthreads();
}

template <typename T2, int EPT, int NT>
__device__ __forceinline__ void fft_stage2(T2* lds, int RP, int r, int log2ncol, int log2ns,
                                           const T2* __restrict__ twR) {
  const int tid = threadIdx.x;
  const int Rh = 1 << (r - 1);
  const int nbf = (1 << log2ncol) * Rh;
  const int ns = 1 << log2ns;
  const int tstep = 1 << (r - log2ns - 1);
  T2 v[EPT];
#pragma unroll
  for (int kk = 0; kk < EPT / 2; ++kk) {
    const int bi = tid + kk * NT;
    if (bi < nbf) {
      const int c = bi >> (r - 1);
      const int j = bi & (Rh - 1);
      const T2* col = lds + c * RP;
      T2 a0 = col[LPAD(j)], a1 = col[LPAD(j + Rh)];
      if (ns > 1) a1 = cmul(a1, twR[(j & (ns - 1)) * tstep]);
      v[2 * kk + 0] = cadd(a0, a1);
      v[2 * kk + 1] = csub(a0, a1);
    }
  }
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < EPT / 2; ++kk) {
    const int bi = tid + kk * NT;
    if (bi < nbf) {
      const int c = bi >> (r - 1);
      const int j = bi & (Rh - 1);
      T2* col = lds + c * RP;
      const int base = ((j >> log2ns) << (log2ns + 1)) + (j & (ns - 1));
      col[LPAD(base)] = v[2 * kk + 0];
      col[LPAD(base + ns)] = v[2 * kk + 1];
    }
  }
  __syncthreads();
}

template <typename Real, int LOAD, int STORE>
__global__ __launch_bounds__((Tile<Real, LOAD>::nt)) void fft_pass_kernel(PassArgs A) {
  using T2 = typename CT<Real>::T2;
  using KT = typename CT<Real>::K;
  constexpr int CAP = Tile<Real, LOAD>::cap, EPT = Tile<Real, LOAD>::ept;
  constexpr int NT = Tile<Real, LOAD>::nt;
  __shared__ T2 lds[CAP + CAP / 16 + FFT_QMAX];
  const int tid = threadIdx.x;
  // The comb-gathering pass runs all comb loops in one grid dimension, loop index fastest,
  // so the same tile of every loop is handled back to back on one XCD: comb loops whose
  // offsets fall in the same 128-byte window read the same lines, fetched once.
  int tile;
  long long b;
  if (LOAD == kLoadComb) {
    const int s = xcd_tile(blockIdx.x, gridDim.x);
    b = s % A.nbatch;
    tile = s / A.nbatch;
  } else {
    tile = xcd_tile(blockIdx.x, gridDim.x);
    b = blockIdx.y;
  }
  const int r = A.r;
  const int R = 1 << r;
  const int RP = R + (R >> 4) + 1;   // padded column stride (see LPAD)
  const int log2Lprev = A.log2Lprev;
  const int log2Lcur = log2Lprev + r;
  const long long Lprev = 1LL << log2Lprev;
  const long long nq = 1LL << (A.log2N - log2Lcur);
  const bool first = (log2Lprev == 0);
  int log2ncol;
  long long q0 = 0, q = 0, t0 = 0;
  if (first) {
    log2ncol = min(A.log2Q, A.log2N - log2Lcur);
    q0 = (long long)tile << log2ncol;
  } else {
    log2ncol = A.log2Q;
    const int log2tpq = log2Lprev - A.log2Q;
    q = (long long)tile >> log2tpq;
    t0 = ((long long)tile & ((1LL << log2tpq) - 1)) << A.log2Q;
  }
  const int ncol = 1 << log2ncol;
  const int TE = ncol << r;

  // ---- load (+ outer twiddle) into LDS, column c at lds[c*RP + v]
  const T2* __restrict__ in = (const T2*)A.in + b * A.in_bstride;
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int e = tid + k * NT;
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
      lds[c * RP + LPAD(v)] = val;
    }
  }
  __syncthreads();

  // ---- sub-FFT of every column: radix-4 stages, then one radix-2 stage if r is odd
  const T2* __restrict__ twR = (const T2*)A.twR;
  int log2ns = 0;
  int rem = r;
  while (rem >= 2) {
    fft_stage4<T2, EPT, NT>(lds, RP, r, log2ncol, log2ns, twR, A.inverse);
    log2ns += 2;
    rem -= 2;
  }
  if (rem == 1) fft_stage2<T2, EPT, NT>(lds, RP, r, log2ncol, log2ns, twR);

  // ---- store: out[q*Lcur + t + Lprev*k]
  T2* __restrict__ out = (T2*)A.out + b * A.out_bstride;
  KT* __restrict__ keys = (KT*)A.keys + b * A.keys_bstride;
  const bool wkeys = (STORE != kStoreCplx) && (b < A.key_batches);
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int e = tid + k * NT;
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
      const T2 val = lds[c * RP + LPAD(kk)];
      if (STORE != kStoreKeys) out[g] = val;
      if (wkeys) {
        const long long kpos = A.keys_tiled ? (long long)tile * TE + e : g;
        const KT key = sample_key(val);
        keys[kpos] = key;
        if (A.digits)
          A.digits[b * A.keys_bstride + kpos] =
              (uint16_t)((key >> (sizeof(KT) * 8 - 13)) & (KT)0xFFF);
      }
    }
  }
}

template <typename Real, int LOAD, int STORE>
static hipError_t launch_pass(const PassArgs& A, int tiles, int batch, hipStream_t st) {
  const dim3 grid = (LOAD == kLoadComb) ? dim3(tiles * batch, 1) : dim3(tiles, batch);
  hipLaunchKernelGGL((fft_pass_kernel<Real, LOAD, STORE>), grid, dim3(Tile<Real, LOAD>::nt), 0,
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

// log2 of the columns per tile of pass p (r bits, after log2Lprev bits): as many as fit the
// tile, up to qmax; it does not change the arithmetic
static int pass_log2q(const FftTables& T, const FftLaunch& L, int p, int r, int log2Lprev) {
  const int qmax = (L.qmax > 0 && L.qmax <= FFT_QMAX) ? L.qmax : FFT_QMAX;
  const int load = (p == 0) ? L.load : kLoadPlain;
  const int cap = tile_cap(T.is_fp32, load);
  int log2Q = 0;
  while ((1 << (log2Q + 1)) <= qmax && ((1 << (log2Q + 1)) << r) <= cap) ++log2Q;
  if (p > 0 && log2Q > log2Lprev) log2Q = log2Lprev;
  return log2Q;
}

KeyMap fft_keys_map(const FftLaunch& L) {
  KeyMap m;
  const FftTables& T = *L.tab;
  m.log2n = T.log2n;
  if (!L.keys_tiled || T.npass < 2) return m;
  const int last = T.npass - 1;
  const int r = T.logs[last];
  m.mid = r;
  m.lo = pass_log2q(T, L, last, r, T.log2n - r);
  return m;
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
    const int load = (p == 0) ? L.load : kLoadPlain;
    const int log2Q = pass_log2q(T, L, p, r, log2Lprev);
    int tiles;
    if (p == 0) {
      const int lnq = log2N - log2Lcur;
      const int lq = log2Q < lnq ? log2Q : lnq;
      tiles = 1 << (lnq - lq);
    } else {
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
    A.nbatch = L.batch;
    A.keys_tiled = (last && p > 0 && L.keys_tiled) ? 1 : 0;
    A.digits = last ? L.digits : nullptr;
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
// The term as the reference writes it (innerLoop.cu:740-741, 770-771): both products
// rounded, combined, then added to the bucket sum.
struct Acc64 {
  double2 s;
  __device__ __forceinline__ void init() { s = make_double2(0.0, 0.0); }
  __device__ __forceinline__ void add(double2 x, double2 f) {
    const double tx = x.x * f.x - x.y * f.y;
    const double ty = x.x * f.y + x.y * f.x;
    s.x = s.x + tx;
    s.y = s.y + ty;
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
    const float tx = x.x * f.x - x.y * f.y;
    const float ty = x.x * f.y + x.y * f.x;
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

// bins[l*B + r] = sum over the taps id of bucket r (canonical order) of v_l(id) * f[id], where
// v_l(id) = x[(id * ai_l) mod n] gathered from the signal (TAPS = false), or read from the
// tap-ordered staging array T[l*w + id] that the streaming pass filled (TAPS = true; same
// values, coalesced reads).  One thread owns bucket r of LPT loops l0 .. l0+LPT-1.  Every
// launch shape (single transform, batch, any LPT or block size) runs this body, so the sums
// are the same bits whatever the shape.
template <typename Real, int LPT, bool TAPS>
__device__ __forceinline__ void pf_body(const typename CT<Real>::T2* __restrict__ x,
                                        const typename CT<Real>::T2* __restrict__ f,
                                        typename CT<Real>::T2* __restrict__ bins,
                                        const int* __restrict__ ai_dev, uint32_t mask, int B,
                                        int w, int loops, int tiled, int r, int l0) {
  using T2 = typename CT<Real>::T2;
  using AC = typename AccOf<Real>::A;
  if (r >= B) return;
  uint32_t ai[LPT];
  const T2* __restrict__ src[LPT];
  AC acc[LPT];
#pragma unroll
  for (int li = 0; li < LPT; ++li) {
    ai[li] = (l0 + li < loops) ? (uint32_t)ai_dev[l0 + li] : 0u;
    src[li] = x + (size_t)min(l0 + li, loops - 1) * (size_t)w;   // staging rows (TAPS)
    acc[li].init();
  }
  auto tap = [&](uint32_t id) {
    const T2 fv = f[id];
    T2 xv[LPT];
#pragma unroll
    for (int li = 0; li < LPT; ++li) xv[li] = TAPS ? src[li][id] : x[(id * ai[li]) & mask];
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

template <typename Real, int LPT, bool TAPS>
__global__ __launch_bounds__(PF_NT) void perm_filter_kernel(
    const typename CT<Real>::T2* __restrict__ x, const typename CT<Real>::T2* __restrict__ f,
    typename CT<Real>::T2* __restrict__ bins, const int* __restrict__ ai_dev, uint32_t mask,
    int B, int w, int loops, int tiled) {
  pf_body<Real, LPT, TAPS>(x, f, bins, ai_dev, mask, B, w, loops, tiled,
                           blockIdx.x * PF_NT + threadIdx.x, blockIdx.y * LPT);
}

template <int LPT, bool TAPS>
static hipError_t pf_launch_lpt(int is_fp32, const void* src, const void* f, void* bins,
                                const int* ai_dev, int n, int B, int w, int loops, int tiled,
                                hipStream_t st) {
  dim3 grid((B + PF_NT - 1) / PF_NT, (loops + LPT - 1) / LPT);
  const uint32_t mask = (uint32_t)n - 1u;
  if (is_fp32)
    hipLaunchKernelGGL((perm_filter_kernel<float, LPT, TAPS>), grid, dim3(PF_NT), 0, st,
                       (const float2*)src, (const float2*)f, (float2*)bins, ai_dev, mask, B, w,
                       loops, tiled);
  else
    hipLaunchKernelGGL((perm_filter_kernel<double, LPT, TAPS>), grid, dim3(PF_NT), 0, st,
                       (const double2*)src, (const double2*)f, (double2*)bins, ai_dev, mask, B,
                       w, loops, tiled);
  return hipGetLastError();
}

// Loops per thread: as many as possible up to 9, split evenly.  One thread then reads each
// filter tap once for all its loops and keeps up to 18 independent gathers in flight (the tap
// loop is unrolled twice); at n = 2^27, k = 16000 (9 loops) this is 7% faster than 3 loops per
// thread.  The sums do not depend on the split.
static int pf_loops_per_thread(int loops) {
  const int groups = (loops + 8) / 9;
  return (loops + groups - 1) / groups;
}

template <bool TAPS>
static hipError_t pf_launch(int is_fp32, const void* src, const void* f, void* bins,
                            const int* ai_dev, int n, int B, int w, int loops, int tiled,
                            hipStream_t st) {
  switch (pf_loops_per_thread(loops)) {
#define SFFT_PF_LPT(k) \
  case k:              \
    return pf_launch_lpt<k, TAPS>(is_fp32, src, f, bins, ai_dev, n, B, w, loops, tiled, st);
    SFFT_PF_LPT(1) SFFT_PF_LPT(2) SFFT_PF_LPT(3) SFFT_PF_LPT(4) SFFT_PF_LPT(5)
    SFFT_PF_LPT(6) SFFT_PF_LPT(7) SFFT_PF_LPT(8) SFFT_PF_LPT(9)
#undef SFFT_PF_LPT
    default:
      return hipErrorInvalidValue;
  }
}

// Batch: the permute+filter of every signal of a batch in one persistent launch.  Work item
// i of job j (item0_j <= i < item0_{j+1}) is bucket block i % nbx of loop group i / nbx;
// the workgroups walk the items with a grid stride (every workgroup runs the same number of
// items +-1, whatever the signal sizes).
template <typename Real, int LPT, int NT>
__global__ __launch_bounds__(NT) void perm_filter_batch_kernel(const PfJob* __restrict__ jobs,
                                                               int njobs, int total) {
  using T2 = typename CT<Real>::T2;
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (jobs[mid].item0 <= item) lo = mid;
      else hi = mid - 1;
    }
    const PfJob& J = jobs[lo];
    const int local = item - J.item0;
    const int bx = local % J.nbx, gy = local / J.nbx;
    pf_body<Real, LPT, false>((const T2*)J.x, (const T2*)J.f, (T2*)J.bins, J.ai, J.mask, J.B,
                              J.w, J.loops, J.tiled, bx * NT + threadIdx.x, gy * LPT);
  }
}

int pf_batch_items(int B, int loops, int lpt, int nt, int* nbx) {
  *nbx = (B + nt - 1) / nt;
  return *nbx * ((loops + lpt - 1) / lpt);
}

template <typename Real, int LPT>
static hipError_t pfb_nt(const PfJob* jobs, int njobs, int total, int nt, int grid,
                         hipStream_t st) {
  if (nt == 128)
    hipLaunchKernelGGL((perm_filter_batch_kernel<Real, LPT, 128>), dim3(grid), dim3(128), 0, st,
                       jobs, njobs, total);
  else if (nt == 256)
    hipLaunchKernelGGL((perm_filter_batch_kernel<Real, LPT, 256>), dim3(grid), dim3(256), 0, st,
                       jobs, njobs, total);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template <typename Real>
static hipError_t pfb_lpt(const PfJob* jobs, int njobs, int total, int lpt, int nt, int grid,
                          hipStream_t st) {
  switch (lpt) {
    case 1: return pfb_nt<Real, 1>(jobs, njobs, total, nt, grid, st);
    case 2: return pfb_nt<Real, 2>(jobs, njobs, total, nt, grid, st);
    case 3: return pfb_nt<Real, 3>(jobs, njobs, total, nt, grid, st);
    case 4: return pfb_nt<Real, 4>(jobs, njobs, total, nt, grid, st);
    case 5: return pfb_nt<Real, 5>(jobs, njobs, total, nt, grid, st);
    case 6: return pfb_nt<Real, 6>(jobs, njobs, total, nt, grid, st);
    case 7: return pfb_nt<Real, 7>(jobs, njobs, total, nt, grid, st);
    case 8: return pfb_nt<Real, 8>(jobs, njobs, total, nt, grid, st);
    case 9: return pfb_nt<Real, 9>(jobs, njobs, total, nt, grid, st);
  }
  return hipErrorInvalidValue;
}

hipError_t perm_filter_batch_launch(int is_fp32, const PfJob* d_jobs, int njobs, int total,
                                    int lpt, int nt, int grid, hipStream_t st) {
  if (njobs <= 0 || total <= 0) return hipSuccess;
  if (grid > total) grid = total;
  return is_fp32 ? pfb_lpt<float>(d_jobs, njobs, total, lpt, nt, grid, st)
                 : pfb_lpt<double>(d_jobs, njobs, total, lpt, nt, grid, st);
}

hipError_t perm_filter_launch(int is_fp32, const void* x, const void* f, void* bins,
                              const int* ai_dev, int n, int B, int w, int loops, int tiled,
                              hipStream_t st) {
  return pf_launch<false>(is_fp32, x, f, bins, ai_dev, n, B, w, loops, tiled, st);
}

hipError_t fold_taps_launch(int is_fp32, const void* T, const void* f, void* bins,
                            const int* ai_dev, int n, int B, int w, int loops, int tiled,
                            hipStream_t st) {
  return pf_launch<true>(is_fp32, T, f, bins, ai_dev, n, B, w, loops, tiled, st);
}

// ============================================================================================
// streaming pass: read the signal once in address order and route every element to the tap
// slots and comb samples that use it:
//   T[l*w + i] = x[m]  for each loop l with i = (m * a_l) mod n < w   (i.e. m = i * ai_l)
//   y[r*W + j] = x[m]  for each comb r with m = off_r + j * sigma
// Each thread handles SS_EPT consecutive elements; i advances by a_l per element (no multiply).
// ============================================================================================
constexpr int SS_NT = 256;
constexpr int SS_EPT = 4;

template <typename T2>
__global__ __launch_bounds__(SS_NT) void stream_scatter_kernel(
    const T2* __restrict__ x, T2* __restrict__ T, T2* __restrict__ y,
    const int* __restrict__ a_dev, const int* __restrict__ offs, uint32_t n, uint32_t w,
    int loops, int Rl, int log2sigma, uint32_t W) {
  const uint32_t mask = n - 1u;
  const uint32_t m0 = (blockIdx.x * SS_NT + threadIdx.x) * SS_EPT;
  if (m0 >= n) return;
  T2 v[SS_EPT];
#pragma unroll
  for (int e = 0; e < SS_EPT; ++e) v[e] = x[m0 + e];
  for (int l = 0; l < loops; ++l) {
    const uint32_t a = (uint32_t)a_dev[l];
    uint32_t i = m0 * a;
    T2* __restrict__ Tl = T + (size_t)l * w;
#pragma unroll
    for (int e = 0; e < SS_EPT; ++e) {
      const uint32_t ii = i & mask;
      if (ii < w) Tl[ii] = v[e];
      i += a;
    }
  }
  const uint32_t smask = (1u << log2sigma) - 1u;
  for (int r = 0; r < Rl; ++r) {
    const uint32_t d0 = m0 - (uint32_t)offs[r];   // wraps for m < off: never a multiple
#pragma unroll
    for (int e = 0; e < SS_EPT; ++e) {
      const uint32_t d = d0 + (uint32_t)e;
      if ((d & smask) == 0u) y[(size_t)r * W + (d >> log2sigma)] = v[e];
    }
  }
}

hipError_t stream_scatter_launch(int is_fp32, const void* x, void* T, void* y, const int* a_dev,
                                 const int* offs, int n, int w, int loops, int Rl, int W,
                                 hipStream_t st) {
  int log2sigma = 0;
  while ((W << log2sigma) < n) ++log2sigma;
  const unsigned blocks = (unsigned)(((long long)n / SS_EPT + SS_NT - 1) / SS_NT);
  if (is_fp32)
    hipLaunchKernelGGL((stream_scatter_kernel<float2>), dim3(blocks), dim3(SS_NT), 0, st,
                       (const float2*)x, (float2*)T, (float2*)y, a_dev, offs, (uint32_t)n,
                       (uint32_t)w, loops, Rl, log2sigma, (uint32_t)W);
  else
    hipLaunchKernelGGL((stream_scatter_kernel<double2>), dim3(blocks), dim3(SS_NT), 0, st,
                       (const double2*)x, (double2*)T, (double2*)y, a_dev, offs, (uint32_t)n,
                       (uint32_t)w, loops, Rl, log2sigma, (uint32_t)W);
  return hipGetLastError();
}

}  // namespace sfft
