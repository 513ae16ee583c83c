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
#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

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

// complex element as a native vector (arrays of HIP's double2/float2 wrappers end up in scratch)
template <typename Real>
struct NV {
  typedef Real type __attribute__((ext_vector_type(2)));
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
// columns per tile: 32, so that the sub-FFTs of 256 points (FP32) or 128 (FP64) and fewer
// also fill a tile
constexpr int FFT_QMAX = 32;
// complex elements per tile (LDS): 8192 complex64 (66 KB) or 4096 complex128 (66 KB) with
// 1024 threads, so that two workgroups (32 waves) fill a CU and one loads while the other
// computes (1024 threads rather than 512: the n = 2^27 comb and bucket FFT passes run 3-10%
// faster).
// The pass that gathers the comb samples from the signal (random lines) uses one 2048-point
// column per tile (32 KB FP64 / 16 KB FP32) with 512 threads (FP64: four workgroups per CU)
// or 256 (FP32: eight), so more gathers are in flight while other workgroups compute
// (measured at n = 2^27: 256 threads cost FP64 10 us, 1024 cost it 65 us; FP32 gains 6 us).
template <typename Real, int LOAD = 0>
struct Tile {   // LOAD: kLoadComb (gathers from x); kLoadPlain, kLoadCombRows, kLoadPerm (plain)
  static constexpr int cap = LOAD == 1 ? 2048 : (sizeof(Real) == 8 ? 4096 : 8192);
  // threads per workgroup; FP32 plain tiles: 512, so that a radix-16 step's 16 complex64
  // values and their twiddles fit the 128 registers of two 512-thread workgroups per CU
  // (FP64 keeps 1024 threads and single radix-4 stages: measured faster at 2^22)
  static constexpr int nt = LOAD == 1 ? (sizeof(Real) == 8 ? 512 : 256) : (sizeof(Real) == 8 ? 1024 : 512);
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
  const void* tw4;      // radix-4 stage twiddle quads (see fft_tables_create)
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
  int log2nlo, log2m;   // kLoadCombRows: g at in[(g mod 2^log2nlo) << log2m | g >> log2nlo]
  const int* perm;      // kLoadPerm: g of batch b at in[g * perm[b] mod 2^log2N]
  const int* perm_inv;  // kLoadPerm: the inverse of perm[b] mod 2^log2N
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

// The twiddles of a stage's butterflies for this thread: radix-4 stage with ns = 4^k > 1 (its
// quads {W^jm, W^2jm, W^3jm}), loaded one stage ahead so that their latency hides behind the
// stage before.  A butterfly's j (and so its slot here) does not depend on the stage.
template <typename T2, int EPT, int NT>
__device__ __forceinline__ void stage4_twiddles(T2 (&w)[EPT / 4][3], int r, int log2ncol,
                                                int log2ns, const T2* __restrict__ tw4) {
  const int Rq = 1 << (r - 2);
  const int nbf = (1 << log2ncol) * Rq;
  const int ns = 1 << log2ns;
  const T2* __restrict__ q4 = tw4 + 4 * ((ns - 4) / 3);   // this stage's quads
#pragma unroll
  for (int kk = 0; kk < EPT / 4; ++kk) {
    const int bi = min((int)threadIdx.x + kk * NT, nbf - 1);
    const T2* q = q4 + 4 * ((bi & (Rq - 1)) & (ns - 1));
    w[kk][0] = q[0];
    w[kk][1] = q[1];
    w[kk][2] = q[2];
  }
}

template <typename T2, int EPT, int NT, bool TWID>
__device__ __forceinline__ void fft_stage4(T2* lds, int RP, int r, int log2ncol, int log2ns,
                                           const T2 (&w)[EPT / 4][3], int inverse) {
  // Butterfly indices past the tile (partial tiles) are clamped to the last butterfly, whose
  // thread stores the same values: no per-butterfly branches, so every LDS read of the thread
  // is in flight together.  TWID: ns > 1 (twiddles w from stage4_twiddles).
  const int tid = threadIdx.x;
  const int Rq = 1 << (r - 2);
  const int nbf = (1 << log2ncol) * Rq;
  const int ns = 1 << log2ns;
  T2 v[EPT];
#pragma unroll
  for (int kk = 0; kk < EPT / 4; ++kk) {
    const int bi = min(tid + kk * NT, nbf - 1);
    const int c = bi >> (r - 2);
    const int j = bi & (Rq - 1);
    const T2* col = lds + c * RP;
    T2 a0 = col[LPAD(j)], a1 = col[LPAD(j + Rq)], a2 = col[LPAD(j + 2 * Rq)],
       a3 = col[LPAD(j + 3 * Rq)];
    if (TWID) {
      a1 = cmul(a1, w[kk][0]);
      a2 = cmul(a2, w[kk][1]);
      a3 = cmul(a3, w[kk][2]);
    }
    bfly4(a0, a1, a2, a3, inverse);
    v[4 * kk + 0] = a0;
    v[4 * kk + 1] = a1;
    v[4 * kk + 2] = a2;
    v[4 * kk + 3] = a3;
  }
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < EPT / 4; ++kk) {
    const int bi = min(tid + kk * NT, nbf - 1);
    const int c = bi >> (r - 2);
    const int j = bi & (Rq - 1);
    T2* col = lds + c * RP;
    const int base = ((j >> log2ns) << (log2ns + 2)) + (j & (ns - 1));
    col[LPAD(base)] = v[4 * kk + 0];
    col[LPAD(base + ns)] = v[4 * kk + 1];
    col[LPAD(base + 2 * ns)] = v[4 * kk + 2];
    col[LPAD(base + 3 * ns)] = v[4 * kk + 3];
  }
  __syncthreads();
}

// Two radix-4 stages (spans ns and 4ns) in registers: task (column c, tix = a0*ns + jm), one
// per thread, reads the 16 elements tix + u*R/16 (u = t + 4m: input m of the first stage's
// butterfly t, whose index is tix + t*R/16), runs the four first-stage butterflies (twiddles
// W_{4ns}^{m*jm}, the same for every t), then the four second-stage butterflies (butterfly m
// takes the first stage's outputs m of t = 0..3; twiddles W_{16ns}^{t*(m*ns + jm)}) and writes
// output m'' of butterfly m to a0*16ns + jm + (m + 4m'')*ns.  These are the two stages' own
// operations in their own order (the same bits), with one LDS round trip instead of two.
// the two stages of a radix-16 step on its 16 elements in registers (see fft_stage16)
template <typename T2, bool TW1>
__device__ __forceinline__ void stage16_regs(T2 (&x)[16], int jm, int ns,
                                             const T2* __restrict__ tw4, int inverse) {
  if (TW1) {
    const T2* w = tw4 + 4 * ((ns - 4) / 3 + jm);
    const T2 w1 = w[0], w2 = w[1], w3 = w[2];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      x[t + 4] = cmul(x[t + 4], w1);
      x[t + 8] = cmul(x[t + 8], w2);
      x[t + 12] = cmul(x[t + 12], w3);
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) bfly4(x[t], x[t + 4], x[t + 8], x[t + 12], inverse);
  // second stage: butterfly m takes x[4m' ... ] -> inputs m' = t: x[t + 4m]
  const T2* q2 = tw4 + 4 * ((4 * ns - 4) / 3 + jm);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const T2* w = q2 + 4 * m * ns;
    x[1 + 4 * m] = cmul(x[1 + 4 * m], w[0]);
    x[2 + 4 * m] = cmul(x[2 + 4 * m], w[1]);
    x[3 + 4 * m] = cmul(x[3 + 4 * m], w[2]);
    bfly4(x[4 * m], x[1 + 4 * m], x[2 + 4 * m], x[3 + 4 * m], inverse);
  }
}

template <typename T2, int NT, bool TW1>
__device__ __forceinline__ void fft_stage16(T2* lds, int RP, int r, int log2ncol, int log2ns,
                                            const T2* __restrict__ tw4, int inverse) {
  const int tid = threadIdx.x;
  const int R16 = 1 << (r - 4);
  const int ntask = (1 << log2ncol) * R16;
  const int ns = 1 << log2ns;
  T2 x[16];
  int base = 0, c = 0;
  if (tid < ntask) {
    c = tid >> (r - 4);
    const int tix = tid & (R16 - 1);
    const int jm = tix & (ns - 1);
    base = ((tix >> log2ns) << (log2ns + 4)) + jm;
    const T2* col = lds + c * RP;
#pragma unroll
    for (int u = 0; u < 16; ++u) x[u] = col[LPAD(tix + u * R16)];
    stage16_regs<T2, TW1>(x, jm, ns, tw4, inverse);
  }
  __syncthreads();
  if (tid < ntask) {
    T2* col = lds + c * RP;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int m2 = 0; m2 < 4; ++m2) col[LPAD(base + (m + 4 * m2) * ns)] = x[m2 + 4 * m];
  }
  __syncthreads();
}

// The last radix-4 stage (span ns = R/8) and the radix-2 stage after it (span R/2) in
// registers: task (column c, jm < R/8) reads the 8 elements jm + i*R/8, runs the radix-4
// butterflies j = jm (inputs i = 0, 2, 4, 6) and j = jm + R/8 (inputs i = 1, 3, 5, 7, with the
// same twiddles W_{4ns}^{m*jm}), whose outputs k belong at jm + k*R/8 and R/2 + jm + k*R/8,
// then the four radix-2 butterflies j' = jm + k*R/8 (twiddle W_R^{j'}) on those pairs, whose
// outputs belong at j' and j' + R/2.  The two stages' own operations in their own order (the
// same bits), with one LDS round trip instead of two.  A task writes exactly the 8 elements it
// read, so it stores each output as soon as it has it (no barrier between its reads and its
// writes).  So every task runs exactly once: a second thread on the same elements (a clamped
// duplicate) could read them after their owner has rewritten them.  A thread takes up to
// EPT/8 tasks (one when EPT < 8; threads past the tasks idle, whole waves at a time).
template <typename T2, int EPT, int NT, bool TW1>
__device__ __forceinline__ void fft_stage8(T2* lds, int RP, int r, int log2ncol,
                                           const T2* __restrict__ tw4,
                                           const T2* __restrict__ twR, int inverse) {
  constexpr int TPT = EPT >= 8 ? EPT / 8 : 1;
  const int tid = threadIdx.x;
  const int R8 = 1 << (r - 3);
  const int ntask = (1 << log2ncol) * R8;
#pragma unroll
  for (int kk = 0; kk < TPT; ++kk) {
    const int ti = tid + kk * NT;
    if (ti < ntask) {
      const int jm = ti & (R8 - 1);
      T2* col = lds + (ti >> (r - 3)) * RP;
      T2 x[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) x[i] = col[LPAD(jm + i * R8)];
      if (TW1) {
        const T2* w = tw4 + 4 * ((R8 - 4) / 3 + jm);
        const T2 w1 = w[0], w2 = w[1], w3 = w[2];
        x[2] = cmul(x[2], w1);
        x[4] = cmul(x[4], w2);
        x[6] = cmul(x[6], w3);
        x[3] = cmul(x[3], w1);
        x[5] = cmul(x[5], w2);
        x[7] = cmul(x[7], w3);
      }
      bfly4(x[0], x[2], x[4], x[6], inverse);
      bfly4(x[1], x[3], x[5], x[7], inverse);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const T2 a1 = cmul(x[2 * k + 1], twR[jm + k * R8]);
        col[LPAD(jm + k * R8)] = cadd(x[2 * k], a1);
        col[LPAD(jm + (4 + k) * R8)] = csub(x[2 * k], a1);
      }
    }
  }
  __syncthreads();
}

template <typename T2, int EPT, int NT>
__device__ __forceinline__ void fft_stage2(T2* lds, int RP, int r, int log2ncol, int log2ns,
                                           const T2* __restrict__ twR) {
  // (clamped butterfly indices as in fft_stage4)
  const int tid = threadIdx.x;
  const int Rh = 1 << (r - 1);
  const int nbf = (1 << log2ncol) * Rh;
  const int ns = 1 << log2ns;
  const int tstep = 1 << (r - log2ns - 1);
  T2 v[EPT];
  if (ns > 1) {
#pragma unroll
    for (int kk = 0; kk < EPT / 2; ++kk) {
      const int bi = min(tid + kk * NT, nbf - 1);
      const int c = bi >> (r - 1);
      const int j = bi & (Rh - 1);
      const T2* col = lds + c * RP;
      T2 a0 = col[LPAD(j)], a1 = col[LPAD(j + Rh)];
      a1 = cmul(a1, twR[(j & (ns - 1)) * tstep]);
      v[2 * kk + 0] = cadd(a0, a1);
      v[2 * kk + 1] = csub(a0, a1);
    }
  } else {
#pragma unroll
    for (int kk = 0; kk < EPT / 2; ++kk) {
      const int bi = min(tid + kk * NT, nbf - 1);
      const int c = bi >> (r - 1);
      const int j = bi & (Rh - 1);
      const T2* col = lds + c * RP;
      const T2 a0 = col[LPAD(j)], a1 = col[LPAD(j + Rh)];
      v[2 * kk + 0] = cadd(a0, a1);
      v[2 * kk + 1] = csub(a0, a1);
    }
  }
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < EPT / 2; ++kk) {
    const int bi = min(tid + kk * NT, nbf - 1);
    const int c = bi >> (r - 1);
    const int j = bi & (Rh - 1);
    T2* col = lds + c * RP;
    const int base = ((j >> log2ns) << (log2ns + 1)) + (j & (ns - 1));
    col[LPAD(base)] = v[2 * kk + 0];
    col[LPAD(base + ns)] = v[2 * kk + 1];
  }
  __syncthreads();
}

// TW: the pass's outer twiddles, known at compile time so that the load phase is straight-line
// code: 0 none (the first pass), 1 direct table entries, 2 products of a high and a low entry
enum PassTw { kTwNone = 0, kTwDirect = 1, kTwSplit = 2 };

template <typename Real, int LOAD, int STORE, int TW>
// two workgroups per CU for the plain tiles: 8 waves per SIMD at 1024 threads, 4 at 512
__global__ __launch_bounds__((Tile<Real, LOAD>::nt),
                             (LOAD == 1 ? 1 : Tile<Real, LOAD>::nt / 128)) void fft_pass_kernel(
    PassArgs A) {
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
  constexpr bool first = (TW == kTwNone);   // log2Lprev == 0
  int log2ncol;
  long long q0 = 0, q = 0, t0 = 0;
  // kLoadCombRowsG (comb rows with nq = 2^d * nlo sub-FFTs, 0 < d <= log2ncol): sub-FFT q
  // reads every 2^d-th element of row q mod nlo, interleaved with its siblings q + j*nlo
  // (j < 2^d).  A tile then takes 2^d sibling groups of ncol/2^d adjacent sub-FFTs, so that
  // it reads whole lines: tile column c is sub-FFT q0 + qoff(c) (the same sub-FFTs and values,
  // grouped differently).
  constexpr bool rows = (LOAD == kLoadCombRows || LOAD == kLoadCombRowsG);
  int log2cpj = 0;
  if (first) {
    log2ncol = min(A.log2Q, A.log2N - log2Lcur);
    log2cpj = log2ncol;
    if (LOAD == kLoadCombRowsG) log2cpj = log2ncol - ((A.log2N - log2Lcur) - A.log2nlo);
    q0 = (long long)tile << log2cpj;
  } else {
    log2ncol = A.log2Q;
    const int log2tpq = log2Lprev - A.log2Q;
    q = (long long)tile >> log2tpq;
    t0 = ((long long)tile & ((1LL << log2tpq) - 1)) << A.log2Q;
  }
  const int ncol = 1 << log2ncol;
  const int TE = ncol << r;
  // element (c, v) of the tile: first pass g = q0 + c + nq*v, later passes
  // g = (q + nq*v)*Lprev + t0 + c -- both gbase + c + vstride*v (no per-element branches)
  const long long gbase = first ? q0 : q * Lprev + t0;
  const long long vstride = first ? nq : nq * Lprev;
  auto qoff = [&](int c) __attribute__((always_inline)) {
    return (long long)(c & ((1 << log2cpj) - 1)) + ((long long)(c >> log2cpj) << A.log2nlo);
  };

  // ---- load (+ outer twiddle) into LDS, column c at lds[c*RP + v].  Every element load of
  // the thread (and every twiddle load) is issued before any is used: in one basic block,
  // with out-of-tile elements clamped to a valid address, so the loads are in flight together
  // (guarding each load with its own branch made the compiler wait for each one in turn).
  using V2 = typename NV<Real>::type;
  const T2* __restrict__ in = (const T2*)A.in + b * A.in_bstride;
  const T2* __restrict__ cx =
      (LOAD == kLoadComb) ? (const T2*)A.comb_x + A.comb_offs[b] : (const T2*)nullptr;
  // kLoadPerm: tile column c takes the input positions p = q0 + c + nq*X (X < R), so that the
  // tile's ncol columns read whole lines.  Position p holds element g = p * pinv mod N of the
  // canonical order (perm[b] * pinv = 1 mod N), which is element v = g / nq of sub-FFT
  // g mod nq = (q0 + c) * pinv mod nq: the column is that sub-FFT, its elements land at their
  // own v (the same sub-FFTs and values as loading g at g * perm[b], grouped differently).
  const uint32_t pinv = (LOAD == kLoadPerm) ? (uint32_t)A.perm_inv[b] : 0u;
  const uint32_t nmask = (1u << A.log2N) - 1u;
  const int log2nq = A.log2N - log2Lcur;
  auto perm_v = [&](int e) __attribute__((always_inline)) {
    const uint32_t pos = (uint32_t)q0 + (uint32_t)(e & (ncol - 1)) +
                         ((uint32_t)(e >> log2ncol) << log2nq);
    return (int)(((pos * pinv) & nmask) >> log2nq);
  };
  // loads in flight per batch: the plain tiles take batches of 8 bytes x 4 (FP32, 512
  // threads) or 16 bytes x 2 (FP64, 1024 threads: within 64 registers); the comb-gathering
  // pass issues all of its loads together
  constexpr int LB = (LOAD == kLoadComb) ? EPT : (sizeof(Real) == 8 ? 2 : 4);
  constexpr bool split = (TW == kTwSplit);   // log2Lcur > kMaxPassLog
  const T2* __restrict__ twR = (const T2*)A.twR;
  const T2* __restrict__ tw4 = (const T2*)A.tw4;
  // batch kb: loads elements e = tid + (kb + k)*NT (k < LB), with the outer twiddle applied
  auto load_batch = [&](int kb, T2 (&val)[LB]) __attribute__((always_inline)) {
    V2 vv[LB];
#pragma unroll
    for (int k = 0; k < LB; ++k) {
      const int e = min(tid + (kb + k) * NT, TE - 1);
      const int c = e & (ncol - 1);
      const int v = e >> log2ncol;
      const long long g = gbase + c + vstride * v;
      const T2* src;
      if (LOAD == kLoadComb)
        src = cx + g * (long long)A.comb_sigma;
      else if (rows) {
        const long long gr = (LOAD == kLoadCombRowsG) ? gbase + qoff(c) + vstride * v : g;
        src = in + (((gr & ((1LL << A.log2nlo) - 1)) << A.log2m) | (gr >> A.log2nlo));
      }
      else if (LOAD == kLoadPerm)
        src = in + ((uint32_t)q0 + (uint32_t)c + ((uint32_t)v << log2nq));
      else
        src = in + g;
      vv[k] = *(const V2*)src;
    }
#pragma unroll
    for (int k = 0; k < LB; ++k) {
      val[k].x = vv[k].x;
      val[k].y = vv[k].y;
    }
    if constexpr (!first) {
      // outer twiddles W_Lcur^((t0+c)*v): a direct table entry, or the product of a high and
      // a low table entry (every table load of the batch issued first, as for the data)
      V2 tw[LB], tl[LB];
      if constexpr (split) {
        const V2* __restrict__ thi = (const V2*)A.thi;
        const V2* __restrict__ tlo = (const V2*)A.tlo;
        const uint32_t lmask = (1u << A.s) - 1u;
#pragma unroll
        for (int k = 0; k < LB; ++k) {
          const int e = min(tid + (kb + k) * NT, TE - 1);
          const uint32_t ex = (uint32_t)((t0 + (e & (ncol - 1))) * (e >> log2ncol));
          tw[k] = thi[ex >> A.s];
          tl[k] = tlo[ex & lmask];
        }
#pragma unroll
        for (int k = 0; k < LB; ++k) {
          T2 hi, lo;
          hi.x = tw[k].x;
          hi.y = tw[k].y;
          lo.x = tl[k].x;
          lo.y = tl[k].y;
          const T2 w = cmul(hi, lo);
          tw[k].x = w.x;
          tw[k].y = w.y;
        }
      } else {
        const V2* __restrict__ tdir = (const V2*)A.tdir;
#pragma unroll
        for (int k = 0; k < LB; ++k) {
          const int e = min(tid + (kb + k) * NT, TE - 1);
          const uint32_t ex = (uint32_t)((t0 + (e & (ncol - 1))) * (e >> log2ncol));
          tw[k] = tdir[ex];
        }
      }
#pragma unroll
      for (int k = 0; k < LB; ++k) {
        const int e = min(tid + (kb + k) * NT, TE - 1);
        T2 w;
        w.x = tw[k].x;
        w.y = tw[k].y;
        const T2 tv = cmul(val[k], w);
        if ((e >> log2ncol) > 0) val[k] = tv;
      }
    }
  };
  // A full plain tile (TE == NT*EPT): the thread's elements e = tid + k*NT are the elements
  // v = j + k*R/EPT of column c = tid mod ncol (j = tid / ncol), the inputs of one task of the
  // first stage -- FP32 (EPT 16): the radix-16 step of fft_stage16 (its task (c, tix = j));
  // FP64 (EPT 4): the radix-4 butterfly j of fft_stage4.  The thread runs that task on the
  // loaded values in registers and writes only its outputs to LDS: the same operations on the
  // same values as storing the loads and running the stage, with one LDS round trip fewer.
  constexpr int FR = sizeof(Real) == 4 ? 16 : 4;
  constexpr bool can_fuse = (LOAD == kLoadPlain || rows) && EPT == FR;
  int log2ns0 = 0;
  bool fused = false;
  if constexpr (can_fuse) {
    if (TE == NT * EPT && (1 << r) >= FR) {
      fused = true;
      T2 x[EPT];
#pragma unroll
      for (int kb = 0; kb < EPT; kb += LB) {
        T2 val[LB];
        load_batch(kb, val);
#pragma unroll
        for (int k = 0; k < LB; ++k) x[kb + k] = val[k];
      }
      const int j = tid >> log2ncol;
      T2* col = lds + (tid & (ncol - 1)) * RP;
      if constexpr (EPT == 16) {
        stage16_regs<T2, false>(x, 0, 1, tw4, A.inverse);
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int m2 = 0; m2 < 4; ++m2) col[LPAD((j << 4) + m + 4 * m2)] = x[m2 + 4 * m];
        log2ns0 = 4;
      } else {
        bfly4(x[0], x[1], x[2], x[3], A.inverse);
#pragma unroll
        for (int m = 0; m < 4; ++m) col[LPAD((j << 2) + m)] = x[m];
        log2ns0 = 2;
      }
    }
  }
  if (!fused) {
#pragma unroll
    for (int kb = 0; kb < EPT; kb += LB) {
      T2 val[LB];
      load_batch(kb, val);
#pragma unroll
      for (int k = 0; k < LB; ++k) {
        // (a clamped element is stored again with its own value: no branch)
        const int e = min(tid + (kb + k) * NT, TE - 1);
        const int c = e & (ncol - 1);
        const int v = (LOAD == kLoadPerm) ? perm_v(e) : e >> log2ncol;
        lds[c * RP + LPAD(v)] = val[k];
      }
    }
  }
  __syncthreads();

  // ---- sub-FFT of every column: radix-4 stages, then one radix-2 stage if r is odd
  // FP32: pairs of radix-4 stages as radix-16 steps in registers; FP64: radix-4 stages.  Then
  // the remaining radix-4 stages, the last one fused with the radix-2 stage when r is odd (the
  // same stages in the same order), after the stage run at load time on a full tile
  int log2ns = log2ns0;
  int rem = r - log2ns0;
  while (sizeof(Real) == 4 && rem >= 4) {
    if (log2ns == 0)
      fft_stage16<T2, NT, false>(lds, RP, r, log2ncol, 0, tw4, A.inverse);
    else
      fft_stage16<T2, NT, true>(lds, RP, r, log2ncol, log2ns, tw4, A.inverse);
    log2ns += 4;
    rem -= 4;
  }
  while (rem >= 2) {
    if (rem == 3) {   // the last radix-4 stage with the radix-2 stage: one LDS round trip
      if (log2ns > 0)
        fft_stage8<T2, EPT, NT, true>(lds, RP, r, log2ncol, tw4, twR, A.inverse);
      else
        fft_stage8<T2, EPT, NT, false>(lds, RP, r, log2ncol, tw4, twR, A.inverse);
      log2ns += 3;
      rem = 0;
      break;
    }
    T2 wc[EPT / 4][3];
    if (log2ns > 0) {
      stage4_twiddles<T2, EPT, NT>(wc, r, log2ncol, log2ns, tw4);
      fft_stage4<T2, EPT, NT, true>(lds, RP, r, log2ncol, log2ns, wc, A.inverse);
    } else {
      fft_stage4<T2, EPT, NT, false>(lds, RP, r, log2ncol, 0, wc, A.inverse);
    }
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
        // kLoadPerm: column c is sub-FFT (q0 + c) * pinv mod nq; kLoadCombRowsG: q0 + qoff(c)
        const long long qc =
            (LOAD == kLoadPerm)
                ? (long long)(((uint32_t)(q0 + c) * pinv) & (((uint32_t)nq) - 1u))
                : (LOAD == kLoadCombRowsG ? q0 + qoff(c) : q0 + c);
        g = qc * R + kk;
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
          A.digits[b * A.keys_bstride + kpos] = (uint16_t)sel_first_digit(key);
      }
    }
  }
}

template <typename Real, int LOAD, int STORE, int TW>
static hipError_t launch_pass(const PassArgs& A, int tiles, int batch, hipStream_t st) {
  const dim3 grid = (LOAD == kLoadComb) ? dim3(tiles * batch, 1) : dim3(tiles, batch);
  hipLaunchKernelGGL((fft_pass_kernel<Real, LOAD, STORE, TW>), grid, dim3(Tile<Real, LOAD>::nt),
                     0, st, A);
  return hipGetLastError();
}

template <typename Real, int LOAD, int TW>
static hipError_t dispatch_store(const PassArgs& A, int store, int tiles, int batch,
                                 hipStream_t st) {
  if (store == kStoreCplx) return launch_pass<Real, LOAD, kStoreCplx, TW>(A, tiles, batch, st);
  if (store == kStoreKeys) return launch_pass<Real, LOAD, kStoreKeys, TW>(A, tiles, batch, st);
  return launch_pass<Real, LOAD, kStoreBoth, TW>(A, tiles, batch, st);
}

// the special loads only occur in first passes (no outer twiddles)
template <typename Real>
static hipError_t dispatch_pass(const PassArgs& A, int load, int store, int tw, int tiles,
                                int batch, hipStream_t st) {
  if (load == kLoadCombRows)
    return dispatch_store<Real, kLoadCombRows, kTwNone>(A, store, tiles, batch, st);
  if (load == kLoadCombRowsG)
    return dispatch_store<Real, kLoadCombRowsG, kTwNone>(A, store, tiles, batch, st);
  if (load == kLoadPerm)
    return dispatch_store<Real, kLoadPerm, kTwNone>(A, store, tiles, batch, st);
  if (load == kLoadComb)
    return dispatch_store<Real, kLoadComb, kTwNone>(A, store, tiles, batch, st);
  if (tw == kTwSplit)
    return dispatch_store<Real, kLoadPlain, kTwSplit>(A, store, tiles, batch, st);
  if (tw == kTwDirect)
    return dispatch_store<Real, kLoadPlain, kTwDirect>(A, store, tiles, batch, st);
  return dispatch_store<Real, kLoadPlain, kTwNone>(A, store, tiles, batch, st);
}

static int pass_log2q(const FftTables& T, const FftLaunch& L, int p, int r, int log2Lprev) {
  const int load = (p == 0) ? L.load : kLoadPlain;
  // a kLoadPerm pass takes as many columns as fit: whole lines of its input per tile
  const int qmax = (load != kLoadPerm && L.qmax > 0 && L.qmax <= FFT_QMAX) ? L.qmax : FFT_QMAX;
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
    int load = (p == 0) ? L.load : kLoadPlain;
    const int log2Q = pass_log2q(T, L, p, r, log2Lprev);
    int tiles;
    if (p == 0) {
      const int lnq = log2N - log2Lcur;
      const int lq = log2Q < lnq ? log2Q : lnq;
      tiles = 1 << (lnq - lq);
      // comb rows holding 2^d sibling sub-FFTs per row (0 < d <= columns per tile): grouped
      const int d = lnq - L.comb_log2nlo;
      if (load == kLoadCombRows && d > 0 && d <= lq) load = kLoadCombRowsG;
    } else {
      tiles = (int)((1LL << (log2N - log2Lcur)) << (log2Lprev - log2Q));
    }
    A.log2Q = log2Q;
    A.inverse = T.inverse;
    A.twR = T.twR[p];
    A.tw4 = T.tw4[p];
    A.tdir = T.tdir[p];
    A.thi = T.thi[p];
    A.tlo = T.tlo[p];
    A.s = T.s[p];
    A.comb_x = L.comb_x;
    A.comb_offs = L.comb_offs;
    A.comb_sigma = L.comb_sigma;
    A.log2nlo = L.comb_log2nlo;
    A.log2m = L.comb_log2m;
    A.perm = L.perm;
    A.perm_inv = L.perm_inv;
    if (load == kLoadPerm && (!L.perm || !L.perm_inv)) return hipErrorInvalidValue;
    A.nbatch = L.batch;
    A.keys_tiled = (last && p > 0 && L.keys_tiled) ? 1 : 0;
    A.digits = last ? L.digits : nullptr;
    const int store = last ? L.store : kStoreCplx;
    const int tw = (p == 0) ? kTwNone : (log2Lcur > kMaxPassLog ? kTwSplit : kTwDirect);
    if (load != kLoadPlain && tw != kTwNone) return hipErrorInvalidValue;
    hipError_t e = T.is_fp32 ? dispatch_pass<float>(A, load, store, tw, tiles, L.batch, st)
                             : dispatch_pass<double>(A, load, store, tw, tiles, L.batch, st);
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
// fold_term(x, f) is the tap's product, accumulate() adds a term: add() = both, in that order.
__device__ __forceinline__ double2 fold_term(double2 x, double2 f) {
  return make_double2(x.x * f.x - x.y * f.y, x.x * f.y + x.y * f.x);
}
__device__ __forceinline__ float2 fold_term(float2 x, float2 f) {
  return make_float2(x.x * f.x - x.y * f.y, x.x * f.y + x.y * f.x);
}

struct Acc64 {
  double2 s;
  __device__ __forceinline__ void init() { s = make_double2(0.0, 0.0); }
  __device__ __forceinline__ void accumulate(double2 t) {
    s.x = s.x + t.x;
    s.y = s.y + t.y;
  }
  __device__ __forceinline__ void add(double2 x, double2 f) { accumulate(fold_term(x, f)); }
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
  __device__ __forceinline__ void accumulate(float2 t) {
    kahan(sx, cx, t.x);
    kahan(sy, cy, t.y);
  }
  __device__ __forceinline__ void add(float2 x, float2 f) { accumulate(fold_term(x, f)); }
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
// column-ordered fold
//
// View x as m = n/B rows of B columns (x[row*B + col]).  Every tap id = t*B + r of bucket r in
// loop l gathers x[(id*ai_l) mod n], whose column is (id*ai_l) mod B = (r*ai_l) mod B: all the
// taps of a bucket lie in ONE column, c = r*ai_l mod B, and conversely column c holds bucket
// r = c*a_l mod B of loop l.  A workgroup that owns a block of adjacent columns for every loop
// therefore gathers only from the 128-byte lines of those columns, and the taps of different
// loops that land in the same line share one fetch (at n = 2^27, k = 16000: 20.1 M taps over
// 8.4 M lines of complex64, 16.8 M of complex128).  The filter taps of bucket r are read from
// a bucket-major copy fT[r*P + t] (one contiguous row per bucket, built by filter_rows_kernel
// when the filter is set), since the buckets of one column block are scattered over r.
// Each thread owns one (loop, column) pair and sums its bucket's taps in the canonical order
// (t ascending, then the reference's tail tap when tiled), so the sums are the bits of
// pf_body.
// The comb samples y_r[j] = x[off_r + j*sigma] (sigma = n/W divides B) are the columns
// c = off_r mod sigma + lo*sigma of the same view: sample j = row*(B/sigma) + lo.  The threads
// of a column block also copy the comb columns that fall in it, while its lines are being
// fetched for the taps, into column-major rows yc[r*W + lo*m + row] (m contiguous samples per
// comb column); the comb FFT's first pass loads them with kLoadCombRows.
// ============================================================================================
constexpr int FC_NT = 256;
constexpr int FC_CHUNK = 8;   // taps whose loads are issued together

// fT[r*P + t]: row r holds the taps of bucket r in summation order: t*B + r for t < T, then at
// t = T the tiled fold's tail tap r*T + B (tiled, r < w - T*B) or the last full-fold tap T*B + r
// (r < w - T*B); zero elsewhere.  T = w / B.
template <typename T2>
__global__ void filter_rows_kernel(const T2* __restrict__ f, T2* __restrict__ fT, int B, int w,
                                   int T, int P, int tiled) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)B * P) return;
  const int r = (int)(i / P), t = (int)(i % P);
  const int residue = w - T * B;
  long long id = -1;
  if (t < T) id = (long long)t * B + r;
  else if (t == T && r < residue) id = tiled ? (long long)r * T + B : (long long)T * B + r;
  T2 v;
  v.x = 0;
  v.y = 0;
  fT[i] = (id >= 0) ? f[id] : v;
}

template <typename Real>
__global__ __launch_bounds__(FC_NT) void fold_columns_kernel(
    const typename CT<Real>::T2* __restrict__ x, const typename CT<Real>::T2* __restrict__ fT,
    typename CT<Real>::T2* __restrict__ bins, const int* __restrict__ a_dev,
    const int* __restrict__ ai_dev, uint32_t mask, int log2B, int w, int loops, int tiled,
    int T, int P, int log2lw, int remap, typename CT<Real>::T2* __restrict__ yc,
    const int* __restrict__ offs, int Rl, int log2sigma, int log2m) {
  using T2 = typename CT<Real>::T2;
  using AC = typename AccOf<Real>::A;
  const int blk = remap ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const uint32_t item = (uint32_t)blk * FC_NT + threadIdx.x;
  const uint32_t B = 1u << log2B;
  if (item >= (uint32_t)loops * B) return;
  // item = ((column block * loops + l) << log2lw) | column within the block
  const uint32_t lw = 1u << log2lw;
  const uint32_t cc = item & (lw - 1u), rest = item >> log2lw;
  const uint32_t l = rest % (uint32_t)loops, cb = rest / (uint32_t)loops;
  const uint32_t c = (cb << log2lw) | cc;
  const uint32_t ai = (uint32_t)ai_dev[l];
  const uint32_t r = (c * (uint32_t)a_dev[l]) & (B - 1u);
  const int residue = w - T * (int)B;
  // main taps t*B + r: T of them, T+1 for the full fold's r < residue
  const int nmain = (!tiled && (int)r < residue) ? T + 1 : T;
  const uint32_t base = (r * ai) & mask, step = (B * ai) & mask;
  const T2* __restrict__ frow = fT + (size_t)r * (size_t)P;
  AC acc;
  acc.init();
  for (int t0 = 0; t0 < nmain; t0 += FC_CHUNK) {
    T2 xv[FC_CHUNK], fv[FC_CHUNK];
#pragma unroll
    for (int k = 0; k < FC_CHUNK; ++k) {
      const int t = t0 + k;
      if (t < nmain) {
        xv[k] = x[(base + (uint32_t)t * step) & mask];
        fv[k] = frow[t];
      }
    }
#pragma unroll
    for (int k = 0; k < FC_CHUNK; ++k)
      if (t0 + k < nmain) acc.add(xv[k], fv[k]);
  }
  if (tiled && (int)r < residue) {   // the reference's tail tap of its tiled fold (n < 2^27): another column
    const uint32_t id = r * (uint32_t)T + B;
    acc.add(x[(id * ai) & mask], frow[T]);
  }
  bins[(size_t)l * B + r] = acc.get();
  // comb columns of this column block: rank-strided rows, contiguous stores
  const uint32_t rank = (l << log2lw) | cc, G = (uint32_t)loops << log2lw;
  const uint32_t c0 = cb << log2lw, smask = (1u << log2sigma) - 1u, m = 1u << log2m;
  const size_t W = (size_t)B << log2m >> log2sigma;
  for (int rr = 0; rr < Rl; ++rr) {
    const uint32_t o = (uint32_t)offs[rr];
    for (uint32_t col = c0 + ((o - c0) & smask); col < c0 + lw; col += smask + 1u) {
      T2* __restrict__ dst = yc + (size_t)rr * W + ((size_t)(col >> log2sigma) << log2m);
#pragma unroll 4
      for (uint32_t row = rank; row < m; row += G) dst[row] = x[((size_t)row << log2B) | col];
    }
  }
}

// LDS-staged column fold.  A slab is SC adjacent columns (SC * e = 256 bytes, two 128-byte
// lines per row: narrower slabs read at a power-of-two row stride reach only half the bandwidth
// on MI355X) over all m rows of the view.  A workgroup streams its slabs (s, s + G, s + 2G, ...
// for a grid of G resident workgroups) through LDS in chunks of SL_ROWS rows with 16-byte loads;
// the next chunk's loads, also across a slab boundary, are in flight in registers while the
// current chunk is used, so the stream does not stop between slabs: across the grid the whole
// signal is read once.  For each slab the workgroup holds the filter rows of the buckets of its
// (loop, column) pairs in an LDS tap array (slot t of pair p at t * pairs + p; the next slab's
// rows are fetched while the current slab is finished).  As the rows of its taps pass (tap t
// lies in row ((r*ai + t*B*ai) mod n) / B), the thread of pair p replaces slot t by the tap's
// product term x * f; after the slab's last chunk it adds its slots up in the canonical order.
// The bucket sums are written in column order (bins[l*B + c] for the bucket r = c*a_l mod B of
// column c: whole segments instead of scattered 8/16-byte stores); the bucket FFT reads them
// back through that permutation (kLoadPerm).  The comb columns of the slab are copied out of
// each chunk.  Same per-bucket terms and order as pf_body: the same bits.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int SL_ROWS = 128;  // rows per chunk
constexpr int SL_PITCH = 17;  // 16-byte units per chunk row: 256 bytes + 16 of padding (banks)
constexpr int SL_U = 4;       // 16-byte loads per thread for one chunk (512 threads)
constexpr int SL_TABLE = 128; // LDS ints after the slots: a [0,32), ai [32,64), comb offsets [64,128)
template <typename Real>
struct SlabFetch {            // filter-row loads per thread per slab (all in flight together)
  static constexpr int fb = sizeof(Real) == 8 ? 8 : 16;
};

template <typename Real, int PMAX, int NT>
__global__ __launch_bounds__(NT, 4) void fold_slab_kernel(
    const typename CT<Real>::T2* __restrict__ x, const typename CT<Real>::T2* __restrict__ fT,
    typename CT<Real>::T2* __restrict__ bins, const int* __restrict__ a_dev,
    const int* __restrict__ ai_dev, uint32_t mask, int log2B, int log2m, int w, int loops,
    int tiled, int T, int P, int log2sc, typename CT<Real>::T2* __restrict__ yc,
    const int* __restrict__ offs, int Rl, int log2sigma, FoldClear clr) {
  using T2 = typename CT<Real>::T2;
  using V2 = typename NV<Real>::type;
  using AC = typename AccOf<Real>::A;
  constexpr int EPU = 16 / (int)sizeof(T2);             // elements per 16-byte unit
  constexpr int FB = SlabFetch<Real>::fb;
  extern __shared__ u32x4 dyn16[];
  u32x4* chunk16 = dyn16;                              // SL_ROWS x SL_PITCH units
  const V2* ldsv = (const V2*)dyn16;                   // element (row, k): row*PITCH*EPU + k
  V2* slotv = (V2*)(dyn16 + SL_ROWS * SL_PITCH);       // P x pairs slots
  const int pairs = loops << log2sc;
  int* atab = (int*)(slotv + (size_t)P * pairs);       // SL_TABLE ints
  const int tid = threadIdx.x;
  const uint32_t B = 1u << log2B, SC = 1u << log2sc;
  const uint32_t nslabs = B >> log2sc;
  const bool active = tid < pairs;
  const uint32_t cc = (uint32_t)tid & (SC - 1u), l = (uint32_t)tid >> log2sc;
  const int residue = w - T * (int)B;
  const int log2rows = log2m < 7 ? log2m : 7;          // SL_ROWS = 2^7
  const int rows = 1 << log2rows;
  const int nchunks = 1 << (log2m - log2rows);
  const int units = rows * 16;                         // 16-byte units of x per chunk
  const u32x4* __restrict__ x16 = (const u32x4*)x;
  const uint32_t row16 = (uint32_t)(((size_t)B * sizeof(T2)) / 16);
  const uint32_t smask = (1u << log2sigma) - 1u;
  const size_t W = (size_t)B << log2m >> log2sigma;
  // comb columns of a slab per comb loop: SC / sigma when sigma < SC, else at most one
  const int log2ncs = log2sc > log2sigma ? log2sc - log2sigma : 0;
  const int ncopy = (Rl << log2ncs) << log2rows;       // comb copy elements per chunk
  const int total = P * pairs;                         // filter slots (at most FB * NT)
  const uint32_t magic = P > 1 ? 0xFFFFFFFFu / (uint32_t)P + 1u : 0u;   // k / P = umulhi(k, magic)
  const V2* __restrict__ fTv = (const V2*)fT;

  // The chunks of the workgroup in order: chunk q is chunk (q mod nchunks) of its slab
  // q / nchunks (slab blockIdx.x + (q / nchunks) * G).  The next chunk is in flight in
  // registers (xa) while one is used from LDS.  q past the last chunk is clamped to it.
  const int log2nch = log2m - log2rows;
  const uint32_t G = gridDim.x;
  const int nsl = (int)((nslabs - blockIdx.x + G - 1) / G);   // slabs of this workgroup
  const int Q = nsl << log2nch;
  u32x4 xa[SL_U];
  // the loads of a chunk in one basic block (clamped addresses): in flight together.  Unit
  // indices in 32 bits (n * e / 16 < 2^32).
  auto load = [&](u32x4* buf, int q) __attribute__((always_inline)) {
    const int qq = min(q, Q - 1);
    const uint32_t c0 = (blockIdx.x + (uint32_t)(qq >> log2nch) * G) << log2sc;
    const int ch = qq & (nchunks - 1);
    const uint32_t col16 = (uint32_t)(((size_t)c0 * sizeof(T2)) / 16);
#pragma unroll
    for (int u = 0; u < SL_U; ++u) {
      const int i = min(u * NT + tid, units - 1);
      const uint32_t unit =
          ((uint32_t)(ch * rows) + (uint32_t)(i >> 4)) * row16 + col16 + (uint32_t)(i & 15);
      buf[u] = __builtin_nontemporal_load(&x16[unit]);
    }
  };
  // every thread stores every one of its units (a clamped one again, with the same value), so
  // that all of the chunk's loads are known to be complete here
  auto store = [&](const u32x4* buf) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < SL_U; ++u) {
      const int i = min(u * NT + tid, units - 1);
      chunk16[(i >> 4) * SL_PITCH + (i & 15)] = buf[u];
    }
  };
  // the filter rows of a slab's pairs: entry k = p * P + t of the fetch (entry t of pair p's
  // row; neighbouring threads read neighbouring entries of a row) goes to slot t * pairs + p;
  // any thread any entry (clamped addresses; the loads in flight together)
  V2 fv[FB];
  auto entry = [&](int u, int tv, int& p, int& t) __attribute__((always_inline)) {
    const int k = min(u * NT + tv, total - 1);
    p = P > 1 ? (int)__umulhi((uint32_t)k, magic) : k;
    t = k - p * P;
  };
  auto fetch_rows = [&](uint32_t c0) __attribute__((always_inline)) {
    int tv = tid;
    asm volatile("" : "+v"(tv));   // entry indices recomputed here, not kept in registers
#pragma unroll
    for (int u = 0; u < FB; ++u) {
      int p, t;
      entry(u, tv, p, t);
      const uint32_t rp =
          ((c0 + ((uint32_t)p & (SC - 1u))) * (uint32_t)atab[p >> log2sc]) & (B - 1u);
      fv[u] = fTv[(size_t)rp * (size_t)P + (size_t)t];
    }
  };
  // (a clamped entry is stored again with its own value: every load is known complete here)
  auto put_rows = [&]() __attribute__((always_inline)) {
    int tv = tid;
    asm volatile("" : "+v"(tv));
#pragma unroll
    for (int u = 0; u < FB; ++u) {
      int p, t;
      entry(u, tv, p, t);
      slotv[t * pairs + p] = fv[u];
    }
  };
  uint32_t r = 0, ai = 0, base = 0, step = 0;
  int nmain = 0;
  auto pair_state = [&](uint32_t c0) __attribute__((always_inline)) {
    if (active) {
      ai = (uint32_t)atab[32 + l];
      r = ((c0 + cc) * (uint32_t)atab[l]) & (B - 1u);
      nmain = (!tiled && (int)r < residue) ? T + 1 : T;
      base = (r * ai) & mask;
      step = (B * ai) & mask;
    }
  };

  // (a, ai) of every loop and the comb offsets into LDS: their loads go out before the first
  // chunk's, so that waiting for them does not wait for the chunk
  // the caller's clear, spread over the grid (one 4-byte store per thread at n = 2^27)
  for (size_t i = (size_t)blockIdx.x * NT + tid; i < clr.nbm; i += (size_t)gridDim.x * NT)
    clr.bm[i] = 0u;
  if (blockIdx.x == 0 && tid < clr.nmax) clr.maxf[tid] = 0xFFFFFFFFu;
  int av = 0, aiv = 0, ov = 0;
  if (tid < loops) {
    av = a_dev[tid];
    aiv = ai_dev[tid];
  }
  if (tid < Rl) ov = offs[tid];
  load(xa, 0);
  if (tid < loops) {
    atab[tid] = av;
    atab[32 + tid] = aiv;
  }
  if (tid < Rl) atab[64 + tid] = ov;
  __syncthreads();
  fetch_rows(blockIdx.x << log2sc);
  put_rows();
  pair_state(blockIdx.x << log2sc);
  store(xa);
  __syncthreads();

  // chunk q from LDS while chunk q + 1 is on its way.  Every load and store is issued on every
  // call (the clamped chunk past the end again): no path on which a load could still be in
  // flight, so the compiler adds no waits for one.
  auto body = [&](int q) __attribute__((always_inline)) {
    const int ch = q & (nchunks - 1);
    const int sl = q >> log2nch;
    const uint32_t c0 = (blockIdx.x + (uint32_t)sl * G) << log2sc;
    const uint32_t cnext = (blockIdx.x + (uint32_t)min(sl + 1, nsl - 1) * G) << log2sc;
    const uint32_t r0 = (uint32_t)(ch * rows);
    // comb columns of the slab: this chunk's rows
    for (int e = tid; e < ncopy; e += NT) {
      const int row = e & (rows - 1);
      const int ci = e >> log2rows;
      const int rr = ci >> log2ncs;
      const uint32_t col = c0 + (((uint32_t)atab[64 + rr] - c0) & smask) +
                           ((uint32_t)(ci & ((1 << log2ncs) - 1)) << log2sigma);
      if (col < c0 + SC) {
        V2* __restrict__ dst = (V2*)yc + (size_t)rr * W +
                               ((size_t)(col >> log2sigma) << log2m) + r0 + (uint32_t)row;
        *dst = ldsv[row * (SL_PITCH * EPU) + (int)(col - c0)];
      }
    }
    load(xa, q + 1);
    if (active) {
      // the taps whose rows fall in this chunk (a bit per tap): slot t becomes x * f (tap
      // rows recomputed per chunk, not kept in registers)
      uint32_t bs = base, sp = step;
      int nm = nmain;
      asm volatile("" : "+v"(bs), "+v"(sp), "+v"(nm));
      uint32_t hits = 0;   // built from the top tap down (no per-tap bit constants)
#pragma unroll
      for (int t = PMAX - 1; t >= 0; --t) {
        const uint32_t row = ((bs + (uint32_t)t * sp) & mask) >> log2B;
        hits = (hits << 1) | (uint32_t)(t < nm && row - r0 < (uint32_t)rows);
      }
      while (hits) {
        const int t = __ffs(hits) - 1;
        hits &= hits - 1u;
        const uint32_t row = ((bs + (uint32_t)t * sp) & mask) >> log2B;
        const V2 xv = ldsv[(row - r0) * (SL_PITCH * EPU) + cc];
        const V2 fw = slotv[t * pairs + tid];
        T2 xt, ft;
        xt.x = xv.x;
        xt.y = xv.y;
        ft.x = fw.x;
        ft.y = fw.y;
        const T2 term = fold_term(xt, ft);
        V2 tv;
        tv.x = term.x;
        tv.y = term.y;
        slotv[t * pairs + tid] = tv;
      }
    }
    __syncthreads();
    if (ch == nchunks - 1) {
      // the slab is done: the next slab's filter rows go out (the current slab's again after
      // the workgroup's last slab), this slab's buckets are summed and written while they
      // arrive
      fetch_rows(cnext);
      if (active) {
        int tv = tid;
        asm volatile("" : "+v"(tv));   // slot addresses recomputed here
        AC acc;
        acc.init();
        // the slots in batches of SB reads in flight together (clamped), added up in order
        constexpr int SB = sizeof(Real) == 8 ? 4 : 8;
#pragma unroll
        for (int t0 = 0; t0 < PMAX; t0 += SB) {
          V2 sv[SB];
#pragma unroll
          for (int j = 0; j < SB; ++j) sv[j] = slotv[min(t0 + j, P - 1) * pairs + tv];
#pragma unroll
          for (int j = 0; j < SB; ++j) {
            if (t0 + j < nmain) {
              T2 term;
              term.x = sv[j].x;
              term.y = sv[j].y;
              acc.accumulate(term);
            }
          }
        }
        if (tiled && (int)r < residue) {   // the reference's tail tap of its tiled fold (n < 2^27): another column
          const uint32_t id = r * (uint32_t)T + B;
          acc.add(x[(id * ai) & mask], fT[(size_t)r * (size_t)P + T]);
        }
        bins[(size_t)l * B + c0 + cc] = acc.get();   // column order: whole segments
      }
      __syncthreads();   // every slot read
      put_rows();
      pair_state(cnext);
    }
    store(xa);
    __syncthreads();
  };
#pragma unroll 1
  for (int q = 0; q < Q; ++q) body(q);
}

int fold_rows_pitch(int B, int w) { return w / B + 1; }

static size_t slab_lds_bytes(int is_fp32, int loops, int lsc, int P) {
  return (size_t)SL_ROWS * SL_PITCH * 16 + (size_t)P * ((size_t)loops << lsc) * (is_fp32 ? 8 : 16) +
         SL_TABLE * sizeof(int);   // chunk, tap slots, (a, ai) of up to 32 loops, comb offsets
}

// columns per slab (log2) for the LDS-staged fold, or -1 when it does not apply: 256-byte row
// segments, at most 512 (loop, column) threads, at most 32 taps per bucket, at most 96 KB of
// LDS (chunk + tap slots)
static int slab_log2sc(int is_fp32, int log2B, int log2m, int loops, int P) {
  const int lsc = is_fp32 ? 5 : 4;
  (void)log2m;
  if (lsc > log2B) return -1;
  if ((loops << lsc) > 512 || P > 32) return -1;
  // every filter slot of a slab fetched by one round of loads
  const int fb = is_fp32 ? SlabFetch<float>::fb : SlabFetch<double>::fb;
  if (P * (loops << lsc) > fb * 512) return -1;
  if (slab_lds_bytes(is_fp32, loops, lsc, P) > 96 * 1024) return -1;
  return lsc;
}

template <typename Real, int PMAX, int NT>
static hipError_t slab_launch(const void* x, const void* fT, void* bins, const int* a_dev,
                              const int* ai_dev, uint32_t mask, int log2B, int log2m, int w,
                              int loops, int tiled, int T, int P, int lsc, void* yc,
                              const int* offs, int Rl, int log2sigma, const FoldClear& clr,
                              hipStream_t st) {
  using T2 = typename CT<Real>::T2;
  const unsigned nslabs = 1u << (log2B - lsc);
  const size_t lds = slab_lds_bytes(sizeof(Real) == 4, loops, lsc, P);
  // one workgroup per resident slot (each streams nslabs / grid slabs); any grid is correct.
  // The resident count is asked of the runtime once per (device, LDS size) and kept.
  static std::mutex mu;
  static std::map<std::pair<int, size_t>, int> resident;
  int dev = 0;
  unsigned blocks = nslabs;
  if (hipGetDevice(&dev) == hipSuccess) {
    std::lock_guard<std::mutex> lock(mu);
    auto it = resident.find({dev, lds});
    if (it == resident.end()) {
      int cus = 0, per_cu = 0;
      if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
          hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &per_cu, reinterpret_cast<const void*>(&fold_slab_kernel<Real, PMAX, NT>), NT,
              lds) != hipSuccess)
        cus = per_cu = 0;
      it = resident.emplace(std::make_pair(dev, lds), cus * per_cu).first;
    }
    if (it->second > 0) blocks = std::min(nslabs, (unsigned)it->second);
  }
  hipLaunchKernelGGL((fold_slab_kernel<Real, PMAX, NT>), dim3(blocks), dim3(NT), lds, st,
                     (const T2*)x, (const T2*)fT, (T2*)bins, a_dev, ai_dev, mask, log2B, log2m,
                     w, loops, tiled, T, P, lsc, (T2*)yc, offs, Rl, log2sigma, clr);
  return hipGetLastError();
}

template <typename Real, int NT>
static hipError_t slab_dispatch_nt(int P, const void* x, const void* fT, void* bins,
                                   const int* a_dev, const int* ai_dev, uint32_t mask, int log2B,
                                   int log2m, int w, int loops, int tiled, int T, int lsc,
                                   void* yc, const int* offs, int Rl, int log2sigma,
                                   const FoldClear& clr, hipStream_t st) {
  if (P <= 16)
    return slab_launch<Real, 16, NT>(x, fT, bins, a_dev, ai_dev, mask, log2B, log2m, w, loops,
                                     tiled, T, P, lsc, yc, offs, Rl, log2sigma, clr, st);
  if (P <= 24)
    return slab_launch<Real, 24, NT>(x, fT, bins, a_dev, ai_dev, mask, log2B, log2m, w, loops,
                                     tiled, T, P, lsc, yc, offs, Rl, log2sigma, clr, st);
  return slab_launch<Real, 32, NT>(x, fT, bins, a_dev, ai_dev, mask, log2B, log2m, w, loops,
                                   tiled, T, P, lsc, yc, offs, Rl, log2sigma, clr, st);
}

// 512 threads per workgroup: two chunks of loads in flight per workgroup at 4 loads per thread
template <typename Real>
static hipError_t slab_dispatch(int P, const void* x, const void* fT, void* bins,
                                const int* a_dev, const int* ai_dev, uint32_t mask, int log2B,
                                int log2m, int w, int loops, int tiled, int T, int lsc, void* yc,
                                const int* offs, int Rl, int log2sigma, const FoldClear& clr,
                                hipStream_t st) {
  return slab_dispatch_nt<Real, 512>(P, x, fT, bins, a_dev, ai_dev, mask, log2B, log2m, w,
                                     loops, tiled, T, lsc, yc, offs, Rl, log2sigma, clr, st);
}

int fold_slab_applies(int is_fp32, int n, int B, int w, int loops) {
  int log2B = 0, log2n = 0;
  while ((1 << log2B) < B) ++log2B;
  while ((1LL << log2n) < n) ++log2n;
  return slab_log2sc(is_fp32, log2B, log2n - log2B, loops, fold_rows_pitch(B, w)) >= 0;
}

hipError_t filter_rows_launch(int is_fp32, const void* f, void* fT, int B, int w, int tiled,
                              hipStream_t st) {
  const int T = w / B, P = fold_rows_pitch(B, w);
  const long long total = (long long)B * P;
  const unsigned blocks = (unsigned)((total + 255) / 256);
  if (is_fp32)
    hipLaunchKernelGGL(filter_rows_kernel<float2>, dim3(blocks), dim3(256), 0, st,
                       (const float2*)f, (float2*)fT, B, w, T, P, tiled);
  else
    hipLaunchKernelGGL(filter_rows_kernel<double2>, dim3(blocks), dim3(256), 0, st,
                       (const double2*)f, (double2*)fT, B, w, T, P, tiled);
  return hipGetLastError();
}

hipError_t fold_columns_launch(int is_fp32, const void* x, const void* fT, void* bins,
                               const int* a_dev, const int* ai_dev, int n, int B, int w,
                               int loops, int tiled, int by_column, int remap, void* yc,
                               const int* offs, int Rl, int W, const FoldClear& clr,
                               hipStream_t st) {
  int log2B = 0, log2n = 0, log2W = 0;
  while ((1 << log2B) < B) ++log2B;
  while ((1LL << log2n) < n) ++log2n;
  while ((1 << log2W) < W) ++log2W;
  const int log2m = log2n - log2B, log2sigma = log2n - log2W;
  if (!yc || log2sigma > log2B) Rl = 0;   // no comb columns to copy (sigma must divide B)
  const int T = w / B, P = fold_rows_pitch(B, w);
  const uint32_t mask = (uint32_t)n - 1u;
  if (by_column) {
    const int lsc = Rl <= 64 ? slab_log2sc(is_fp32, log2B, log2m, loops, P) : -1;
    if (lsc < 0) return hipErrorInvalidValue;
    return is_fp32 ? slab_dispatch<float>(P, x, fT, bins, a_dev, ai_dev, mask, log2B, log2m, w,
                                          loops, tiled, T, lsc, yc, offs, Rl, log2sigma, clr, st)
                   : slab_dispatch<double>(P, x, fT, bins, a_dev, ai_dev, mask, log2B, log2m, w,
                                           loops, tiled, T, lsc, yc, offs, Rl, log2sigma, clr,
                                           st);
  }
  if (clr.nbm || clr.nmax) return hipErrorInvalidValue;   // only the slab fold clears
  // columns per block: one 128-byte line of x (8 complex128 / 16 complex64), fewer if B is
  // smaller
  int log2lw = is_fp32 ? 4 : 3;
  if (log2lw > log2B) log2lw = log2B;
  const long long total = (long long)loops * B;
  const unsigned blocks = (unsigned)((total + FC_NT - 1) / FC_NT);
  if (remap < 0) remap = 1;
  if (is_fp32)
    hipLaunchKernelGGL(fold_columns_kernel<float>, dim3(blocks), dim3(FC_NT), 0, st,
                       (const float2*)x, (const float2*)fT, (float2*)bins, a_dev, ai_dev, mask,
                       log2B, w, loops, tiled, T, P, log2lw, remap, (float2*)yc, offs, Rl,
                       log2sigma, log2m);
  else
    hipLaunchKernelGGL(fold_columns_kernel<double>, dim3(blocks), dim3(FC_NT), 0, st,
                       (const double2*)x, (const double2*)fT, (double2*)bins, a_dev, ai_dev,
                       mask, log2B, w, loops, tiled, T, P, log2lw, remap, (double2*)yc, offs,
                       Rl, log2sigma, log2m);
  return hipGetLastError();
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
