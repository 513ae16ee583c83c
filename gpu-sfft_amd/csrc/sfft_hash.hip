// sfft_hash.hip -- hand-written CDNA4 (gfx950) kernels of the sparse-FFT hot path, part 2.
//
//  * hash_*_kernel        reverse hash: approved-residue bitmaps, deterministic hit list
//                         (loop, bucket, walk order); the votes of earlier loops are tested
//                         directly from their bucket selections (no vote table)
//  * estimate_kernel      per-location FP64 filter division and stable lower median
//
// Floating-point operations that must match the CPU oracle bit for bit are written out
// explicitly (fma where fused); the library is built with -ffp-contract=off.
#include "sfft_internal.h"

namespace sfft {

template <typename R>
struct CT;
template <>
struct CT<double> {
  using T2 = double2;
};
template <>
struct CT<float> {
  using T2 = float2;
};

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

// the per-transform clear in one launch: approved-residue bitmaps to 0, the largest approved
// residue of each location loop to -1 (the bucket selection bitmaps are rewritten whole by the
// cutoff)
__global__ void hash_clear_kernel(uint32_t* __restrict__ bm, size_t nbm,
                                  uint32_t* __restrict__ maxfirst, int nmax) {
  const size_t s = (size_t)gridDim.x * blockDim.x;
  const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (size_t i = i0; i < nbm; i += s) bm[i] = 0u;
  if (i0 < (size_t)nmax) maxfirst[i0] = 0xFFFFFFFFu;
}

hipError_t hash_clear_launch(const HashArgs& h, hipStream_t st) {
  const size_t bm_words = (size_t)h.loops_loc * h.bm_words;
  size_t blocks = (bm_words + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(hash_clear_kernel, dim3((unsigned)blocks), dim3(256), 0, st,
                     (uint32_t*)h.bitmap, bm_words, (uint32_t*)h.maxfirst, h.loops_loc);
  return hipGetLastError();
}

// approved residue first = J2 * ai mod W of every comb location, as a bitmap per location
// loop, and the largest one (reduced across the block first: one atomic per block on the
// loop's single max word -- same-address atomics serialise)
constexpr int BM_NT = 1024;
__global__ __launch_bounds__(BM_NT) void hash_bitmap_kernel(HashArgs h) {
  __shared__ int wmax[BM_NT / 64];
  const int i = blockIdx.x * BM_NT + threadIdx.x;
  const int l = blockIdx.y;
  int v = -1;
  if (i < h.nJ2) {
    const uint64_t j2 = (uint64_t)h.J2[i];
    const uint32_t first =
        (uint32_t)((j2 * (uint64_t)(uint32_t)h.ai_dev[l]) & (uint64_t)(h.W - 1));
    atomicOr(&h.bitmap[(size_t)l * h.bm_words + (first >> 5)], 1u << (first & 31));
    v = (int)first;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = -1;
    for (int w = 0; w < BM_NT / 64; ++w) m = max(m, wmax[w]);
    if (m >= 0) atomicMax(&h.maxfirst[l], m);
  }
}

hipError_t hash_bitmap_launch(const HashArgs& h, hipStream_t st) {
  dim3 grid((h.nJ2 + BM_NT - 1) / BM_NT, h.loops_loc);
  hipLaunchKernelGGL(hash_bitmap_kernel, grid, dim3(BM_NT), 0, st, h);
  return hipGetLastError();
}

// Lanes per bucket window: one per 32 positions (a bitmap word), at most a wave, at least one.
// m = n/B is a power of two, so the count is a power of two and divides 64.
__host__ __device__ __forceinline__ int hash_lpj(long long m) {
  long long l = m / 128;
  if (l < 1) l = 1;
  if (l > 64) l = 64;
  return (int)l;
}

// Enumerate, in increasing circular order, the approved permuted positions p in the slice
// [low + d_begin, low + d_begin + d_len) of bucket J's window [J m - m/2, J m + m/2) (mod n);
// the whole window is empty when (low mod W) exceeds the largest approved residue (the
// reference's findUpperBound returns the last entry and the walk stops without wrapping).
// Calls fn(loc, rel) with loc = p * a mod n and rel = p - (low + d_begin), the position in
// the slice.
template <typename F>
__device__ __forceinline__ void hash_walk(const HashArgs& h, int l, int i, long long d_begin,
                                          long long d_len, F&& fn) {
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
  const long long d_end = d_begin + d_len;
  for (long long d = d_begin; d < d_end;) {
    const uint64_t p = (uint64_t)((low + d) & (n - 1));
    const uint32_t bpos = (uint32_t)(p & (uint64_t)(h.W - 1));
    const int off = (int)(bpos & 31u);
    const long long left = d_end - d;
    const int take = (left < 32 - off) ? (int)left : (32 - off);
    uint32_t word = bm[bpos >> 5] >> off;
    if (take < 32) word &= (1u << take) - 1u;
    while (word) {
      const int bit = __ffs(word) - 1;
      word &= word - 1u;
      const uint64_t pp = p + (uint64_t)bit;
      fn((uint32_t)((pp * a) & nmask), (int)(d - d_begin) + bit);
    }
    d += take;
  }
}

constexpr int HASH_NT = 256;

int hash_blocks(int loops_loc, int K, long long m) {
  return (int)(((long long)loops_loc * K * hash_lpj(m) + HASH_NT - 1) / HASH_NT);
}

// 32-bit words of one slice's hit mask (a slice covers m / lpj window positions)
__host__ __device__ __forceinline__ int hash_slice_words(long long m) {
  return (int)((m / hash_lpj(m) + 31) / 32);
}
size_t hash_mask_bytes(int loops_loc, int K, long long m) {
  return (size_t)hash_blocks(loops_loc, K, m) * HASH_NT * hash_slice_words(m) * 4;
}

// A location is voted at most once per loop (the J windows of a loop are disjoint), so its
// vote count after loop l is the number of loops 0..l that walked it.  Hence (1) only loops
// l >= loops_thresh-1 can produce hits, and (2) the hit test of loop l needs only whether each
// loop l' < l walked the location (its own walk is implied).  Loop l' walked loc exactly when
// loc's position p = loc * ai_l' mod n lies in the window of a bucket J that l' selected,
// J = ((p + m/2) mod n) / m, that window is not cut off by the largest approved residue (the
// walk's early exit), and p's residue p mod W is approved: three table reads instead of a
// vote pass over every window and a vote table.  The count and write passes cover the loops
// [thresh-1, loops_loc) (loops_thresh >= 1 is checked when the plan is made).
static int hash_hit_count(const HashArgs& h) { return h.loops_loc - (h.loops_thresh - 1); }

// thread -> (location loop l, bucket index i, slice sub of lpj)
struct HashSlot {
  int l, i, sub, lpj;
  long long jidx, d0, dl;
  bool valid;
};

__device__ __forceinline__ int hash_hit_first_d(const HashArgs& h) { return h.loops_thresh - 1; }
__device__ __forceinline__ int hash_hit_count_d(const HashArgs& h) {
  return h.loops_loc - (h.loops_thresh - 1);
}

// threads cover loops [l_first, l_first + nloops)
__device__ __forceinline__ HashSlot hash_slot(const HashArgs& h, int l_first, int nloops) {
  HashSlot s;
  const long long m = (long long)h.n / h.B;
  s.lpj = hash_lpj(m);
  const long long gid = (long long)blockIdx.x * HASH_NT + threadIdx.x;
  const long long local = gid / s.lpj;
  s.sub = (int)(gid % s.lpj);
  s.valid = local < (long long)nloops * h.K;
  s.jidx = local + (long long)l_first * h.K;
  s.l = (int)(s.jidx / h.K);
  s.i = (int)(s.jidx - (long long)s.l * h.K);
  s.dl = m / s.lpj;
  s.d0 = s.sub * s.dl;
  return s;
}

// whether location loop lp walked loc: its bucket J was selected, J's window is not cut off
// by the largest approved residue, and loc's residue is approved
__device__ __forceinline__ bool hash_walked(const HashArgs& h, int lp, uint32_t loc) {
  const long long n = h.n;
  const long long m = n / h.B;
  const uint64_t p = ((uint64_t)loc * (uint64_t)(uint32_t)h.ai_dev[lp]) & (uint64_t)(n - 1);
  const int lm = 63 - __clzll(m);   // m = n / B, both powers of two
  const uint32_t J = (uint32_t)((((long long)p + (m >> 1)) & (n - 1)) >> lm);
  const long long low = ((long long)J * m - (m >> 1) + n) & (n - 1);
  const uint32_t r = (uint32_t)(p & (uint64_t)(h.W - 1));
  // the three table reads are independent: issued together, tested together
  const unsigned js = h.jsel[(size_t)lp * h.jsel_words + (J >> 5)];
  const int mf = h.maxfirst[lp];
  const unsigned bw = h.bitmap[(size_t)lp * h.bm_words + (r >> 5)];
  return (((js >> (J & 31)) & (bw >> (r & 31))) & 1u) && (int)(low & (h.W - 1)) <= mf;
}

// a location walked by loop l is a hit of loop l when exactly loops_thresh of the loops 0..l
// voted for it, loop l included (the loop in which the reference's count reaches the
// threshold, innerLoop.cu:809-814): loops_thresh - 1 of the loops before l
__device__ __forceinline__ bool hash_is_hit(const HashArgs& h, int l, uint32_t loc) {
  int v = 1;   // loop l's own walk
  for (int lp = 0; lp < l; ++lp) v += hash_walked(h, lp, loc) ? 1 : 0;
  return v == h.loops_thresh;
}

// hit count of a slot's slice, and its hit mask (bit rel of word rel/32 set for a hit at
// slice position rel) into mk[0 .. words)
__device__ __forceinline__ int hash_count_slice(const HashArgs& h, const HashSlot& s,
                                                unsigned* __restrict__ mk, int words) {
  int c = 0;
  if (!s.valid) return 0;
  unsigned cur = 0u;
  int cw = 0;
  hash_walk(h, s.l, s.i, s.d0, s.dl, [&](uint32_t loc, int rel) {
    if (hash_is_hit(h, s.l, loc)) {
      for (; cw < (rel >> 5); ++cw) {
        mk[cw] = cur;
        cur = 0u;
      }
      cur |= 1u << (rel & 31);
      ++c;
    }
  });
  for (; cw < words; ++cw) {
    mk[cw] = cur;
    cur = 0u;
  }
  return c;
}

__global__ __launch_bounds__(HASH_NT) void hash_count_kernel(HashArgs h) {
  __shared__ int sh[16];
  const HashSlot s = hash_slot(h, hash_hit_first_d(h), hash_hit_count_d(h));
  const int words = hash_slice_words((long long)h.n / h.B);
  const long long gid = (long long)blockIdx.x * HASH_NT + threadIdx.x;
  const int c = hash_count_slice(h, s, h.hmask + gid * words, words);
  int tot = c;   // total of the bucket's lanes (aligned groups of lpj lanes)
  for (int o = 1; o < s.lpj; o <<= 1) tot += __shfl_xor(tot, o, 64);
  const int v = (s.valid && s.sub == 0) ? tot : 0;
  const int ex = block_excl_scan(v, sh);
  if (s.valid && s.sub == 0) h.hitoff[s.jidx] = ex;
  if (threadIdx.x == HASH_NT - 1) h.blocksum[blockIdx.x] = ex + v;
}

// the hits of each slice, from the count pass's hit mask (no second walk).  Each block first
// sums the hit totals of the blocks before it (a few KB, read by every block in parallel: no
// separate scan launch); the last block stores the total hit count.
__global__ __launch_bounds__(HASH_NT) void hash_write_kernel(HashArgs h) {
  __shared__ int sh[16];
  int pre = 0;
  for (int b = threadIdx.x; b < (int)blockIdx.x; b += HASH_NT) pre += h.blocksum[b];
  const int boff = block_sum(pre, sh);
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *h.count = boff + h.blocksum[blockIdx.x];
  const HashSlot s = hash_slot(h, hash_hit_first_d(h), hash_hit_count_d(h));
  const int words = hash_slice_words((long long)h.n / h.B);
  const long long gid = (long long)blockIdx.x * HASH_NT + threadIdx.x;
  const unsigned* __restrict__ mk = h.hmask + gid * words;
  int c = 0;
  if (s.valid)
    for (int w = 0; w < words; ++w) c += __popc(mk[w]);
  // exclusive scan over the bucket's lanes (in window order)
  const int lane = threadIdx.x & 63;
  int x = c;
  for (int o = 1; o < s.lpj; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if ((lane & (s.lpj - 1)) >= o) x += y;
  }
  if (!s.valid || c == 0) return;
  int pos = boff + h.hitoff[s.jidx] + (x - c);
  const long long n = h.n;
  const long long m = n / h.B;
  const long long J = h.J[(long long)s.l * h.K + s.i];
  const long long start = ((J * m - (m >> 1)) + n + s.d0) & (n - 1);   // low + d0
  const uint64_t a = (uint64_t)(uint32_t)h.a_dev[s.l];
  for (int w = 0; w < words; ++w) {
    unsigned word = mk[w];
    while (word) {
      const int rel = w * 32 + __ffs(word) - 1;
      word &= word - 1u;
      const uint64_t p = (uint64_t)((start + rel) & (n - 1));
      if (pos < h.capacity) h.hits[pos] = (int)((p * a) & ((uint64_t)n - 1));
      ++pos;
    }
  }
}

hipError_t hash_hits_launch(const HashArgs& h, hipStream_t st) {
  const int nb = hash_blocks(hash_hit_count(h), h.K, (long long)h.n / h.B);
  hipLaunchKernelGGL(hash_count_kernel, dim3(nb), dim3(HASH_NT), 0, st, h);
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

// One thread per hit.  EXACT: the loop count is the template constant MAXL (every address
// first, then every load, so the 2*MAXL reads of a hit are in flight together); otherwise
// loops <= MAXL at run time.  The arithmetic per loop is the same either way.  A block's
// nodes are staged in LDS and stored as consecutive 8-byte words (coalesced, also when `out`
// is pinned host memory written over the bus); count_copy (optional) receives the count.
constexpr int EST_NT = 256;
constexpr int EST_GRID = 512;   // blocks of the estimation launch (grid-stride over the hits)
template <typename Real, int MAXL, bool EXACT>
__global__ __launch_bounds__(EST_NT) void estimate_kernel(
    const int* __restrict__ hits, const int* __restrict__ count, int capacity,
    const typename CT<Real>::T2* __restrict__ bins, int loops_rt, uint32_t n,
    const int* __restrict__ ai_dev, int B, const double2* __restrict__ win,
    Node* __restrict__ out, uint8_t* __restrict__ sel, int* __restrict__ count_copy) {
  static_assert(sizeof(Node) == 24, "Node: key, 4 bytes of padding, two doubles");
  __shared__ unsigned long long stage[EST_NT * 3];
  const int loops = EXACT ? MAXL : loops_rt;
  const int cnt_all = *count;
  const int total = min(cnt_all, capacity);
  if (count_copy && blockIdx.x == 0 && threadIdx.x == 0) *count_copy = cnt_all;
  const long long m = (long long)(n / (uint32_t)B);
  const int lm = 63 - __clzll(m);   // m = n / B, both powers of two
  const long long nn = n;
  const int med = (loops - 1) / 2;
  for (int h0 = blockIdx.x * EST_NT; h0 < total; h0 += gridDim.x * EST_NT) {
    const int h = h0 + (int)threadIdx.x;
    const bool valid = h < total;
    const uint32_t I = (uint32_t)hits[valid ? h : h0];
    long long bidx[MAXL], widx[MAXL];
#pragma unroll
    for (int l = 0; l < MAXL; ++l) {
      const int lc = (EXACT || l < loops) ? l : 0;   // a valid address past the loop count
      const long long p = (long long)(((uint64_t)(uint32_t)ai_dev[lc] * I) & (uint64_t)(n - 1));
      long long hashed = p >> lm;
      long long dist = p & (m - 1);
      if (dist > m / 2) {
        hashed = (hashed + 1) & (B - 1);
        dist -= m;
      }
      dist = (nn - dist) & (nn - 1);
      widx[l] = (dist + m / 2) & (nn - 1);
      bidx[l] = (long long)lc * B + hashed;
    }
    double2 F[MAXL], bb[MAXL];
#pragma unroll
    for (int l = 0; l < MAXL; ++l) {
      F[l] = win[widx[l]];
      bb[l] = to_d2(bins[bidx[l]]);
    }
    double vr[MAXL], vi[MAXL];
#pragma unroll
    for (int l = 0; l < MAXL; ++l) {
      // innerLoop.cu:848-852: products rounded, then combined
      const double abs2 = F[l].x * F[l].x + F[l].y * F[l].y;
      const double mr = bb[l].x * F[l].x + bb[l].y * F[l].y;
      const double mi = bb[l].y * F[l].x - bb[l].x * F[l].y;
      vr[l] = mr / abs2;
      vi[l] = mi / abs2;
    }
    // stable rank selection: position of j in an insertion sort = #(v_i < v_j) + #(i<j, v_i==v_j)
    double rr = 0.0, ri = 0.0;
    int sr = 0, si = 0;
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
        if (pr == med) {
          rr = vr[j];
          sr = j;
        }
        if (pi == med) {
          ri = vi[j];
          si = j;
        }
      }
    }
    stage[3 * threadIdx.x] = (unsigned long long)I;   // key, then 4 zero bytes of padding
    stage[3 * threadIdx.x + 1] = (unsigned long long)__double_as_longlong(rr);
    stage[3 * threadIdx.x + 2] = (unsigned long long)__double_as_longlong(ri);
    if (sel && valid) {
      sel[2 * h] = (uint8_t)sr;
      sel[2 * h + 1] = (uint8_t)si;
    }
    __syncthreads();
    const int words = 3 * min(EST_NT, total - h0);
    unsigned long long* o = reinterpret_cast<unsigned long long*>(out + h0);
    for (int j = threadIdx.x; j < words; j += EST_NT) o[j] = stage[j];
    __syncthreads();
  }
}

template <typename Real, int L, bool EXACT>
static void estimate_go(const int* hits, const int* count, int capacity, const void* bins,
                        int loops, int n, const int* ai_dev, int B, const double2* win,
                        Node* out, uint8_t* sel, int* count_copy, hipStream_t st) {
  using T2 = typename CT<Real>::T2;
  hipLaunchKernelGGL((estimate_kernel<Real, L, EXACT>), dim3(EST_GRID), dim3(EST_NT), 0, st,
                     hits,
                     count, capacity, (const T2*)bins, loops, (uint32_t)n, ai_dev, B, win, out,
                     sel, count_copy);
}

template <typename Real>
static hipError_t estimate_dispatch(const int* hits, const int* count, int capacity,
                                    const void* bins, int loops, int n, const int* ai_dev, int B,
                                    const double2* win, Node* out, uint8_t* sel,
                                    int* count_copy, hipStream_t st) {
#define SFFT_EST(L)                                                                         \
  case L:                                                                                   \
    estimate_go<Real, L, true>(hits, count, capacity, bins, loops, n, ai_dev, B, win, out, \
                               sel, count_copy, st);                                        \
    break;
  switch (loops) {
    SFFT_EST(1) SFFT_EST(2) SFFT_EST(3) SFFT_EST(4) SFFT_EST(5) SFFT_EST(6) SFFT_EST(7)
    SFFT_EST(8) SFFT_EST(9) SFFT_EST(10) SFFT_EST(11) SFFT_EST(12) SFFT_EST(13) SFFT_EST(14)
    SFFT_EST(15) SFFT_EST(16)
    default:
      estimate_go<Real, 32, false>(hits, count, capacity, bins, loops, n, ai_dev, B, win, out,
                                   sel, count_copy, st);
  }
#undef SFFT_EST
  return hipGetLastError();
}

hipError_t estimate_launch(int is_fp32, const int* hits, const int* count, int capacity,
                           const void* bins_f, int loops, int n, const int* ai_dev, int B,
                           const double2* filter_win, Node* out, uint8_t* sel,
                           hipStream_t st, int* count_copy) {
  if (is_fp32)
    return estimate_dispatch<float>(hits, count, capacity, bins_f, loops, n, ai_dev, B,
                                    filter_win, out, sel, count_copy, st);
  return estimate_dispatch<double>(hits, count, capacity, bins_f, loops, n, ai_dev, B,
                                   filter_win, out, sel, count_copy, st);
}

}  // namespace sfft
