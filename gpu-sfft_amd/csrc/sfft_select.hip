// sfft_select.hip -- top-K cutoff on the GPU: the K largest |X|^2 of each segment, ties broken
// by the lower index (the reference's cutoff, innerLoop.cu:551-642), without sorting.
//
// Every element gets a unique composite key  c = (key without sign bit) . (N-1-index)  so that
// "larger c" means "larger |X|^2, or equal and lower index"; the top-K set is {c >= T}.
//   1. hist0   : histogram of the top 12 bits of c, one row per block (wave-aggregated LDS
//                atomics), then colsum: the column sums of the rows.
//   2. split   : every block picks the digit d0 holding the K-th largest from the column
//                sums, then one pass over N: digit > d0 -> selected (emitted at once in
//                unordered mode), digit == d0 -> appended to a candidate list (typically ~1%
//                of N).
//   3. refine  : one workgroup per segment: MSD radix select on the remaining bits of c over
//                the candidates only, with early exit once a digit bin fills the remainder;
//                yields T and (unordered mode) emits the selected candidates.
//   4. ordered mode only: write, an order-preserving compaction of {c >= T} that produces
//                the ascending index list the reverse hash walks in order.  The selected
//                count of each block (its digit > d0 count from the split plus its candidates
//                >= T) and their exclusive prefix come out of the refine.
#include "sfft_internal.h"

namespace sfft {

typedef unsigned __int128 u128;

struct SelState {
  unsigned long long thr_hi, thr_lo;   // T as two 64-bit halves
  int d0;                              // top digit holding the K-th largest
  int need0;                           // K - #(digit > d0)
  int cand_count;
  int out_count;
};

constexpr int SEL_NT = 1024;   // threads of the streaming select kernels (512: +11 us comb
                               // cutoff, +3 us bucket cutoff at n = 2^27, k = 16000)
constexpr int SEL_DIG = 12;
constexpr int SEL_BINS = 1 << SEL_DIG;

template <typename K>
struct KeyBits;
template <>
struct KeyBits<unsigned long long> {
  static constexpr int KB = 63;  // bits below the (always zero) sign bit
  // the first digit is not the key's top bits (sel_first_digit): the refine starts at bit 0
  static constexpr bool digit_is_prefix = false;
};
template <>
struct KeyBits<unsigned int> {
  static constexpr int KB = 31;
  static constexpr bool digit_is_prefix = true;
};

// bits of the index field of a composite key: enough for N - 1 (any N, not only powers of 2)
__device__ __forceinline__ int index_bits(int N) { return N > 1 ? 32 - __clz(N - 1) : 0; }


// elements per block: about 256 blocks per segment for large N, at least 2048
static int sel_chunk(int N) {
  int c = 2048;
  while ((long long)c * 256 < N) c <<= 1;
  return c;
}

size_t select_work_bytes(int segs, int N) {
  const int nch = (N + sel_chunk(N) - 1) / sel_chunk(N);
  return (size_t)segs * sizeof(SelState) + (size_t)segs * (nch + 1) * SEL_BINS * 4 +
         (size_t)segs * 4 + (size_t)segs * nch * 4 * 3 + (size_t)segs * N * 8 * 2 +
         (size_t)segs * N * 4 * 3 + 4096;
}

void select_work_bind(SelectWork* w, void* base, int segs, int N) {
  const int nch = (N + sel_chunk(N) - 1) / sel_chunk(N);
  char* p = (char*)base;
  auto take = [&](size_t bytes) {
    char* r = p;
    p += (bytes + 255) / 256 * 256;
    return r;
  };
  w->state = take((size_t)segs * sizeof(SelState));
  w->hist = (unsigned*)take((size_t)segs * (nch + 1) * SEL_BINS * 4);
  w->done = (unsigned*)take((size_t)segs * 4);
  w->cnt = (int*)take((size_t)segs * nch * 4);
  w->cand_key = take((size_t)segs * N * 8);
  w->cand_idx = (int*)take((size_t)segs * N * 4);
  w->emit_idx = (int*)take((size_t)segs * N * 4);
  w->blk_key = take((size_t)segs * N * 8);
  w->blk_idx = (int*)take((size_t)segs * N * 4);
  w->ccnt = (int*)take((size_t)segs * nch * 4);
  w->ecnt = (int*)take((size_t)segs * nch * 4);
  w->max_segs = segs;
  w->max_chunks = nch;
}

__device__ __forceinline__ int block_excl_scan_s(int v, int* sh) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
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

// increment h[bin] for the active lanes: one atomic for the whole wave when every active lane
// holds the same bin (flat or tied data), else one LDS atomic per lane (spread data, where
// per-bin aggregation would loop once per distinct bin)
__device__ __forceinline__ void wave_hist_add(unsigned* h, unsigned bin, bool active) {
  const int lane = threadIdx.x & 63;
  const unsigned long long act = __ballot(active);
  if (!act) return;
  const int first = __ffsll((long long)act) - 1;
  const unsigned b0 = __shfl(bin, first, 64);
  const unsigned long long same = __ballot(active && bin == b0);
  if (same == act) {
    if (lane == first) atomicAdd(&h[b0], (unsigned)__popcll(act));
  } else if (active) {
    atomicAdd(&h[bin], 1u);
  }
}

// wave-aggregated append: returns the slot of this lane (valid only where pred)
__device__ __forceinline__ int wave_append(int* counter, bool pred) {
  const int lane = threadIdx.x & 63;
  const unsigned long long m = __ballot(pred);
  if (!m) return -1;
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, __popcll(m));
  base = __shfl(base, leader, 64);
  return base + __popcll(m & ((1ull << lane) - 1ull));
}

template <typename K>
__device__ __forceinline__ unsigned top_digit(K key) {
  return sel_first_digit(key);
}

template <typename K>
__device__ __forceinline__ u128 composite(K key, unsigned idx, int N, int IB) {
  const K km = key & (K)(~((K)1 << KeyBits<K>::KB));
  return ((u128)km << IB) | (u128)(unsigned)((unsigned)(N - 1) - idx);
}

// ---- 1. top-digit histogram: one row of SEL_BINS counts per block (plain stores) ---------
template <typename K>
__global__ __launch_bounds__(SEL_NT) void sel_hist0_kernel(const K* __restrict__ keys,
                                                           long long seg_stride, int N, int chunk,
                                                           int nchunks,
                                                           unsigned* __restrict__ rows,
                                                           const uint16_t* __restrict__ digits) {
  __shared__ unsigned h[SEL_BINS];
  const int tid = threadIdx.x, seg = blockIdx.y;
  for (int i = tid; i < SEL_BINS; i += SEL_NT) h[i] = 0u;
  __syncthreads();
  const K* kp = keys + (long long)seg * seg_stride;
  const int begin = blockIdx.x * chunk, end = min(N, begin + chunk);
  constexpr int U = 8;   // independent loads in flight per thread
  if (digits) {          // the first digits only (2 bytes per key instead of 4 or 8)
    const uint16_t* dp = digits + (long long)seg * seg_stride;
    for (int base = begin; base < end; base += SEL_NT * U) {
      unsigned dd[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = base + u * SEL_NT + tid;
        dd[u] = i < end ? (unsigned)dp[i] : 0u;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = base + u * SEL_NT + tid;
        wave_hist_add(h, dd[u], i < end);
      }
    }
  } else {
    for (int base = begin; base < end; base += SEL_NT * U) {
      K kk[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = base + u * SEL_NT + tid;
        kk[u] = i < end ? kp[i] : (K)0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = base + u * SEL_NT + tid;
        wave_hist_add(h, top_digit<K>(kk[u]), i < end);
      }
    }
  }
  __syncthreads();
  unsigned* row = rows + ((size_t)seg * nchunks + blockIdx.x) * SEL_BINS;
  for (int i = tid; i < SEL_BINS; i += SEL_NT) row[i] = h[i];
}

// ---- 1b. column sums of the block rows ----------------------------------------------------
// a block sums 16 bins, 16 threads per bin each over every 16th row, then reduces in LDS
constexpr int PICK_NT = 256;
constexpr int PICK_BINS = 16;
__global__ __launch_bounds__(PICK_NT) void sel_colsum_kernel(int nchunks,
                                                             const unsigned* __restrict__ rows,
                                                             unsigned* __restrict__ ghist) {
  __shared__ unsigned part[PICK_NT / PICK_BINS][PICK_BINS];
  const int tid = threadIdx.x, seg = blockIdx.y;
  const int bl = tid % PICK_BINS, g = tid / PICK_BINS;
  constexpr int G = PICK_NT / PICK_BINS;
  const int bin0 = blockIdx.x * PICK_BINS;
  const unsigned* r = rows + (size_t)seg * nchunks * SEL_BINS + bin0 + bl;
  unsigned a = 0u;
#pragma unroll 4
  for (int c = g; c < nchunks; c += G) a += r[(size_t)c * SEL_BINS];
  part[g][bl] = a;
  __syncthreads();
  if (tid < PICK_BINS) {
    unsigned acc = 0u;
#pragma unroll
    for (int j = 0; j < G; ++j) acc += part[j][tid];
    ghist[(size_t)seg * SEL_BINS + bin0 + tid] = acc;
  }
}

// ---- 2. split: collect digit == d0 (candidates), count digit > d0 (and emit them, unordered
// mode) ---------------------------------------------------------------------------------------
// Block b writes its candidates to [b*chunk, b*chunk + count) of the segment's candidate
// arrays, its emitted indices to the same range of the emitted array, and the two counts to
// ccnt/ecnt[seg][b]: no global atomics (the refine pass compacts).  Keys stored in a KeyMap
// order are recorded with their natural indices (the tie rule and the output use those).
template <typename K>
__global__ __launch_bounds__(SEL_NT) void sel_split_kernel(const K* __restrict__ keys,
                                                           long long seg_stride, int N, int chunk,
                                                           int nchunks, int emit, int Ksel,
                                                           const unsigned* __restrict__ ghist,
                                                           SelState* __restrict__ st,
                                                           K* __restrict__ ckey,
                                                           int* __restrict__ cidx,
                                                           int* __restrict__ eidx,
                                                           int* __restrict__ ccnt,
                                                           int* __restrict__ ecnt,
                                                           KeyMap km,
                                                           const uint16_t* __restrict__ digits) {
  static_assert(SEL_NT * 4 == SEL_BINS, "the pick gives each thread four bins");
  __shared__ int s_nc, s_ne, s_d0;
  __shared__ int sh[16];
  const int tid = threadIdx.x, seg = blockIdx.y;
  if (tid == 0) {
    s_nc = 0;
    s_ne = 0;
    s_d0 = 0;
  }
  // the top digit d0 holding the K-th largest key, from the column sums: every block finds
  // the same one (16 KB of sums, no separate pick launch); block 0 records it and the count
  // still needed inside it for the refine
  {
    const unsigned* gh = ghist + (size_t)seg * SEL_BINS;
    unsigned v[4];
    int S = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v[j] = gh[SEL_BINS - 1 - (4 * tid + j)];
      S += (int)v[j];
    }
    const int E = block_excl_scan_s(S, sh);   // (its barriers also publish the defaults)
    if (E < Ksel && Ksel <= E + S) {
      int cum = E;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (cum + (int)v[j] >= Ksel) {
          const int dd = SEL_BINS - 1 - (4 * tid + j);
          s_d0 = dd;
          if (blockIdx.x == 0) {
            st[seg].d0 = dd;
            st[seg].need0 = Ksel - cum;
          }
          break;
        }
        cum += (int)v[j];
      }
    }
  }
  __syncthreads();
  const unsigned d0 = (unsigned)s_d0;
  const K* kp = keys + (long long)seg * seg_stride;
  const long long rb = (long long)seg * N + (long long)blockIdx.x * chunk;
  K* ck = ckey + rb;
  int* ci = cidx + rb;
  int* ce = eidx + rb;
  const int begin = blockIdx.x * chunk, end = min(N, begin + chunk);
  const uint16_t* dp = digits ? digits + (long long)seg * seg_stride : nullptr;
  // SU elements per thread at a time: their digit (or key) loads are in flight together, then
  // the candidates' key loads, so that the loop waits twice per SU elements, not twice per
  // element (the order of a block's candidates and winners is set by the wave appends either
  // way; every later pass is independent of it)
  constexpr int SU = 8;
  for (int base = begin; base < end; base += SEL_NT * SU) {
    unsigned d[SU];
    K k[SU];
#pragma unroll
    for (int u = 0; u < SU; ++u) {
      const int i = base + u * SEL_NT + tid;
      const bool ok = i < end;
      if (dp) {
        d[u] = ok ? (unsigned)dp[i] : 0u;
      } else {
        k[u] = ok ? kp[i] : (K)0;
        d[u] = top_digit<K>(k[u]);
      }
    }
    // with the first digits at hand, a key is read only when it is a candidate
    if (dp) {
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        const int i = base + u * SEL_NT + tid;
        k[u] = (i < end && d[u] == d0) ? kp[i] : (K)0;
      }
    }
#pragma unroll
    for (int u = 0; u < SU; ++u) {
      const int i = base + u * SEL_NT + tid;
      const bool ok = i < end;
      const bool isc = ok && d[u] == d0;
      const int pc = wave_append(&s_nc, isc);
      if (isc) {
        ck[pc] = k[u];
        ci[pc] = (int)key_natural(km, i);
      }
      const bool ise = ok && d[u] > d0;
      const int pe = wave_append(&s_ne, ise);
      if (emit && ise) ce[pe] = (int)key_natural(km, i);
    }
  }
  __syncthreads();
  if (tid == 0) {
    ccnt[(long long)seg * nchunks + blockIdx.x] = s_nc;
    ecnt[(long long)seg * nchunks + blockIdx.x] = s_ne;
  }
}

constexpr int SEL_MAX_CHUNKS = 1024;   // blocks per segment
constexpr int CMP_NT = 256;

// ---- 2b. compaction: block b of the split moves its candidates to the contiguous list at
// the sum of the counts of blocks < b, and (unordered mode) its emitted indices to out --------
template <typename K>
__global__ __launch_bounds__(CMP_NT) void sel_compact_kernel(int N, int Ksel, int chunk,
                                                             int nchunks, int emit,
                                                             SelState* __restrict__ st,
                                                             const K* __restrict__ bkey,
                                                             const int* __restrict__ bidx,
                                                             const int* __restrict__ eidx,
                                                             const int* __restrict__ ccnt,
                                                             const int* __restrict__ ecnt,
                                                             K* __restrict__ ckey,
                                                             int* __restrict__ cidx,
                                                             int* __restrict__ out) {
  __shared__ int sh[16];
  const int tid = threadIdx.x, seg = blockIdx.y, b = blockIdx.x;
  const int* cc = ccnt + (long long)seg * nchunks;
  const int* ec = ecnt + (long long)seg * nchunks;
  int pc = 0, pe = 0;
  for (int c = tid; c < b; c += CMP_NT) {
    pc += cc[c];
    pe += ec[c];
  }
  const int oc = block_excl_scan_s(pc, sh), oe = block_excl_scan_s(pe, sh);
  // the block totals: exclusive scan of the last thread plus its own value
  __shared__ int s_oc, s_oe;
  if (tid == CMP_NT - 1) {
    s_oc = oc + pc;
    s_oe = oe + pe;
  }
  __syncthreads();
  const int offc = s_oc, offe = s_oe;
  const int nc = cc[b], ne = ec[b];
  const long long src = (long long)seg * N + (long long)b * chunk;
  const long long dst = (long long)seg * N + offc;
  for (int j = tid; j < nc; j += CMP_NT) {
    ckey[dst + j] = bkey[src + j];
    cidx[dst + j] = bidx[src + j];
  }
  if (emit) {
    int* o = out + (long long)seg * Ksel;
    for (int j = tid; j < ne; j += CMP_NT)
      if (offe + j < Ksel) o[offe + j] = eidx[src + j];
  }
  if (b == nchunks - 1 && tid == 0) {
    st[seg].cand_count = offc + nc;
    st[seg].out_count = offe + ne;
  }
}

// minimum of v over a 1024-thread block (every thread gets it); s: 16 entries of LDS
__device__ __forceinline__ u128 block_min_u128(u128 v, u128* s) {
  unsigned long long hi = (unsigned long long)(v >> 64), lo = (unsigned long long)v;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long h2 = __shfl_xor(hi, o, 64), l2 = __shfl_xor(lo, o, 64);
    if (h2 < hi || (h2 == hi && l2 < lo)) {
      hi = h2;
      lo = l2;
    }
  }
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s[w] = ((u128)hi << 64) | (u128)lo;
  __syncthreads();
  u128 m = s[0];
  for (int i = 1; i < (int)(blockDim.x >> 6); ++i)
    if (s[i] < m) m = s[i];
  __syncthreads();
  return m;
}

// ---- 3. refine over the candidates (one workgroup per segment) ----------------------------
// The first level scans the candidate list; from then on only the survivors of the previous
// level (the composite keys under the chosen prefix, usually a handful) are scanned, from LDS
// (back to the full list if they outgrow RF_CAP).  With ccnt set the candidates are read
// where the split left them (block b's at [b*chunk, b*chunk + ccnt[b])) instead of from the
// compacted list: the ordered mode needs no compaction launch.  With base set (ordered mode),
// base[b] = the number of selected elements in blocks < b, from the split's digit > d0 counts
// ecnt and the candidates >= T of each block: the write pass needs no count pass.
constexpr int RF_CAP = 2048;
template <typename K>
__global__ __launch_bounds__(1024) void sel_refine_kernel(int N, int Ksel, int emit,
                                                          SelState* __restrict__ st,
                                                          const K* __restrict__ ckey,
                                                          const int* __restrict__ cidx,
                                                          int* __restrict__ out,
                                                          const int* __restrict__ ccnt,
                                                          int nchunks, int chunk,
                                                          const int* __restrict__ ecnt,
                                                          int* __restrict__ base) {
  __shared__ unsigned h[SEL_BINS];
  __shared__ u128 lst[2][RF_CAP];
  __shared__ int s_len[2];
  __shared__ int sh[16];
  __shared__ int s_digit, s_above, s_exit;
  __shared__ u128 s_min[16];
  __shared__ int s_out;
  __shared__ int s_pre[SEL_MAX_CHUNKS + 1];
  __shared__ int s_bc[SEL_MAX_CHUNKS];
  const int tid = threadIdx.x, seg = blockIdx.x;
  const int IB = index_bits(N);
  const int TB = KeyBits<K>::KB + IB;
  int nc;
  if (ccnt) {   // exclusive prefix of the per-block candidate counts (nchunks <= 1024)
    const int v = tid < nchunks ? ccnt[(long long)seg * nchunks + tid] : 0;
    const int ex = block_excl_scan_s(v, sh);
    if (tid < nchunks) s_pre[tid] = ex;
    if (tid == nchunks - 1) s_pre[nchunks] = ex + v;
    __syncthreads();
    nc = s_pre[nchunks];
  } else {
    nc = st[seg].cand_count;
  }
  // storage position of candidate j: the last block b with s_pre[b] <= j holds it
  auto at = [&](int j) -> long long {
    if (!ccnt) return j;
    int lo = 0, hi = nchunks - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_pre[mid] <= j) lo = mid;
      else hi = mid - 1;
    }
    return (long long)lo * chunk + (j - s_pre[lo]);
  };
  const K* ck = ckey + (long long)seg * N;
  const int* ci = cidx + (long long)seg * N;
  if (tid == 0) s_out = 0;
  int r = st[seg].need0;
  // the candidates share the first digit d0; where it is the top SEL_DIG bits of c the radix
  // continues below it, else it starts over at the top of c (one more level)
  u128 prefix = KeyBits<K>::digit_is_prefix ? (u128)(unsigned)st[seg].d0 : (u128)0;
  int consumed = KeyBits<K>::digit_is_prefix ? SEL_DIG : 0;
  u128 T = 0;
  bool done = false;
  int cur = -1;   // survivors: -1 = the candidate list, else lst[cur][0 .. s_len[cur])
  // at most ceil(TB / SEL_DIG) levels; the bound only guards against a broken invariant
  for (int level = 0; !done && level < (TB + SEL_DIG - 1) / SEL_DIG; ++level) {
    const int width = min(SEL_DIG, TB - consumed);
    const int shift = TB - consumed - width;
    for (int i = tid; i < SEL_BINS; i += 1024) h[i] = 0u;
    if (tid == 0) {
      s_digit = 0;
      s_above = 0;
      s_exit = 1;
    }
    __syncthreads();
    const unsigned dmask = (1u << width) - 1u;
    if (cur < 0) {
      for (int j = tid; j < nc; j += 1024) {
        const long long q = at(j);
        const u128 c = composite<K>(ck[q], (unsigned)ci[q], N, IB);
        if ((c >> (shift + width)) == prefix) atomicAdd(&h[(unsigned)(c >> shift) & dmask], 1u);
      }
    } else {
      for (int j = tid; j < s_len[cur]; j += 1024)
        atomicAdd(&h[(unsigned)(lst[cur][j] >> shift) & dmask], 1u);
    }
    __syncthreads();
    const int nb = 1 << width;
    unsigned v[4];
    int S = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int rank = 4 * tid + j;
      v[j] = rank < nb ? h[nb - 1 - rank] : 0u;
      S += (int)v[j];
    }
    const int E = block_excl_scan_s(S, sh);   // (contains the barrier after the defaults)
    if (E < r && r <= E + S) {
      int cum = E;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (cum + (int)v[j] >= r) {
          s_digit = nb - 1 - (4 * tid + j);
          s_above = cum;
          s_exit = ((cum + (int)v[j] == r) || (shift == 0)) ? 1 : 0;
          break;
        }
        cum += (int)v[j];
      }
    }
    __syncthreads();
    const unsigned d = (unsigned)s_digit;
    prefix = (prefix << width) | (u128)d;
    r -= s_above;
    consumed += width;
    if (s_exit) {
      T = prefix << shift;
      if (!KeyBits<K>::digit_is_prefix && shift > 0) {
        // keys outside the candidate list can share these top bits (their first digit is
        // below d0, so they are smaller than every candidate): the threshold is the smallest
        // candidate under the prefix instead of the prefix's lower bound
        u128 mn = ~(u128)0;
        if (cur < 0) {
          for (int j = tid; j < nc; j += 1024) {
            const long long q = at(j);
            const u128 c = composite<K>(ck[q], (unsigned)ci[q], N, IB);
            if ((c >> shift) == prefix && c < mn) mn = c;
          }
        } else {
          for (int j = tid; j < s_len[cur]; j += 1024) {
            const u128 c = lst[cur][j];
            if ((c >> shift) == prefix && c < mn) mn = c;
          }
        }
        T = block_min_u128(mn, s_min);
      }
      done = true;
    } else {
      // keep the survivors: the keys whose bits above `shift` equal the new prefix
      const int nxt = (cur < 0) ? 0 : 1 - cur;
      if (tid == 0) s_len[nxt] = 0;
      __syncthreads();
      if (cur < 0) {
        for (int j = tid; j < nc; j += 1024) {
          const long long q = at(j);
          const u128 c = composite<K>(ck[q], (unsigned)ci[q], N, IB);
          if ((c >> shift) == prefix) {
            const int p = atomicAdd(&s_len[nxt], 1);
            if (p < RF_CAP) lst[nxt][p] = c;
          }
        }
      } else {
        for (int j = tid; j < s_len[cur]; j += 1024) {
          const u128 c = lst[cur][j];
          if ((c >> shift) == prefix) {
            const int p = atomicAdd(&s_len[nxt], 1);
            if (p < RF_CAP) lst[nxt][p] = c;
          }
        }
      }
      __syncthreads();
      cur = (s_len[nxt] <= RF_CAP) ? nxt : -1;
      if (cur >= 0 && s_len[cur] <= 1024) {
        // few survivors: the r-th largest of them directly, by rank (composites are distinct)
        const int S = s_len[cur];
        if (tid < S) {
          const u128 c = lst[cur][tid];
          int above = 0;
          for (int j = 0; j < S; ++j) above += lst[cur][j] > c ? 1 : 0;
          if (above == r - 1) s_min[0] = c;
        }
        __syncthreads();
        T = s_min[0];
        done = true;
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    st[seg].thr_hi = (unsigned long long)(T >> 64);
    st[seg].thr_lo = (unsigned long long)T;
  }
  if (base) {
    // selected per block: digit > d0 (split) + candidates >= T; then the exclusive prefix
    for (int b = tid; b < nchunks; b += 1024) s_bc[b] = 0;
    __syncthreads();
    for (int j = tid; j < nc; j += 1024) {
      const long long q = at(j);
      if (composite<K>(ck[q], (unsigned)ci[q], N, IB) >= T) atomicAdd(&s_bc[(int)(q / chunk)], 1);
    }
    __syncthreads();
    const int v = tid < nchunks ? s_bc[tid] + ecnt[(long long)seg * nchunks + tid] : 0;
    const int ex = block_excl_scan_s(v, sh);
    if (tid < nchunks) base[(long long)seg * nchunks + tid] = ex;
  }
  if (emit) {
    // after the emitted indices (placed by the compaction): the candidates >= T
    int* o = out + (long long)seg * Ksel;
    const int ne = st[seg].out_count;
    for (int j0 = 0; j0 < nc; j0 += 1024) {
      const int j = j0 + tid;
      const long long q = j < nc ? at(j) : 0;
      const bool sel = j < nc && composite<K>(ck[q], (unsigned)ci[q], N, IB) >= T;
      const int po = ne + wave_append(&s_out, sel);
      if (sel && po < Ksel) o[po] = ci[q];
    }
  }
}

// ---- 4. ordered compaction of {c >= T} -----------------------------------------------------
template <typename K>
__device__ __forceinline__ bool sel_pred(K k, int i, int N, int IB, u128 T) {
  return composite<K>(k, (unsigned)i, N, IB) >= T;
}

template <typename K>
__global__ __launch_bounds__(SEL_NT) void sel_write_kernel(const K* __restrict__ keys,
                                                           long long seg_stride, int N, int chunk,
                                                           int nchunks, int Ksel,
                                                           const SelState* __restrict__ st,
                                                           const int* __restrict__ base0,
                                                           int* __restrict__ out,
                                                           unsigned* __restrict__ bits) {
  __shared__ int sh[16];
  __shared__ int s_tot;
  const int seg = blockIdx.y, tid = threadIdx.x;
  const int IB = index_bits(N);
  const u128 T = ((u128)st[seg].thr_hi << 64) | (u128)st[seg].thr_lo;
  const K* kp = keys + (long long)seg * seg_stride;
  int base = base0[(long long)seg * nchunks + blockIdx.x];
  int* o = out + (long long)seg * Ksel;
  const int begin = blockIdx.x * chunk, end = min(N, begin + chunk);
  for (int b0 = begin; b0 < end; b0 += SEL_NT) {
    const int i = b0 + tid;
    const bool p = i < end && sel_pred<K>(kp[i], i, N, IB, T);
    if (bits) {
      // selection bitmap: each wave owns the two words of its 64 indices (b0 is a multiple of
      // 64), so every word of the segment is stored once, with plain stores
      const unsigned long long bal = __ballot(p);
      const int lane = tid & 63;
      const int i0 = i - lane;
      unsigned* bw = bits + (long long)seg * ((N + 31) / 32);
      if (lane == 0 && i0 < N) bw[i0 >> 5] = (unsigned)bal;
      if (lane == 32 && i0 + 32 < N) bw[(i0 + 32) >> 5] = (unsigned)(bal >> 32);
    }
    const int ex = block_excl_scan_s(p ? 1 : 0, sh);
    if (p && base + ex < Ksel) o[base + ex] = i;
    if (tid == SEL_NT - 1) s_tot = ex + (p ? 1 : 0);   // total of this round
    __syncthreads();
    base += s_tot;
    __syncthreads();
  }
}

template <typename K>
static hipError_t select_impl(const K* keys, long long seg_stride, int N, int Ksel, int segs,
                              int* out, int ordered, const SelectWork& w, hipStream_t st,
                              unsigned* bits, const KeyMap* map, const uint16_t* digits) {
  const int chunk = sel_chunk(N);
  const int nchunks = (N + chunk - 1) / chunk;
  if (nchunks > SEL_MAX_CHUNKS || nchunks > w.max_chunks || segs > w.max_segs)
    return hipErrorInvalidValue;
  KeyMap km;
  if (map && map->mid != 0) {
    // the ordered compaction walks storage order: natural order only
    if (ordered || (1LL << map->log2n) != (long long)N) return hipErrorInvalidValue;
    km = *map;
  }
  dim3 grid(nchunks, segs);
  SelState* S = (SelState*)w.state;
  // work layout: column sums, then the per-block histogram rows
  unsigned* rows = w.hist + (size_t)segs * SEL_BINS;
  hipLaunchKernelGGL((sel_hist0_kernel<K>), grid, dim3(SEL_NT), 0, st, keys, seg_stride, N,
                     chunk, nchunks, rows, digits);
  hipLaunchKernelGGL(sel_colsum_kernel, dim3(SEL_BINS / PICK_BINS, segs), dim3(PICK_NT), 0, st,
                     nchunks, (const unsigned*)rows, w.hist);
  // (the split blocks pick d0 from the column sums themselves)
  hipLaunchKernelGGL((sel_split_kernel<K>), grid, dim3(SEL_NT), 0, st, keys, seg_stride, N,
                     chunk, nchunks, ordered ? 0 : 1, Ksel, (const unsigned*)w.hist, S,
                     (K*)w.blk_key,
                     w.blk_idx, w.emit_idx, w.ccnt, w.ecnt, km, digits);
  if (ordered) {
    // the refine reads the candidates from the split's per-block regions
    hipLaunchKernelGGL((sel_refine_kernel<K>), dim3(segs), dim3(1024), 0, st, N, Ksel, 0, S,
                       (const K*)w.blk_key, (const int*)w.blk_idx, out, (const int*)w.ccnt,
                       nchunks, chunk, (const int*)w.ecnt, w.cnt);
  } else {
    hipLaunchKernelGGL((sel_compact_kernel<K>), grid, dim3(CMP_NT), 0, st, N, Ksel, chunk,
                       nchunks, 1, S, (const K*)w.blk_key, (const int*)w.blk_idx,
                       (const int*)w.emit_idx, (const int*)w.ccnt, (const int*)w.ecnt,
                       (K*)w.cand_key, w.cand_idx, out);
    hipLaunchKernelGGL((sel_refine_kernel<K>), dim3(segs), dim3(1024), 0, st, N, Ksel, 1, S,
                       (const K*)w.cand_key, (const int*)w.cand_idx, out, (const int*)nullptr,
                       nchunks, chunk, (const int*)nullptr, (int*)nullptr);
  }
  if (ordered) {
    hipLaunchKernelGGL((sel_write_kernel<K>), grid, dim3(SEL_NT), 0, st, keys, seg_stride, N,
                       chunk, nchunks, Ksel, (const SelState*)S, (const int*)w.cnt, out, bits);
  }
  return hipGetLastError();
}

hipError_t select_topk(int is_fp32, const void* keys, long long seg_stride, int N, int K,
                       int segs, int* out, int ordered, const SelectWork& w, hipStream_t st,
                       unsigned* bits, const KeyMap* map, const uint16_t* digits) {
  if (ordered && digits) return hipErrorInvalidValue;   // the write pass reads the keys
  if (N < 2 || K < 1 || K > N || segs < 1) return hipErrorInvalidValue;
  if (is_fp32)
    return select_impl<unsigned int>((const unsigned int*)keys, seg_stride, N, K, segs, out,
                                     ordered, w, st, bits, map, digits);
  return select_impl<unsigned long long>((const unsigned long long*)keys, seg_stride, N, K, segs,
                                         out, ordered, w, st, bits, map, digits);
}

}  // namespace sfft
