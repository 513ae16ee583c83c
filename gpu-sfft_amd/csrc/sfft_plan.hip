// sfft_plan.hip -- host orchestration of one sparse-FFT transform on the GPU (the native plan
// API of include/sfft.h) and the canonical FFT tables shared by every FFT launch.
//
// Per transform, on the plan's stream, with no host round trip until the final copy:
//   clear bitmaps -> comb FFT (gather fused into pass 0, |Y|^2 keys out of the last
//   pass) -> top-B_thresh of each comb -> approved-residue bitmaps per location loop ->
//   permute+filter of every loop -> batched bucket FFT (bins_f + keys of location loops) ->
//   top-B_thresh per location loop (+ selection bitmaps) -> deterministic hit list ->
//   estimation.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "sfft_internal.h"

namespace sfft {

// ---- canonical twiddles: exp(-+2 pi i e/T), octant reduction so cos/sin see [0, pi/4] ----
void canon_twiddle(uint64_t e, uint64_t T, int inverse, double* cr, double* sr) {
  double c, s;
  e %= T;
  if (T < 8) {
    const uint64_t e4 = e * (4 / T);
    static const double cq[4] = {1.0, 0.0, -1.0, 0.0};
    static const double sq[4] = {0.0, 1.0, 0.0, -1.0};
    c = cq[e4 & 3];
    s = sq[e4 & 3];
  } else {
    const uint64_t quarter = T / 4;
    const uint64_t qd = e / quarter;
    const uint64_t r = e % quarter;
    const bool swap = 2 * r > quarter;
    const uint64_t rr = swap ? (quarter - r) : r;
    const double ang = (2.0 * M_PI * (double)rr) / (double)T;
    double c0, s0;
    sincos(ang, &s0, &c0);   // one libm call: sin()/cos() may round differently from it
    if (swap) {
      const double t = c0;
      c0 = s0;
      s0 = t;
    }
    switch (qd) {
      case 0: c = c0; s = s0; break;
      case 1: c = -s0; s = c0; break;
      case 2: c = -c0; s = -s0; break;
      default: c = s0; s = -c0; break;
    }
  }
  *cr = c;
  *sr = inverse ? s : -s;
}

int fft_plan_passes(int log2n, int* pass_logs) {
  if (log2n <= 0) return 0;
  const int p = (log2n + kMaxPassLog - 1) / kMaxPassLog;
  const int base = log2n / p, extra = log2n % p;
  for (int i = 0; i < p; ++i) pass_logs[i] = base + (i < extra ? 1 : 0);
  return p;
}

hipError_t fft_tables_create(int log2n, int inverse, int is_fp32, FftTables* t) {
  *t = FftTables();
  t->log2n = log2n;
  t->inverse = inverse;
  t->is_fp32 = is_fp32;
  t->npass = fft_plan_passes(log2n, t->logs);
  const size_t esz = is_fp32 ? 8 : 16;
  // layout: per pass [twR (R)] [tdir (Lcur) | thi (Lcur>>s) tlo (1<<s)] [tw4]; tw4: for
  // each radix-4 stage with ns = 4^k > 1 (k = 1, 2, ... while 2k <= r - 2), ns quads
  // {twR[jm*t], twR[2*jm*t], twR[3*jm*t], 0} (t = R / (4*ns), jm < ns), stage k at quad
  // (4^k - 4) / 3: a stage's three twiddles of neighbouring butterflies are neighbours
  size_t cnt[kMaxPasses][5] = {};
  size_t total = 0;
  int log2Lprev = 0;
  for (int p = 0; p < t->npass; ++p) {
    const int r = t->logs[p];
    const int log2Lcur = log2Lprev + r;
    cnt[p][0] = (size_t)1 << r;
    if (p > 0) {
      if (log2Lcur <= kMaxPassLog) {
        cnt[p][1] = (size_t)1 << log2Lcur;
      } else {
        // low table of at most 256 entries: read at scattered indices, it stays in L1
        t->s[p] = (log2Lcur + 1) / 2 < 8 ? (log2Lcur + 1) / 2 : 8;
        cnt[p][2] = ((size_t)1 << log2Lcur) >> t->s[p];
        cnt[p][3] = (size_t)1 << t->s[p];
      }
    }
    for (int k = 1; 2 * k <= r - 2; ++k) cnt[p][4] += (size_t)4 << (2 * k);
    for (int j = 0; j < 5; ++j) total += cnt[p][j];
    log2Lprev = log2Lcur;
  }
  t->bytes = total * esz;
  std::vector<unsigned char> host(t->bytes ? t->bytes : 1);
  size_t off = 0;
  std::vector<size_t> offs;
  auto put = [&](uint64_t e, uint64_t T) {
    double c, s;
    canon_twiddle(e, T, inverse, &c, &s);
    if (is_fp32) {
      float v[2] = {(float)c, (float)s};
      memcpy(&host[off], v, 8);
    } else {
      double v[2] = {c, s};
      memcpy(&host[off], v, 16);
    }
    off += esz;
  };
  auto put_zero = [&]() {
    memset(&host[off], 0, esz);
    off += esz;
  };
  log2Lprev = 0;
  size_t start[kMaxPasses][5] = {};
  for (int p = 0; p < t->npass; ++p) {
    const int r = t->logs[p];
    const int log2Lcur = log2Lprev + r;
    const uint64_t R = 1ull << r, Lcur = 1ull << log2Lcur;
    start[p][0] = off;
    for (uint64_t e = 0; e < R; ++e) put(e, R);
    start[p][1] = off;
    for (size_t e = 0; e < cnt[p][1]; ++e) put(e, Lcur);
    start[p][2] = off;
    for (size_t e = 0; e < cnt[p][2]; ++e) put((uint64_t)e << t->s[p], Lcur);
    start[p][3] = off;
    for (size_t e = 0; e < cnt[p][3]; ++e) put(e, Lcur);
    start[p][4] = off;
    for (int k = 1; 2 * k <= r - 2; ++k) {
      const uint64_t ns = 1ull << (2 * k), ts = R / (4 * ns);
      for (uint64_t jm = 0; jm < ns; ++jm) {
        put(jm * ts, R);
        put(2 * jm * ts, R);
        put(3 * jm * ts, R);
        put_zero();
      }
    }
    log2Lprev = log2Lcur;
  }
  void* d = nullptr;
  hipError_t err = hipMalloc(&d, t->bytes ? t->bytes : 16);
  if (err != hipSuccess) return err;
  if (t->bytes) {
    err = hipMemcpy(d, host.data(), t->bytes, hipMemcpyHostToDevice);
    if (err != hipSuccess) {
      (void)hipFree(d);
      return err;
    }
  }
  t->storage = d;
  for (int p = 0; p < t->npass; ++p) {
    unsigned char* base = (unsigned char*)d;
    t->twR[p] = base + start[p][0];
    t->tdir[p] = cnt[p][1] ? base + start[p][1] : nullptr;
    t->thi[p] = cnt[p][2] ? base + start[p][2] : nullptr;
    t->tlo[p] = cnt[p][3] ? base + start[p][3] : nullptr;
    t->tw4[p] = cnt[p][4] ? base + start[p][4] : nullptr;
  }
  return hipSuccess;
}

void fft_tables_destroy(FftTables* t) {
  if (t->storage) (void)hipFree(t->storage);
  t->storage = nullptr;
}

// cache of tables for the standalone device FFT entry point
struct TabKey {
  int log2n, inverse, fp32, device;
};
static std::mutex g_tab_mu;
static std::vector<std::pair<TabKey, FftTables*>> g_tabs;

const FftTables* cached_tables(int log2n, int inverse, int fp32, hipError_t* err) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_tab_mu);
  for (auto& kv : g_tabs)
    if (kv.first.log2n == log2n && kv.first.inverse == inverse && kv.first.fp32 == fp32 &&
        kv.first.device == dev)
      return kv.second;
  FftTables* t = new FftTables();
  *err = fft_tables_create(log2n, inverse, fp32, t);
  if (*err != hipSuccess) {
    delete t;
    return nullptr;
  }
  g_tabs.push_back({TabKey{log2n, inverse, fp32, dev}, t});
  return t;
}

}  // namespace sfft

using namespace sfft;

#include "sfft_plan_impl.h"

static int ilog2_exact(long long v) {
  if (v <= 0) return -1;
  int l = 0;
  while ((1LL << l) < v) ++l;
  return ((1LL << l) == v) ? l : -1;
}

extern "C" const char* sfft_status_string(sfft_status s) {
  switch (s) {
    case SFFT_OK: return "ok";
    case SFFT_ERR_INVALID: return "invalid parameters";
    case SFFT_ERR_ALLOC: return "device allocation failed";
    case SFFT_ERR_HIP: return "HIP runtime error";
    case SFFT_ERR_HIT_OVERFLOW: return "located indices exceed hit capacity";
    case SFFT_ERR_NO_FILTER: return "plan has no filter";
    case SFFT_ERR_NO_DEVICE: return "no HIP device";
  }
  return "unknown";
}

extern "C" sfft_status sfft_params_validate(const sfft_params* p) {
  if (!p) return SFFT_ERR_INVALID;
  const int ln = ilog2_exact(p->n), lB = ilog2_exact(p->B), lW = ilog2_exact(p->W_RLFC);
  if (ln < 5 || ln > 30 || lB < 1 || lW < 5) return SFFT_ERR_INVALID;
  if (p->B > p->n || p->W_RLFC > p->n) return SFFT_ERR_INVALID;
  if (!(p->B_thresh >= 1 && p->B_thresh < p->B)) return SFFT_ERR_INVALID;       // experiment.c:83
  if (!(p->B_thresh < p->W_RLFC)) return SFFT_ERR_INVALID;                       // experiment.c:82
  if (!(p->loops_thresh >= 1 && p->loops_thresh <= p->loops_loc)) return SFFT_ERR_INVALID;
  if (p->loops_loc < 1 || p->loops_loc > 32 || p->loops < p->loops_loc || p->loops > 32)
    return SFFT_ERR_INVALID;
  if (p->RLFC_loops < 1 || p->RLFC_loops > 64) return SFFT_ERR_INVALID;
  if (p->w < 1 || p->w > p->n) return SFFT_ERR_INVALID;                          // cuExperiment.cu:283
  if (p->fold_mode == SFFT_FOLD_REFERENCE && p->n < (1 << 27) && p->w < p->B)
    return SFFT_ERR_INVALID;   // the tiled fold's tail taps would index past w
  if (p->fold_mode != SFFT_FOLD_REFERENCE && p->fold_mode != SFFT_FOLD_ALL_TAPS)
    return SFFT_ERR_INVALID;
  if (p->hit_capacity < 0) return SFFT_ERR_INVALID;
  if (p->front_end < SFFT_FRONT_AUTO || p->front_end > SFFT_FRONT_COLUMNS) return SFFT_ERR_INVALID;
  if (p->W_RLFC < p->n / p->B) return SFFT_ERR_INVALID;   // the comb block must span a window
  return SFFT_OK;
}

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

extern "C" sfft_status sfft_plan_create(const sfft_params* prm, sfft_precision prec,
                                        void* stream, sfft_plan** out) {
  if (!out) return SFFT_ERR_INVALID;
  *out = nullptr;
  sfft_status v = sfft_params_validate(prm);
  if (v != SFFT_OK) return v;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SFFT_ERR_NO_DEVICE;
  sfft_plan* P = new sfft_plan();
  P->p = *prm;
  P->fp32 = (prec == SFFT_FP32);
  P->st = (hipStream_t)stream;
  P->log2n = ilog2_exact(prm->n);
  P->log2B = ilog2_exact(prm->B);
  P->log2W = ilog2_exact(prm->W_RLFC);
  P->m = prm->n / prm->B;
  P->cap = prm->hit_capacity > 0 ? prm->hit_capacity : MAXSIZE_HITS;
  const int L = prm->loops, Ll = prm->loops_loc, Rl = prm->RLFC_loops, K = prm->B_thresh;
  const size_t W = prm->W_RLFC, B = prm->B;
  const size_t e = P->fp32 ? 8 : 16, ke = P->fp32 ? 4 : 8;
  P->bm_words = (int)((W + 31) / 32);
  const int nhb = hash_blocks(Ll, K, (long long)prm->n / prm->B);
  // carve one workspace
  struct Item {
    void** dst;
    size_t bytes;
  };
  std::vector<Item> items;
  items.push_back({(void**)&P->d_a, (size_t)(2 * L + Rl) * 4});   // a | ai | offsets
  P->stream = prm->front_end == SFFT_FRONT_STREAM;
  // AUTO: the column front end wherever its LDS-staged slab fold applies (it streams x once,
  // 1 GiB / 2 GiB at n = 2^27 for FP32 / FP64, and hands the comb samples to the comb FFT, against
  // one random 128-byte line per tap and per comb sample for the gathers; measured at
  // n = 2^27, k = 16000: FP64 0.90 ms against 0.99 ms per transform)
  P->columns = prm->front_end == SFFT_FRONT_COLUMNS ||
               (prm->front_end == SFFT_FRONT_AUTO &&
                fold_slab_applies(P->fp32, prm->n, prm->B, prm->w, L));
  P->bins_perm = P->columns && fold_slab_applies(P->fp32, prm->n, prm->B, prm->w, L);
  if (P->stream) {
    items.push_back({&P->taps, (size_t)L * prm->w * e});
    items.push_back({&P->comb_y, (size_t)Rl * W * e});
  }
  if (P->columns) {
    items.push_back({&P->frows, B * (size_t)fold_rows_pitch((int)B, prm->w) * e});
    items.push_back({&P->comb_y, (size_t)Rl * W * e});   // comb columns copied by the fold
  }
  items.push_back({&P->comb0, (size_t)Rl * W * e});
  items.push_back({&P->comb1, (size_t)Rl * W * e});
  items.push_back({&P->comb_keys, (size_t)Rl * W * ke});
  items.push_back({(void**)&P->comb_digits, (size_t)Rl * W * 2});
  items.push_back({&P->bins_t, (size_t)L * B * e});
  items.push_back({&P->bins_scr, (size_t)L * B * e});
  items.push_back({&P->bins_f, (size_t)L * B * e});
  items.push_back({&P->bkeys, (size_t)Ll * B * ke});
  items.push_back({(void**)&P->J, (size_t)Ll * K * 4});
  items.push_back({(void**)&P->J2, (size_t)Rl * K * 4});
  void* selW_base = nullptr;
  void* selB_base = nullptr;
  items.push_back({&selW_base, select_work_bytes(Rl, (int)W)});
  items.push_back({&selB_base, select_work_bytes(Ll, (int)B)});
  items.push_back({(void**)&P->bitmap, (size_t)Ll * P->bm_words * 4});
  items.push_back({(void**)&P->maxfirst, (size_t)Ll * 4});
  items.push_back({(void**)&P->jsel, (size_t)Ll * ((B + 31) / 32) * 4});
  items.push_back({(void**)&P->hitoff, (size_t)Ll * K * 4});
  items.push_back({(void**)&P->blocksum, (size_t)nhb * 4});
  items.push_back({(void**)&P->hmask, hash_mask_bytes(Ll, K, (long long)prm->n / prm->B)});
  items.push_back({(void**)&P->hits, (size_t)P->cap * 4});
  items.push_back({(void**)&P->count, 16});
  items.push_back({(void**)&P->win, (size_t)P->m * 16});
  items.push_back({(void**)&P->nodes, (size_t)P->cap * sizeof(Node)});
  items.push_back({(void**)&P->sel, (size_t)P->cap * 2});
  items.push_back({(void**)&P->adj_head, (size_t)L * B * 4});
  items.push_back({(void**)&P->adj_next, (size_t)P->cap * 2 * 4});
  size_t total = 0;
  for (auto& it : items) total = align_up(total, 256) + align_up(it.bytes, 256);
  if (hipMalloc(&P->ws, total) != hipSuccess) {
    delete P;
    return SFFT_ERR_ALLOC;
  }
  P->ws_bytes = total;
  size_t off = 0;
  for (auto& it : items) {
    off = align_up(off, 256);
    *it.dst = (char*)P->ws + off;
    off += align_up(it.bytes, 256);
  }
  P->d_ai = P->d_a + L;
  P->d_offs = P->d_a + 2 * L;
  select_work_bind(&P->selW, selW_base, Rl, (int)W);
  select_work_bind(&P->selB, selB_base, Ll, (int)B);
  // select state must start at zero (hist/done are reset by the kernels themselves)
  if (hipMemset(P->ws, 0, total) != hipSuccess ||
      fft_tables_create(P->log2B, 0, P->fp32, &P->tabB) != hipSuccess ||
      fft_tables_create(P->log2W, 0, P->fp32, &P->tabW) != hipSuccess ||
      hipHostMalloc((void**)&P->h_draws, (size_t)(2 * L + Rl) * 4, hipHostMallocDefault) !=
          hipSuccess ||
      hipHostMalloc((void**)&P->h_count, 16, hipHostMallocDefault) != hipSuccess ||
      hipHostGetDevicePointer((void**)&P->h_count_dev, P->h_count, 0) != hipSuccess ||
      hipEventCreateWithFlags(&P->up_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&P->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipStreamCreateWithFlags(&P->st2, hipStreamNonBlocking) != hipSuccess) {
    sfft_plan_destroy(P);
    return SFFT_ERR_ALLOC;
  }
  *out = P;
  return SFFT_OK;
}

static void drop_plan_graph(sfft_plan* P);

extern "C" void sfft_plan_destroy(sfft_plan* P) {
  if (!P) return;
  if (P->st) (void)hipStreamSynchronize(P->st);
  else (void)hipDeviceSynchronize();
  if (P->st2) (void)hipStreamSynchronize(P->st2);
  drop_plan_graph(P);
  if (P->own_st) (void)hipStreamDestroy(P->st);
  if (P->st2) (void)hipStreamDestroy(P->st2);
  if (P->ev_fork) (void)hipEventDestroy(P->ev_fork);
  if (P->ev_join) (void)hipEventDestroy(P->ev_join);
  for (auto& e : P->ev)
    if (e) (void)hipEventDestroy(e);
  if (P->up_ev) (void)hipEventDestroy(P->up_ev);
  if (P->h_draws) (void)hipHostFree(P->h_draws);
  if (P->h_count) (void)hipHostFree(P->h_count);
  fft_tables_destroy(&P->tabB);
  fft_tables_destroy(&P->tabW);
  if (P->ws) (void)hipFree(P->ws);
  delete P;
}

extern "C" size_t sfft_plan_workspace_bytes(const sfft_plan* P) { return P ? P->ws_bytes : 0; }

extern "C" sfft_status sfft_plan_set_filter(sfft_plan* P, const void* d_filter_time,
                                            const complex_t* h_filter_win) {
  if (!P || !d_filter_time || !h_filter_win) return SFFT_ERR_INVALID;
  // A transform still running on the plan's streams reads the old taps and window: let it
  // finish before the window is rewritten (and before the caller may free the old taps).  A
  // captured graph holds the old taps pointer as a kernel argument: drop it, so the next
  // graph-mode execute captures the new one.
  if (P->st) HIPCK(hipStreamSynchronize(P->st));
  else HIPCK(hipDeviceSynchronize());
  if (P->st2) HIPCK(hipStreamSynchronize(P->st2));
  drop_plan_graph(P);
  if (hipMemcpy(P->win, h_filter_win, (size_t)P->m * 16, hipMemcpyHostToDevice) != hipSuccess)
    return SFFT_ERR_HIP;
  if (P->columns) {   // the column fold's bucket-major copy of the taps
    const int tiled = (P->p.fold_mode == SFFT_FOLD_REFERENCE) && (P->p.n < (1 << 27));
    HIPCK(filter_rows_launch(P->fp32, d_filter_time, P->frows, P->p.B, P->p.w, tiled, P->st));
    HIPCK(hipStreamSynchronize(P->st));
  }
  P->d_ft = d_filter_time;
  P->has_filter = 1;
  return SFFT_OK;
}

sfft_status draws_check(const sfft_plan* P, const sfft_draws* d) {
  if (!d || !d->a || !d->ai || !d->offsets) return SFFT_ERR_INVALID;
  const long long n = P->p.n;
  const int sigma = P->p.n / P->p.W_RLFC;
  for (int l = 0; l < P->p.loops; ++l) {
    if (d->a[l] <= 0 || d->a[l] >= n || (d->a[l] & 1) == 0) return SFFT_ERR_INVALID;
    if (d->ai[l] <= 0 || d->ai[l] >= n) return SFFT_ERR_INVALID;
    if ((((long long)d->a[l] * d->ai[l]) & (n - 1)) != 1) return SFFT_ERR_INVALID;
  }
  for (int r = 0; r < P->p.RLFC_loops; ++r)
    if (d->offsets[r] < 0 || d->offsets[r] >= sigma) return SFFT_ERR_INVALID;
  return SFFT_OK;
}

HashArgs hash_args(const sfft_plan* P, int* d_count) {
  const sfft_params& p = P->p;
  HashArgs h;
  h.n = p.n;
  h.B = p.B;
  h.K = p.B_thresh;
  h.W = p.W_RLFC;
  h.loops_loc = p.loops_loc;
  h.loops_thresh = p.loops_thresh;
  h.J = P->J;
  h.J2 = P->J2;
  h.nJ2 = p.RLFC_loops * p.B_thresh;
  h.a_dev = P->d_a;
  h.ai_dev = P->d_ai;
  h.bitmap = P->bitmap;
  h.bm_words = P->bm_words;
  h.maxfirst = P->maxfirst;
  h.jsel = P->jsel;
  h.jsel_words = (p.B + 31) / 32;
  h.hitoff = P->hitoff;
  h.blocksum = P->blocksum;
  h.hmask = P->hmask;
  h.hits = P->hits;
  h.capacity = P->cap;
  h.count = d_count ? d_count : P->count;
  return h;
}

// The transform in three stage groups, each enqueued on an explicit stream (the plan's own
// for a single transform; a batch spreads its signals over several streams and runs the
// permute+filter of all signals as one launch).  The stage-boundary events of the profiling
// mode are recorded by the front and back groups when `marks` is set.
// front: bitmap clear, streaming pass (option), comb FFT (run_comb), then the comb
// cutoff and the approved bitmaps (run_cut)
static FftLaunch comb_fft(const sfft_plan* P, const void* d_x);
static sfft_status run_comb(sfft_plan* P, const void* d_x, const HashArgs& h, hipStream_t st,
                            bool marks) {
  const sfft_params& p = P->p;
  const int L = p.loops, Rl = p.RLFC_loops;
  auto mark = [&](int i) -> hipError_t {
    return (marks && P->profile) ? hipEventRecord(P->ev[i], st) : hipSuccess;
  };
  HIPCK(mark(0));
  HIPCK(hash_clear_launch(h, st));
  HIPCK(mark(1));
  // streaming front end: one pass over x fills the tap staging array and the comb samples
  if (P->stream)
    HIPCK(stream_scatter_launch(P->fp32, d_x, P->taps, P->comb_y, P->d_a, P->d_offs, p.n, p.w,
                                L, Rl, p.W_RLFC, st));
  HIPCK(mark(2));
  HIPCK(fft_run(comb_fft(P, d_x), st));
  return SFFT_OK;
}

// RLFC: comb samples (gathered in pass 0, or from the streaming pass) + W-point FFT, keys of
// |Y|^2 out of the last pass in its tile order (whole-line stores; the comb cutoff maps the
// storage positions back to the bins)
static FftLaunch comb_fft(const sfft_plan* P, const void* d_x) {
  const sfft_params& p = P->p;
  const int Rl = p.RLFC_loops;
  const long long W = p.W_RLFC;
  FftLaunch c;
  c.tab = &P->tabW;
  c.out = P->comb0;
  c.out_bstride = W;
  c.scratch = P->comb1;
  c.keys = P->comb_keys;
  c.keys_bstride = W;
  c.key_batches = Rl;
  c.batch = Rl;
  c.store = kStoreKeys;
  c.qmax = P->qmax_W;
  c.keys_tiled = 1;
  c.digits = P->comb_digits;
  if (P->stream) {
    c.in = P->comb_y;
    c.in_bstride = W;
    c.load = kLoadPlain;
  } else if (P->columns) {
    // the fold copied the comb samples column-major: sample j = row * (B/sigma) + lo at
    // comb_y[lo * m + row]
    c.in = P->comb_y;
    c.in_bstride = W;
    c.load = kLoadCombRows;
    c.comb_log2nlo = P->log2W - (P->log2n - P->log2B);
    c.comb_log2m = P->log2n - P->log2B;
  } else {
    c.load = kLoadComb;
    c.comb_x = d_x;
    c.comb_offs = P->d_offs;
    c.comb_sigma = (int)(p.n / p.W_RLFC);
  }
  return c;
}

// comb cutoff: the comb lists only feed the order-free bitmap build (unordered output)
static sfft_status run_cut(sfft_plan* P, const HashArgs& h, hipStream_t st, bool marks) {
  const sfft_params& p = P->p;
  const long long W = p.W_RLFC;
  const KeyMap km = fft_keys_map(comb_fft(P, nullptr));
  HIPCK(select_topk(P->fp32, P->comb_keys, W, (int)W, p.B_thresh, p.RLFC_loops, P->J2, 0,
                    P->selW, st, nullptr, &km, P->comb_digits));
  HIPCK(hash_bitmap_launch(h, st));
  if (marks && P->profile) HIPCK(hipEventRecord(P->ev[3], st));
  return SFFT_OK;
}

sfft_status run_front(sfft_plan* P, const void* d_x, const HashArgs& h, hipStream_t st,
                      bool marks) {
  sfft_status v = run_comb(P, d_x, h, st, marks);
  return v != SFFT_OK ? v : run_cut(P, h, st, marks);
}

// fold: permutation + filtering of every loop into bins_t
sfft_status run_fold(sfft_plan* P, const void* d_x, hipStream_t st, const FoldClear* clr) {
  const sfft_params& p = P->p;
  if (clr && !P->columns) return SFFT_ERR_INVALID;   // only the slab fold clears
  const int tiled = (p.fold_mode == SFFT_FOLD_REFERENCE) && (p.n < (1 << 27));
  if (P->stream)
    HIPCK(fold_taps_launch(P->fp32, P->taps, P->d_ft, P->bins_t, P->d_ai, p.n, p.B, p.w, p.loops,
                           tiled, st));
  else if (P->columns)
    HIPCK(fold_columns_launch(P->fp32, d_x, P->frows, P->bins_t, P->d_a, P->d_ai, p.n, p.B, p.w,
                              p.loops, tiled, P->bins_perm, 1, P->comb_y, P->d_offs,
                              p.RLFC_loops, p.W_RLFC, clr ? *clr : FoldClear(), st));
  else
    HIPCK(perm_filter_launch(P->fp32, d_x, P->d_ft, P->bins_t, P->d_ai, p.n, p.B, p.w, p.loops,
                             tiled, st));
  return SFFT_OK;
}

// bucket: bucket FFT + cutoff (after the fold)
sfft_status run_bucket(sfft_plan* P, hipStream_t st, bool marks) {
  const sfft_params& p = P->p;
  const int L = p.loops, Ll = p.loops_loc, K = p.B_thresh;
  const long long B = p.B;
  if (marks && P->profile) HIPCK(hipEventRecord(P->ev[4], st));
  FftLaunch f;
  f.tab = &P->tabB;
  f.in = P->bins_t;
  f.in_bstride = B;
  f.out = P->bins_f;
  f.out_bstride = B;
  f.scratch = P->bins_scr;
  f.keys = P->bkeys;
  f.keys_bstride = B;
  f.key_batches = Ll;
  f.batch = L;
  f.store = kStoreBoth;
  f.qmax = P->qmax_B;
  if (P->bins_perm) {   // bucket r of loop l at column r * ai_l mod B
    f.load = kLoadPerm;
    f.perm = P->d_ai;
    f.perm_inv = P->d_a;
  }
  HIPCK(fft_run(f, st));
  // ascending bucket lists: the hit list is ordered by (loop, bucket, walk)
  HIPCK(select_topk(P->fp32, P->bkeys, B, (int)B, K, Ll, P->J, 1, P->selB, st, P->jsel));
  if (marks && P->profile) HIPCK(hipEventRecord(P->ev[5], st));
  return SFFT_OK;
}

// locate: reverse hash (needs the comb bitmaps and the bucket cutoffs) + estimation
sfft_status run_locate(sfft_plan* P, const HashArgs& h, Node* d_nodes, hipStream_t st,
                       bool marks) {
  const sfft_params& p = P->p;
  HIPCK(hash_hits_launch(h, st));
  if (marks && P->profile) HIPCK(hipEventRecord(P->ev[6], st));
  HIPCK(estimate_launch(P->fp32, P->hits, h.count, P->cap, P->bins_f, p.loops, p.n, P->d_ai,
                        p.B, P->win, d_nodes ? d_nodes : P->nodes, P->sel, st, P->cnt_copy));
  if (marks && P->profile) HIPCK(hipEventRecord(P->ev[7], st));
  return SFFT_OK;
}

sfft_status run_back(sfft_plan* P, const HashArgs& h, Node* d_nodes, hipStream_t st,
                     bool marks) {
  sfft_status v = run_bucket(P, st, marks);
  return v != SFFT_OK ? v : run_locate(P, h, d_nodes, st, marks);
}

// Column front end: the fold also copies the comb samples, so it runs first (and does the
// transform's clear of the approved-residue bitmaps on its way: no clear launch); then the
// comb chain (comb FFT from the copied samples, cutoff, bitmaps) on the plan stream and the
// bucket chain (bucket FFT, cutoff) on the second stream, joined before the reverse hash.  The
// comb FFT is enqueued before the bucket chain: in a captured graph it is then the fold's
// first dependent, which the graph replay keeps on the fold's queue, so the longer chain
// crosses no queue.  Profiling runs everything serially; its "clear" interval is then empty,
// its "streaming pass" interval is the fold (with the comb copy) and its "fold" interval is
// empty.
static sfft_status run_all_columns(sfft_plan* P, const void* d_x, const HashArgs& h,
                                   Node* d_nodes) {
  sfft_status v;
  const bool prof = P->profile != 0;
  auto mark = [&](int i) -> hipError_t {
    return prof ? hipEventRecord(P->ev[i], P->st) : hipSuccess;
  };
  FoldClear clr;
  clr.bm = h.bitmap;
  clr.nbm = (size_t)h.loops_loc * (size_t)h.bm_words;
  clr.maxf = (uint32_t*)h.maxfirst;
  clr.nmax = h.loops_loc;
  HIPCK(mark(0));
  HIPCK(mark(1));
  if ((v = run_fold(P, d_x, P->st, &clr)) != SFFT_OK) return v;
  HIPCK(mark(2));
  if (prof) {
    HIPCK(fft_run(comb_fft(P, d_x), P->st));
    if ((v = run_cut(P, h, P->st, true)) != SFFT_OK) return v;
    if ((v = run_bucket(P, P->st, true)) != SFFT_OK) return v;
    return run_locate(P, h, d_nodes, P->st, true);
  }
  HIPCK(hipEventRecord(P->ev_fork, P->st));
  HIPCK(fft_run(comb_fft(P, d_x), P->st));
  HIPCK(hipStreamWaitEvent(P->st2, P->ev_fork, 0));
  if ((v = run_bucket(P, P->st2, false)) != SFFT_OK) return v;
  HIPCK(hipEventRecord(P->ev_join, P->st2));
  if ((v = run_cut(P, h, P->st, false)) != SFFT_OK) return v;
  HIPCK(hipStreamWaitEvent(P->st, P->ev_join, 0));
  return run_locate(P, h, d_nodes, P->st, false);
}

// One transform as two concurrent chains joined before the reverse hash: the comb chain
// (clear, comb FFT, cutoff, bitmaps) on the plan stream and the bucket chain (permute+filter,
// bucket FFT, cutoff) on a second stream, so the latency-bound kernels of one chain fill the
// gaps of the other's gathers.  Profiling (and the streaming front end, whose pass feeds the
// fold) runs everything serially on the plan stream.
static sfft_status run_all(sfft_plan* P, const void* d_x, const HashArgs& h, Node* d_nodes) {
  sfft_status v;
  if (P->columns) return run_all_columns(P, d_x, h, d_nodes);
  if (P->profile || P->stream) {   // (the streaming pass feeds the fold: one chain)
    if ((v = run_front(P, d_x, h, P->st, true)) != SFFT_OK) return v;
    if ((v = run_fold(P, d_x, P->st)) != SFFT_OK) return v;
    return run_back(P, h, d_nodes, P->st, true);
  }
  if ((v = run_comb(P, d_x, h, P->st, false)) != SFFT_OK) return v;
  HIPCK(hipEventRecord(P->ev_fork, P->st));
  HIPCK(hipStreamWaitEvent(P->st2, P->ev_fork, 0));
  if ((v = run_fold(P, d_x, P->st2)) != SFFT_OK) return v;
  if ((v = run_bucket(P, P->st2, false)) != SFFT_OK) return v;
  HIPCK(hipEventRecord(P->ev_join, P->st2));
  if ((v = run_cut(P, h, P->st, false)) != SFFT_OK) return v;
  HIPCK(hipStreamWaitEvent(P->st, P->ev_join, 0));
  return run_locate(P, h, d_nodes, P->st, false);
}

static void drop_plan_graph(sfft_plan* P) {
  if (P->gexec) (void)hipGraphExecDestroy(P->gexec);
  if (P->graph) (void)hipGraphDestroy(P->graph);
  P->gexec = nullptr;
  P->graph = nullptr;
}

// One transform enqueued on the plan stream: nodes to d_nodes (device memory, or the device
// view of pinned host memory; nullptr: the plan's buffer), the count to d_count (nullptr: the
// plan's) and, when cnt_copy is set, also to *cnt_copy.
static sfft_status enqueue(sfft_plan* P, const void* d_x, const sfft_draws* dr, Node* d_nodes,
                           int* d_count, int* cnt_copy) {
  if (!P || !d_x) return SFFT_ERR_INVALID;
  if (!P->has_filter) return SFFT_ERR_NO_FILTER;
  sfft_status v = upload_draws(P, dr);
  if (v != SFFT_OK) return v;
  const HashArgs h = hash_args(P, d_count);
  P->cnt_copy = cnt_copy;   // read by run_locate while this transform is enqueued / captured
  struct Reset {
    sfft_plan* P;
    ~Reset() { P->cnt_copy = nullptr; }
  } reset{P};
  if (!P->use_graph || P->profile) return run_all(P, d_x, h, d_nodes);   // eager
  if (P->gexec && (P->g_x != d_x || P->g_nodes != d_nodes || P->g_count != d_count ||
                   P->g_cnt_copy != cnt_copy))
    drop_plan_graph(P);
  if (!P->gexec) {
    HIPCK(hipStreamBeginCapture(P->st, hipStreamCaptureModeThreadLocal));
    v = run_all(P, d_x, h, d_nodes);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(P->st, &g);
    if (v != SFFT_OK || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return v != SFFT_OK ? v : SFFT_ERR_HIP;
    }
    P->graph = g;
    if (hipGraphInstantiate(&P->gexec, g, nullptr, nullptr, 0) != hipSuccess) {
      drop_plan_graph(P);
      return SFFT_ERR_HIP;
    }
    P->g_x = d_x;
    P->g_nodes = d_nodes;
    P->g_count = d_count;
    P->g_cnt_copy = cnt_copy;
  }
  HIPCK(hipGraphLaunch(P->gexec, P->st));
  return SFFT_OK;
}

extern "C" sfft_status sfft_execute_async(sfft_plan* P, const void* d_x, const sfft_draws* dr,
                                          Node* d_nodes, int* d_count) {
  return enqueue(P, d_x, dr, d_nodes, d_count, nullptr);
}

// The device view of h when h is pinned host memory with room for `bytes` from h on, else
// nullptr (pageable memory, or too short an allocation).  Asked on every call: an address
// freed and reused for pageable memory must never be written through a stale view.
static void* pinned_view(void* h, size_t bytes) {
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, h) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  if (at.type != hipMemoryTypeHost || !at.devicePointer) return nullptr;
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)at.devicePointer) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  const size_t off = (size_t)((char*)at.devicePointer - (char*)base);
  return (off <= size && size - off >= bytes) ? at.devicePointer : nullptr;
}

// Graph mode: the transform's launch sequence is captured once and replayed (the draws are
// uploaded before each replay, outside the graph).  A plan created without a stream gets a
// stream of its own for this (the legacy default stream cannot be captured).
extern "C" sfft_status sfft_plan_set_graph(sfft_plan* P, int on) {
  if (!P) return SFFT_ERR_INVALID;
  drop_plan_graph(P);
  if (on && !P->st) {
    HIPCK(hipStreamSynchronize(nullptr));
    if (P->up_pending) HIPCK(hipEventSynchronize(P->up_ev));
    HIPCK(hipStreamCreateWithFlags(&P->st, hipStreamNonBlocking));
    P->own_st = 1;
  }
  P->use_graph = on ? 1 : 0;
  return SFFT_OK;
}

// upload one transform's draws (after the previous upload out of the staging has completed)
sfft_status upload_draws(sfft_plan* P, const sfft_draws* dr) {
  sfft_status v = draws_check(P, dr);
  if (v != SFFT_OK) return v;
  const int L = P->p.loops, Rl = P->p.RLFC_loops;
  if (P->up_pending) HIPCK(hipEventSynchronize(P->up_ev));
  memcpy(P->h_draws, dr->a, (size_t)L * 4);
  memcpy(P->h_draws + L, dr->ai, (size_t)L * 4);
  memcpy(P->h_draws + 2 * L, dr->offsets, (size_t)Rl * 4);
  HIPCK(hipMemcpyAsync(P->d_a, P->h_draws, (size_t)(2 * L + Rl) * 4, hipMemcpyHostToDevice,
                       P->st));
  HIPCK(hipEventRecord(P->up_ev, P->st));
  P->up_pending = 1;
  return SFFT_OK;
}

static sfft_status value_map_check(const sfft_plan* P, int nh, const int* d_hits,
                                   const unsigned char* d_sel) {
  if (P->fp32) return SFFT_ERR_INVALID;                 // FP64 plans only
  if (!P->has_filter) return SFFT_ERR_NO_FILTER;
  if (nh < 0 || nh > P->cap || (nh > 0 && (!d_hits || !d_sel))) return SFFT_ERR_INVALID;
  if (P->p.n / P->p.B > 8192) return SFFT_ERR_INVALID;  // bounds the per-bucket entry lists
  return SFFT_OK;
}

extern "C" sfft_status sfft_value_apply(sfft_plan* P, const void* d_x, const sfft_draws* dr,
                                        int nh, const int* d_hits, const unsigned char* d_sel,
                                        complex_t* d_values) {
  if (!P || !d_x || (nh > 0 && !d_values)) return SFFT_ERR_INVALID;
  sfft_status v = value_map_check(P, nh, d_hits, d_sel);
  if (v == SFFT_OK) v = upload_draws(P, dr);
  if (v != SFFT_OK) return v;
  const sfft_params& p = P->p;
  const int tiled = (p.fold_mode == SFFT_FOLD_REFERENCE) && (p.n < (1 << 27));
  HIPCK(perm_filter_launch(0, d_x, P->d_ft, P->bins_t, P->d_ai, p.n, p.B, p.w, p.loops, tiled,
                           P->st));
  FftLaunch f;
  f.tab = &P->tabB;
  f.in = P->bins_t;
  f.in_bstride = p.B;
  f.out = P->bins_f;
  f.out_bstride = p.B;
  f.scratch = P->bins_scr;
  f.batch = p.loops;
  f.store = kStoreCplx;
  HIPCK(fft_run(f, P->st));
  HIPCK(value_pick_launch(d_hits, nh, d_sel, (const double2*)P->bins_f, p.n, P->d_ai, p.B,
                          P->win, (double2*)d_values, P->st));
  return SFFT_OK;
}

extern "C" sfft_status sfft_value_adjoint(sfft_plan* P, const complex_t* d_grad,
                                          const sfft_draws* dr, int nh, const int* d_hits,
                                          const unsigned char* d_sel, complex_t* d_grad_x) {
  if (!P || !d_grad_x || (nh > 0 && !d_grad)) return SFFT_ERR_INVALID;
  sfft_status v = value_map_check(P, nh, d_hits, d_sel);
  if (v == SFFT_OK) v = upload_draws(P, dr);
  if (v != SFFT_OK) return v;
  const sfft_params& p = P->p;
  hipError_t err = hipSuccess;
  const FftTables* inv = cached_tables(P->log2B, 1, 0, &err);
  if (!inv) return SFFT_ERR_HIP;
  // 1. per-bucket seeds, 2. their unnormalised inverse FFT_B, 3. scatter through the taps
  HIPCK(adjoint_seed_launch(d_hits, nh, d_sel, (const double2*)d_grad, p.n, P->d_ai, p.B,
                            p.loops, P->win, P->adj_head, P->adj_next, (double2*)P->bins_t,
                            P->st));
  FftLaunch f;
  f.tab = inv;
  f.in = P->bins_t;
  f.in_bstride = p.B;
  f.out = P->bins_f;
  f.out_bstride = p.B;
  f.scratch = P->bins_scr;
  f.batch = p.loops;
  f.store = kStoreCplx;
  HIPCK(fft_run(f, P->st));
  const int tiled = (p.fold_mode == SFFT_FOLD_REFERENCE) && (p.n < (1 << 27));
  HIPCK(adjoint_taps_launch((const double2*)P->d_ft, (const double2*)P->bins_f, P->d_ai, p.n,
                            p.B, p.w, p.loops, tiled, (double2*)d_grad_x, P->st));
  return SFFT_OK;
}

extern "C" sfft_status sfft_plan_set_profiling(sfft_plan* P, int on) {
  if (!P) return SFFT_ERR_INVALID;
  if (on)
    for (auto& e : P->ev)
      if (!e && hipEventCreate(&e) != hipSuccess) return SFFT_ERR_HIP;
  P->profile = on ? 1 : 0;
  return SFFT_OK;
}

// Seconds between the plan's stage events of the last profiled transform: clear, streaming
// pass (0 with gathers), comb FFT + cutoff + bitmaps, permute+filter (or tap fold), bucket
// FFT + cutoff, reverse hash, estimation.
extern "C" sfft_status sfft_plan_interval_times(sfft_plan* P, double* out7) {
  if (!P || !out7 || !P->profile) return SFFT_ERR_INVALID;
  for (int i = 0; i < 7; ++i) {
    float t = 0;
    if (hipEventElapsedTime(&t, P->ev[i], P->ev[i + 1]) != hipSuccess) return SFFT_ERR_HIP;
    out7[i] = t * 1e-3;
  }
  return SFFT_OK;
}

// Seconds per stage in the reference's table order (outerLoop.cu:174-187): votetable,
// restr-loc-index, perm+filter, reverse-hash, estimation, FFT+Cutoff, total.
extern "C" sfft_status sfft_plan_stage_times(sfft_plan* P, double* out7) {
  double t[7];
  const sfft_status s = sfft_plan_interval_times(P, t);
  if (s != SFFT_OK) return s;
  out7[0] = t[0];            // votetable: bitmap and residue-maximum clear
  out7[1] = t[2];            // restr-loc-index: comb FFT + cutoff + bitmaps
  out7[2] = t[1] + t[3];     // perm+filter (streaming pass + fold, or gathers)
  out7[3] = t[5];            // reverse-hash
  out7[4] = t[6];            // estimation
  out7[5] = t[4];            // FFT+Cutoff: bucket FFT + location cutoffs
  float tot = 0;
  if (hipEventElapsedTime(&tot, P->ev[0], P->ev[7]) != hipSuccess) return SFFT_ERR_HIP;
  out7[6] = tot * 1e-3;
  return SFFT_OK;
}

extern "C" sfft_status sfft_execute(sfft_plan* P, const void* d_x, const sfft_draws* dr,
                                    Node* h_nodes, int* h_count) {
  if (!P) return SFFT_ERR_INVALID;
  // pinned h_nodes: the estimation kernel stores the nodes and the count straight into host
  // memory (no copies after the transform)
  Node* zc = h_nodes ? (Node*)pinned_view(h_nodes, (size_t)P->cap * sizeof(Node)) : nullptr;
  if (zc) {
    sfft_status s = enqueue(P, d_x, dr, zc, P->count, P->h_count_dev);
    if (s != SFFT_OK) return s;
    HIPCK(hipStreamSynchronize(P->st));
    const int cnt = *(volatile int*)P->h_count;
    P->last_count = cnt;
    if (h_count) *h_count = cnt;
    return cnt > P->cap ? SFFT_ERR_HIT_OVERFLOW : SFFT_OK;
  }
  sfft_status s = enqueue(P, d_x, dr, P->nodes, P->count, nullptr);
  if (s != SFFT_OK) return s;
  // the count and, speculatively, as many nodes as the previous transform had plus 1/32 go
  // back in one round trip; the rest (if the count grew past that) in a second one
  HIPCK(hipMemcpyAsync(P->h_count, P->count, 4, hipMemcpyDeviceToHost, P->st));
  long long guess = (long long)P->last_count + P->last_count / 32;
  if (guess > P->cap) guess = P->cap;
  if (h_nodes && guess > 0)
    HIPCK(hipMemcpyAsync(h_nodes, P->nodes, (size_t)guess * sizeof(Node), hipMemcpyDeviceToHost,
                         P->st));
  HIPCK(hipStreamSynchronize(P->st));
  const int cnt = *P->h_count;
  P->last_count = cnt;
  if (h_count) *h_count = cnt;
  const long long ncopy = cnt < P->cap ? cnt : P->cap;
  if (h_nodes && ncopy > guess) {
    HIPCK(hipMemcpyAsync(h_nodes + guess, P->nodes + guess, (size_t)(ncopy - guess) * sizeof(Node),
                         hipMemcpyDeviceToHost, P->st));
    HIPCK(hipStreamSynchronize(P->st));
  }
  return cnt > P->cap ? SFFT_ERR_HIT_OVERFLOW : SFFT_OK;
}

extern "C" sfft_status sfft_plan_comb_key_order(const sfft_plan* P, int* lo, int* mid) {
  if (!P || !lo || !mid) return SFFT_ERR_INVALID;
  const KeyMap km = fft_keys_map(comb_fft(P, nullptr));
  *lo = km.lo;
  *mid = km.mid;
  return SFFT_OK;
}

extern "C" int sfft_plan_front_end(const sfft_plan* P) {
  if (!P) return -1;
  return P->stream ? SFFT_FRONT_STREAM : P->columns ? SFFT_FRONT_COLUMNS : SFFT_FRONT_GATHER;
}

extern "C" const void* sfft_plan_debug_buffer(const sfft_plan* P, int which) {
  if (!P) return nullptr;
  switch (which) {
    case 0: return P->bins_f;
    case 1: return P->J;
    case 2: return P->J2;
    case 3: return P->hits;
    case 4: return P->bins_t;
    case 5: return P->nodes;
    case 6: return P->count;
    case 7: return P->sel;
    case 8: return P->comb_keys;
  }
  return nullptr;
}

// ============================================================================================
// standalone canonical FFT
// ============================================================================================
extern "C" sfft_status sfft_fft_device(const void* d_in, void* d_out, int log2n, int count,
                                       int inverse, sfft_precision prec, void* stream) {
  if (log2n < 0 || log2n > 30 || count < 1 || !d_in || !d_out) return SFFT_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  const size_t N = (size_t)1 << log2n;
  const size_t esz = prec == SFFT_FP32 ? 8 : 16;
  if (log2n == 0) {
    if (d_in != d_out)
      HIPCK(hipMemcpyAsync(d_out, d_in, (size_t)count * esz, hipMemcpyDeviceToDevice, st));
    return SFFT_OK;
  }
  hipError_t err = hipSuccess;
  const FftTables* T = cached_tables(log2n, inverse ? 1 : 0, prec == SFFT_FP32, &err);
  if (!T) return SFFT_ERR_ALLOC;
  void* scratch = nullptr;
  const void* in = d_in;
  void* tmp_in = nullptr;
  if (T->npass > 1) {
    if (hipMallocAsync(&scratch, N * count * esz, st) != hipSuccess) return SFFT_ERR_ALLOC;
    if (d_in == d_out) {
      // multi-pass needs in != out: stage the input
      if (hipMallocAsync(&tmp_in, N * count * esz, st) != hipSuccess) return SFFT_ERR_ALLOC;
      HIPCK(hipMemcpyAsync(tmp_in, d_in, N * count * esz, hipMemcpyDeviceToDevice, st));
      in = tmp_in;
    }
  }
  FftLaunch L;
  L.tab = T;
  L.in = in;
  L.in_bstride = (long long)N;
  L.out = d_out;
  L.out_bstride = (long long)N;
  L.scratch = scratch;
  L.batch = count;
  L.load = kLoadPlain;
  L.store = kStoreCplx;
  err = fft_run(L, st);
  if (scratch) (void)hipFreeAsync(scratch, st);
  if (tmp_in) (void)hipFreeAsync(tmp_in, st);
  return err == hipSuccess ? SFFT_OK : SFFT_ERR_HIP;
}

extern "C" void sfft_twiddle(uint64_t e, uint64_t T, int inverse, double* re, double* im) {
  canon_twiddle(e, T, inverse, re, im);
}

extern "C" int sfft_fft_plan_passes(int log2n, int* pass_logs) {
  return fft_plan_passes(log2n, pass_logs);
}

extern "C" const char* sfft_version(void) { return "sfft-mi355x 0.1.0 (gfx950)"; }

extern "C" int sfft_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

namespace sfft {
hipError_t fft_plain(const void* in, void* out, void* scratch, int log2n, int count, int inverse,
                     int is_fp32, hipStream_t st) {
  hipError_t err = hipSuccess;
  const FftTables* T = cached_tables(log2n, inverse ? 1 : 0, is_fp32, &err);
  if (!T) return err == hipSuccess ? hipErrorOutOfMemory : err;
  const long long N = 1LL << log2n;
  FftLaunch L;
  L.tab = T;
  L.in = in;
  L.in_bstride = N;
  L.out = out;
  L.out_bstride = N;
  L.scratch = scratch;
  L.batch = count;
  return fft_run(L, st);
}
}  // namespace sfft
