// sfft_batch.hip -- many signals of uneven n and k transformed as one batch.
//
// Each signal keeps its own plan (parameters, workspace, draws).  One batch execution is:
//   * the permute+filter of ALL signals as one persistent launch (perm_filter_batch_kernel:
//     a work table of (signal, loop group, bucket block) items walked with a grid stride, so
//     small and large signals share the chip evenly);
//   * every signal's other stage groups (clear + comb restriction, then bucket FFT + cutoff,
//     reverse hash, estimation) on one of several worker streams, joined back to the batch
//     stream;
// either enqueued eagerly or replayed from hipGraphs: each worker stream's front groups and
// its back groups are captured once as two linear chains (on that stream), and a replay
// launches them around the directly launched permute+filter (one graph over all streams
// replayed slower than the eager launches: its cross-stream dependencies serialise the
// ~1,600 nodes).  The draws are uploaded before the replay, outside the graphs; the graphs
// hold no memset nodes (buffers are cleared by kernels).  The graphs are re-captured when
// the signal pointers or the tuned shapes change.
//
// Autotuner: times, on the batch itself, only launch shapes that compute the same bits --
// the permute+filter loops per thread, block size and persistent grid (every shape runs the
// same per-bucket body in the same tap order), and the FFT columns per tile of the comb and
// bucket FFTs (columns are independent transforms) -- and keeps the fastest.
#include <algorithm>
#include <cstring>
#include <vector>

#include "sfft_plan_impl.h"

namespace {
constexpr int kWorkers = 4;
}

struct sfft_batch {
  int count = 0;
  int fp32 = 0;
  hipStream_t st = nullptr;
  int own_st = 0;                    // st created here (a graph cannot capture stream 0)
  std::vector<sfft_plan*> plans;
  std::vector<hipStream_t> workers;
  hipEvent_t ev_fork = nullptr, ev_pf = nullptr;
  std::vector<hipEvent_t> ev_join;
  // batched permute+filter
  std::vector<PfJob> jobs;
  PfJob* d_jobs = nullptr;
  int total_items = 0;
  int lpt = 3, nt = 256, grid = 2048;
  std::vector<const void*> jobs_x;   // signal pointers the device job table holds
  // per-signal hit counts, contiguous
  int* d_counts = nullptr;
  int* h_counts = nullptr;
  // graphs: per worker stream, one of its signals' front groups and one of their back groups
  // (linear chains, each captured on its own worker stream); the persistent permute+filter
  // between them is launched directly
  int use_graph = 0;
  std::vector<hipGraph_t> graphs;
  std::vector<hipGraphExec_t> execs;   // [2*s] front, [2*s+1] back of worker s (may be null)
  bool captured = false;
  std::vector<const void*> graph_x;
};

static void drop_graph(sfft_batch* b) {
  for (auto e : b->execs)
    if (e) (void)hipGraphExecDestroy(e);
  for (auto g : b->graphs)
    if (g) (void)hipGraphDestroy(g);
  b->execs.clear();
  b->graphs.clear();
  b->captured = false;
  b->graph_x.clear();
}

extern "C" void sfft_batch_destroy(sfft_batch* b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->st);
  for (auto w : b->workers) (void)hipStreamSynchronize(w);
  drop_graph(b);
  for (auto p : b->plans) sfft_plan_destroy(p);
  for (auto w : b->workers) (void)hipStreamDestroy(w);
  for (auto e : b->ev_join) (void)hipEventDestroy(e);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_pf) (void)hipEventDestroy(b->ev_pf);
  if (b->own_st && b->st) (void)hipStreamDestroy(b->st);
  if (b->d_jobs) (void)hipFree(b->d_jobs);
  if (b->d_counts) (void)hipFree(b->d_counts);
  if (b->h_counts) (void)hipHostFree(b->h_counts);
  delete b;
}

extern "C" sfft_status sfft_batch_create(const sfft_params* params, int count,
                                         sfft_precision prec, void* stream, sfft_batch** out) {
  if (!params || count <= 0 || !out) return SFFT_ERR_INVALID;
  *out = nullptr;
  for (int i = 0; i < count; ++i) {
    const sfft_status s = sfft_params_validate(&params[i]);
    if (s != SFFT_OK) return s;
    // the batch folds every signal with the gathering permute+filter (one persistent launch)
    if (params[i].front_end != SFFT_FRONT_AUTO && params[i].front_end != SFFT_FRONT_GATHER)
      return SFFT_ERR_INVALID;
  }
  sfft_batch* b = new sfft_batch();
  b->count = count;
  b->fp32 = prec == SFFT_FP32;
  b->st = (hipStream_t)stream;
  if (!b->st) {
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
      delete b;
      return SFFT_ERR_HIP;
    }
    b->own_st = 1;
  }
  for (int i = 0; i < count; ++i) {
    sfft_plan* p = nullptr;
    sfft_params pi = params[i];
    pi.front_end = SFFT_FRONT_GATHER;   // AUTO: the batch's own front end
    const sfft_status s = sfft_plan_create(&pi, prec, b->st, &p);
    if (s != SFFT_OK) {
      sfft_batch_destroy(b);
      return s;
    }
    b->plans.push_back(p);
  }
  b->workers.resize(kWorkers, nullptr);
  b->ev_join.resize(kWorkers, nullptr);
  bool ok = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->ev_pf, hipEventDisableTiming) == hipSuccess &&
            hipMalloc((void**)&b->d_jobs, sizeof(PfJob) * (size_t)count) == hipSuccess &&
            hipMalloc((void**)&b->d_counts, sizeof(int) * (size_t)count) == hipSuccess &&
            hipHostMalloc((void**)&b->h_counts, sizeof(int) * (size_t)count,
                          hipHostMallocDefault) == hipSuccess;
  for (int s = 0; ok && s < kWorkers; ++s)   // default priority, non-blocking
    ok = hipStreamCreateWithFlags(&b->workers[s], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&b->ev_join[s], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    sfft_batch_destroy(b);
    return SFFT_ERR_ALLOC;
  }
  *out = b;
  return SFFT_OK;
}

extern "C" sfft_plan* sfft_batch_plan(sfft_batch* b, int i) {
  return (b && i >= 0 && i < b->count) ? b->plans[i] : nullptr;
}

extern "C" sfft_status sfft_batch_set_filter(sfft_batch* b, int i, const void* d_filter_time,
                                             const complex_t* h_filter_win) {
  if (!b || i < 0 || i >= b->count) return SFFT_ERR_INVALID;
  drop_graph(b);
  b->jobs_x.clear();   // the job table holds the taps pointer
  return sfft_plan_set_filter(b->plans[i], d_filter_time, h_filter_win);
}

// (re)builds the permute+filter work table for the current shape and signal pointers
static sfft_status build_jobs(sfft_batch* b, const void* const* d_x) {
  b->jobs.resize(b->count);
  int item = 0;
  for (int i = 0; i < b->count; ++i) {
    const sfft_plan* P = b->plans[i];
    PfJob& J = b->jobs[i];
    J.x = d_x[i];
    J.f = P->d_ft;
    J.bins = P->bins_t;
    J.ai = P->d_ai;
    J.mask = (uint32_t)P->p.n - 1u;
    J.B = P->p.B;
    J.w = P->p.w;
    J.loops = P->p.loops;
    J.tiled = (P->p.fold_mode == SFFT_FOLD_REFERENCE) && (P->p.n < (1 << 27));
    J.item0 = item;
    item += pf_batch_items(J.B, J.loops, b->lpt, b->nt, &J.nbx);
  }
  b->total_items = item;
  HIPCK(hipMemcpyAsync(b->d_jobs, b->jobs.data(), sizeof(PfJob) * (size_t)b->count,
                       hipMemcpyHostToDevice, b->st));
  b->jobs_x.assign(d_x, d_x + b->count);
  return SFFT_OK;
}

// the stage groups of worker s's signals (i = s, s + kWorkers, ...): the front groups (clear,
// comb FFT, comb cutoff, bitmaps) or the back groups (bucket FFT, cutoff, reverse hash,
// estimation; after the fold)
static sfft_status enqueue_worker(sfft_batch* b, const void* const* d_x, int s, bool back) {
  for (int i = s; i < b->count; i += kWorkers) {
    sfft_plan* P = b->plans[i];
    const HashArgs h = hash_args(P, b->d_counts + i);
    const sfft_status v = back ? run_back(P, h, nullptr, b->workers[s], false)
                               : run_front(P, d_x[i], h, b->workers[s], false);
    if (v != SFFT_OK) return v;
  }
  return SFFT_OK;
}

// one batch execution from the batch stream (fork/join): the persistent permute+filter of
// all signals and the workers' front groups run concurrently, then the back groups.  Eager
// launches go round-robin over the signals (so every worker starts at once); with graphs,
// each worker's groups are replayed from its two captured chains.
static sfft_status enqueue(sfft_batch* b, const void* const* d_x, bool graphs) {
  HIPCK(hipEventRecord(b->ev_fork, b->st));
  HIPCK(perm_filter_batch_launch(b->fp32, b->d_jobs, b->count, b->total_items, b->lpt, b->nt,
                                 b->grid, b->st));
  HIPCK(hipEventRecord(b->ev_pf, b->st));
  for (int s = 0; s < kWorkers; ++s) HIPCK(hipStreamWaitEvent(b->workers[s], b->ev_fork, 0));
  for (int back = 0; back < 2; ++back) {
    if (back)
      for (int s = 0; s < kWorkers; ++s) HIPCK(hipStreamWaitEvent(b->workers[s], b->ev_pf, 0));
    if (graphs) {
      for (int s = 0; s < kWorkers; ++s)
        if (b->execs[2 * s + back]) HIPCK(hipGraphLaunch(b->execs[2 * s + back], b->workers[s]));
      continue;
    }
    for (int i = 0; i < b->count; ++i) {
      sfft_plan* P = b->plans[i];
      const HashArgs h = hash_args(P, b->d_counts + i);
      hipStream_t w = b->workers[i % kWorkers];
      const sfft_status v = back ? run_back(P, h, nullptr, w, false)
                                 : run_front(P, d_x[i], h, w, false);
      if (v != SFFT_OK) return v;
    }
  }
  for (int s = 0; s < kWorkers; ++s) {
    HIPCK(hipEventRecord(b->ev_join[s], b->workers[s]));
    HIPCK(hipStreamWaitEvent(b->st, b->ev_join[s], 0));
  }
  return SFFT_OK;
}

// capture every worker's front and back chains (each on its own worker stream)
static sfft_status capture(sfft_batch* b, const void* const* d_x) {
  b->graphs.assign(2 * kWorkers, nullptr);
  b->execs.assign(2 * kWorkers, nullptr);
  for (int s = 0; s < kWorkers && s < b->count; ++s) {
    for (int part = 0; part < 2; ++part) {
      HIPCK(hipStreamBeginCapture(b->workers[s], hipStreamCaptureModeThreadLocal));
      const sfft_status v = enqueue_worker(b, d_x, s, part == 1);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(b->workers[s], &g);
      if (v != SFFT_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        drop_graph(b);
        return v != SFFT_OK ? v : SFFT_ERR_HIP;
      }
      b->graphs[2 * s + part] = g;
      if (hipGraphInstantiate(&b->execs[2 * s + part], g, nullptr, nullptr, 0) != hipSuccess) {
        drop_graph(b);
        return SFFT_ERR_HIP;
      }
    }
  }
  b->captured = true;
  b->graph_x.assign(d_x, d_x + b->count);
  return SFFT_OK;
}

static sfft_status prepare(sfft_batch* b, const void* const* d_x, const sfft_draws* draws) {
  for (int i = 0; i < b->count; ++i) {
    if (!d_x[i]) return SFFT_ERR_INVALID;
    if (!b->plans[i]->has_filter) return SFFT_ERR_NO_FILTER;
  }
  for (int i = 0; i < b->count; ++i) {
    const sfft_status v = upload_draws(b->plans[i], &draws[i]);
    if (v != SFFT_OK) return v;
  }
  if (b->jobs_x.size() != (size_t)b->count ||
      !std::equal(b->jobs_x.begin(), b->jobs_x.end(), d_x)) {
    drop_graph(b);
    return build_jobs(b, d_x);
  }
  return SFFT_OK;
}

extern "C" sfft_status sfft_batch_execute_async(sfft_batch* b, const void* const* d_x,
                                                const sfft_draws* draws) {
  if (!b || !d_x || !draws) return SFFT_ERR_INVALID;
  sfft_status v = prepare(b, d_x, draws);
  if (v != SFFT_OK) return v;
  if (!b->use_graph) return enqueue(b, d_x, false);
  if (!b->captured) {
    // nothing in flight may still use the worker streams' captured work
    for (int s = 0; s < kWorkers; ++s) HIPCK(hipStreamSynchronize(b->workers[s]));
    v = capture(b, d_x);
    if (v != SFFT_OK) return v;
  }
  return enqueue(b, d_x, true);
}

extern "C" sfft_status sfft_batch_execute(sfft_batch* b, const void* const* d_x,
                                          const sfft_draws* draws, Node* const* h_nodes,
                                          int* h_counts) {
  sfft_status v = sfft_batch_execute_async(b, d_x, draws);
  if (v != SFFT_OK) return v;
  HIPCK(hipMemcpyAsync(b->h_counts, b->d_counts, sizeof(int) * (size_t)b->count,
                       hipMemcpyDeviceToHost, b->st));
  HIPCK(hipStreamSynchronize(b->st));
  sfft_status res = SFFT_OK;
  for (int i = 0; i < b->count; ++i) {
    sfft_plan* P = b->plans[i];
    const int cnt = b->h_counts[i];
    P->last_count = cnt;
    if (h_counts) h_counts[i] = cnt;
    if (cnt > P->cap) res = SFFT_ERR_HIT_OVERFLOW;
    const int ncopy = std::min(cnt, P->cap);
    if (h_nodes && h_nodes[i] && ncopy > 0)
      HIPCK(hipMemcpyAsync(h_nodes[i], P->nodes, sizeof(Node) * (size_t)ncopy,
                           hipMemcpyDeviceToHost, b->st));
  }
  HIPCK(hipStreamSynchronize(b->st));
  return res;
}

extern "C" sfft_status sfft_batch_set_graph(sfft_batch* b, int on) {
  if (!b) return SFFT_ERR_INVALID;
  if (!on) drop_graph(b);
  b->use_graph = on ? 1 : 0;
  return SFFT_OK;
}

extern "C" const int* sfft_batch_counts_device(const sfft_batch* b) {
  return b ? b->d_counts : nullptr;
}

// mean milliseconds of `reps` runs of fn on the batch stream (one warm-up run first)
template <typename F>
static sfft_status time_ms(sfft_batch* b, int reps, F fn, float* ms) {
  hipEvent_t e0, e1;
  HIPCK(hipEventCreate(&e0));
  HIPCK(hipEventCreate(&e1));
  sfft_status v = fn();
  if (v == SFFT_OK && hipEventRecord(e0, b->st) != hipSuccess) v = SFFT_ERR_HIP;
  for (int r = 0; v == SFFT_OK && r < reps; ++r) v = fn();
  if (v == SFFT_OK && (hipEventRecord(e1, b->st) != hipSuccess ||
                       hipEventSynchronize(e1) != hipSuccess ||
                       hipEventElapsedTime(ms, e0, e1) != hipSuccess))
    v = SFFT_ERR_HIP;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (v == SFFT_OK) *ms /= (float)reps;
  return v;
}

extern "C" sfft_status sfft_batch_tune(sfft_batch* b, const void* const* d_x,
                                       const sfft_draws* draws, int reps, int* choice4) {
  if (!b || !d_x || !draws || reps < 1) return SFFT_ERR_INVALID;
  drop_graph(b);
  sfft_status v = prepare(b, d_x, draws);
  if (v != SFFT_OK) return v;
  // 1. permute+filter: loops per thread x block size x persistent grid
  const int lpts[] = {1, 2, 3, 4, 5, 6, 7, 8, 9}, nts[] = {128, 256},
            grids[] = {1024, 2048, 4096};
  float best = 1e30f;
  int bl = b->lpt, bn = b->nt, bg = b->grid;
  for (int lpt : lpts)
    for (int nt : nts)
      for (int grid : grids) {
        b->lpt = lpt;
        b->nt = nt;
        b->grid = grid;
        if ((v = build_jobs(b, d_x)) != SFFT_OK) return v;
        float ms = 0;
        v = time_ms(b, reps, [&]() -> sfft_status {
          HIPCK(perm_filter_batch_launch(b->fp32, b->d_jobs, b->count, b->total_items, b->lpt,
                                         b->nt, b->grid, b->st));
          return SFFT_OK;
        }, &ms);
        if (v != SFFT_OK) return v;
        if (ms < best) {
          best = ms;
          bl = lpt;
          bn = nt;
          bg = grid;
        }
      }
  b->lpt = bl;
  b->nt = bn;
  b->grid = bg;
  if ((v = build_jobs(b, d_x)) != SFFT_OK) return v;
  // 2. FFT columns per tile, bucket FFTs then comb FFTs, one choice for the whole batch
  const int qs[] = {1, 2, 4, 8, 16, 32};
  int best_q[2] = {0, 0};
  for (int which = 0; which < 2; ++which) {
    float bq = 1e30f;
    for (int q : qs) {
      float ms = 0;
      v = time_ms(b, reps, [&]() -> sfft_status {
        for (int i = 0; i < b->count; ++i) {
          sfft_plan* P = b->plans[i];
          FftLaunch f;
          if (which == 0) {
            f.tab = &P->tabB;
            f.in = P->bins_t;
            f.in_bstride = P->p.B;
            f.out = P->bins_f;
            f.out_bstride = P->p.B;
            f.scratch = P->bins_scr;
            f.keys = P->bkeys;
            f.keys_bstride = P->p.B;
            f.key_batches = P->p.loops_loc;
            f.batch = P->p.loops;
            f.store = kStoreBoth;
          } else {
            f.tab = &P->tabW;
            f.out = P->comb0;
            f.out_bstride = P->p.W_RLFC;
            f.scratch = P->comb1;
            f.keys = P->comb_keys;
            f.keys_bstride = P->p.W_RLFC;
            f.key_batches = P->p.RLFC_loops;
            f.batch = P->p.RLFC_loops;
            f.store = kStoreKeys;
            f.keys_tiled = 1;
            f.digits = P->comb_digits;
            f.load = kLoadComb;
            f.comb_x = d_x[i];
            f.comb_offs = P->d_offs;
            f.comb_sigma = P->p.n / P->p.W_RLFC;
          }
          f.qmax = q;
          HIPCK(fft_run(f, b->st));
        }
        return SFFT_OK;
      }, &ms);
      if (v != SFFT_OK) return v;
      if (ms < bq) {
        bq = ms;
        best_q[which] = q;
      }
    }
  }
  for (auto P : b->plans) {
    P->qmax_B = best_q[0];
    P->qmax_W = best_q[1];
  }
  if (choice4) {
    choice4[0] = b->lpt;
    choice4[1] = b->nt;
    choice4[2] = b->grid;
    choice4[3] = best_q[0] | (best_q[1] << 8);
  }
  return SFFT_OK;
}

// explicit shape (for tests that compare shapes): lpt in 1..9, nt 128/256, grid >= 1,
// FFT columns per tile (0 = default) for the bucket and comb FFTs
extern "C" sfft_status sfft_batch_set_shape(sfft_batch* b, int lpt, int nt, int grid,
                                            int qmax_bucket, int qmax_comb) {
  if (!b || lpt < 1 || lpt > 9 || (nt != 128 && nt != 256) || grid < 1 || qmax_bucket < 0 ||
      qmax_bucket > 32 || qmax_comb < 0 || qmax_comb > 32)
    return SFFT_ERR_INVALID;
  drop_graph(b);
  b->jobs_x.clear();
  b->lpt = lpt;
  b->nt = nt;
  b->grid = grid;
  for (auto P : b->plans) {
    P->qmax_B = qmax_bucket;
    P->qmax_W = qmax_comb;
  }
  return SFFT_OK;
}
