// sfft_plan_impl.h -- the plan object and its stage groups, shared by the single-transform
// API (sfft_plan.hip) and the batch executor (sfft_batch.hip).  Internal: not installed.
#pragma once
#include "sfft_internal.h"

using namespace sfft;

// plan: one transform configuration and its workspace
struct sfft_plan {
  sfft_params p;
  int fp32 = 0;
  hipStream_t st = nullptr;
  hipStream_t st2 = nullptr;    // second chain (permute+filter, bucket FFT, cutoff)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int log2n = 0, log2B = 0, log2W = 0, m = 0;
  int cap = MAXSIZE_HITS;
  FftTables tabB, tabW;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  int* d_a = nullptr;
  int* d_ai = nullptr;
  int* d_offs = nullptr;
  void* comb0 = nullptr;
  void* comb1 = nullptr;
  void* comb_keys = nullptr;
  uint16_t* comb_digits = nullptr;   // first select digit of each comb key (same order)
  int stream = 0;               // streaming front end (else gathers)
  void* taps = nullptr;         // loops x w tap values (stream)
  void* comb_y = nullptr;       // RLFC_loops x W comb samples (stream)
  int columns = 0;              // column-ordered fold (SFFT_FRONT_COLUMNS)
  int bins_perm = 0;            // the slab fold: bins_t in column order, read via kLoadPerm
  void* frows = nullptr;        // bucket-major filter taps of the column fold
  void* bins_t = nullptr;
  void* bins_scr = nullptr;
  void* bins_f = nullptr;
  void* bkeys = nullptr;
  int* J = nullptr;
  int* J2 = nullptr;
  SelectWork selW, selB;
  unsigned* bitmap = nullptr;
  int bm_words = 0;
  int* maxfirst = nullptr;
  unsigned* jsel = nullptr;     // buckets selected per location loop (bitmaps)
  int* hitoff = nullptr;
  int* blocksum = nullptr;
  unsigned* hmask = nullptr;
  int* hits = nullptr;
  int* count = nullptr;
  double2* win = nullptr;
  Node* nodes = nullptr;
  uint8_t* sel = nullptr;       // median-selected loops of the last transform (2 per hit)
  // FFT columns per tile (0 = as many as fit, up to 32); the bucket FFTs take 16: full tiles
  // in their FP64 second pass (at n = 2^20, k = 1000, where 4 columns would give four times
  // the workgroups, both measure the same; 32, full FP32 tiles too, measured the same at
  // n = 2^27, k = 16000); the batch autotuner may set both
  int qmax_W = 0, qmax_B = 16;
  int* adj_head = nullptr;      // adjoint scratch: loops*B list heads, 2*cap links
  int* adj_next = nullptr;
  const void* d_ft = nullptr;
  int has_filter = 0;
  int last_count = 0;
  // pinned host staging: draws (a, ai, offsets packed like d_a..d_offs) and the hit count
  int* h_draws = nullptr;
  int* h_count = nullptr;
  int* h_count_dev = nullptr;   // device view of h_count
  // where the estimation kernel also stores the hit count (nullptr: nowhere); part of the
  // captured graph, like the node and count pointers
  int* cnt_copy = nullptr;
  hipEvent_t up_ev = nullptr;   // last upload out of h_draws
  int up_pending = 0;
  // optional stage timing (hipEvents at the reference's stage boundaries)
  int profile = 0;
  // optional hipGraph of the whole transform (front, fold, back), replayed while the signal
  // and output pointers stay the same; own_st: a stream made for it when the plan had none
  int use_graph = 0;
  int own_st = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  const void* g_x = nullptr;
  Node* g_nodes = nullptr;
  int* g_count = nullptr;
  int* g_cnt_copy = nullptr;
  hipEvent_t ev[9] = {nullptr, nullptr, nullptr, nullptr, nullptr,
                      nullptr, nullptr, nullptr, nullptr};
};


#define HIPCK(x)                                \
  do {                                          \
    if ((x) != hipSuccess) return SFFT_ERR_HIP; \
  } while (0)

sfft_status draws_check(const sfft_plan* P, const sfft_draws* d);
// packs the draws into the plan's pinned staging and copies them to the device on its stream
sfft_status upload_draws(sfft_plan* P, const sfft_draws* dr);
HashArgs hash_args(const sfft_plan* P, int* d_count);
// the transform in three stage groups on an explicit stream (marks: record the plan's
// profiling events)
sfft_status run_front(sfft_plan* P, const void* d_x, const HashArgs& h, hipStream_t st,
                      bool marks);
// clr: cleared by the slab fold first (column front end only; nullptr: nothing)
sfft_status run_fold(sfft_plan* P, const void* d_x, hipStream_t st,
                     const FoldClear* clr = nullptr);
sfft_status run_back(sfft_plan* P, const HashArgs& h, Node* d_nodes, hipStream_t st, bool marks);
sfft_status run_bucket(sfft_plan* P, hipStream_t st, bool marks);
sfft_status run_locate(sfft_plan* P, const HashArgs& h, Node* d_nodes, hipStream_t st,
                       bool marks);
