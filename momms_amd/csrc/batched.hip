// Thousands of independent GEMMs of uneven shape in one launch.
//
// No reference counterpart: the batching addition of the engine. A batch is a
// table of column-major problems (momms_gemm_desc: m, n, k, element offsets
// into shared A/B/C arenas, leading dimensions). A plan, built once per table,
//   - gives every problem a tile shape by its smaller side (plain FP64:
//     128 x 128 from T2 up, 64 x 64 from T up, one-wavefront 32 x 32 below;
//     (T, T2) chosen by timing the candidates on the batch itself, the
//     autotuner) or a slice count (split-BF16: the smallest one whose sampled
//     error stays within 2x plain FP64's);
//   - places every problem whole on one of eight queues, one per XCD,
//     balancing their work, and sorts each queue by k, largest first
//     (longest job first keeps the tail of a queue short);
//   - runs persistent workgroups that pull items from the queues' device
//     counters until all are empty: no per-problem launch.
// The whole sequence (counter reset by a kernel, slicing for the split mode,
// the persistent multiplies) can be captured once into a hipGraph and then
// replayed with one hipGraphLaunch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "batch_tiles.hpp"
#include "batched.hpp"
#include "crt_bf16.hpp"
#include "crt_recon.hpp"
#include "dgemm.hpp"
#include "dgemm_tile.hpp"
#include "mfma.hpp"
#include "momms_hip.h"
#include "split_bf16.hpp"
#include "split_tile.hpp"
#include "status.hpp"

namespace momms {


// Slicing task: 64 vectors (rows of A: kind 0, columns of B: kind 1) of one
// problem, and for the slice pass one 64-wide block of k.
struct SliceTask {
  int32_t prob, kind, vblock, pblock;
};

constexpr int kBatchPanel = 64;  // split mode: 64 x 64 tiles, panels of 64

__global__ void reset_counters(unsigned* c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0u;
}


// Plain mode, one persistent kernel over kBatchQueues queues, one per XCD.
// An item with prob >= 0 is a 64 x 64 workgroup tile (batch_tile<64, 64>); an item
// with prob = -1 - q is a quad: wavefront w of the workgroup computes 32 x 32
// tile 4q + w of the wave list (batch_wave_tile32; entries with prob < 0 pad
// a queue's wave list to whole quads), so small problems run four tiles per
// workgroup with no LDS and no barrier inside a tile. Every problem sits in
// one queue, whole, and workgroup b serves queue b % 8 first (workgroups are
// dispatched to the 8 XCDs in turn), so the A and B panels a problem's
// tiles share are fetched into one XCD's L2. A workgroup whose queue is
// empty takes items from the next queues. Each queue is sorted by k, largest
// first, so its tail is short and a quad's tiles take about equally long.
// Only problems with lda, ldb, ldc <= kBatchLeanLd and a smaller side below
// the plan's threshold T2 come here; the others go to batched_wide_kernel
// and batched_big_kernel.
// Three workgroups per CU (at most 168 registers a lane, 39 KB of LDS
// each): more resident waves hide the latency of short tiles; a fourth would
// cost the tiles their deeper prefetch or spill.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
batched_plain_kernel(const momms_gemm_desc* __restrict__ d, const BatchItem* __restrict__ items,
                     QueueBounds qb, const BatchItem* __restrict__ waves, int nwaves,
                     unsigned* counters, double alpha, const double* __restrict__ A,
                     const double* __restrict__ B, double beta, double* __restrict__ C) {
  run_queues(qb, counters, [&](int i) {
    const BatchItem w = items[uniform(i)];
    const int prob = uniform(w.prob);
    if (prob >= 0) {
      const momms_gemm_desc p = d[prob];
      batch_tile<64, 64, 3>(p.m, p.n, p.k, alpha, A + p.a_off, p.lda, B + p.b_off, p.ldb, beta,
                            C + p.c_off, p.ldc, (int64_t)uniform(w.tm) * 64,
                            (int64_t)uniform(w.tn) * 64);
    } else {
      // the same for all lanes of a wavefront
      const int wi = uniform(4 * (-1 - prob) + (int)(threadIdx.x >> 6));
      if (wi < nwaves) {
        const BatchItem t = waves[wi];
        const int tp = uniform(t.prob);
        if (tp >= 0) {
          const momms_gemm_desc p = d[tp];
          batch_wave_tile32<4>(p.m, p.n, p.k, alpha, A + p.a_off, p.lda, B + p.b_off, p.ldb,
                               beta, C + p.c_off, p.ldc, (int64_t)uniform(t.tm) * 32,
                               (int64_t)uniform(t.tn) * 32);
        }
      }
    }
  });
}

// Exponents of 64 vectors per task: 4 threads per vector share its k range.
__global__ void __launch_bounds__(256)
batched_exponents(const momms_gemm_desc* __restrict__ d, const SplitProb* __restrict__ sp,
                  const SliceTask* __restrict__ tasks, const double* __restrict__ A,
                  const double* __restrict__ B, char* __restrict__ ws) {
  __shared__ uint64_t part[4][64];
  const SliceTask t = tasks[blockIdx.x];
  const momms_gemm_desc p = d[t.prob];
  const int tid = threadIdx.x;
  // kind 0 (rows of A): adjacent rows are adjacent in memory -> vector = tid % 64
  // kind 1 (columns of B): k is contiguous -> 4 adjacent lanes per vector
  const int vv = t.kind == 0 ? (tid & 63) : (tid >> 2);
  const int q = t.kind == 0 ? (tid >> 6) : (tid & 3);
  const int64_t v = (int64_t)t.vblock * 64 + vv;
  const int64_t count = t.kind == 0 ? p.m : p.n;
  uint64_t best = 0;
  if (v < count) {
    const double* x = t.kind == 0 ? A + p.a_off + v : B + p.b_off + v * p.ldb;
    const int64_t se = t.kind == 0 ? p.lda : 1;
    // eight loads in flight per step (indices past k read element k-1 again)
    for (int64_t e0 = q; e0 < p.k; e0 += 32) {
      uint64_t b[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int64_t e = e0 + 4 * u;
        b[u] = (uint64_t)__double_as_longlong(x[(e < p.k ? e : p.k - 1) * se]) & 0x7fffffffffffffffull;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) best = b[u] > best ? b[u] : best;
    }
  }
  part[q][vv] = best;
  __syncthreads();
  if (tid < 64) {
    uint64_t b = part[0][tid];
    for (int i = 1; i < 4; ++i) b = part[i][tid] > b ? part[i][tid] : b;
    const int64_t w = (int64_t)t.vblock * 64 + tid;
    const SplitProb s = sp[t.prob];
    int* ex = reinterpret_cast<int*>(ws + (t.kind == 0 ? s.ea : s.eb));
    const int64_t vpad = t.kind == 0 ? s.mpad : s.npad;
    if (w < vpad) {
      int E = 0;
      if (w < count) {
        if (b >= 0x7ff0000000000000ull) E = kNonFinite;
        else if (b != 0) frexp(__longlong_as_double((long long)b), &E);
      }
      ex[w] = E;
    }
  }
}

template <int S>
__global__ void __launch_bounds__(256)
batched_slice(const momms_gemm_desc* __restrict__ d, const SplitProb* __restrict__ sp,
              const SliceTask* __restrict__ tasks, const double* __restrict__ A,
              const double* __restrict__ B, char* __restrict__ ws) {
  const SliceTask t = tasks[blockIdx.x];
  const momms_gemm_desc p = d[t.prob];
  const SplitProb s = sp[t.prob];
  if (t.kind == 0) {
    slice_block<S, kBatchPanel>(A + p.a_off, p.m, p.k, 1, p.lda,
                                reinterpret_cast<const int*>(ws + s.ea), s.mpad, s.kpad,
                                reinterpret_cast<uint16_t*>(ws + s.aq), (int64_t)t.vblock * 64,
                                (int64_t)t.pblock * 64);
  } else {
    slice_block<S, kBatchPanel>(B + p.b_off, p.n, p.k, p.ldb, 1,
                                reinterpret_cast<const int*>(ws + s.eb), s.npad, s.kpad,
                                reinterpret_cast<uint16_t*>(ws + s.bq), (int64_t)t.vblock * 64,
                                (int64_t)t.pblock * 64);
  }
}

template <int S>
__global__ void __launch_bounds__(256)
batched_split_kernel(const momms_gemm_desc* __restrict__ d, const SplitProb* __restrict__ sp,
                     const BatchItem* __restrict__ items, int nitems, unsigned* counter,
                     double alpha, double beta, double* __restrict__ C,
                     const char* __restrict__ ws, int slabs_per_chunk) {
  run_queue(counter, nitems, [&](int it) {
    const BatchItem w = items[it];
    const momms_gemm_desc p = d[w.prob];
    const SplitProb s = sp[w.prob];
    split_tile<S, 2, 2>(p.m, p.n, p.k > 0 ? s.kpad : 0, alpha, beta, C + p.c_off, 1, p.ldc,
                        reinterpret_cast<const uint16_t*>(ws + s.aq),
                        reinterpret_cast<const uint16_t*>(ws + s.bq),
                        reinterpret_cast<const int*>(ws + s.ea),
                        reinterpret_cast<const int*>(ws + s.eb), s.mpad / kBatchPanel,
                        s.npad / kBatchPanel, slabs_per_chunk, w.tm, w.tn);
  });
}

// ---------------------------------------------------------------------------
// host side

// Plain mode: a problem whose smaller side is at least the plan's threshold
// T2 takes 128 x 128 tiles (batched_big_kernel), one with it at least T
// 64 x 64 workgroup tiles, the others 32 x 32 one-wavefront tiles (both in
// batched_plain_kernel). The autotuner picks (T, T2) by timing the whole
// batch with each pair of candidates.
static const int kThresholds[] = {32, 48, 64, 96, 1 << 30};
static const int kBigThresholds[] = {128, 192, 256, 1 << 30};
constexpr int kDefaultThreshold = 48;
constexpr int kDefaultBigThreshold = 256;

// The autotuner's choices, remembered for the process per (kind, class of
// the batch: problem count and mean k), so plans for similar batches reuse
// them; exported and imported as a momms_tuning table.
static std::mutex g_tune_mu;
static std::map<std::pair<int, int>, std::array<int32_t, 4>> g_tuned;

static bool tuned_lookup(int kind, int key, std::array<int32_t, 4>& v) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto it = g_tuned.find({kind, key});
  if (it == g_tuned.end()) return false;
  v = it->second;
  return true;
}

static void tuned_store(int kind, int key, std::array<int32_t, 4> v) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tuned[{kind, key}] = v;
}

// The class key of a batch: 16 * min(15, floor(log2(1 + P))) +
// min(8, floor(log2(1 + K))) for P problems with m, n > 0 of mean k K.
static int batch_class(const momms_gemm_desc* d, int64_t count) {
  int64_t ksum = 0, live = 0;
  for (int64_t i = 0; i < count; ++i)
    if (d[i].m && d[i].n) {
      ksum += d[i].k;
      ++live;
    }
  const int kcls = live ? (int)std::min<int64_t>(8, (int64_t)std::log2(1.0 + (double)ksum / live)) : 0;
  const int ccls = (int)std::min<int64_t>(15, (int64_t)std::log2(1.0 + (double)live));
  return ccls * 16 + kcls;
}

// Whether a table entry's values are in range for its kind.
static bool tuning_entry_ok(const momms_tuning_entry& t) {
  if (t.key < 0 || t.key > 255) return false;
  const int32_t* v = t.value;
  switch (t.kind) {
    case MOMMS_TUNE_BATCH_FP64:
      return v[0] >= 16 && v[0] <= v[1] && v[1] <= (1 << 30) && v[2] == 0 && v[3] == 0;
    case MOMMS_TUNE_BATCH_SPLIT:
      return v[0] >= 1 && v[0] <= kMaxSlices && v[1] == 0 && v[2] == 0 && v[3] == 0;
    case MOMMS_TUNE_BATCH_CRT:
      return v[0] >= 1 && v[0] <= kCrtMaxModuli && v[1] == 0 && v[2] == 0 && v[3] == 0;
    default:
      return false;
  }
}

static int occupancy(const void* kernel, int threads, int cus) {
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per < 1)
    per = 1;
  return per * cus;
}

struct Launchable {
  int cfg;  // unused
  int nitems;
  BatchItem* items;
  int grid;
};

}  // namespace momms

struct momms_batch_plan {
  int mode = 0, slices = 0;
  int64_t count = 0;
  int cus = 256;
  std::vector<momms_gemm_desc> descs;
  momms_gemm_desc* d_desc = nullptr;
  unsigned* d_counters = nullptr;
  momms::Launchable plain{};                 // plain mode: the queue of 64 x 64 tiles and quads
  momms::Launchable wide{};                  // plain mode: tiles of problems with a large ld
  momms::Launchable big{};                   // plain mode: 128 x 128 tiles
  momms::QueueBounds qb_big{};               // plain mode: the per-XCD queues in big.items
  int threshold_big = 0;                     // plain mode: min(m, n) for 128 x 128 tiles
  int64_t tiles_big = 0;
  momms::BatchItem* d_waves = nullptr;       // plain mode: the 32 x 32 wave tiles
  int nwaves = 0;                            // plain mode: wave list length (whole quads)
  int64_t nwave_tiles = 0;                   // plain mode: 32 x 32 tiles in it
  momms::QueueBounds qb{};                   // plain mode: the per-XCD queues in plain.items
  int threshold = 0;                         // plain mode: min(m, n) for 64 x 64 tiles
  // split mode
  std::vector<momms::SplitProb> sp;
  momms::SplitProb* d_sp = nullptr;
  char* d_ws = nullptr;
  size_t ws_bytes = 0;
  momms::SliceTask* d_exp_tasks = nullptr;
  int n_exp_tasks = 0;
  momms::SliceTask* d_slice_tasks = nullptr;
  int n_slice_tasks = 0;
  momms::Launchable mul{};
  // residue mode (and the exponent tables above)
  int moduli = 0;
  momms::Launchable crt{};
  momms::QueueBounds qb_crt{};
  int beta_min = 0, beta_max = 0;
  // graph
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string info;
  int64_t total_tiles = 0;
  double flops = 0;
};

namespace momms {

static hipError_t upload(void** dst, const void* src, size_t bytes) {
  hipError_t e = hipMalloc(dst, std::max<size_t>(bytes, 16));
  if (e != hipSuccess) return e;
  if (bytes) e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
  return e;
}

static std::vector<BatchItem> tile_items(const std::vector<momms_gemm_desc>& d,
                                         const std::vector<int>& probs, int edge) {
  std::vector<std::pair<int64_t, BatchItem>> v;
  for (int p : probs) {
    const int64_t tm = ceil_div(d[p].m, edge), tn = ceil_div(d[p].n, edge);
    for (int64_t i = 0; i < tm; ++i)
      for (int64_t j = 0; j < tn; ++j) v.push_back({d[p].k, BatchItem{p, (int32_t)i, (int32_t)j, 0}});
  }
  std::stable_sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
  std::vector<BatchItem> out;
  out.reserve(v.size());
  for (auto& x : v) out.push_back(x.second);
  return out;
}

// The plain-mode queue for threshold T (replacing an earlier one).
// Problems to queues, largest work first onto the least loaded queue: the
// work of a problem is its count of 32 x 32 tile-units (a 64 x 64 tile is 4,
// a 128 x 128 one 16), each with a fixed cost of about 32 k-steps (prologue
// and C update).
static void assign_queues(const std::vector<momms_gemm_desc>& d, const std::vector<int>& probs,
                          const std::vector<int>& edge, std::vector<int> (&out)[kBatchQueues]) {
  std::vector<std::pair<double, int>> work;
  for (size_t i = 0; i < probs.size(); ++i) {
    const int p = probs[i], e = edge[i];
    const double units = (double)(ceil_div(d[p].m, e) * ceil_div(d[p].n, e)) * (e / 32) * (e / 32);
    work.push_back({units * (double)(d[p].k + 32), (int)i});
  }
  std::stable_sort(work.begin(), work.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
  double load[kBatchQueues] = {};
  for (const auto& w : work) {
    const int q = (int)(std::min_element(load, load + kBatchQueues) - load);
    load[q] += w.first;
    out[q].push_back((int)w.second);
  }
}

// The plain-mode queues for thresholds T and T2 (replacing earlier ones): a
// problem whose smaller side is at least T2 takes 128 x 128 tiles
// (batched_big_kernel), one with it at least T 64 x 64 tiles, the others
// 32 x 32 one-wavefront tiles (both in batched_plain_kernel).
static hipError_t build_plain(momms_batch_plan* P, int T, int T2) {
  (void)hipFree(P->plain.items);
  (void)hipFree(P->big.items);
  (void)hipFree(P->wide.items);
  (void)hipFree(P->d_waves);
  P->plain = Launchable{};
  P->big = Launchable{};
  P->wide = Launchable{};
  P->d_waves = nullptr;
  P->threshold = T;
  P->threshold_big = T2;
  const auto& d = P->descs;
  std::vector<int> big, mid, wide;  // mid: 64 x 64 or wave tiles
  std::vector<int> mid_edge;
  for (size_t i = 0; i < d.size(); ++i) {
    if (!d[i].m || !d[i].n) continue;
    const int64_t side = std::min(d[i].m, d[i].n);
    if (std::max(d[i].lda, std::max(d[i].ldb, d[i].ldc)) > kBatchLeanLd) {
      wide.push_back((int)i);
    } else if (side >= T2) {
      big.push_back((int)i);
    } else {
      mid.push_back((int)i);
      mid_edge.push_back(side >= T ? 64 : 32);
    }
  }
  hipError_t e;
  // 128 x 128 tiles, per queue by k
  {
    std::vector<int> qi_[kBatchQueues];
    assign_queues(d, big, std::vector<int>(big.size(), 128), qi_);
    std::vector<BatchItem> items;
    for (int qi = 0; qi < kBatchQueues; ++qi) {
      std::vector<int> probs;
      for (int i : qi_[qi]) probs.push_back(big[i]);
      std::sort(probs.begin(), probs.end());
      const auto tiles = tile_items(d, probs, 128);
      P->qb_big.off[qi] = (int)items.size();
      items.insert(items.end(), tiles.begin(), tiles.end());
    }
    P->qb_big.off[kBatchQueues] = (int)items.size();
    if (!items.empty()) {
      if ((e = upload((void**)&P->big.items, items.data(), items.size() * sizeof(BatchItem))) != hipSuccess)
        return e;
      P->big.nitems = (int)items.size();
      const int slots = occupancy(batched_big_kernel_ptr(), 256, P->cus);
      P->big.grid = std::max(1, std::min(P->big.nitems, slots));
    }
  }
  // per queue: its 64 x 64 tiles and quads of its wave tiles, by k
  std::vector<int> qmid[kBatchQueues];
  assign_queues(d, mid, mid_edge, qmid);
  std::vector<BatchItem> items, waves;
  int64_t ntiles = 0, nwave_tiles = 0;
  for (int qi = 0; qi < kBatchQueues; ++qi) {
    std::vector<int> p64, p32;
    for (int i : qmid[qi]) (mid_edge[i] == 64 ? p64 : p32).push_back(mid[i]);
    std::sort(p64.begin(), p64.end());
    std::sort(p32.begin(), p32.end());
    const auto tiles = tile_items(d, p64, 64);
    auto wq = tile_items(d, p32, 32);
    ntiles += (int64_t)tiles.size();
    nwave_tiles += (int64_t)wq.size();
    std::vector<std::pair<int64_t, BatchItem>> q;
    for (const auto& t : tiles) q.push_back({d[t.prob].k, t});
    const int64_t w0 = (int64_t)waves.size();
    for (size_t w = 0; w < wq.size(); w += 4)
      q.push_back({d[wq[w].prob].k, BatchItem{-1 - (int32_t)((w0 + (int64_t)w) / 4), 0, 0, 0}});
    while (wq.size() % 4) wq.push_back(BatchItem{-1, 0, 0, 0});
    waves.insert(waves.end(), wq.begin(), wq.end());
    std::stable_sort(q.begin(), q.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    P->qb.off[qi] = (int)items.size();
    for (auto& x : q) items.push_back(x.second);
  }
  P->qb.off[kBatchQueues] = (int)items.size();
  if (!items.empty()) {
    if ((e = upload((void**)&P->plain.items, items.data(), items.size() * sizeof(BatchItem))) != hipSuccess)
      return e;
    if ((e = upload((void**)&P->d_waves, waves.data(), waves.size() * sizeof(BatchItem))) != hipSuccess)
      return e;
    P->plain.nitems = (int)items.size();
    P->nwaves = (int)waves.size();
    const int slots = occupancy(reinterpret_cast<const void*>(&batched_plain_kernel), 256, P->cus);
    P->plain.grid = std::max(1, std::min(P->plain.nitems, slots));
  }
  const auto wide_tiles = tile_items(d, wide, 64);
  if (!wide_tiles.empty()) {
    if ((e = upload((void**)&P->wide.items, wide_tiles.data(), wide_tiles.size() * sizeof(BatchItem))) !=
        hipSuccess)
      return e;
    P->wide.nitems = (int)wide_tiles.size();
    const int wslots = occupancy(batched_wide_kernel_ptr(), 256, P->cus);
    P->wide.grid = std::max(1, std::min(P->wide.nitems, wslots));
  }
  P->tiles_big = P->big.nitems;
  P->total_tiles = P->big.nitems + ntiles + nwave_tiles + (int64_t)wide_tiles.size();
  P->nwave_tiles = nwave_tiles;
  return hipSuccess;
}

#define MOMMS_SPLIT_SWITCH(S_, CALL) \
  switch (S_) {                      \
    case 1: CALL(1); break;          \
    case 2: CALL(2); break;          \
    case 3: CALL(3); break;          \
    case 4: CALL(4); break;          \
    case 5: CALL(5); break;          \
    case 6: CALL(6); break;          \
    case 7: CALL(7); break;          \
    default: CALL(8); break;         \
  }

static hipError_t enqueue(const momms_batch_plan* P, double alpha, const double* A,
                          const double* B, double beta, double* C, hipStream_t st) {
  // plain mode: counters 0..7 the queues, 8 the wide problems, 9..16 the
  // queues of 128 x 128 tiles
  const int nctr = P->mode == 0 ? 2 * kBatchQueues + 1 : (P->mode == 2 ? kBatchQueues : 1);
  hipLaunchKernelGGL(reset_counters, dim3(1), dim3(64), 0, st, P->d_counters, nctr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (P->mode == 0) {
    if (P->big.nitems)
      launch_batched_big(P->big.grid, st, P->d_desc, P->big.items, P->qb_big,
                         P->d_counters + kBatchQueues + 1, alpha, A, B, beta, C);
    if (P->plain.nitems)
      hipLaunchKernelGGL(batched_plain_kernel, dim3((unsigned)P->plain.grid), dim3(256), 0, st,
                         P->d_desc, P->plain.items, P->qb, P->d_waves, P->nwaves,
                         P->d_counters, alpha, A, B, beta, C);
    if (P->wide.nitems)
      launch_batched_wide(P->wide.grid, st, P->d_desc, P->wide.items, P->wide.nitems,
                          P->d_counters + kBatchQueues, alpha, A, B, beta, C);
    return hipGetLastError();
  }
  if (P->mode == 2) {
    if (P->n_exp_tasks)
      hipLaunchKernelGGL(batched_exponents, dim3((unsigned)P->n_exp_tasks), dim3(256), 0, st,
                         P->d_desc, P->d_sp, P->d_exp_tasks, A, B, P->d_ws);
    if (P->crt.nitems)
      launch_batched_crt(P->moduli, P->crt.grid, st, P->d_desc, P->d_sp, P->crt.items, P->qb_crt,
                         P->d_counters, alpha, A, B, beta, C, P->d_ws);
    return hipGetLastError();
  }
  if (P->n_exp_tasks) {
    hipLaunchKernelGGL(batched_exponents, dim3((unsigned)P->n_exp_tasks), dim3(256), 0, st,
                       P->d_desc, P->d_sp, P->d_exp_tasks, A, B, P->d_ws);
  }
  if (P->n_slice_tasks) {
#define MOMMS_BS(S) hipLaunchKernelGGL((batched_slice<S>), dim3((unsigned)P->n_slice_tasks), dim3(256), 0, st, P->d_desc, P->d_sp, P->d_slice_tasks, A, B, P->d_ws)
    MOMMS_SPLIT_SWITCH(P->slices, MOMMS_BS)
#undef MOMMS_BS
  }
  if (P->mul.nitems) {
    const int spc = split_chunk(P->slices) / kSlab;
#define MOMMS_BM(S) hipLaunchKernelGGL((batched_split_kernel<S>), dim3((unsigned)P->mul.grid), dim3(256), 0, st, P->d_desc, P->d_sp, P->mul.items, P->mul.nitems, P->d_counters, alpha, beta, C, P->d_ws, spc)
    MOMMS_SPLIT_SWITCH(P->slices, MOMMS_BM)
#undef MOMMS_BM
  }
  return hipGetLastError();
}

static const void* split_kernel_ptr(int s) {
#define MOMMS_KP(S) return reinterpret_cast<const void*>(&batched_split_kernel<S>)
  MOMMS_SPLIT_SWITCH(s, MOMMS_KP)
#undef MOMMS_KP
  return nullptr;
}

static void free_plan(momms_batch_plan* P) {
  if (!P) return;
  (void)hipFree(P->plain.items);
  (void)hipFree(P->big.items);
  (void)hipFree(P->wide.items);
  (void)hipFree(P->d_waves);
  (void)hipFree(P->mul.items);
  (void)hipFree(P->d_desc);
  (void)hipFree(P->d_counters);
  (void)hipFree(P->d_sp);
  (void)hipFree(P->d_ws);
  (void)hipFree(P->d_exp_tasks);
  (void)hipFree(P->d_slice_tasks);
  (void)hipFree(P->crt.items);
  if (P->exec) (void)hipGraphExecDestroy(P->exec);
  if (P->graph) (void)hipGraphDestroy(P->graph);
  if (P->cap_stream) (void)hipStreamDestroy(P->cap_stream);
  delete P;
}

// Builds the split-mode tables for slice count s (replacing earlier ones).
static hipError_t build_split(momms_batch_plan* P, int s) {
  (void)hipFree(P->d_sp);
  (void)hipFree(P->d_ws);
  (void)hipFree(P->d_exp_tasks);
  (void)hipFree(P->d_slice_tasks);
  (void)hipFree(P->mul.items);
  P->d_sp = nullptr;
  P->d_ws = nullptr;
  P->d_exp_tasks = P->d_slice_tasks = nullptr;
  P->mul = Launchable{};
  P->slices = s;
  const auto& d = P->descs;
  P->sp.assign(d.size(), SplitProb{});
  size_t off = 0;
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  std::vector<SliceTask> et, st;
  std::vector<int> probs;
  for (size_t i = 0; i < d.size(); ++i) {
    SplitProb& q = P->sp[i];
    q.mpad = (int32_t)(ceil_div(d[i].m, kBatchPanel) * kBatchPanel);
    q.npad = (int32_t)(ceil_div(d[i].n, kBatchPanel) * kBatchPanel);
    q.kpad = (int32_t)(ceil_div(std::max<int64_t>(d[i].k, 1), kSlab) * kSlab);
    q.aq = (int64_t)off;
    off = al(off + (size_t)s * q.mpad * q.kpad * 2);
    q.bq = (int64_t)off;
    off = al(off + (size_t)s * q.npad * q.kpad * 2);
    q.ea = (int64_t)off;
    off = al(off + (size_t)q.mpad * 4);
    q.eb = (int64_t)off;
    off = al(off + (size_t)q.npad * 4);
    if (d[i].m == 0 || d[i].n == 0) continue;
    probs.push_back((int)i);
    for (int kind = 0; kind < 2; ++kind) {
      const int vb = (kind == 0 ? q.mpad : q.npad) / 64;
      const int pb = (int)ceil_div(q.kpad, 64);
      for (int v = 0; v < vb; ++v) {
        et.push_back({(int32_t)i, kind, v, 0});
        for (int pq = 0; pq < pb; ++pq) st.push_back({(int32_t)i, kind, v, pq});
      }
    }
  }
  P->ws_bytes = std::max<size_t>(off, 256);
  hipError_t e;
  if ((e = hipMalloc((void**)&P->d_ws, P->ws_bytes)) != hipSuccess) return e;
  if ((e = upload((void**)&P->d_sp, P->sp.data(), P->sp.size() * sizeof(SplitProb))) != hipSuccess)
    return e;
  if ((e = upload((void**)&P->d_exp_tasks, et.data(), et.size() * sizeof(SliceTask))) != hipSuccess)
    return e;
  if ((e = upload((void**)&P->d_slice_tasks, st.data(), st.size() * sizeof(SliceTask))) != hipSuccess)
    return e;
  P->n_exp_tasks = (int)et.size();
  P->n_slice_tasks = (int)st.size();
  auto items = tile_items(d, probs, kBatchPanel);
  P->mul.nitems = (int)items.size();
  if ((e = upload((void**)&P->mul.items, items.data(), items.size() * sizeof(BatchItem))) != hipSuccess)
    return e;
  P->mul.grid = std::max(1, std::min(P->mul.nitems, occupancy(split_kernel_ptr(s), 256, P->cus)));
  return hipSuccess;
}

// Builds the residue-mode tables for N moduli (replacing earlier ones):
// the exponent workspace, beta of every problem and the per-XCD queues of
// 64 x 64 tiles. Fails (hipErrorInvalidValue) if N moduli cannot hold some
// problem's k-term products.
static hipError_t build_crt(momms_batch_plan* P, int N) {
  (void)hipFree(P->d_sp);
  (void)hipFree(P->d_ws);
  (void)hipFree(P->d_exp_tasks);
  (void)hipFree(P->crt.items);
  P->d_sp = nullptr;
  P->d_ws = nullptr;
  P->d_exp_tasks = nullptr;
  P->crt = Launchable{};
  P->moduli = N;
  const auto& d = P->descs;
  P->sp.assign(d.size(), SplitProb{});
  size_t off = 0;
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  std::vector<SliceTask> et;
  std::vector<int> probs;
  P->beta_min = 1 << 30;
  P->beta_max = 0;
  for (size_t i = 0; i < d.size(); ++i) {
    SplitProb& q = P->sp[i];
    q.mpad = (int32_t)(ceil_div(d[i].m, 64) * 64);
    q.npad = (int32_t)(ceil_div(d[i].n, 64) * 64);
    q.kpad = (int32_t)ceil_div(d[i].k, 16) * 16;
    q.ea = (int64_t)off;
    off = al(off + (size_t)q.mpad * 4);
    q.eb = (int64_t)off;
    off = al(off + (size_t)q.npad * 4);
    if (d[i].m == 0 || d[i].n == 0) continue;
    q.beta = crt_beta(d[i].k, N);
    if (q.beta < 1) return hipErrorInvalidValue;
    P->beta_min = std::min(P->beta_min, (int)q.beta);
    P->beta_max = std::max(P->beta_max, (int)q.beta);
    probs.push_back((int)i);
    for (int kind = 0; kind < 2; ++kind) {
      const int vb = (kind == 0 ? q.mpad : q.npad) / 64;
      for (int v = 0; v < vb; ++v) et.push_back({(int32_t)i, kind, v, 0});
    }
  }
  if (probs.empty()) P->beta_min = 0;
  P->ws_bytes = std::max<size_t>(off, 256);
  hipError_t e;
  if ((e = hipMalloc((void**)&P->d_ws, P->ws_bytes)) != hipSuccess) return e;
  if ((e = upload((void**)&P->d_sp, P->sp.data(), P->sp.size() * sizeof(SplitProb))) != hipSuccess)
    return e;
  if ((e = upload((void**)&P->d_exp_tasks, et.data(), et.size() * sizeof(SliceTask))) != hipSuccess)
    return e;
  P->n_exp_tasks = (int)et.size();
  // per queue: its problems' 64 x 64 tiles, by k
  std::vector<int> qp[kBatchQueues];
  assign_queues(d, probs, std::vector<int>(probs.size(), 64), qp);
  std::vector<BatchItem> items;
  for (int qi = 0; qi < kBatchQueues; ++qi) {
    std::vector<int> ps;
    for (int i : qp[qi]) ps.push_back(probs[i]);
    std::sort(ps.begin(), ps.end());
    const auto tiles = tile_items(d, ps, 64);
    P->qb_crt.off[qi] = (int)items.size();
    items.insert(items.end(), tiles.begin(), tiles.end());
  }
  P->qb_crt.off[kBatchQueues] = (int)items.size();
  P->crt.nitems = (int)items.size();
  if (!items.empty()) {
    if ((e = upload((void**)&P->crt.items, items.data(), items.size() * sizeof(BatchItem))) != hipSuccess)
      return e;
    P->crt.grid = std::max(1, std::min(P->crt.nitems, occupancy(batched_crt_kernel_ptr(N), 512, P->cus)));
  }
  P->total_tiles = P->crt.nitems;
  return hipSuccess;
}

// The smallest moduli count that can hold every problem's k-term products.
static int crt_min_moduli(const std::vector<momms_gemm_desc>& d) {
  int64_t kmax = 0;
  for (const auto& p : d)
    if (p.m && p.n) kmax = std::max(kmax, p.k);
  for (int N = 1; N <= kCrtMaxModuli; ++N)
    if (crt_beta(kmax, N) >= 1) return N;
  return kCrtMaxModuli + 1;
}

static float time_enqueue(momms_batch_plan* P, const double* A, const double* B, double* C,
                          hipStream_t st, int reps) {
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  float best = 1e30f;
  for (int r = 0; r < reps; ++r) {
    (void)hipEventRecord(e0, st);
    (void)enqueue(P, 1.0, A, B, 0.0, C, st);
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    best = std::min(best, ms);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

static size_t c_extent(const std::vector<momms_gemm_desc>& d) {
  size_t e = 0;
  for (auto& p : d)
    if (p.m && p.n) e = std::max(e, (size_t)(p.c_off + (p.n - 1) * p.ldc + p.m));
  return e;
}

}  // namespace momms

using namespace momms;

extern "C" int momms_hip_batch_plan_create(const momms_gemm_desc* descs, int64_t count, int mode,
                                           int slices, int autotune, const double* A,
                                           const double* B, momms_batch_plan** out,
                                           void* stream) {
  static const char* fn = "momms_hip_batch_plan_create";
  if (descs == nullptr && count > 0) return bad_arg(fn, 1, "descs is null");
  if (count < 0 || count > (1 << 24)) return bad_arg(fn, 2, "count out of range");
  if (mode < 0 || mode > 2)
    return bad_arg(fn, 3, "mode must be 0 (FP64), 1 (split-BF16 slices) or 2 (BF16 residues)");
  if (mode == 1 && (slices < 0 || slices > kMaxSlices))
    return bad_arg(fn, 4, "slices must be 0 (tuned) or 1..8");
  if (mode == 2 && (slices < 0 || slices > kCrtMaxModuli))
    return bad_arg(fn, 4, "moduli must be 0 (tuned) or 1..16");
  if (autotune && (A == nullptr || B == nullptr)) return bad_arg(fn, 6, "autotune needs A and B");
  if (out == nullptr) return bad_arg(fn, 8, "out is null");
  for (int64_t i = 0; i < count; ++i) {
    const auto& p = descs[i];
    if (p.m < 0 || p.n < 0 || p.k < 0 || p.m > (1 << 24) || p.n > (1 << 24) ||
        p.a_off < 0 || p.b_off < 0 || p.c_off < 0 || p.lda < std::max<int64_t>(p.m, 1) ||
        p.ldb < std::max<int64_t>(p.k, 1) || p.ldc < std::max<int64_t>(p.m, 1)) {
      char buf[160];
      std::snprintf(buf, sizeof(buf), "descs[%lld] is invalid", (long long)i);
      return bad_arg(fn, 1, buf);
    }
    // the residue mode's fused kernel takes 32-bit byte offsets from a slab's corner
    if (mode == 2 && p.m && p.n && std::max(p.lda, p.ldb) > kBatchLeanLd)
      return bad_arg(fn, 1, "residue mode: lda and ldb must be at most 2^21");
  }
  hipStream_t st = (hipStream_t)stream;
  auto* P = new momms_batch_plan();
  P->mode = mode;
  P->count = count;
  P->descs.assign(descs, descs + count);
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&P->cus, hipDeviceAttributeMultiprocessorCount, dev);
  for (auto& p : P->descs) P->flops += 2.0 * (double)p.m * (double)p.n * (double)p.k;
  hipError_t e;
#define MOMMS_FAIL(E)        \
  do {                       \
    free_plan(P);            \
    return hip_status(fn, E); \
  } while (0)
  if ((e = upload((void**)&P->d_desc, P->descs.data(), P->descs.size() * sizeof(momms_gemm_desc))) != hipSuccess)
    MOMMS_FAIL(e);
  if ((e = hipMalloc((void**)&P->d_counters, 64 * sizeof(unsigned))) != hipSuccess) MOMMS_FAIL(e);

  double* scratchC = nullptr;
  const size_t cext = c_extent(P->descs);
  if (autotune && cext) {
    if ((e = hipMalloc((void**)&scratchC, cext * sizeof(double))) != hipSuccess) MOMMS_FAIL(e);
  }

  char line[384];
  if (mode == 0) {
    int64_t live = 0;
    for (auto& p : P->descs) live += (p.m && p.n) ? 1 : 0;
    const int key = batch_class(P->descs.data(), count);
    int T = kDefaultThreshold, T2 = kDefaultBigThreshold;
    std::array<int32_t, 4> tv{};
    const bool known = tuned_lookup(MOMMS_TUNE_BATCH_FP64, key, tv);
    if (known) {
      T = tv[0];
      T2 = tv[1];
    }
    if (!known && autotune && live) {
      float best = 1e30f;
      for (int cand2 : kBigThresholds)
        for (int cand : kThresholds) {
          if (cand > cand2 && cand != (1 << 30)) continue;
          if ((e = build_plain(P, cand, cand2)) != hipSuccess) {
            (void)hipFree(scratchC);
            MOMMS_FAIL(e);
          }
          (void)time_enqueue(P, A, B, scratchC, st, 1);  // warm up
          const float ms = time_enqueue(P, A, B, scratchC, st, 3);
          if (ms < best) {
            best = ms;
            T = cand;
            T2 = cand2;
          }
        }
      tuned_store(MOMMS_TUNE_BATCH_FP64, key, {T, T2, 0, 0});
    }
    if ((e = build_plain(P, T, T2)) != hipSuccess) {
      (void)hipFree(scratchC);
      MOMMS_FAIL(e);
    }
    std::snprintf(line, sizeof(line),
                  "{\"mode\":\"fp64\",\"threshold\":%d,\"threshold_128\":%d,\"tiles_128x128\":%lld,"
                  "\"tiles_64x64\":%lld,\"tiles_32x32_wavefront\":%lld,\"queue\":%d,\"queues\":%d,"
                  "\"workgroups\":%d,\"workgroups_128\":%d}",
                  T, T2, (long long)P->tiles_big,
                  (long long)(P->total_tiles - P->nwave_tiles - P->tiles_big),
                  (long long)P->nwave_tiles, P->plain.nitems, kBatchQueues, P->plain.grid,
                  P->big.grid);
    P->info = line;
  } else {
    // mode 1: slice count; mode 2: moduli count (the `slices` argument)
    const bool crt = mode == 2;
    const int kind = crt ? MOMMS_TUNE_BATCH_CRT : MOMMS_TUNE_BATCH_SPLIT;
    const int cmin = crt ? crt_min_moduli(P->descs) : 1;
    if (cmin > kCrtMaxModuli) {
      (void)hipFree(scratchC);
      free_plan(P);
      return bad_arg(fn, 1, "a problem's k is too large for 16 moduli");
    }
    if (crt && slices != 0 && slices < cmin) {
      (void)hipFree(scratchC);
      free_plan(P);
      return bad_arg(fn, 4, "too few moduli for the largest k of the batch");
    }
    auto build = [&](int c) { return crt ? build_crt(P, c) : build_split(P, c); };
    int s = slices;
    const int key = batch_class(P->descs.data(), count);
    std::array<int32_t, 4> tv{};
    if (s == 0 && tuned_lookup(kind, key, tv) && tv[0] >= cmin) s = tv[0];
    if (s == 0 && autotune) {
      // the smallest count whose sampled error is within 2x plain FP64's
      const int first = crt ? std::max(cmin, 8) : 4, last = crt ? kCrtMaxModuli : kMaxSlices;
      momms_batch_plan T;  // the plain FP64 batch, for the reference error
      T.mode = 0;
      T.cus = P->cus;
      T.descs = P->descs;
      T.d_desc = P->d_desc;
      T.d_counters = P->d_counters;
      std::vector<int> all;
      for (int64_t i = 0; i < count; ++i)
        if (P->descs[i].m && P->descs[i].n) all.push_back((int)i);
      e = build_plain(&T, kDefaultThreshold, kDefaultBigThreshold);
      double* plainC = nullptr;
      if (e == hipSuccess) e = hipMalloc((void**)&plainC, std::max<size_t>(cext, 1) * sizeof(double));
      if (e == hipSuccess) e = enqueue(&T, 1.0, A, B, 0.0, plainC, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      double plain_err = 0;
      const int sample_probs = (int)std::min<size_t>(all.size(), 64);
      auto max_err = [&](const double* Cm) {
        double worst = 0;
        for (int q = 0; q < sample_probs; ++q) {
          const auto& p = P->descs[all[(size_t)q * all.size() / sample_probs]];
          const double one = 1.0, zero = 0.0;
          double ms = 0, mr = 0;
          if (momms_hip_gemm_check_sampled(p.m, p.n, p.k, &one, A + p.a_off, 1, p.lda, B + p.b_off,
                                           1, p.ldb, &zero, nullptr, Cm + p.c_off, 1, p.ldc, 64,
                                           0x6D6F6D6D73ull + q, nullptr, &ms, &mr, st) == 0)
            worst = std::max(worst, ms);
        }
        return worst;
      };
      if (e == hipSuccess) plain_err = max_err(plainC);
      s = last;
      for (int cand = first; cand <= last && e == hipSuccess; ++cand) {
        if ((e = build(cand)) != hipSuccess) break;
        if ((e = enqueue(P, 1.0, A, B, 0.0, scratchC, st)) != hipSuccess) break;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) break;
        if (max_err(scratchC) <= 2.0 * plain_err) {
          s = cand;
          break;
        }
      }
      (void)hipFree(T.plain.items);
      (void)hipFree(T.big.items);
      (void)hipFree(T.wide.items);
      (void)hipFree(T.d_waves);
      T.plain.items = nullptr;
      T.big.items = nullptr;
      T.wide.items = nullptr;
      T.d_waves = nullptr;
      T.d_desc = nullptr;  // owned by P
      T.d_counters = nullptr;
      (void)hipFree(plainC);
      if (e != hipSuccess) {
        (void)hipFree(scratchC);
        MOMMS_FAIL(e);
      }
      tuned_store(kind, key, {s, 0, 0, 0});
    } else if (s == 0) {
      s = crt ? std::max(cmin, 13) : 7;
    }
    if ((e = build(s)) != hipSuccess) {
      (void)hipFree(scratchC);
      MOMMS_FAIL(e);
    }
    if (crt) {
      std::snprintf(line, sizeof(line),
                    "{\"mode\":\"crt_bf16\",\"moduli\":%d,\"tile\":64,\"tiles\":%d,\"beta_min\":%d,"
                    "\"beta_max\":%d,\"queues\":%d,\"workgroups\":%d,\"workspace_bytes\":%zu}",
                    s, P->crt.nitems, P->beta_min, P->beta_max, kBatchQueues, P->crt.grid, P->ws_bytes);
    } else {
      P->total_tiles = P->mul.nitems;
      std::snprintf(line, sizeof(line),
                    "{\"mode\":\"split_bf16\",\"slices\":%d,\"tile\":%d,\"tiles\":%d,\"workspace_bytes\":%zu}",
                    s, kBatchPanel, P->mul.nitems, P->ws_bytes);
    }
    P->info = line;
  }
  (void)hipFree(scratchC);
  if ((e = hipStreamCreate(&P->cap_stream)) != hipSuccess) MOMMS_FAIL(e);
#undef MOMMS_FAIL
  *out = P;
  return 0;
}

extern "C" int momms_hip_batch_run(momms_batch_plan* plan, const double* alpha, const double* A,
                                   const double* B, const double* beta, double* C, void* stream) {
  static const char* fn = "momms_hip_batch_run";
  if (!plan) return bad_arg(fn, 1, "plan is null");
  if (!alpha) return bad_arg(fn, 2, "alpha is null");
  if (!A && plan->count) return bad_arg(fn, 3, "A is null");
  if (!B && plan->count) return bad_arg(fn, 4, "B is null");
  if (!beta) return bad_arg(fn, 5, "beta is null");
  if (!C && plan->count) return bad_arg(fn, 6, "C is null");
  return hip_status(fn, enqueue(plan, *alpha, A, B, *beta, C, (hipStream_t)stream));
}

extern "C" int momms_hip_batch_capture(momms_batch_plan* plan, const double* alpha,
                                       const double* A, const double* B, const double* beta,
                                       double* C) {
  static const char* fn = "momms_hip_batch_capture";
  if (!plan) return bad_arg(fn, 1, "plan is null");
  if (!alpha) return bad_arg(fn, 2, "alpha is null");
  if (!A && plan->count) return bad_arg(fn, 3, "A is null");
  if (!B && plan->count) return bad_arg(fn, 4, "B is null");
  if (!beta) return bad_arg(fn, 5, "beta is null");
  if (!C && plan->count) return bad_arg(fn, 6, "C is null");
  if (plan->exec) {
    (void)hipGraphExecDestroy(plan->exec);
    plan->exec = nullptr;
  }
  if (plan->graph) {
    (void)hipGraphDestroy(plan->graph);
    plan->graph = nullptr;
  }
  MOMMS_HIP_TRY(fn, hipStreamBeginCapture(plan->cap_stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = enqueue(plan, *alpha, A, B, *beta, C, plan->cap_stream);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(plan->cap_stream, &g);
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return hip_status(fn, e);
  }
  if (e2 != hipSuccess) return hip_status(fn, e2);
  plan->graph = g;
  MOMMS_HIP_TRY(fn, hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0));
  return 0;
}

extern "C" int momms_hip_batch_launch(momms_batch_plan* plan, void* stream) {
  static const char* fn = "momms_hip_batch_launch";
  if (!plan) return bad_arg(fn, 1, "plan is null");
  if (!plan->exec) {
    set_error("momms_hip_batch_launch: plan has no captured graph (call momms_hip_batch_capture)");
    return -1;
  }
  return hip_status(fn, hipGraphLaunch(plan->exec, (hipStream_t)stream));
}

extern "C" const char* momms_hip_batch_plan_describe(const momms_batch_plan* plan) {
  return plan ? plan->info.c_str() : "";
}

extern "C" int momms_hip_batch_plan_destroy(momms_batch_plan* plan) {
  free_plan(plan);
  return 0;
}

extern "C" int momms_hip_dgemm_batched_loop(const momms_gemm_desc* descs, int64_t count,
                                            const double* alpha, const double* A,
                                            const double* B, const double* beta, double* C,
                                            void* stream) {
  static const char* fn = "momms_hip_dgemm_batched_loop";
  if (descs == nullptr && count > 0) return bad_arg(fn, 1, "descs is null");
  if (count < 0) return bad_arg(fn, 2, "count < 0");
  if (!alpha) return bad_arg(fn, 3, "alpha is null");
  if (!beta) return bad_arg(fn, 6, "beta is null");
  for (int64_t i = 0; i < count; ++i) {
    const auto& p = descs[i];
    const int st = momms_hip_dgemm(p.m, p.n, p.k, alpha, A + p.a_off, 1, p.lda, B + p.b_off, 1,
                                   p.ldb, beta, C + p.c_off, 1, p.ldc, stream);
    if (st != 0) return st;
  }
  return 0;
}

extern "C" int momms_hip_batch_class(const momms_gemm_desc* descs, int64_t count, int32_t* key) {
  static const char* fn = "momms_hip_batch_class";
  if (descs == nullptr && count > 0) return bad_arg(fn, 1, "descs is null");
  if (count < 0) return bad_arg(fn, 2, "count < 0");
  if (key == nullptr) return bad_arg(fn, 3, "key is null");
  *key = batch_class(descs, count);
  return 0;
}

extern "C" int momms_hip_tuning_export(momms_tuning* out) {
  static const char* fn = "momms_hip_tuning_export";
  if (out == nullptr) return bad_arg(fn, 1, "out is null");
  std::memset(out, 0, sizeof(*out));
  out->version = MOMMS_TUNING_VERSION;
  std::lock_guard<std::mutex> g(g_tune_mu);
  for (const auto& kv : g_tuned) {
    if (out->count == MOMMS_TUNING_MAX) break;
    momms_tuning_entry& t = out->entry[out->count++];
    t.kind = kv.first.first;
    t.key = kv.first.second;
    for (int i = 0; i < 4; ++i) t.value[i] = kv.second[i];
  }
  return 0;
}

extern "C" int momms_hip_tuning_import(const momms_tuning* in, int replace) {
  static const char* fn = "momms_hip_tuning_import";
  if (in == nullptr) return bad_arg(fn, 1, "in is null");
  if (in->version != MOMMS_TUNING_VERSION) return bad_arg(fn, 1, "table version differs from MOMMS_TUNING_VERSION");
  if (in->count > MOMMS_TUNING_MAX) return bad_arg(fn, 1, "count exceeds MOMMS_TUNING_MAX");
  for (uint32_t i = 0; i < in->count; ++i)
    if (!tuning_entry_ok(in->entry[i])) {
      char buf[96];
      std::snprintf(buf, sizeof(buf), "entry[%u] has an unknown kind or a value out of range", i);
      return bad_arg(fn, 1, buf);
    }
  std::lock_guard<std::mutex> g(g_tune_mu);
  if (replace) g_tuned.clear();
  for (uint32_t i = 0; i < in->count; ++i) {
    const momms_tuning_entry& t = in->entry[i];
    g_tuned[{t.kind, t.key}] = {t.value[0], t.value[1], t.value[2], t.value[3]};
  }
  return 0;
}
