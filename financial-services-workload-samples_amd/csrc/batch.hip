// Mixed serving batch: GBM European and Heston-Asian requests of uneven
// sizes, each returning price, delta, vega and rho (with standard errors).
//
// Requests are served in micro-batches of `requests_per_launch` in arrival
// order.  Each micro-batch is three launches over a table of (request, chunk)
// work items: one grid of GBM items (pathwise Greeks), one grid of Heston
// items (heston_tangent_path: the three Greeks carried forward with each
// path), and a finalize grid that combines each request's item
// partials in item order (bitwise reproducible) into mean and standard error.
// A HIP event after each micro-batch gives the per-request completion time;
// an autotuner times chunk size x micro-batch size and keeps the fastest.
// With replay_graph the whole launch sequence is captured once per tuning
// into a hipGraph and replayed by every run: one launch from the host per
// run whatever the micro-batch count.  Completion times then come from a
// one-thread kernel after each micro-batch that stores the GPU's constant-
// rate wall clock (HIP events recorded by graph nodes cannot be timed
// against each other here).  No node clears memory (every partial, result
// and stamp is written before it is read), so replays need no reset.
// A request's normals and per-path arithmetic are those of the standalone
// entry point with the request index as the option index.
#include <algorithm>
#include <initializer_list>
#include <map>
#include <vector>

#include "fsmc_device.hpp"
#include "fsmc_host.hpp"
#include "path_models.hpp"

namespace fsmc {
namespace {

constexpr int kBlock = 256;
constexpr int kNV = 8;  // sums of price, delta, vega, rho and their squares

struct ReqDev {
  int kind, n_paths, n_steps, pad;
  double S0, K, T, r, sigma, v0, kappa, theta, xi, rho;
};

struct BatchArgs {
  const ReqDev* req;
  const int2* items;    // (request, chunk), GBM items then Heston items per micro-batch
  double* partials;     // [item][kNV]
  int chunk;
  uint32_t k0, k1, w_gbm, w_heston;
};

// One GBM item (a chunk of one request's paths) by the whole workgroup:
// pathwise price, delta, vega and rho sums into the item's partials.
template <typename Real>
__device__ __forceinline__ void gbm_item(const BatchArgs& a, int item, double* lds) {
  const int2 e = a.items[item];
  const ReqDev q = a.req[e.x];
  const int p0 = e.y * a.chunk;
  const int p1 = min(q.n_paths, p0 + a.chunk);
  const double disc = exp(-q.r * q.T), fd = disc / q.S0, kt = disc * q.K * q.T;
  const GbmConsts<Real> g(q.S0, q.K, q.T, q.r, q.sigma);
  double s[kNV] = {0, 0, 0, 0, 0, 0, 0, 0};
  // Philox layout of the GBM entry point: counter (path/4, 0, request)
  for (int base = p0 + 4 * (int)threadIdx.x; base < p1; base += 4 * kBlock) {
    Real z[4];
    normals4<Real>(U4{(uint32_t)(base >> 2), 0u, (uint32_t)e.x, a.w_gbm}, a.k0, a.k1, z[0], z[1], z[2], z[3]);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (base + j < p1) {
        Real st, dlog;
        const Real c = gbm_path(g, z[j], st, dlog);
        if (c > Real(0)) {
          const double pr = disc * (double)c, de = fd * (double)st, ve = disc * (double)(st * dlog);
          s[0] += pr;
          s[1] += pr * pr;
          s[2] += de;
          s[3] += de * de;
          s[4] += ve;
          s[5] += ve * ve;
          s[6] += kt;
          s[7] += kt * kt;
        }
      }
    }
  }
  block_sum<kNV>(s, lds);
  if (threadIdx.x == 0) {
    double* dst = a.partials + (size_t)item * kNV;
#pragma unroll
    for (int k = 0; k < kNV; ++k) dst[k] = s[k];
  }
}

// One Heston item: the tangent path of each of the chunk's paths (price,
// delta, vega, rho as the standalone tangent entry point computes them).
template <typename Real>
__device__ __forceinline__ void heston_item(const BatchArgs& a, int item, double* lds) {
  const int2 e = a.items[item];
  const ReqDev q = a.req[e.x];
  const int p0 = e.y * a.chunk;
  const int p1 = min(q.n_paths, p0 + a.chunk);
  const double disc = exp(-q.r * q.T);
  const HestonConsts<Real> c(q.T, q.n_steps, q.r, q.kappa, q.xi, q.theta, q.rho);
  double s[kNV] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int p = p0 + (int)threadIdx.x; p < p1; p += kBlock) {
    double g[3];
    const double x = heston_tangent_path<Real>(c, q.S0, q.K, q.T, q.v0, q.r, q.n_steps, (uint32_t)p, (uint32_t)e.x,
                                               a.w_heston, a.k0, a.k1, g);
    const double pr = disc * x;
    s[0] += pr;
    s[1] += pr * pr;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      s[2 + 2 * k] += g[k];
      s[3 + 2 * k] += g[k] * g[k];
    }
  }
  block_sum<kNV>(s, lds);
  if (threadIdx.x == 0) {
    double* dst = a.partials + (size_t)item * kNV;
#pragma unroll
    for (int k = 0; k < kNV; ++k) dst[k] = s[k];
  }
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void batch_gbm_kernel(BatchArgs a, int first_item) {
  __shared__ double lds[kNV * (kBlock / 64)];
  gbm_item<Real>(a, first_item + blockIdx.x, lds);
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void batch_heston_kernel(BatchArgs a, int first_item) {
  __shared__ double lds[kNV * (kBlock / 64)];
  heston_item<Real>(a, first_item + blockIdx.x, lds);
}

// Persistent variant: gridDim.x = CUs x workgroups per CU; workgroup g runs
// the items sched[off[g]] .. sched[off[g + 1] - 1] of this micro-batch, both
// kinds, in the order the host's longest-first schedule put them.  The
// schedule is static: no queue counter, no fence, no workgroup waits on
// another, and every loop is bounded by the schedule.
template <typename Real>
__global__ __launch_bounds__(kBlock) void batch_persistent_kernel(BatchArgs a, const int* sched, const int* off) {
  __shared__ double lds[kNV * (kBlock / 64)];
  const int j1 = off[blockIdx.x + 1];
  for (int j = off[blockIdx.x]; j < j1; ++j) {
    const int item = sched[j];
    if (a.req[a.items[item].x].kind == FSMC_REQ_GBM_EUROPEAN)  // uniform per workgroup
      gbm_item<Real>(a, item, lds);
    else
      heston_item<Real>(a, item, lds);
    __syncthreads();  // thread 0's reads of lds in block_sum before the next item's writes
  }
}

// Work-queue variant: gridDim.x = CUs x workgroups per CU; each workgroup
// takes the next item of this micro-batch's longest-first order from the
// shared counter *head (one relaxed device-scope increment per item, by
// thread 0, broadcast through LDS; the increment for the following item is
// issued when an item starts, so its latency hides behind the item) and
// runs it, until the index it takes is past the end.  Every index below
// n_items is taken exactly once, so a workgroup takes at most n_items + 1
// indices: the loop ends for every workgroup whatever the others do.  Nothing waits on another workgroup and
// nothing needs a fence: the items' inputs are written before the launch
// and their partials are read by the next kernel.  *head is zeroed by
// batch_reset_kernel earlier on the same stream.  One queue per request
// kind (kHeston): a kernel that could run either kind carries the union of
// their registers (100 VGPRs, four waves per SIMD, where the Heston items
// alone need 68 and run at seven), so a micro-batch runs its Heston queue
// and then its GBM queue.
template <typename Real, bool kHeston>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(sizeof(Real) == 4 ? 7 : 2))) void
batch_queue_kernel(BatchArgs a, const int* order, int n_items, int* head) {  // FP32: <= 72 VGPRs, seven waves
                                                                             // per SIMD as the grouped Heston
                                                                             // grid gets
  __shared__ double lds[kNV * (kBlock / 64)];
  __shared__ int next;
  if (threadIdx.x == 0) next = atomicAdd(head, 1);
  __syncthreads();
  // the index is uniform: readfirstlane keeps it (and everything the item
  // derives from it: addresses, request parameters) in scalar registers
  int j = __builtin_amdgcn_readfirstlane(next);
  __syncthreads();  // every thread has read next before thread 0 replaces it
  while (j < n_items) {  // uniform per workgroup
    // thread 0 takes the following index now; its latency overlaps the item
    const int ahead = threadIdx.x == 0 ? atomicAdd(head, 1) : 0;
    const int item = __builtin_amdgcn_readfirstlane(order[j]);
    if (kHeston)
      heston_item<Real>(a, item, lds);
    else
      gbm_item<Real>(a, item, lds);
    if (threadIdx.x == 0) next = ahead;
    __syncthreads();  // next is written, and thread 0's reads of lds in block_sum precede the next item's writes
    j = __builtin_amdgcn_readfirstlane(next);
    __syncthreads();  // every thread has read next before thread 0 replaces it
  }
}

__global__ void batch_reset_kernel(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Wall-clock stamp i (graph mode: i = 0 before the first micro-batch, w + 1
// after micro-batch w).
__global__ void batch_stamp_kernel(unsigned long long* stamps, int i) {
  if (threadIdx.x == 0) stamps[i] = wall_clock64();
}

// One thread per request of the micro-batch [r0, r0 + n): its items are
// first_item[r] .. first_item[r] + n_items[r] - 1 (contiguous per kind).
__global__ void batch_finalize_kernel(const ReqDev* req, const int* first_item, const int* n_items,
                                      const double* partials, int r0, int n, double* results) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = r0 + i;
  double t[kNV] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int it = 0; it < n_items[r]; ++it)
    for (int k = 0; k < kNV; ++k) t[k] += partials[(size_t)(first_item[r] + it) * kNV + k];
  const double N = (double)req[r].n_paths;
  double* res = results + (size_t)r * 8;
  for (int k = 0; k < 4; ++k) {
    res[k] = t[2 * k] / N;
    const double var = (N * t[2 * k + 1] - t[2 * k] * t[2 * k]) / (N * (N - 1.0));
    res[4 + k] = sqrt(var > 0.0 ? var : 0.0) / sqrt(N);
  }
}

}  // namespace

struct MicroBatch {
  int r0, n;                 // requests [r0, r0 + n)
  int gbm_first, gbm_count;  // items
  int hes_first, hes_count;
  int off_first;             // persistent mode: this micro-batch's G + 1 schedule offsets
  int q_first, q_count;      // work-queue mode: its items' order in qorder (Heston items, then GBM items)
  int q_hes;                 // the Heston items among them
};

struct BatchImpl {
  int n = 0;
  fsmc_mode mode;
  std::vector<fsmc_request> req;
  fsmc_batch_tuning tuning = {4096, 0, 0, 0, 0};
  std::vector<MicroBatch> waves;
  int cus = 0;  // compute units (persistent grid = cus x workgroups_per_cu)
  DeviceBuffer ws;
  size_t off_req = 0, off_items = 0, off_first = 0, off_nitems = 0, off_part = 0, off_res = 0,
         off_stamp = 0, off_sched = 0, off_soff = 0, off_qorder = 0, off_heads = 0;
  std::vector<hipEvent_t> done;  // one per micro-batch (stream mode)
  hipEvent_t start = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;  // autotuner's timing around a run
  hipEvent_t ran = nullptr;  // recorded by fsmc_batch_run after the last launch
  double wall_khz = 0.0;
  hipStream_t capture = nullptr;  // private stream the graph is captured on
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static std::map<std::vector<long long>, fsmc_batch_tuning>& tuning_cache() {
  static std::map<std::vector<long long>, fsmc_batch_tuning> c;
  return c;
}
static std::mutex g_cache_mu;

static std::vector<long long> batch_key(const BatchImpl& b) {
  // batches of the same size and mix share a tuning
  std::vector<long long> k = {b.mode.precision, b.n};
  long long gp = 0, hp = 0, hs = 0;
  for (const auto& r : b.req) {
    if (r.kind == FSMC_REQ_GBM_EUROPEAN) {
      gp += r.n_paths;
    } else {
      hp += r.n_paths;
      hs += (long long)r.n_paths * r.n_steps;
    }
  }
  k.push_back(gp >> 16);
  k.push_back(hp >> 16);
  k.push_back(hs >> 20);
  return k;
}

static void batch_release_events(BatchImpl& b) {
  if (b.exec) (void)hipGraphExecDestroy(b.exec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  b.exec = nullptr;
  b.graph = nullptr;
  for (hipEvent_t e : b.done) (void)hipEventDestroy(e);
  b.done.clear();
}

template <typename Real>
static int batch_enqueue(BatchImpl& b, hipStream_t s, bool stamps);

// Captures the launch sequence of the current tuning into b.exec.
static int batch_capture(BatchImpl& b) {
  if (!b.capture) FSMC_HIP(hipStreamCreateWithFlags(&b.capture, hipStreamNonBlocking));
  FSMC_HIP(hipStreamBeginCapture(b.capture, hipStreamCaptureModeThreadLocal));
  const int st = b.mode.precision == FSMC_FP64 ? batch_enqueue<double>(b, b.capture, true)
                                               : batch_enqueue<float>(b, b.capture, true);
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(b.capture, &g);
  if (st != FSMC_OK || ce != hipSuccess || g == nullptr) {
    if (g) (void)hipGraphDestroy(g);
    if (st != FSMC_OK) return st;
    set_error("graph capture of the batch failed: %s", hipGetErrorString(ce));
    return FSMC_ERR_DEVICE;
  }
  b.graph = g;
  FSMC_HIP(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
  return FSMC_OK;
}

// Builds the item tables and workspace for a tuning (caller holds ctx->mu).
static int batch_configure(fsmc_ctx* ctx, BatchImpl& b, fsmc_batch_tuning t) {
  FSMC_REQUIRE(t.chunk_paths >= 4 * kBlock && t.chunk_paths % (4 * kBlock) == 0,
               "chunk_paths %d must be a positive multiple of %d", t.chunk_paths, 4 * kBlock);
  FSMC_REQUIRE(t.requests_per_launch >= 0, "requests_per_launch %d < 0", t.requests_per_launch);
  FSMC_REQUIRE(t.replay_graph == 0 || t.replay_graph == 1, "replay_graph %d is not 0 or 1", t.replay_graph);
  FSMC_REQUIRE(t.workgroups_per_cu >= 0 && t.workgroups_per_cu <= 16, "workgroups_per_cu %d outside [0, 16]",
               t.workgroups_per_cu);
  FSMC_REQUIRE(t.work_queue == 0 || (t.work_queue == 1 && t.workgroups_per_cu >= 1),
               "work_queue %d must be 0, or 1 with workgroups_per_cu >= 1", t.work_queue);
  FSMC_HIP(hipSetDevice(ctx->device));
  FSMC_HIP(hipStreamSynchronize(ctx->stream));
  batch_release_events(b);
  b.tuning = t;
  const int per = t.requests_per_launch > 0 ? std::min(t.requests_per_launch, b.n) : b.n;
  std::vector<ReqDev> rd(b.n);
  std::vector<int2> items;
  std::vector<int> first(b.n), nitems(b.n);
  b.waves.clear();
  for (int r0 = 0; r0 < b.n; r0 += per) {
    MicroBatch mb;
    mb.r0 = r0;
    mb.n = std::min(per, b.n - r0);
    for (int pass = 0; pass < 2; ++pass) {  // GBM items, then Heston items
      const int kind = pass == 0 ? FSMC_REQ_GBM_EUROPEAN : FSMC_REQ_HESTON_ASIAN;
      const int begin = (int)items.size();
      // longest requests first (the dispatcher hands out workgroups in
      // index order, so the longest items start first and the tail is short);
      // a request's own items stay contiguous and in chunk order
      std::vector<int> order;
      for (int i = r0; i < r0 + mb.n; ++i)
        if (b.req[i].kind == kind) order.push_back(i);
      std::stable_sort(order.begin(), order.end(),
                       [&](int x, int y) { return b.req[x].n_steps > b.req[y].n_steps; });
      for (int i : order) {
        first[i] = (int)items.size();
        nitems[i] = (b.req[i].n_paths + t.chunk_paths - 1) / t.chunk_paths;
        for (int c = 0; c < nitems[i]; ++c) items.push_back(make_int2(i, c));
      }
      if (pass == 0) {
        mb.gbm_first = begin;
        mb.gbm_count = (int)items.size() - begin;
      } else {
        mb.hes_first = begin;
        mb.hes_count = (int)items.size() - begin;
      }
    }
    b.waves.push_back(mb);
  }
  // Persistent mode: per micro-batch, items longest first (cost = paths x
  // work per path) onto the least-loaded of G = cus x workgroups_per_cu
  // workgroup lists; the lists are concatenated into sched, with G + 1
  // offsets per micro-batch in soff.
  const int G = t.workgroups_per_cu > 0 ? b.cus * t.workgroups_per_cu : 0;
  std::vector<int> sched, soff, qorder;
  if (G > 0) {
    for (MicroBatch& mb : b.waves) {
      std::vector<std::pair<double, int>> cost;
      for (int it = mb.gbm_first; it < mb.hes_first + mb.hes_count; ++it) {
        const fsmc_request& r = b.req[items[it].x];
        const int paths = std::min(t.chunk_paths, r.n_paths - items[it].y * t.chunk_paths);
        // one GBM path ~ 35 instructions, one Heston adjoint step ~ 110
        cost.push_back({(double)paths * (r.kind == FSMC_REQ_GBM_EUROPEAN ? 1.0 : 1.7 * r.n_steps), it});
      }
      std::stable_sort(cost.begin(), cost.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
      // the work queues hand the items out in this order, one queue per kind:
      // the Heston items longest first, then the GBM items longest first
      mb.q_first = (int)qorder.size();
      mb.q_count = (int)cost.size();
      for (const auto& c : cost)
        if (b.req[items[c.second].x].kind != FSMC_REQ_GBM_EUROPEAN) qorder.push_back(c.second);
      mb.q_hes = (int)qorder.size() - mb.q_first;
      for (const auto& c : cost)
        if (b.req[items[c.second].x].kind == FSMC_REQ_GBM_EUROPEAN) qorder.push_back(c.second);
      {
        // the order is a permutation of the micro-batch's items (checked
        // here, so the kernel's reads order[j], j < q_count, stay in bounds)
        std::vector<char> seen(cost.size(), 0);
        for (int q = 0; q < mb.q_count; ++q) {
          const int it = qorder[mb.q_first + q] - mb.gbm_first;
          FSMC_REQUIRE(it >= 0 && it < mb.q_count && !seen[it], "internal: work-queue order is not a permutation");
          seen[it] = 1;
        }
      }
      std::vector<std::vector<int>> lists(G);
      std::vector<std::pair<double, int>> load;  // min-heap of (load, workgroup)
      for (int g = 0; g < G; ++g) load.push_back({0.0, g});
      auto cmp = [](const auto& x, const auto& y) { return x.first > y.first || (x.first == y.first && x.second > y.second); };
      std::make_heap(load.begin(), load.end(), cmp);
      for (const auto& c : cost) {
        std::pop_heap(load.begin(), load.end(), cmp);
        auto& least = load.back();
        lists[least.second].push_back(c.second);
        least.first += c.first;
        std::push_heap(load.begin(), load.end(), cmp);
      }
      mb.off_first = (int)soff.size();
      for (int g = 0; g < G; ++g) {
        soff.push_back((int)sched.size());
        sched.insert(sched.end(), lists[g].begin(), lists[g].end());
      }
      soff.push_back((int)sched.size());
    }
  }
  for (int i = 0; i < b.n; ++i) {
    const fsmc_request& r = b.req[i];
    ReqDev& d = rd[i];
    d.kind = r.kind;
    d.n_paths = r.n_paths;
    d.n_steps = r.n_steps;
    d.pad = 0;
    d.S0 = r.S0;
    d.K = r.K;
    d.T = r.T;
    d.r = r.r;
    d.sigma = r.sigma;
    d.v0 = r.v0;
    d.kappa = r.kappa;
    d.theta = r.theta;
    d.xi = r.xi;
    d.rho = r.rho;
  }
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = align_up(off + bytes, 256);
    return at;
  };
  b.off_req = take(sizeof(ReqDev) * b.n);
  b.off_items = take(sizeof(int2) * items.size());
  b.off_first = take(sizeof(int) * b.n);
  b.off_nitems = take(sizeof(int) * b.n);
  b.off_part = take(sizeof(double) * kNV * items.size());
  b.off_res = take(sizeof(double) * 8 * b.n);
  b.off_stamp = take(sizeof(unsigned long long) * (b.waves.size() + 1));
  b.off_sched = take(sizeof(int) * std::max<size_t>(1, sched.size()));
  b.off_soff = take(sizeof(int) * std::max<size_t>(1, soff.size()));
  b.off_qorder = take(sizeof(int) * std::max<size_t>(1, qorder.size()));
  b.off_heads = take(sizeof(int) * 2 * b.waves.size());  // a Heston and a GBM queue per micro-batch
  FSMC_TRY(b.ws.reserve(off));
  char* ws = (char*)b.ws.ptr;
  hipStream_t s = ctx->stream;
  FSMC_HIP(hipMemcpyAsync(ws + b.off_req, rd.data(), sizeof(ReqDev) * b.n, hipMemcpyHostToDevice, s));
  FSMC_HIP(hipMemcpyAsync(ws + b.off_items, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice, s));
  FSMC_HIP(hipMemcpyAsync(ws + b.off_first, first.data(), sizeof(int) * b.n, hipMemcpyHostToDevice, s));
  FSMC_HIP(hipMemcpyAsync(ws + b.off_nitems, nitems.data(), sizeof(int) * b.n, hipMemcpyHostToDevice, s));
  if (G > 0) {
    FSMC_HIP(hipMemcpyAsync(ws + b.off_sched, sched.data(), sizeof(int) * sched.size(), hipMemcpyHostToDevice, s));
    FSMC_HIP(hipMemcpyAsync(ws + b.off_soff, soff.data(), sizeof(int) * soff.size(), hipMemcpyHostToDevice, s));
    FSMC_HIP(hipMemcpyAsync(ws + b.off_qorder, qorder.data(), sizeof(int) * qorder.size(), hipMemcpyHostToDevice,
                            s));
  }
  FSMC_HIP(hipStreamSynchronize(s));
  if (t.replay_graph) {
    FSMC_TRY(batch_capture(b));
  } else {
    b.done.resize(b.waves.size(), nullptr);
    for (auto& e : b.done) FSMC_HIP(hipEventCreate(&e));
  }
  return FSMC_OK;
}

template <typename Real>
static int batch_enqueue(BatchImpl& b, hipStream_t s, bool stamps) {
  char* ws = (char*)b.ws.ptr;
  BatchArgs a;
  a.req = (const ReqDev*)(ws + b.off_req);
  a.items = (const int2*)(ws + b.off_items);
  a.partials = (double*)(ws + b.off_part);
  a.chunk = b.tuning.chunk_paths;
  a.k0 = (uint32_t)b.mode.seed;
  a.k1 = (uint32_t)(b.mode.seed >> 32);
  a.w_gbm = counter_w(kTagGbm, b.mode.stream);
  a.w_heston = counter_w(kTagHeston, b.mode.stream);
  unsigned long long* stamp = (unsigned long long*)(ws + b.off_stamp);
  int* heads = (int*)(ws + b.off_heads);
  if (b.tuning.work_queue) {  // two counters per micro-batch, zeroed by a kernel (no memset node in a graph)
    const int nw = 2 * (int)b.waves.size();
    hipLaunchKernelGGL(batch_reset_kernel, dim3((nw + 255) / 256), dim3(256), 0, s, heads, nw);
  }
  if (stamps)
    hipLaunchKernelGGL(batch_stamp_kernel, dim3(1), dim3(64), 0, s, stamp, 0);
  else
    FSMC_HIP(hipEventRecord(b.start, s));
  for (size_t w = 0; w < b.waves.size(); ++w) {
    const MicroBatch& mb = b.waves[w];
    if (b.tuning.workgroups_per_cu > 0 && b.tuning.work_queue) {
      const int G = b.cus * b.tuning.workgroups_per_cu;
      const int* q = (const int*)(ws + b.off_qorder) + mb.q_first;
      if (mb.q_hes > 0)
        hipLaunchKernelGGL((batch_queue_kernel<Real, true>), dim3(G), dim3(kBlock), 0, s, a, q, mb.q_hes,
                           heads + 2 * w);
      if (mb.q_count > mb.q_hes)
        hipLaunchKernelGGL((batch_queue_kernel<Real, false>), dim3(G), dim3(kBlock), 0, s, a, q + mb.q_hes,
                           mb.q_count - mb.q_hes, heads + 2 * w + 1);
    } else if (b.tuning.workgroups_per_cu > 0) {
      const int G = b.cus * b.tuning.workgroups_per_cu;
      hipLaunchKernelGGL(batch_persistent_kernel<Real>, dim3(G), dim3(kBlock), 0, s, a,
                         (const int*)(ws + b.off_sched), (const int*)(ws + b.off_soff) + mb.off_first);
    } else {
      if (mb.gbm_count > 0)
        hipLaunchKernelGGL(batch_gbm_kernel<Real>, dim3(mb.gbm_count), dim3(kBlock), 0, s, a, mb.gbm_first);
      if (mb.hes_count > 0)
        hipLaunchKernelGGL(batch_heston_kernel<Real>, dim3(mb.hes_count), dim3(kBlock), 0, s, a, mb.hes_first);
    }
    hipLaunchKernelGGL(batch_finalize_kernel, dim3((mb.n + 63) / 64), dim3(64), 0, s, a.req,
                       (const int*)(ws + b.off_first), (const int*)(ws + b.off_nitems), a.partials, mb.r0, mb.n,
                       (double*)(ws + b.off_res));
    if (stamps)
      hipLaunchKernelGGL(batch_stamp_kernel, dim3(1), dim3(64), 0, s, stamp, (int)w + 1);
    FSMC_HIP(hipGetLastError());
    if (!stamps) FSMC_HIP(hipEventRecord(b.done[w], s));
  }
  return FSMC_OK;
}

}  // namespace fsmc

using namespace fsmc;

struct fsmc_batch {
  fsmc_ctx* ctx;
  BatchImpl impl;
};

extern "C" {

int fsmc_batch_create(fsmc_ctx* ctx, const fsmc_request* req, int n, const fsmc_mode* m,
                      const fsmc_batch_tuning* tuning, fsmc_batch** out) {
  FSMC_REQUIRE(ctx && req && out, "NULL argument");
  *out = nullptr;
  FSMC_REQUIRE(n >= 1, "empty batch");
  FSMC_TRY(validate_mode(m, false));
  FSMC_REQUIRE(m->quirks == 0, "quirks are not supported in batches");
  for (int i = 0; i < n; ++i) {
    const fsmc_request& r = req[i];
    FSMC_REQUIRE(r.kind == FSMC_REQ_GBM_EUROPEAN || r.kind == FSMC_REQ_HESTON_ASIAN, "request %d: bad kind %d", i,
                 r.kind);
    FSMC_REQUIRE(r.n_paths >= 2 && r.n_paths <= (1 << 30), "request %d: n_paths %d outside [2, 2^30]", i,
                 r.n_paths);
    FSMC_REQUIRE(r.S0 > 0 && r.K >= 0 && r.T > 0, "request %d: S0, T must be > 0 and K >= 0", i);
    if (r.kind == FSMC_REQ_GBM_EUROPEAN) {
      FSMC_REQUIRE(r.sigma > 0, "request %d: sigma must be > 0", i);
    } else {
      FSMC_REQUIRE(r.n_steps >= 1 && r.v0 > 0 && r.rho >= -1 && r.rho <= 1,
                   "request %d: Heston needs n_steps >= 1, v0 > 0, |rho| <= 1", i);
    }
  }
  fsmc_batch* b = new fsmc_batch();
  b->ctx = ctx;
  BatchImpl& bi = b->impl;
  bi.n = n;
  bi.mode = *m;
  bi.req.assign(req, req + n);
  fsmc_batch_tuning t = {4096, 0, 0, 0, 0};
  if (tuning) {
    t = *tuning;
  } else {
    std::lock_guard<std::mutex> cl(g_cache_mu);
    auto it = tuning_cache().find(batch_key(bi));
    if (it != tuning_cache().end()) t = it->second;
  }
  int st;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    int khz = 0;
    st = hipSetDevice(ctx->device) == hipSuccess &&
                 hipDeviceGetAttribute(&bi.cus, hipDeviceAttributeMultiprocessorCount, ctx->device) == hipSuccess &&
                 bi.cus > 0 && hipEventCreate(&bi.start) == hipSuccess &&
                 hipEventCreate(&bi.t0) == hipSuccess && hipEventCreate(&bi.t1) == hipSuccess &&
                 hipEventCreateWithFlags(&bi.ran, hipEventDisableTiming) == hipSuccess &&
                 hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device) == hipSuccess && khz > 0
             ? FSMC_OK
             : FSMC_ERR_DEVICE;
    if (st != FSMC_OK) set_error("event creation or wall-clock rate query failed");
    bi.wall_khz = khz;
    if (st == FSMC_OK) st = batch_configure(ctx, bi, t);
  }
  if (st != FSMC_OK) {
    fsmc_batch_destroy(b);
    return st;
  }
  *out = b;
  return FSMC_OK;
}

int fsmc_batch_run(fsmc_batch* b, void* stream) {
  FSMC_REQUIRE(b != nullptr, "batch is NULL");
  FSMC_HIP(hipSetDevice(b->ctx->device));
  hipStream_t s = stream ? (hipStream_t)stream : b->ctx->stream;
  if (b->impl.exec) {
    FSMC_HIP(hipGraphLaunch(b->impl.exec, s));
  } else {
    FSMC_TRY(b->impl.mode.precision == FSMC_FP64 ? batch_enqueue<double>(b->impl, s, false)
                                                 : batch_enqueue<float>(b->impl, s, false));
  }
  // fsmc_batch_results may be called with another stream: it waits on this.
  FSMC_HIP(hipEventRecord(b->impl.ran, s));
  return FSMC_OK;
}

int fsmc_batch_results(fsmc_batch* b, void* stream, fsmc_result* out) {
  FSMC_REQUIRE(b && out, "NULL argument");
  BatchImpl& bi = b->impl;
  hipStream_t s = stream ? (hipStream_t)stream : b->ctx->stream;
  std::vector<double> res((size_t)bi.n * 8);
  FSMC_HIP(hipSetDevice(b->ctx->device));
  FSMC_HIP(hipStreamWaitEvent(s, bi.ran, 0));  // the last run, whatever stream it was on
  FSMC_HIP(hipMemcpyAsync(res.data(), (char*)bi.ws.ptr + bi.off_res, res.size() * sizeof(double),
                          hipMemcpyDeviceToHost, s));
  std::vector<unsigned long long> stamp(bi.exec ? bi.waves.size() + 1 : 0);
  if (bi.exec)
    FSMC_HIP(hipMemcpyAsync(stamp.data(), (char*)bi.ws.ptr + bi.off_stamp, stamp.size() * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, s));
  FSMC_HIP(hipStreamSynchronize(s));
  for (size_t w = 0; w < bi.waves.size(); ++w) {
    double us = 0.0;
    if (bi.exec) {
      us = (double)(stamp[w + 1] - stamp[0]) / bi.wall_khz * 1e3;
    } else {
      float ms = 0.0f;
      FSMC_HIP(hipEventElapsedTime(&ms, bi.start, bi.done[w]));
      us = 1e3 * ms;
    }
    const MicroBatch& mb = bi.waves[w];
    for (int i = mb.r0; i < mb.r0 + mb.n; ++i) out[i].latency_us = us;
  }
  for (int i = 0; i < bi.n; ++i) {
    fsmc_result& r = out[i];
    const double* v = res.data() + (size_t)i * 8;
    r.price = v[0];
    r.delta = v[1];
    r.vega = v[2];
    r.rho = v[3];
    r.price_se = v[4];
    r.delta_se = v[5];
    r.vega_se = v[6];
    r.rho_se = v[7];
  }
  return FSMC_OK;
}

}  // extern "C"

// Configures tuning t and returns the best of two timed runs (after one
// untimed run) from the batch's start event to its last completion event.
static int batch_time(fsmc_batch* b, fsmc_batch_tuning t, float* best) {
  BatchImpl& bi = b->impl;
  {
    std::lock_guard<std::mutex> lock(b->ctx->mu);
    FSMC_TRY(batch_configure(b->ctx, bi, t));
  }
  *best = 1e30f;
  hipStream_t s = b->ctx->stream;
  for (int rep = 0; rep < 3; ++rep) {
    FSMC_HIP(hipEventRecord(bi.t0, s));
    FSMC_TRY(fsmc_batch_run(b, nullptr));
    FSMC_HIP(hipEventRecord(bi.t1, s));
    FSMC_HIP(hipEventSynchronize(bi.t1));
    float ms = 0.0f;
    FSMC_HIP(hipEventElapsedTime(&ms, bi.t0, bi.t1));
    if (rep > 0) *best = std::min(*best, ms);  // first run untimed
  }
  return FSMC_OK;
}

extern "C" {

// The search of both autotune entry points over the micro-batch sizes pers
// (0 = the whole batch in one launch); cache: remember the winner for
// batches of this shape (the throughput search only, so that a later
// fsmc_batch_create with no tuning gets the throughput choice).
static int batch_autotune(fsmc_batch* b, std::initializer_list<int> pers, bool cache, fsmc_batch_tuning* chosen) {
  BatchImpl& bi = b->impl;
  fsmc_ctx* ctx = b->ctx;
  const int chunks[] = {1024, 2048, 4096, 8192, 16384};
  fsmc_batch_tuning best = bi.tuning;
  float best_ms = 1e30f;
  int st = FSMC_OK;
  for (int c : chunks) {
    for (int per : pers) {
      if (pers.size() > 1 && per != 0 && per >= bi.n) continue;  // the same as one launch
      float ms = 0.0f;
      st = batch_time(b, fsmc_batch_tuning{c, per, 0, 0, 0}, &ms);
      if (st != FSMC_OK) break;
      if (ms < best_ms) {
        best_ms = ms;
        best = fsmc_batch_tuning{c, per, 0, 0, 0};
      }
    }
    if (st != FSMC_OK) break;
  }
  if (st != FSMC_OK) return st;
  const int chunk = best.chunk_paths, per = best.requests_per_launch;
  for (int queue : {0, 1}) {
    for (int wpc : {1, 2, 3, 4, 6, 8}) {  // the persistent kernels at this chunk and micro-batch size
      float ms = 0.0f;
      const fsmc_batch_tuning p = {chunk, per, 0, wpc, queue};
      FSMC_TRY(batch_time(b, p, &ms));
      if (ms < best_ms) {
        best_ms = ms;
        best = p;
      }
    }
  }
  {
    float ms = 0.0f;
    const fsmc_batch_tuning g = {best.chunk_paths, best.requests_per_launch, 1, best.workgroups_per_cu,
                                 best.work_queue};
    FSMC_TRY(batch_time(b, g, &ms));
    if (ms < best_ms) best = g;
  }
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    FSMC_TRY(batch_configure(ctx, bi, best));
  }
  if (cache) {
    std::lock_guard<std::mutex> cl(g_cache_mu);
    tuning_cache()[batch_key(bi)] = best;
  }
  if (chosen) *chosen = best;
  return FSMC_OK;
}

int fsmc_batch_autotune(fsmc_batch* b, fsmc_batch_tuning* chosen) {
  FSMC_REQUIRE(b != nullptr, "batch is NULL");
  return batch_autotune(b, {0, 1024, 256}, true, chosen);
}

int fsmc_batch_autotune_microbatch(fsmc_batch* b, int requests_per_launch, fsmc_batch_tuning* chosen) {
  FSMC_REQUIRE(b != nullptr, "batch is NULL");
  FSMC_REQUIRE(requests_per_launch > 0, "requests_per_launch must be positive");
  return batch_autotune(b, {requests_per_launch}, false, chosen);
}

int fsmc_batch_tuning_get(const fsmc_batch* b, fsmc_batch_tuning* out) {
  FSMC_REQUIRE(b && out, "NULL argument");
  *out = b->impl.tuning;
  return FSMC_OK;
}

int fsmc_batch_destroy(fsmc_batch* b) {
  if (!b) return FSMC_OK;
  (void)hipSetDevice(b->ctx->device);
  // runs may sit on any caller stream: wait for all work before the graph,
  // events and workspace go
  (void)hipDeviceSynchronize();
  batch_release_events(b->impl);
  if (b->impl.start) (void)hipEventDestroy(b->impl.start);
  if (b->impl.t0) (void)hipEventDestroy(b->impl.t0);
  if (b->impl.t1) (void)hipEventDestroy(b->impl.t1);
  if (b->impl.ran) (void)hipEventDestroy(b->impl.ran);
  if (b->impl.capture) (void)hipStreamDestroy(b->impl.capture);
  b->impl.ws.release();
  delete b;
  return FSMC_OK;
}

}  // extern "C"
