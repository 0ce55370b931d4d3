// Host orchestration behind the C ABI of include/cvmatch_hip.h.
//
// One engine per process owns the HIP stream, the per-size twiddle tables
// (generated on the device), grow-only scratch buffers and the profiling
// events.  Every pipeline runs stream-ordered with no host round trip inside
// a call: the channel count (Match's alpha rule, cvmatch.go:82-85) and the
// flat-template short cut (impl.go:1137) are decided on the device and read
// by the kernels from a CallState.  The host only plans the tile geometry,
// which is pure integer arithmetic (newGoPlan impl.go:246-295).
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cvmatch_hip.h"
#include "cvm_internal.h"
#include "cvm_score.h"

namespace cvm {

// ---------------------------------------------------------------- errors --

static thread_local char g_err[512];

static int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(x)                                                                                      \
  do {                                                                                                 \
    hipError_t e_ = (x);                                                                               \
    if (e_ != hipSuccess)                                                                              \
      return fail(CVM_E_HIP, "cvmatch: HIP error '%s' at %s:%d", hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

#define CHK(x)          \
  do {                  \
    int rc_ = (x);      \
    if (rc_) return rc_; \
  } while (0)

// ------------------------------------------------------------------ plan --

struct Plan {
  int dftW, dftH, hw, blockW, blockH;
  long long cost;
};

static int next_pow2(int v) {
  int n = 1;
  while (n < v) n <<= 1;
  return n;
}
static int bit_len(unsigned long long n) {
  int l = 0;
  while (n) {
    l++;
    n >>= 1;
  }
  return l;
}
static int ilog2(int n) { return bit_len((unsigned)n) - 1; }

// newGoPlan impl.go:246-295: integer cost-model argmin over power-of-two
// transform sizes (areaCap 2^22), ties to the smaller area.
static Plan make_plan_search(int tw, int th, int rw, int rh) {
  const long long areaCap = 1LL << 22;
  int minW = std::max(2, next_pow2(tw)), minH = std::max(2, next_pow2(th));
  int maxW = std::max(minW, next_pow2(rw + tw - 1)), maxH = std::max(minH, next_pow2(rh + th - 1));
  int bestW = minW, bestH = minH;
  long long bestCost = -1, bestArea = 0;
  for (int dftW = minW; dftW <= maxW; dftW <<= 1) {
    for (int dftH = minH; dftH <= maxH; dftH <<= 1) {
      long long area = (long long)dftW * dftH;
      if (area > areaCap) continue;
      int blockW = std::min(dftW - tw + 1, rw), blockH = std::min(dftH - th + 1, rh);
      long long ntx = (rw + blockW - 1) / blockW;
      long long hw = dftW / 2 + 1;
      long long rowFFT = 5LL * dftW * ilog2(dftW);
      long long colPass = 5LL * dftH * ilog2(dftH) * hw;
      long long specN = (long long)dftH * hw;
      long long cost = 0;
      for (int y0 = 0; y0 < rh; y0 += blockH) {
        int bh = std::min(blockH, rh - y0);
        int loadH = std::min(bh + th - 1, dftH);
        cost += ntx * ((long long)((loadH + 1) / 2 + (bh + 1) / 2) * rowFFT + 2 * colPass + 16 * specN);
      }
      if (bestCost < 0 || cost < bestCost || (cost == bestCost && area < bestArea)) {
        bestCost = cost;
        bestArea = area;
        bestW = dftW;
        bestH = dftH;
      }
    }
  }
  Plan p;
  p.dftW = bestW;
  p.dftH = bestH;
  p.hw = bestW / 2 + 1;
  p.blockW = std::min(bestW - tw + 1, rw);
  p.blockH = std::min(bestH - th + 1, rh);
  p.cost = bestCost;
  return p;
}

// The plan of a shape, searched once per process: the search walks every
// size pair and, for each, every band of the map (a thousand bands for a
// 1080p map at the smallest heights), tens of microseconds of host time
// that a repeated call would otherwise spend before its first launch.
static Plan make_plan(int tw, int th, int rw, int rh) {
  static std::mutex mu;
  static std::map<std::array<int, 4>, Plan> memo;
  const std::array<int, 4> key = {tw, th, rw, rh};
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = memo.find(key);
    if (it != memo.end()) return it->second;
  }
  const Plan p = make_plan_search(tw, th, rw, rh);
  std::lock_guard<std::mutex> lk(mu);
  if (memo.size() >= 4096) memo.clear();
  memo.emplace(key, p);
  return p;
}

// shrinkDH impl.go:573-581
static int shrink_dh(const Plan& p, int rh, int th) {
  int lastY0 = ((rh - 1) / p.blockH) * p.blockH;
  if (lastY0 > 0) {
    int dh2 = std::max(2, next_pow2(rh - lastY0 + th - 1));
    if (dh2 < p.dftH) return dh2;
  }
  return 0;
}

static int64_t stats_cap(int cn) {
  static const int64_t caps[5] = {0, 11909805, 8421504, 6876129, 5954902};
  return (cn >= 1 && cn <= 4) ? caps[cn] : 0;
}

// Kernel limits: row transforms keep two dftW rows in LDS and
// column slabs dftH x CG complex, both <= 64 KiB.  Templates are at most
// dftW wide, which the normalisation's strips take (up to 256 + 4096 - 1
// columns).
constexpr int kMaxFftW = 4096;
constexpr int kMaxFftH = 8192;

static int round_up(int v, int m) { return (v + m - 1) / m * m; }

// --------------------------------------------------------------- engine --

// Bumped on every device reallocation: captured call graphs hold raw
// pointers and are stale once any buffer moved.
static std::atomic<unsigned long long> g_alloc_epoch{0};

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t n) {
    if (n <= cap) return 0;
    g_alloc_epoch.fetch_add(1);
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max(n, (size_t)256);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(CVM_E_HIP, "cvmatch: device allocation of %zu bytes failed: %s", want, hipGetErrorString(e));
    }
    cap = want;
    return 0;
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

struct Res6 {
  float min_v;
  int32_t min_x, min_y;
  float max_v;
  int32_t max_x, max_y;
};

struct ProfStage {
  std::string name;
  hipEvent_t a, b;
};

// Process-wide state: the per-size twiddle tables (generated on the device
// once, read-only afterwards), the pool of call contexts and the stage
// timings.  Nothing here is touched by a call's kernels except the tables.
struct Ctx;
struct Engine {
  std::mutex mu;  // ready, tables, the context pool
  bool ready = false;
  hipStream_t tstream = nullptr;  // table generation
  std::map<int, FftTables> tables;
  std::vector<Ctx*> idle;  // most recently released last
  int created = 0;
  std::condition_variable cv;
  std::atomic<bool> prof{false};
  std::mutex pmu;
  std::map<std::string, std::pair<double, long long>> prof_acc;

  int init() {
    if (ready) return 0;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev < 1) return fail(CVM_E_HIP, "cvmatch: no HIP device available");
    HIPCHK(hipStreamCreateWithFlags(&tstream, hipStreamNonBlocking));
    ready = true;
    return 0;
  }

  int tab(int n, const FftTables** out) {
    std::lock_guard<std::mutex> lk(mu);
    CHK(init());
    auto it = tables.find(n);
    if (it == tables.end()) {
      FftTables t;
      t.n = n;
      int len = tri_len(n);
      HIPCHK(hipMalloc(&t.tri, sizeof(float2) * (size_t)std::max(len, 1)));
      HIPCHK(hipMalloc(&t.brev, sizeof(int32_t) * (size_t)n));
      launch_twiddles(t.tri, t.brev, n, tstream);
      HIPCHK(hipGetLastError());
      // tables are shared by every stream a caller may pass in
      HIPCHK(hipStreamSynchronize(tstream));
      it = tables.emplace(n, t).first;
    }
    *out = &it->second;
    return 0;
  }

  int acquire(Ctx** out);
  void release(Ctx* x);
  void add_times(const std::vector<std::pair<std::string, float>>& t) {
    std::lock_guard<std::mutex> lk(pmu);
    for (auto& p : t) {
      auto& acc = prof_acc[p.first];
      acc.first += p.second;
      acc.second += 1;
    }
  }
};

static Engine& engine() {
  static Engine* e = new Engine();
  return *e;
}

// Captured calls: up to kGraphKeys argument keys (a caller that cycles
// through a few frame buffers keeps each one's graph), each with its
// sightings and, from the second one on, its executable graph; the least
// recently used key goes first.
constexpr int kGraphKeys = 4;
struct GraphSlot {
  struct Entry {
    std::vector<uint64_t> key;
    int seen = 0;
    hipGraphExec_t exec = nullptr;
    unsigned long long used = 0;
    // a synchronous call site's choice between replaying the graph and
    // launching eagerly (GraphMode), from the wall time of a few calls of
    // each once the graph exists
    int mode = 0;
    int last = -1;  // how the last call ran: 0 graph, 1 eager, -1 neither timed
    double t[2] = {0.0, 0.0};
    int n[2] = {0, 0};
  };
  std::vector<Entry> e;
  unsigned long long tick = 0;
  Entry& find(const std::vector<uint64_t>& key) {
    for (auto& x : e)
      if (x.key == key) {
        x.used = ++tick;
        return x;
      }
    Entry* slot;
    if ((int)e.size() < kGraphKeys) {
      e.emplace_back();
      slot = &e.back();
    } else {
      slot = &e[0];
      for (auto& x : e)
        if (x.used < slot->used) slot = &x;
      if (slot->exec) (void)hipGraphExecDestroy(slot->exec);
      *slot = Entry();
    }
    slot->key = key;
    slot->used = ++tick;
    return *slot;
  }
  void drop() {
    for (auto& x : e)
      if (x.exec) (void)hipGraphExecDestroy(x.exec);
    e.clear();
  }
};

// One call context: the streams, events and grow-only scratch of one call in
// flight.  A call takes a context from the engine's pool for its duration
// (calls return with their work complete, so the next user of the context
// finds its buffers free); a frame stream owns one of its own, since its
// frames stay in flight on it between calls.  Concurrent calls therefore
// share no scratch and overlap on the device.
struct Ctx {
  hipStream_t stream = nullptr;
  // side streams of a MatchMap call, forked from the caller's stream: side
  // builds the call state (joined back before the column pass) and then the
  // normalisation partials (joined before the normalisation); side2 builds
  // the template spectrum meanwhile, which needs no call state (joined
  // before the column pass)
  hipStream_t side = nullptr, side2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, join2 = nullptr, join3 = nullptr;
  // a host-buffer call's parent upload, overlapped with the call's own
  // transforms on the stream (run_match_overlap)
  hipStream_t cstream = nullptr;
  // the captured device-resident call (run_graphed)
  GraphSlot graph;
  DevBuf img, tpl, spec, tspec, tspec2, corr, parts, gray0, gray1, mapbuf, nscratch, dbf, pkscr, pkout;
  DevBuf gate;  // the mixed form's two trial maps
  CallState* st = nullptr;
  // the compact form's call state: its accumulators are zero between
  // compact calls (zeroed here, and cleared again by each call's final
  // reduction once nothing reads them), so a call needs no reset launch
  CallState* cst = nullptr;
  Res6* res_d = nullptr;
  Res6* res_h = nullptr;
  // a single-result call's six-tuple (Match family, Find) in coherent host
  // memory mapped for the device: the final reduction writes it there, so
  // no copy follows the call (null when the mapping is unavailable: res_d
  // and a copy instead)
  Res6* rmap_h = nullptr;
  Res6* rmap_d = nullptr;
  std::vector<hipEvent_t> evpool;
  std::vector<ProfStage> pending;
  // a context that never records stage events (the allocation-free call,
  // which has no streams, events or buffers of its own)
  bool noprof = false;
  std::thread::id owner;  // the thread that acquired it last

  int init() {
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&join, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&join2, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&join3, hipEventDisableTiming));
    HIPCHK(hipMalloc(&st, sizeof(CallState)));
    HIPCHK(hipMalloc(&cst, sizeof(CallState)));
    HIPCHK(hipMemset(cst, 0, sizeof(CallState)));
    HIPCHK(hipMalloc(&res_d, sizeof(Res6) * 2));
    HIPCHK(hipHostMalloc(&res_h, sizeof(Res6) * 2, hipHostMallocDefault));
    if (hipHostMalloc((void**)&rmap_h, sizeof(Res6), hipHostMallocCoherent | hipHostMallocMapped) == hipSuccess) {
      void* dp = nullptr;
      if (hipHostGetDevicePointer(&dp, rmap_h, 0) == hipSuccess && dp) {
        rmap_d = (Res6*)dp;
      } else {
        (void)hipHostFree(rmap_h);
        rmap_h = nullptr;
      }
    }
    (void)hipGetLastError();
    return 0;
  }

  // Frees everything; the context's streams must be idle.
  void destroy() {
    graph.drop();
    for (DevBuf* b : {&img, &tpl, &spec, &tspec, &tspec2, &corr, &parts, &gray0, &gray1, &mapbuf, &nscratch, &dbf,
                      &pkscr, &pkout, &gate})
      b->release();
    if (st) (void)hipFree(st);
    if (cst) (void)hipFree(cst);
    if (res_d) (void)hipFree(res_d);
    if (res_h) (void)hipHostFree(res_h);
    if (rmap_h) (void)hipHostFree(rmap_h);
    for (hipStream_t q : {stream, side, side2, cstream})
      if (q) (void)xfer_release_stream(q);
    for (hipEvent_t e : evpool) (void)hipEventDestroy(e);
    for (hipEvent_t e : {fork, join, join2, join3})
      if (e) (void)hipEventDestroy(e);
    for (hipStream_t q : {stream, side, side2, cstream})
      if (q) (void)hipStreamDestroy(q);
  }

  // The side streams and the copy stream are created on first use (an eager
  // call, before any capture): every stream takes one of the device's few
  // hardware queues in turn, and a context that never forks should not
  // push another context's streams onto a shared queue.
  int ensure_side() {
    if (!side) HIPCHK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    if (!side2) HIPCHK(hipStreamCreateWithFlags(&side2, hipStreamNonBlocking));
    return 0;
  }
  int ensure_copy() {
    if (!cstream) HIPCHK(hipStreamCreateWithFlags(&cstream, hipStreamNonBlocking));
    return 0;
  }

  int tab(int n, const FftTables** out) { return engine().tab(n, out); }

  hipEvent_t ev() {
    if (!evpool.empty()) {
      hipEvent_t e = evpool.back();
      evpool.pop_back();
      return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }

  int stage_begin(const char* name, hipStream_t s) {
    if (noprof || !engine().prof.load()) return -1;
    ProfStage p;
    p.name = name;
    p.a = ev();
    p.b = ev();
    if (!p.a || !p.b) return -1;
    (void)hipEventRecord(p.a, s);
    pending.push_back(p);
    return (int)pending.size() - 1;
  }
  void stage_end(int i, hipStream_t s) {
    if (i < 0) return;
    (void)hipEventRecord(pending[i].b, s);
  }
  // after the stream has been synchronised
  void stage_collect() {
    if (pending.empty()) return;
    std::vector<std::pair<std::string, float>> t;
    for (auto& p : pending) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) t.emplace_back(p.name, ms);
      evpool.push_back(p.a);
      evpool.push_back(p.b);
    }
    pending.clear();
    engine().add_times(t);
  }
};

// Up to kMaxCtx contexts exist at once; a caller beyond that waits for one.
constexpr int kMaxCtx = 8;

int Engine::acquire(Ctx** out) {
  std::unique_lock<std::mutex> lk(mu);
  CHK(init());
  const std::thread::id me = std::this_thread::get_id();
  for (;;) {
    if (!idle.empty()) {
      // the context this thread used last (its buffers and captured call
      // are sized and keyed for this caller), else the most recently used
      size_t k = idle.size() - 1;
      for (size_t i = 0; i < idle.size(); i++)
        if (idle[i]->owner == me) k = i;
      *out = idle[k];
      idle.erase(idle.begin() + (long)k);
      (*out)->owner = me;
      return 0;
    }
    if (created < kMaxCtx) break;
    cv.wait(lk);
  }
  created++;
  lk.unlock();
  std::unique_ptr<Ctx> x(new Ctx());
  int rc = x->init();
  if (rc) {
    x->destroy();
    std::lock_guard<std::mutex> g(mu);
    created--;
    cv.notify_one();
    return rc;
  }
  x->owner = me;
  *out = x.release();
  return 0;
}

void Engine::release(Ctx* x) {
  {
    std::lock_guard<std::mutex> lk(mu);
    idle.push_back(x);
  }
  cv.notify_one();
}

// A pool context held for one call.
struct Lease {
  Ctx* x = nullptr;
  int rc = 0;
  Lease() { rc = engine().acquire(&x); }
  ~Lease() {
    if (x) engine().release(x);
  }
  Lease(const Lease&) = delete;
  Lease& operator=(const Lease&) = delete;
};

#define LEASE(name)      \
  Lease lease_;          \
  CHK(lease_.rc);        \
  Ctx& name = *lease_.x

struct Stage {
  Ctx& X;
  int i;
  hipStream_t s;
  Stage(Ctx& x, const char* name, hipStream_t st) : X(x), s(st) { i = X.stage_begin(name, st); }
  ~Stage() { X.stage_end(i, s); }
};

// Runs `run`, which enqueues one call's work on stream s, through a graph
// slot: eagerly the first time its key is seen (that run allocates whatever
// the call needs), captured into a hipGraph on the second consecutive
// sighting, and replayed from then on, which removes the host cost of every
// launch in the call.  The key must name everything the launches depend on
// (pointers, shapes, stream, context); any device reallocation since the
// capture changes the allocation epoch appended to it.  A call that cannot be
// captured (it allocates or synchronises) runs eagerly instead.  Nothing
// inside a call fills memory with hipMemsetAsync (fills are kernels).
// CVMATCH_GRAPHS=0 runs every call eagerly.
static bool graphs_enabled() {
  static const bool on = [] {
    const char* v = getenv("CVMATCH_GRAPHS");
    return !(v && v[0] == '0');
  }();
  return on;
}

// A synchronous call site (one that waits for its result right after) may
// pass `timed`: once its graph exists, its next calls alternate between the
// replay and eager launches, the caller reports each call's wall time
// (graph_note), and after kGraphTrials of each the faster way is kept for the
// key.  A replay saves the host work of every launch, but the host submits
// the graph's nodes one after another: a short first kernel can finish
// before the second node is in the queue, which eager launches of the same
// few kernels avoid.  CVMATCH_GRAPH_MODE=1 (graph) or 2 (eager) fixes the
// choice.
constexpr int kGraphTrials = 4;
static int graph_mode_env() {
  const char* v = getenv("CVMATCH_GRAPH_MODE");
  const int m = v ? atoi(v) : 0;
  return m == 1 || m == 2 ? m : 0;
}

static void graph_note(GraphSlot::Entry* g, double seconds) {
  if (!g || g->last < 0 || g->mode != 0) return;
  g->t[g->last] += seconds;
  if (++g->n[g->last] < kGraphTrials || g->n[1 - g->last] < kGraphTrials) return;
  g->mode = g->t[1] / g->n[1] < g->t[0] / g->n[0] ? 2 : 1;
}

template <class F>
static int run_graphed(GraphSlot& slot, std::vector<uint64_t> key, hipStream_t s, F&& run,
                       GraphSlot::Entry** timed = nullptr) {
  key.push_back(g_alloc_epoch.load());
  key.push_back((uint64_t)(uintptr_t)s);
  if (timed) *timed = nullptr;
  if (engine().prof.load() || !graphs_enabled()) {  // stage events need the eager launches
    slot.drop();
    return run();
  }
  GraphSlot::Entry& g = slot.find(key);
  g.last = -1;
  if (g.exec) {
    int mode = graph_mode_env();
    if (!mode) mode = timed ? g.mode : 1;
    if (!mode) {
      // trials: the way with fewer samples so far (graph first)
      g.last = g.n[1] < g.n[0] ? 1 : 0;
      *timed = &g;
      if (g.last == 1) return run();
      HIPCHK(hipGraphLaunch(g.exec, s));
      return 0;
    }
    if (mode == 2) return run();
    HIPCHK(hipGraphLaunch(g.exec, s));
    return 0;
  }
  if (++g.seen < 2) return run();
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int rc = run();
  const hipError_t ce = hipStreamEndCapture(s, &graph);
  const bool moved = g_alloc_epoch.load() != key[key.size() - 2];
  if (rc != 0 || ce != hipSuccess || moved) {
    if (graph) (void)hipGraphDestroy(graph);
    g.seen = 0;
    (void)hipGetLastError();
    return run();  // eagerly instead (and again with its own error, if any)
  }
  const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ie != hipSuccess) {
    g.exec = nullptr;
    g.seen = 0;
    (void)hipGetLastError();
    return run();
  }
  HIPCHK(hipGraphLaunch(g.exec, s));
  return 0;
}

static uint64_t u64p(const void* p) { return (uint64_t)(uintptr_t)p; }

// ------------------------------------------------------------ pipeline --

// Template side of one correlation on a plan: its spectrum (all cnMax
// channels) and the shrunk-band spectrum, device resident.
struct TplSpec {
  Plan p;
  int dh2 = 0;
  int hwPad = 0;
  int cnMax = 0;
  float2* spec = nullptr;
  float2* spec2 = nullptr;
};

static int plan_tables(Ctx& X, const Plan& p, int dh2, PlanTables& pt) {
  const FftTables *tw, *th, *th2 = nullptr;
  CHK(X.tab(p.dftW, &tw));
  CHK(X.tab(p.dftH, &th));
  if (dh2) CHK(X.tab(dh2, &th2));
  pt.triW = tw->tri;
  pt.brevW = tw->brev;
  pt.triH = th->tri;
  pt.brevH = th->brev;
  pt.triH2 = th2 ? th2->tri : th->tri;
  pt.brevH2 = th2 ? th2->brev : th->brev;
  return 0;
}

static int check_limits(const Plan& p, int tw, int rw, int rh) {
  (void)tw;
  if (p.dftW > kMaxFftW || p.dftH > kMaxFftH)
    return fail(CVM_E_UNSUPPORTED, "cvmatch: transform %dx%d exceeds the kernels' LDS limits (%d x %d)", p.dftW,
                p.dftH, kMaxFftW, kMaxFftH);
  if ((long long)rw * rh >= (1LL << 31))
    return fail(CVM_E_UNSUPPORTED, "cvmatch: result map of %dx%d exceeds 2^31 elements", rw, rh);
  return 0;
}

// The template as the single tile of plan p (its spectrum's grid).
static TileGrid tpl_grid(const Plan& p, int tw, int th) {
  TileGrid gt{};
  gt.ntx = gt.nty = 1;
  gt.blockW = tw;
  gt.blockH = th;
  gt.extW = tw;
  gt.extH = th;
  gt.padW = gt.padH = 0;
  gt.lastY0 = 0;
  gt.dftW = p.dftW;
  gt.log2W = ilog2(p.dftW);
  gt.hw = p.hw;
  gt.hwPad = round_up(p.hw, 16);
  gt.dftH = p.dftH;
  gt.log2H = ilog2(p.dftH);
  gt.dftH2 = 0;
  gt.log2H2 = 0;
  return gt;
}

// The parent's tiles on plan p (crossCorrGo impl.go:616-693), shrunk band dh2.
static TileGrid parent_grid(const Plan& p, int tw, int th, int rw, int rh, int hwPad, int dh2) {
  TileGrid g{};
  g.ntx = (rw + p.blockW - 1) / p.blockW;
  g.nty = (rh + p.blockH - 1) / p.blockH;
  g.blockW = p.blockW;
  g.blockH = p.blockH;
  g.extW = rw;
  g.extH = rh;
  g.padW = tw - 1;
  g.padH = th - 1;
  g.lastY0 = ((rh - 1) / p.blockH) * p.blockH;
  g.dftW = p.dftW;
  g.log2W = ilog2(p.dftW);
  g.hw = p.hw;
  g.hwPad = hwPad;
  g.dftH = p.dftH;
  g.log2H = ilog2(p.dftH);
  g.dftH2 = dh2;
  g.log2H2 = dh2 ? ilog2(dh2) : 0;
  return g;
}

// Template spectrum on plan p (buildTSpecSet impl.go:520-561): forward
// transform of the template at (0,0), scaled by 1/(dftW*dftH), plus the
// shrunk band (shrinkInto impl.go:591-610).  dst buffers sized by caller.
static int build_tspec(Ctx& X, hipStream_t s, const CallState* st, const uint8_t* d_tpl, int64_t tstride, int tw,
                       int th, int step, int cnMax, const Plan& p, int dh2, const PlanTables& pt, float2* spec,
                       float2* spec2) {
  Stage S(X, "tspec", s);
  int hwPad = round_up(p.hw, 16);
  int64_t tchan = (int64_t)p.dftH * hwPad, tchan2 = (int64_t)dh2 * hwPad;
  const TileGrid gt = tpl_grid(p, tw, th);
  launch_row_fwd(d_tpl, tstride, step, cnMax, st, gt, pt, spec, 0, tchan, 1, s);
  float scale = 1.0f / ((float)p.dftW * (float)p.dftH);
  launch_cols(COL_TSPEC, spec, spec, 0, 0, tchan, cnMax, st, gt, pt, nullptr, nullptr, 0, 0, scale, 1, s);
  if (dh2) {
    int shift = ilog2(p.dftH) - ilog2(dh2);
    launch_shrink(spec, spec2, cnMax, dh2, shift, hwPad, tchan, tchan2, s);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// Where a single-result call (Match family, Find) leaves its six-tuple: the
// mapped host result when there is one, else the device one.
static Res6* res_out(Ctx& X) { return X.rmap_d ? X.rmap_d : X.res_d; }

// Correlation + normalisation + extrema for one parent against a prepared
// template spectrum (crossCorrGo + normalizeParallelGo, impl.go:1150-1154).
// CallState must be finalised on the stream before this runs.
// With join set, the call state, the template spectrum and the
// normalisation partials are still being built on another stream that
// records join when done: the parent's row transforms (which need none of
// them) run first, then s waits for join.  Otherwise all of them are ready
// on s, except the partials, built here.
// side (may be null): enqueues the call's side-stream work on X.side and
// X.side2 and records X.join once the call state is ready, X.join3 once the
// template spectrum is and X.join2 once the normalisation partials are.  It
// runs after the parent's row pass has been submitted on s, so that pass
// (which needs none of it) is first in line and overlaps the side work.
// Partials of the widest range of strip forms a call's normalisation may take.
static int norm_partials_max(int rw, int rh, int tw, int th, int cnMax) {
  int n = norm_partials_count(rw, rh, tw, th, cnMax);
  for (int q : {4, 2, 1, 4 | NORM_HALF, 2 | NORM_HALF, 1 | NORM_HALF})
    if (norm_q_ok(tw, th, cnMax, q)) n = std::max(n, norm_partials_count(rw, rh, tw, th, cnMax, q));
  return n;
}

// Strip form of a call's wave-form normalisation, the strip width (result
// columns per lane) in full or half bands: each form that holds the template
// is timed on the call's own correlation -- extrema only, so the map is left
// alone -- the first time the shape comes up (a small map fills the chip only
// with narrow strips, and its waves are latency-bound row chains that half
// bands shorten).  0 (the default form) for the int64 form and for the
// allocation-free call.
static int norm_q_for(Ctx& X, hipStream_t s, const uint8_t* d_img, int64_t istride, int iw, int ih, int step, int tw,
                      int th, int cnMax, const CallState* st, const float* corr, Extrema* parts, void* nscratch,
                      bool exact) {
  if (X.noprof || !norm_wave_ok(tw, th, cnMax)) return 0;
  std::vector<int> cands;
  for (int q : {4, 2, 1, 4 | NORM_HALF, 2 | NORM_HALF, 1 | NORM_HALF})
    if (norm_q_ok(tw, th, cnMax, q)) cands.push_back(q);
  if (cands.size() < 2) return cands.empty() ? 0 : cands[0];
  const int rw = iw - tw + 1, rh = ih - th + 1;
  char what[112];
  snprintf(what, sizeof(what), "norm.q/%dx%d/%dx%d/c%d/%s", iw, ih, tw, th, cnMax, exact ? "exact" : "plain");
  return tune_choice(what, cands, "CVMATCH_NORM_Q", [&](int q) {
    return time_launches(
        s,
        [&] {
          launch_normalize(d_img, istride, iw, ih, step, tw, th, rw, rh, cnMax, st, corr, nullptr, parts, nscratch, s,
                           exact, q);
          return 0;
        },
        2);
  });
}

using SideWork = std::function<int()>;

static int corr_and_normalize(Ctx& X, hipStream_t s, const CallState* st, const uint8_t* d_img, int64_t istride,
                              int iw, int ih, int tw, int th, int step, int cnMax, const TplSpec& ts,
                              const PlanTables& pt, float* d_map, Res6* d_res, const SideWork* side) {
  const Plan& p = ts.p;
  int rw = iw - tw + 1, rh = ih - th + 1;
  int hwPad = ts.hwPad;
  const TileGrid g = parent_grid(p, tw, th, rw, rh, hwPad, ts.dh2);
  const int ntiles = g.ntx * g.nty;
  int64_t chan = (int64_t)p.dftH * hwPad, tileStride = chan * cnMax;
  int64_t tchan = chan, tchan2 = (int64_t)ts.dh2 * hwPad;
  // every buffer first: no allocation lands between fork and join
  CHK(X.spec.ensure(sizeof(float2) * (size_t)tileStride * ntiles));
  float2* spec = X.spec.as<float2>();
  float* corr = d_map;
  if (!corr) {
    CHK(X.corr.ensure(sizeof(float) * (size_t)rw * rh));
    corr = X.corr.as<float>();
  }
  // room for the partials of any strip width the tuner may try
  CHK(X.parts.ensure(sizeof(Extrema) * (size_t)norm_partials_max(rw, rh, tw, th, cnMax)));
  if (!side) CHK(X.nscratch.ensure(norm_scratch_bytes(iw, ih, tw, th, cnMax)));
  if (side) {
    HIPCHK(hipEventRecord(X.fork, s));
    HIPCHK(hipStreamWaitEvent(X.side, X.fork, 0));
  }
  {
    Stage S(X, "row_fwd", s);
    launch_row_fwd(d_img, istride, step, cnMax, side ? nullptr : st, g, pt, spec, tileStride, chan, ntiles, s);
  }
  if (side) {
    CHK((*side)());
    HIPCHK(hipStreamWaitEvent(s, X.join, 0));
    HIPCHK(hipStreamWaitEvent(s, X.join3, 0));
  }
  {
    Stage S(X, "cols", s);
    launch_cols(COL_CORR, spec, spec, tileStride, tileStride, chan, cnMax, st, g, pt, ts.spec, ts.spec2, tchan,
                tchan2, 1.0f, ntiles, s);
  }
  {
    Stage S(X, "row_inv", s);
    launch_row_inv(spec, tileStride, chan, st, cnMax, g, pt, corr, rw, ntiles, s);
  }
  if (side) {
    HIPCHK(hipStreamWaitEvent(s, X.join2, 0));
  } else {
    Stage S(X, "norm_prep", s);
    launch_norm_prep(d_img, istride, iw, ih, step, tw, th, cnMax, st, X.nscratch.p, s);
  }
  const int q = norm_q_for(X, s, d_img, istride, iw, ih, step, tw, th, cnMax, st, corr, X.parts.as<Extrema>(),
                           X.nscratch.p, false);
  {
    Stage S(X, "normalize", s);
    launch_normalize(d_img, istride, iw, ih, step, tw, th, rw, rh, cnMax, st, corr, d_map, X.parts.as<Extrema>(),
                     X.nscratch.p, s, false, q);
  }
  {
    Stage S(X, "reduce", s);
    launch_reduce(X.parts.as<Extrema>(), norm_partials_count(rw, rh, tw, th, cnMax, q), st, rw, d_res, nullptr, rh,
                  s);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// matchU8's argument checks impl.go:1127-1133: the shape, then the
// exact-statistics bound of the channel count.
static int check_shape(int64_t istride, int iw, int ih, int64_t tstride, int tw, int th, int cn, int step) {
  if (cn < 1 || cn > 4 || step < cn || tw < 1 || th < 1 || tw > iw || th > ih || istride < (int64_t)iw * step ||
      tstride < (int64_t)tw * step)
    return fail(CVM_E_BAD_ARGS, "cvmatch: bad match arguments (%dx%d in %dx%d, cn=%d step=%d)", tw, th, iw, ih, cn,
                step);
  return 0;
}

static int check_cap(int tw, int th, int cn) {
  if ((int64_t)tw * th > stats_cap(cn))
    return fail(CVM_E_STATS_CAP, "cvmatch: template area %dx%d exceeds the exact-statistics bound for cn=%d", tw, th,
                cn);
  return 0;
}

static int check_args(int64_t istride, int iw, int ih, int64_t tstride, int tw, int th, int cn, int step) {
  CHK(check_shape(istride, iw, ih, tstride, tw, th, cn, step));
  return check_cap(tw, th, cn);
}

// Match's alpha rule (cvmatch.go:82-86) for templates above statsCap(4):
// only cn = 3 (both alpha planes constant) can take them (impl.go:1131), so
// the rule is settled here, before the call, from device scans of both
// alpha planes; *cn_mode becomes 3, or the call fails with the bound of the
// channel count the rule picked.  Smaller templates keep cn_mode 0 and the
// rule is decided on the device inside the call.
static int alpha_rule_cn(Ctx& X, hipStream_t s, const uint8_t* d_img, int64_t istride, int iw, int ih,
                         const uint8_t* d_tpl, int64_t tstride, int tw, int th, int* cn_mode) {
  if ((int64_t)tw * th <= stats_cap(4)) return 0;
  launch_state_init(X.st, s);
  launch_alpha_scan(d_img, istride, iw, ih, X.st, 0, s);
  launch_alpha_scan(d_tpl, tstride, tw, th, X.st, 1, s);
  CallState hs;
  HIPCHK(hipMemcpyAsync(&hs, X.st, sizeof(CallState), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const int cn = (hs.alpha_bad[0] == 0 && hs.alpha_bad[1] == 0) ? 3 : 4;
  CHK(check_cap(tw, th, cn));
  *cn_mode = cn;
  return 0;
}

// Parents of at most this many pixels take the compact form of a one-shot
// call (launch_compact_fft); CVMATCH_COMPACT=<pixels> moves the bound, 0
// turns the form off (read at every call).
constexpr long long kCompactPixels = 2500000;
static long long compact_max_pixels() {
  const char* e = getenv("CVMATCH_COMPACT");
  return e ? atoll(e) : kCompactPixels;
}

// A small one-shot call in one stream: the transforms, the call state and
// the chunk sums in five launches whose independent jobs share a grid
// (launch_compact_fft), then the normalisation and the reduction.  The
// same kernels' code as the forked form, so the same bits.
static int run_compact(Ctx& X, hipStream_t s, const uint8_t* d_img, int64_t istride, int iw, int ih,
                       const uint8_t* d_tpl, int64_t tstride, int tw, int th, int cn_mode, int step, int cnMax,
                       const TplSpec& ts, const PlanTables& pt, const TileGrid& g, float* d_map) {
  const Plan& p = ts.p;
  const int rw = iw - tw + 1, rh = ih - th + 1, ntiles = g.ntx * g.nty;
  const int64_t chan = (int64_t)p.dftH * ts.hwPad, tileStride = chan * cnMax;
  CHK(X.spec.ensure(sizeof(float2) * (size_t)tileStride * ntiles));
  float* corr = d_map;
  if (!corr) {
    CHK(X.corr.ensure(sizeof(float) * (size_t)rw * rh));
    corr = X.corr.as<float>();
  }
  CHK(X.parts.ensure(sizeof(Extrema) * (size_t)norm_partials_max(rw, rh, tw, th, cnMax)));
  CompactCall c{};
  c.img = d_img;
  c.istride = istride;
  c.iw = iw;
  c.ih = ih;
  c.step = step;
  c.cnMax = cnMax;
  c.cn_mode = cn_mode;
  c.tw = tw;
  c.th = th;
  c.rw = rw;
  c.ntiles = ntiles;
  c.g = g;
  c.gt = tpl_grid(p, tw, th);
  c.spec = X.spec.as<float2>();
  c.tileStride = tileStride;
  c.chanStride = chan;
  c.tpl = d_tpl;
  c.tstride = tstride;
  c.tspec = ts.spec;
  c.tchan = chan;
  c.scale = 1.0f / ((float)p.dftW * (float)p.dftH);
  c.pt = pt;
  c.st = X.cst;
  c.csa = call_state_args(d_img, istride, iw, ih, d_tpl, tstride, tw, th, cnMax, step, cn_mode, X.cst, true);
  c.corr = corr;
  c.nscratch = X.nscratch.p;
  {
    Stage S(X, "compact", s);
    launch_compact_fft(c, s);
  }
  const int q = norm_q_for(X, s, d_img, istride, iw, ih, step, tw, th, cnMax, X.cst, corr, X.parts.as<Extrema>(),
                           X.nscratch.p, false);
  {
    Stage S(X, "normalize", s);
    launch_normalize(d_img, istride, iw, ih, step, tw, th, rw, rh, cnMax, X.cst, corr, d_map, X.parts.as<Extrema>(),
                     X.nscratch.p, s, false, q);
  }
  {
    Stage S(X, "reduce", s);
    launch_reduce_clear(X.parts.as<Extrema>(), norm_partials_count(rw, rh, tw, th, cnMax, q), rw, res_out(X), X.cst,
                        s);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// matchU8 on device buffers.  cn_mode 0 = Match's RGBA alpha rule.
static int run_match(Ctx& X, hipStream_t s, const uint8_t* d_img, int64_t istride, int iw, int ih,
                     const uint8_t* d_tpl, int64_t tstride, int tw, int th, int cn_mode, int step, float* d_map) {
  int rw = iw - tw + 1, rh = ih - th + 1;
  Plan p = make_plan(tw, th, rw, rh);
  CHK(check_limits(p, tw, rw, rh));
  int dh2 = shrink_dh(p, rh, th);
  PlanTables pt;
  CHK(plan_tables(X, p, dh2, pt));
  int cnMax = cn_mode ? cn_mode : 4;
  // side-stream buffers first: no allocation lands between fork and join
  CHK(X.nscratch.ensure(norm_scratch_bytes(iw, ih, tw, th, cnMax)));
  TplSpec ts;
  ts.p = p;
  ts.dh2 = dh2;
  ts.hwPad = round_up(p.hw, 16);
  ts.cnMax = cnMax;
  size_t tchan = (size_t)p.dftH * ts.hwPad, tchan2 = (size_t)dh2 * ts.hwPad;
  CHK(X.tspec.ensure(sizeof(float2) * tchan * cnMax));
  ts.spec = X.tspec.as<float2>();
  if (dh2) {
    CHK(X.tspec2.ensure(sizeof(float2) * tchan2 * cnMax));
    ts.spec2 = X.tspec2.as<float2>();
  }
  const TileGrid g = parent_grid(p, tw, th, rw, rh, ts.hwPad, dh2);
  if ((long long)iw * ih <= compact_max_pixels() && compact_ok(g) && norm_wave_ok(tw, th, cnMax))
    return run_compact(X, s, d_img, istride, iw, ih, d_tpl, tstride, tw, th, cn_mode, step, cnMax, ts, pt, g, d_map);
  CHK(X.ensure_side());
  // while s transforms the parent rows, the call state and then the
  // normalisation partials are built on one side stream and the template
  // spectrum (every channel: it needs no call state) on the other
  SideWork side = [&]() -> int {
    hipStream_t q2 = X.side2;
    HIPCHK(hipStreamWaitEvent(q2, X.fork, 0));
    CHK(build_tspec(X, q2, nullptr, d_tpl, tstride, tw, th, step, cnMax, p, dh2, pt, ts.spec, ts.spec2));
    HIPCHK(hipEventRecord(X.join3, q2));
    hipStream_t q = X.side;
    {
      Stage S(X, "stats", q);
      launch_state_init(X.st, q);
      launch_call_state(d_img, istride, iw, ih, d_tpl, tstride, tw, th, cnMax, step, cn_mode, X.st, q);
    }
    HIPCHK(hipEventRecord(X.join, q));
    {
      Stage S(X, "norm_prep", q);
      launch_norm_prep(d_img, istride, iw, ih, step, tw, th, cnMax, X.st, X.nscratch.p, q);
    }
    HIPCHK(hipEventRecord(X.join2, q));
    return 0;
  };
  return corr_and_normalize(X, s, X.st, d_img, istride, iw, ih, tw, th, step, cnMax, ts, pt, d_map,
                            res_out(X), &side);
}

// The six-tuple of a call that wrote it to res_out(X).
static int finish_result(Ctx& X, hipStream_t s, cvm_result* out) {
  if (X.rmap_d) {
    HIPCHK(hipStreamSynchronize(s));
    X.stage_collect();
    memcpy(out, X.rmap_h, sizeof(cvm_result));
    return 0;
  }
  HIPCHK(hipMemcpyAsync(X.res_h, X.res_d, sizeof(Res6), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  X.stage_collect();
  memcpy(out, X.res_h, sizeof(cvm_result));
  return 0;
}

static int64_t img_bytes(int64_t stride, int w, int h, int step) { return (int64_t)(h - 1) * stride + (int64_t)w * step; }

// Host-buffer front end shared by the one-shot entry points.
static int run_match_direct(Ctx& X, hipStream_t s, const uint8_t* d_img, int64_t istride, int iw, int ih,
                            const uint8_t* d_tpl, int64_t tstride, int tw, int th, int cn_mode, int step,
                            float* d_map);

// Every occurrence (threshold + local-maximum suppression, the recipe of
// TestFindAllOccurrences cvmatch_test.go:346-365) of a call's map.
struct PeakReq {
  float thresh;
  int rx, ry, cap;
  int32_t* xy;  // host, 2 * cap
  float* vals;  // host, cap
  int32_t* count;
};

static int check_peaks(int rw, int rh, int& rx, int& ry, int cap, const void* xy, const void* vals,
                       const void* count) {
  if (!count || (cap > 0 && (!xy || !vals))) return fail(CVM_E_BAD_ARGS, "cvmatch: nil peak output");
  if (rw <= 0 || rh <= 0) return fail(CVM_E_EMPTY_IMAGE, "cvmatch: empty image");
  if ((long long)rw * rh >= (1LL << 31))
    return fail(CVM_E_UNSUPPORTED, "cvmatch: result map of %dx%d exceeds 2^31 elements", rw, rh);
  if (rx < 0 || ry < 0 || cap < 0)
    return fail(CVM_E_BAD_ARGS, "cvmatch: bad peak arguments (rx=%d ry=%d cap=%d)", rx, ry, cap);
  // a window wider than the map sees the same elements as one just as wide
  rx = std::min(rx, rw - 1);
  ry = std::min(ry, rh - 1);
  if (rx > peaks_max_rx())
    return fail(CVM_E_UNSUPPORTED, "cvmatch: peak window half-width %d exceeds the staged row's limit (%d)", rx,
                peaks_max_rx());
  return 0;
}

// Peaks of a device map into d_xy / d_vals; the total stays in the scratch.
static int peaks_dev(Ctx& X, hipStream_t s, const float* d_map, int rw, int rh, float thresh, int rx, int ry,
                     int cap, int32_t* d_xy, float* d_vals) {
  CHK(X.pkscr.ensure(peaks_scratch_bytes(rw, rh)));
  Stage S(X, "peaks", s);
  launch_peaks(d_map, rw, rh, thresh, rx, ry, X.pkscr.p, cap, d_xy, d_vals, s);
  HIPCHK(hipGetLastError());
  return 0;
}

static int peaks_total_host(Ctx& X, hipStream_t s, int rw, int rh, int* tot) {
  HIPCHK(hipMemcpyAsync(tot, peaks_total(X.pkscr.p, rw, rh), sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

// Pageable host parents of at least this many bytes take the overlapped
// form of a one-shot call (run_match_overlap); CVMATCH_OVERLAP=0 turns it
// off.
constexpr size_t kOverlapBytes = (size_t)8 << 20;
static bool overlap_enabled() {
  const char* e = getenv("CVMATCH_OVERLAP");
  return !(e && e[0] == '0');
}

// matchU8 on a host parent whose upload overlaps the call's own transforms:
// the parent moves on the context's copy stream chunk by chunk, and as soon
// as the rows of a tile row are in, that tile row's forward rows and column
// pass run on s (every channel, as the forked form's do: the channel count
// of the alpha rule is not known before the whole parent is).  The template
// (already on the device) gets its spectrum first; once the whole parent is
// in, the call state, the inverse rows, the chunk sums, the normalisation
// and the reduction follow on s.  Every kernel is the one the other forms
// run on the same data, so the bits are theirs.
static int run_match_overlap(Ctx& X, hipStream_t s, const uint8_t* img, size_t ib, int64_t istride, int iw, int ih,
                             const uint8_t* d_tpl, int64_t tstride, int tw, int th, int cn_mode, int step,
                             float* d_map) {
  const int rw = iw - tw + 1, rh = ih - th + 1;
  const Plan p = make_plan(tw, th, rw, rh);
  CHK(check_limits(p, tw, rw, rh));
  const int dh2 = shrink_dh(p, rh, th);
  PlanTables pt;
  CHK(plan_tables(X, p, dh2, pt));
  const int cnMax = cn_mode ? cn_mode : 4;
  TplSpec ts;
  ts.p = p;
  ts.dh2 = dh2;
  ts.hwPad = round_up(p.hw, 16);
  ts.cnMax = cnMax;
  const int64_t tchan = (int64_t)p.dftH * ts.hwPad, tchan2 = (int64_t)dh2 * ts.hwPad;
  CHK(X.tspec.ensure(sizeof(float2) * (size_t)tchan * cnMax));
  ts.spec = X.tspec.as<float2>();
  if (dh2) {
    CHK(X.tspec2.ensure(sizeof(float2) * (size_t)tchan2 * cnMax));
    ts.spec2 = X.tspec2.as<float2>();
  }
  const TileGrid g = parent_grid(p, tw, th, rw, rh, ts.hwPad, dh2);
  const int ntiles = g.ntx * g.nty;
  const int64_t chan = tchan, tileStride = chan * cnMax;
  CHK(X.spec.ensure(sizeof(float2) * (size_t)tileStride * ntiles));
  float2* spec = X.spec.as<float2>();
  float* corr = d_map;
  if (!corr) {
    CHK(X.corr.ensure(sizeof(float) * (size_t)rw * rh));
    corr = X.corr.as<float>();
  }
  CHK(X.parts.ensure(sizeof(Extrema) * (size_t)norm_partials_max(rw, rh, tw, th, cnMax)));
  CHK(X.nscratch.ensure(norm_scratch_bytes(iw, ih, tw, th, cnMax)));
  const uint8_t* d_img = X.img.as<uint8_t>();
  CHK(X.ensure_copy());
  CHK(build_tspec(X, s, nullptr, d_tpl, tstride, tw, th, step, cnMax, p, dh2, pt, ts.spec, ts.spec2));
  int next = 0;  // the first tile row not yet started
  auto on_chunk = [&](size_t done) -> int {
    const int64_t rows = done >= ib ? (int64_t)ih : (int64_t)(done / (size_t)istride);
    int ready = next;
    while (ready < g.nty) {
      const int y0 = ready * g.blockH;
      const int bh = ready == g.nty - 1 ? rh - y0 : g.blockH;
      if (std::min<int64_t>(ih, (int64_t)y0 + bh + th - 1) > rows) break;
      ready++;
    }
    if (ready == next) return 0;
    hipEvent_t e = X.ev();
    if (!e) return 1;
    if (hipEventRecord(e, X.cstream) != hipSuccess || hipStreamWaitEvent(s, e, 0) != hipSuccess) return 1;
    X.evpool.push_back(e);  // reusable once recorded: the wait above already captured it
    for (int r = next; r < ready; r++) {
      launch_row_fwd_range(d_img, istride, step, cnMax, nullptr, g, pt, spec, tileStride, chan, r * g.ntx, g.ntx, s);
      launch_cols_range(COL_CORR, spec, spec, tileStride, chan, cnMax, nullptr, g, pt, ts.spec, ts.spec2, tchan,
                        tchan2, r * g.ntx, g.ntx, s);
    }
    next = ready;
    return 0;
  };
  {
    Stage S(X, "h2d+fft", s);
    HIPCHK(xfer_h2d_chunks(X.img.p, img, ib, X.cstream, on_chunk));
  }
  if (next != g.nty) return fail(CVM_E_HIP, "cvmatch: overlapped upload left tile rows unstarted");
  {
    Stage S(X, "stats", s);
    launch_state_init(X.st, s);
    launch_call_state(d_img, istride, iw, ih, d_tpl, tstride, tw, th, cnMax, step, cn_mode, X.st, s);
  }
  {
    Stage S(X, "row_inv", s);
    launch_row_inv(spec, tileStride, chan, X.st, cnMax, g, pt, corr, rw, ntiles, s);
  }
  {
    Stage S(X, "norm_prep", s);
    launch_norm_prep(d_img, istride, iw, ih, step, tw, th, cnMax, X.st, X.nscratch.p, s);
  }
  const int q = norm_q_for(X, s, d_img, istride, iw, ih, step, tw, th, cnMax, X.st, corr, X.parts.as<Extrema>(),
                           X.nscratch.p, false);
  {
    Stage S(X, "normalize", s);
    launch_normalize(d_img, istride, iw, ih, step, tw, th, rw, rh, cnMax, X.st, corr, d_map, X.parts.as<Extrema>(),
                     X.nscratch.p, s, false, q);
  }
  {
    Stage S(X, "reduce", s);
    launch_reduce(X.parts.as<Extrema>(), norm_partials_count(rw, rh, tw, th, cnMax, q), X.st, rw, res_out(X),
                  nullptr, rh, s);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

static int match_host(const uint8_t* img, int64_t istride, int iw, int ih, const uint8_t* tpl, int64_t tstride,
                      int tw, int th, int cn_mode, int step, float* map, cvm_result* out, bool direct = false,
                      const PeakReq* pk = nullptr) {
  LEASE(X);
  hipStream_t s = X.stream;
  int rw = iw - tw + 1, rh = ih - th + 1;
  size_t ib = (size_t)img_bytes(istride, iw, ih, step), tb = (size_t)img_bytes(tstride, tw, th, step);
  CHK(X.img.ensure(ib));
  CHK(X.tpl.ensure(tb));
  // (a pinned parent moves in one copy, so nothing would overlap it)
  if (!direct && !pk && ib >= kOverlapBytes && overlap_enabled() && !xfer_h2d_source_busy(img, ib) &&
      (cn_mode != 0 || (int64_t)tw * th <= stats_cap(4))) {
    // the overlapped form: the template first, the parent chunk by chunk
    HIPCHK(xfer_h2d(X.tpl.p, tpl, tb, s));
    float* d_map = nullptr;
    if (map) {
      CHK(X.mapbuf.ensure(sizeof(float) * (size_t)rw * rh));
      d_map = X.mapbuf.as<float>();
    }
    CHK(run_match_overlap(X, s, img, ib, istride, iw, ih, X.tpl.as<uint8_t>(), tstride, tw, th, cn_mode, step,
                          d_map));
    if (map) {
      Stage S(X, "d2h", s);
      HIPCHK(xfer_d2h(map, d_map, sizeof(float) * (size_t)rw * rh, s));
    }
    return finish_result(X, s, out);
  }
  {
    Stage S(X, "h2d", s);
    HIPCHK(xfer_h2d(X.img.p, img, ib, s));
    HIPCHK(xfer_h2d(X.tpl.p, tpl, tb, s));
  }
  float* d_map = nullptr;
  if (map || pk) {
    CHK(X.mapbuf.ensure(sizeof(float) * (size_t)rw * rh));
    d_map = X.mapbuf.as<float>();
  }
  int32_t* d_xy = nullptr;
  float* d_vals = nullptr;
  if (pk) {
    CHK(X.pkscr.ensure(peaks_scratch_bytes(rw, rh)));
    CHK(X.pkout.ensure(sizeof(int32_t) * 3 * (size_t)pk->cap + 16));
    d_xy = X.pkout.as<int32_t>();
    d_vals = (float*)(d_xy + 2 * (size_t)pk->cap);
  }
  if (cn_mode == 0)
    CHK(alpha_rule_cn(X, s, X.img.as<uint8_t>(), istride, iw, ih, X.tpl.as<uint8_t>(), tstride, tw, th, &cn_mode));
  if (direct)
    CHK(run_match_direct(X, s, X.img.as<uint8_t>(), istride, iw, ih, X.tpl.as<uint8_t>(), tstride, tw, th, cn_mode,
                         step, d_map));
  else
    CHK(run_match(X, s, X.img.as<uint8_t>(), istride, iw, ih, X.tpl.as<uint8_t>(), tstride, tw, th, cn_mode, step,
                  d_map));
  if (map) {
    Stage S(X, "d2h", s);
    HIPCHK(xfer_d2h(map, d_map, sizeof(float) * (size_t)rw * rh, s));
  }
  if (pk) {
    CHK(peaks_dev(X, s, d_map, rw, rh, pk->thresh, pk->rx, pk->ry, pk->cap, d_xy, d_vals));
    int tot = 0;
    CHK(peaks_total_host(X, s, rw, rh, &tot));
    int k = std::min(tot, pk->cap);
    if (k > 0) {
      HIPCHK(xfer_d2h(pk->xy, d_xy, sizeof(int32_t) * 2 * (size_t)k, s));
      HIPCHK(xfer_d2h(pk->vals, d_vals, sizeof(float) * (size_t)k, s));
    }
    *pk->count = tot;
  }
  return finish_result(X, s, out);
}


// --------------------------------------------------------- state helpers --

__global__ void k_state_set(CallState* st, TemplStats ts, int tpl_alpha_bad) {
  if (threadIdx.x != 0) return;
  for (int k = 0; k < 4; k++) {
    st->tsum[k] = ts.tsum[k];
    st->tsq[k] = ts.tsq[k];
  }
  st->varSum = 0;
  st->cn = 0;
  st->flat = 0;
  st->alpha_bad[0] = 0;
  st->alpha_bad[1] = (unsigned)tpl_alpha_bad;
  st->templNorm = st->numScale = st->varScale = 0.f;
}

void launch_state_set(CallState* st, const TemplStats& ts, int tpl_alpha_bad, hipStream_t s) {
  hipLaunchKernelGGL(k_state_set, dim3(1), dim3(64), 0, s, st, ts, tpl_alpha_bad);
}

// ------------------------------------------------------------- Matcher --

// Per-parent-geometry template state (matcherGeom matcher.go:36-40): the
// plan and the template spectrum of every channel the template has, so the
// device-side channel choice (3 or 4) never needs a rebuild.
struct MatcherGeom {
  int rw = 0, rh = 0;
  TplSpec ts;
  DevBuf spec, spec2;
  ~MatcherGeom() {
    spec.release();
    spec2.release();
  }
};

}  // namespace cvm

struct cvm_matcher {
  int w = 0, h = 0, step = 0;
  int aConst = 0;
  cvm::TemplStats stats{};
  long long varSum = 0;
  uint8_t* d_pix = nullptr;  // compact copy, stride w*step
  std::mutex mu;
  std::shared_ptr<cvm::MatcherGeom> geom;
  // direct-correlation B fragments per channel count (built on first use)
  cvm::DevBuf dbf[5];
  bool dbf_ok[5] = {false, false, false, false, false};
  std::atomic<int> refs{1};
};

struct cvm_fleet;

namespace cvm {

static void matcher_release(cvm_matcher* m) {
  if (!m) return;
  if (m->refs.fetch_sub(1) == 1) {
    // a later matcher may reuse this address: no captured call may replay
    // with this one's buffers
    g_alloc_epoch.fetch_add(1);
    if (m->d_pix) (void)hipFree(m->d_pix);
    for (auto& b : m->dbf) b.release();
    delete m;
  }
}

// newMatcher matcher.go:66-76 (+ alphaConst of the template, matcher.go:48):
// private compact device copy and exact statistics over cn = step channels.
static int matcher_create(const uint8_t* tpl, int64_t stride, int w, int h, int step, cvm_matcher** out) {
  if (!tpl) return fail(CVM_E_NIL_IMAGE, "cvmatch: nil image");
  if (w <= 0 || h <= 0) return fail(CVM_E_EMPTY_IMAGE, "cvmatch: empty image");
  if (step != 1 && step != 4) return fail(CVM_E_BAD_ARGS, "cvmatch: matcher step must be 1 (gray) or 4 (RGBA)");
  if (stride < (int64_t)w * step) return fail(CVM_E_BAD_ARGS, "cvmatch: stride below width*step");
  if ((int64_t)w * h > stats_cap(step))
    return fail(CVM_E_STATS_CAP, "cvmatch: template area %dx%d exceeds the exact-statistics bound", w, h);
  LEASE(X);
  hipStream_t s = X.stream;
  std::unique_ptr<cvm_matcher> m(new cvm_matcher());
  m->w = w;
  m->h = h;
  m->step = step;
  HIPCHK(hipMalloc(&m->d_pix, (size_t)w * h * step));
  HIPCHK(hipMemcpy2DAsync(m->d_pix, (size_t)w * step, tpl, (size_t)stride, (size_t)w * step, h, hipMemcpyHostToDevice,
                          s));
  launch_state_init(X.st, s);
  launch_tpl_stats(m->d_pix, (int64_t)w * step, w, h, step, step, X.st, s);
  if (step == 4) launch_alpha_scan(m->d_pix, (int64_t)w * 4, w, h, X.st, 1, s);
  CallState hs;
  HIPCHK(hipMemcpyAsync(&hs, X.st, sizeof(CallState), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  long long area = (long long)w * h, var = 0;
  for (int k = 0; k < 4; k++) {
    m->stats.tsum[k] = hs.tsum[k];
    m->stats.tsq[k] = hs.tsq[k];
  }
  for (int k = 0; k < step; k++) var += area * (long long)hs.tsq[k] - (long long)hs.tsum[k] * (long long)hs.tsum[k];
  m->varSum = var;
  m->aConst = step == 4 ? (hs.alpha_bad[1] == 0) : 0;
  *out = m.release();
  return 0;
}

// Sets the call state from a matcher's cached statistics and decides cn on
// the device (Find matcher.go:103-112: 3 only when both alpha planes are
// constant; the parent is scanned only when the template's alpha is).
static void matcher_state(Ctx& X, hipStream_t s, CallState* st, const cvm_matcher* m, const uint8_t* d_img,
                          int64_t istride, int iw, int ih) {
  launch_state_set(st, m->stats, m->step == 4 && !m->aConst, s);
  if (m->step == 4 && m->aConst) launch_alpha_scan(d_img, istride, iw, ih, st, 0, s);
  launch_state_finalize(st, m->w, m->h, m->step == 4 ? 0 : 1, s);
}

// Template spectrum of m on plan p built for all `step` channels.
static int matcher_spectrum(Ctx& X, hipStream_t s, CallState* st, const cvm_matcher* m, const Plan& p, int dh2,
                            const PlanTables& pt, TplSpec& ts, DevBuf& spec, DevBuf& spec2) {
  ts.p = p;
  ts.dh2 = dh2;
  ts.hwPad = round_up(p.hw, 16);
  ts.cnMax = m->step;
  size_t tchan = (size_t)p.dftH * ts.hwPad, tchan2 = (size_t)dh2 * ts.hwPad;
  CHK(spec.ensure(sizeof(float2) * tchan * m->step));
  ts.spec = spec.as<float2>();
  ts.spec2 = nullptr;
  if (dh2) {
    CHK(spec2.ensure(sizeof(float2) * tchan2 * m->step));
    ts.spec2 = spec2.as<float2>();
  }
  // build every channel: a state whose cn is the template's channel count
  launch_state_set(st, m->stats, 0, s);
  launch_state_finalize(st, m->w, m->h, m->step, s);
  return build_tspec(X, s, st, m->d_pix, (int64_t)m->w * m->step, m->w, m->h, m->step, m->step, p, dh2, pt, ts.spec,
                     ts.spec2);
}

static int matcher_find_dev(Ctx& X, hipStream_t s, cvm_matcher* m, const uint8_t* d_img, int64_t istride, int iw,
                            int ih, Res6* d_res) {
  int rw = iw - m->w + 1, rh = ih - m->h + 1;
  std::shared_ptr<MatcherGeom> g;
  {
    std::lock_guard<std::mutex> lk(m->mu);
    g = m->geom;
  }
  Plan p;
  int dh2;
  PlanTables pt;
  if (!g || g->rw != rw || g->rh != rh) {
    p = make_plan(m->w, m->h, rw, rh);
    CHK(check_limits(p, m->w, rw, rh));
    dh2 = shrink_dh(p, rh, m->h);
    CHK(plan_tables(X, p, dh2, pt));
    auto ng = std::make_shared<MatcherGeom>();
    ng->rw = rw;
    ng->rh = rh;
    CHK(matcher_spectrum(X, s, X.st, m, p, dh2, pt, ng->ts, ng->spec, ng->spec2));
    // published only once built: a call on another stream may use it next
    HIPCHK(hipStreamSynchronize(s));
    std::lock_guard<std::mutex> lk(m->mu);
    m->geom = ng;
    g = ng;
  } else {
    CHK(plan_tables(X, g->ts.p, g->ts.dh2, pt));
  }
  matcher_state(X, s, X.st, m, d_img, istride, iw, ih);
  return corr_and_normalize(X, s, X.st, d_img, istride, iw, ih, m->w, m->h, m->step, m->step, g->ts, pt, nullptr,
                            d_res, nullptr);
}

// --------------------------------------------------------------- Fleet --

struct FleetBucket {
  int tw = 0, th = 0;
  std::vector<int> members;  // global member indices, ascending
};

// bucketize fleet.go:111-171: pow2 size classes of the non-flat members,
// largest area first (then wider, then taller); a class joins the first
// bucket that dominates it in both axes; bucket dims are its members' maxima.
static std::vector<FleetBucket> bucketize(const std::vector<cvm_matcher*>& ms) {
  std::vector<std::pair<int, int>> keys;
  for (auto* m : ms) {
    if (m->varSum == 0) continue;
    std::pair<int, int> k(next_pow2(m->w), next_pow2(m->h));
    if (std::find(keys.begin(), keys.end(), k) == keys.end()) keys.push_back(k);
  }
  std::sort(keys.begin(), keys.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
    long long aa = (long long)a.first * a.second, ab = (long long)b.first * b.second;
    if (aa != ab) return aa > ab;
    if (a.first != b.first) return a.first > b.first;
    return a.second > b.second;
  });
  std::vector<std::pair<std::pair<int, int>, std::vector<int>>> slots;
  for (auto& k : keys) {
    std::vector<int> mem;
    for (size_t i = 0; i < ms.size(); i++)
      if (ms[i]->varSum != 0 && next_pow2(ms[i]->w) == k.first && next_pow2(ms[i]->h) == k.second)
        mem.push_back((int)i);
    bool placed = false;
    for (auto& s : slots)
      if (s.first.first >= k.first && s.first.second >= k.second) {
        s.second.insert(s.second.end(), mem.begin(), mem.end());
        placed = true;
        break;
      }
    if (!placed) slots.push_back({k, mem});
  }
  std::vector<FleetBucket> out;
  for (auto& s : slots) {
    FleetBucket b;
    b.members = s.second;
    std::sort(b.members.begin(), b.members.end());
    for (int mi : b.members) {
      b.tw = std::max(b.tw, ms[mi]->w);
      b.th = std::max(b.th, ms[mi]->h);
    }
    out.push_back(b);
  }
  return out;
}

// Per-bucket geometry for one parent size (buildGeom fleet.go:337-390).
struct BucketGeom {
  Plan p;
  int dh2 = 0, hwPad = 0, cnMax = 0;
  std::vector<int> cns;            // per bucket position
  std::vector<TplSpec> ts;         // per bucket position
  std::vector<DevBuf> specs, specs2;
  ~BucketGeom() {
    for (auto& b : specs) b.release();
    for (auto& b : specs2) b.release();
  }
};

// The plain FindAll's normalisations, batched: members (in loop order) whose
// correlation maps together fit the batch buffer are normalised by one
// launch per strip form and reduced by one launch once the last of them has
// its map.  Members past the wave form are normalised on their own.
struct PlainNormPlan {
  std::vector<int64_t> corr_off;  // per member: float offset in the batch buffer, -1: on its own
  std::vector<int> last_of;       // per member: batch index it closes, -1 otherwise
  struct Batch {
    std::vector<std::array<int, 4>> groups;  // (q, sq32, first entry in the wg table, workgroups)
    int job0 = 0, njobs = 0;
  };
  std::vector<Batch> batches;
  DevBuf jobs, wg;
  bool used_cn[5] = {false, false, false, false, false};
  int prep_member[5] = {-1, -1, -1, -1, -1};  // a batched member of each channel count
  std::array<void*, 7> ptrs{};  // batch corr, batch parts, results, scratch per cn at build time
  bool built = false;
  ~PlainNormPlan() {
    jobs.release();
    wg.release();
  }
};

struct FleetGeom {
  int pw = 0, ph = 0, pConst = 0;
  std::vector<std::unique_ptr<BucketGeom>> bg;
  CallState* states = nullptr;  // per member
  CallState* frame_state = nullptr;
  PlainNormPlan norm;
  ~FleetGeom() {
    if (states) (void)hipFree(states);
    if (frame_state) (void)hipFree(frame_state);
  }
};

}  // namespace cvm

struct cvm_fleet {
  std::vector<cvm_matcher*> ms;
  int step = 0;
  std::vector<cvm::FleetBucket> buckets;
  std::mutex mu;
  std::mutex call_mu;  // held by FindAll
  std::shared_ptr<cvm::FleetGeom> geom;
  cvm::DevBuf frame, mspec, corr, res;
  // plain FindAll: correlation maps and extrema partials of one batch, and
  // the second member's product spectra of a paired column launch
  cvm::DevBuf pcorr, pparts, mspec2;
  // direct form: per-member call states for the parent alpha key dpConst,
  // and the normalisation partials per channel count
  cvm::DevBuf dstates, dscr[5];
  int dpConst = -1;
  // batched direct launches for one frame shape and parent-alpha key:
  // job tables (device), all members' C' maps and extrema partials
  cvm::DevBuf bj_direct, bj_dwg, bj_norm, bj_nwg, bj_cmaps, bj_parts;
  int bj_pw = -1, bj_ph = -1, bj_pConst = -1, bj_nnorm = 0;
  // the per-channel-count scratch the tables point into (the plain FindAll
  // grows the same buffers for larger frames): tables are rebuilt when any
  // of them has moved
  void* bj_ps[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  std::vector<std::array<int, 4>> bj_groups;   // normalisation: (q, sq32, first wg entry, wg count)
  std::vector<std::array<int, 6>> bj_dgroups;  // direct: (kstep, cmax, first wg entry, wg count, max pitch, waves)
  // the fused form (correlation and scoring in one launch per group, no C'
  // maps): job tables, reduction jobs, per-block extrema partials, groups
  // (kstep, cmax, sq32, first wg entry, wg count, max pitch, waves)
  cvm::DevBuf bj_fused, bj_fwg, bj_fred, bj_fparts;
  std::vector<std::array<int, 7>> bj_fgroups;
  int bj_nfred = 0, bj_use_fused = 0;
  cvm::GraphSlot gplain, gdirect, gmixed;  // the captured FindAll of each form
  // The mixed form's routing for one frame geometry (pw, ph, pConst): the
  // non-flat members split into a direct fleet of their own (didx: their
  // indices here) and the members routed to their plain Find (ridx); gate[i] is
  // member i's measured max |FFT score - exact score| on the tuning frame
  // (-1: not a candidate), crossover the template side from which members
  // were candidates (0: none).
  struct Mixed {
    int pw = -1, ph = -1, pConst = -1, crossover = 0;
    cvm_fleet* dpart = nullptr;
    std::vector<int> didx, ridx;
    std::vector<float> gate;
    cvm::DevBuf idx;  // didx on the device
  };
  std::unique_ptr<Mixed> mixed;
};

namespace cvm {

__global__ void k_state_frame(CallState* st, int cn) {
  if (threadIdx.x != 0) return;
  for (int k = 0; k < 4; k++) {
    st->tsum[k] = 0;
    st->tsq[k] = 0;
  }
  st->varSum = 1;
  st->cn = cn;
  st->flat = 0;
  st->alpha_bad[0] = st->alpha_bad[1] = 0;
  st->templNorm = st->numScale = st->varScale = 0.f;
}

static int fleet_build_geom(Ctx& X, hipStream_t s, cvm_fleet* f, int pw, int ph, int pConst,
                            std::shared_ptr<FleetGeom>& out) {
  auto g = std::make_shared<FleetGeom>();
  g->pw = pw;
  g->ph = ph;
  g->pConst = pConst;
  size_t n = f->ms.size();
  HIPCHK(hipMalloc(&g->states, sizeof(CallState) * std::max(n, (size_t)1)));
  HIPCHK(hipMalloc(&g->frame_state, sizeof(CallState) * std::max(f->buckets.size(), (size_t)1)));
  for (size_t bi = 0; bi < f->buckets.size(); bi++) {
    const FleetBucket& b = f->buckets[bi];
    auto bg = std::make_unique<BucketGeom>();
    int rwF = pw - b.tw + 1, rhF = ph - b.th + 1;
    bg->p = make_plan(b.tw, b.th, rwF, rhF);
    CHK(check_limits(bg->p, b.tw, rwF, rhF));
    bg->dh2 = shrink_dh(bg->p, rhF, b.th);
    bg->hwPad = round_up(bg->p.hw, 16);
    PlanTables pt;
    CHK(plan_tables(X, bg->p, bg->dh2, pt));
    size_t nm = b.members.size();
    bg->cns.resize(nm);
    bg->ts.resize(nm);
    bg->specs.resize(nm);
    bg->specs2.resize(nm);
    for (size_t j = 0; j < nm; j++) {
      cvm_matcher* m = f->ms[b.members[j]];
      int cn = m->step == 4 ? ((m->aConst && pConst) ? 3 : 4) : 1;
      bg->cns[j] = cn;
      bg->cnMax = std::max(bg->cnMax, cn);
      TplSpec& ts = bg->ts[j];
      ts.p = bg->p;
      ts.dh2 = bg->dh2;
      ts.hwPad = bg->hwPad;
      ts.cnMax = cn;
      size_t tchan = (size_t)bg->p.dftH * bg->hwPad, tchan2 = (size_t)bg->dh2 * bg->hwPad;
      CHK(bg->specs[j].ensure(sizeof(float2) * tchan * cn));
      ts.spec = bg->specs[j].as<float2>();
      ts.spec2 = nullptr;
      if (bg->dh2) {
        CHK(bg->specs2[j].ensure(sizeof(float2) * tchan2 * cn));
        ts.spec2 = bg->specs2[j].as<float2>();
      }
      // member spectrum on the bucket plan, shrunk with the bucket th
      // (buildTSpecSet(..., rhF, b.th, ...) fleet.go:373-374)
      CallState* mst = g->states + b.members[j];
      launch_state_set(mst, m->stats, 0, s);
      launch_state_finalize(mst, m->w, m->h, cn, s);
      CHK(build_tspec(X, s, mst, m->d_pix, (int64_t)m->w * m->step, m->w, m->h, m->step, cn, bg->p, bg->dh2, pt,
                      ts.spec, ts.spec2));
    }
    hipLaunchKernelGGL(k_state_frame, dim3(1), dim3(64), 0, s, g->frame_state + bi, bg->cnMax);
    g->bg.push_back(std::move(bg));
  }
  HIPCHK(hipGetLastError());
  out = g;
  return 0;
}

// The plain FindAll's batched normalisation plan for geometry g: built on
// first use, rebuilt when a buffer its job tables point into has moved.
static int plain_norm_plan(Ctx& X, hipStream_t s, cvm_fleet* f, FleetGeom& g, int pw, int ph, Res6* d_res) {
  PlainNormPlan& P = g.norm;
  constexpr size_t kBatchFloats = (size_t)48 << 20;  // 192 MiB of correlation maps per batch
  const size_t n = f->ms.size();
  struct Item {
    int mi, cn, q, sq, rw, rh;
    int64_t off;
  };
  std::vector<std::vector<Item>> bat;
  std::vector<int> order, cnOf(n, 0);
  for (size_t bi = 0; bi < f->buckets.size(); bi++)
    for (size_t j = 0; j < f->buckets[bi].members.size(); j++) {
      int mi = f->buckets[bi].members[j];
      order.push_back(mi);
      cnOf[mi] = g.bg[bi]->cns[j];
    }
  size_t maxMap = 0;
  for (int mi : order)
    maxMap = std::max(maxMap, (size_t)(pw - f->ms[mi]->w + 1) * (size_t)(ph - f->ms[mi]->h + 1));
  const size_t capF = std::max(kBatchFloats, (maxMap + 63) / 64 * 64);
  std::vector<int64_t> off(n, -1);
  std::vector<int> last(n, -1);
  bool usedCn[5] = {false, false, false, false, false};
  int prepM[5] = {-1, -1, -1, -1, -1};
  size_t fill = 0, maxFill = 64;  // the region is sized to the largest batch, not the cap
  for (int mi : order) {
    cvm_matcher* m = f->ms[mi];
    const int cn = cnOf[mi];
    if (!norm_wave_ok(m->w, m->h, cn)) continue;  // normalised on its own (int64 form)
    Item it{mi, cn, 0, 0, pw - m->w + 1, ph - m->h + 1, 0};
    norm_form(m->w, m->h, cn, &it.q, &it.sq);
    const size_t sz = ((size_t)it.rw * it.rh + 63) / 64 * 64;
    if (bat.empty() || fill + sz > capF) {
      bat.emplace_back();
      fill = 0;
    }
    it.off = (int64_t)fill;
    off[mi] = it.off;
    fill += sz;
    maxFill = std::max(maxFill, fill);
    bat.back().push_back(it);
    usedCn[cn] = true;
    if (prepM[cn] < 0) prepM[cn] = mi;
  }
  size_t maxParts = 1;
  for (auto& b : bat) {
    size_t t = 0;
    for (auto& it : b) t += (size_t)norm_grid_x(it.rw, it.q) * norm_grid_y(it.rh);
    maxParts = std::max(maxParts, t);
  }
  CHK(f->pcorr.ensure(sizeof(float) * maxFill));
  CHK(f->pparts.ensure(sizeof(Extrema) * maxParts));
  for (int cn = 1; cn <= 4; cn++)
    if (usedCn[cn]) CHK(f->dscr[cn].ensure(norm_scratch_bytes(pw, ph, 1, 1, cn)));
  const std::array<void*, 7> now = {f->pcorr.p, f->pparts.p, (void*)d_res, f->dscr[1].p, f->dscr[2].p,
                                    f->dscr[3].p, f->dscr[4].p};
  if (P.built && P.ptrs == now) return 0;
  std::vector<NormJob> jobs;
  std::vector<int> wgs;
  P.batches.clear();
  for (size_t b = 0; b < bat.size(); b++) {
    PlainNormPlan::Batch B;
    B.job0 = (int)jobs.size();
    size_t poff = 0;
    for (int q : {4, 2})
      for (int sq : {1, 0}) {
        const int first = (int)wgs.size();
        for (auto& it : bat[b]) {
          if (it.q != q || it.sq != sq) continue;
          cvm_matcher* m = f->ms[it.mi];
          NormJob J{};
          J.st = g.states + it.mi;
          J.ps = f->dscr[it.cn].as<uint32_t>();
          J.corr = f->pcorr.as<float>() + it.off;
          J.parts = f->pparts.as<Extrema>() + poff;
          J.res = d_res + it.mi;
          J.tw = m->w;
          J.th = m->h;
          J.rw = it.rw;
          J.rh = it.rh;
          J.gx = norm_grid_x(it.rw, q);
          J.gy = norm_grid_y(it.rh);
          J.wg0 = (int)wgs.size() - first;
          poff += (size_t)J.gx * J.gy;
          for (int w = 0; w < J.gx * J.gy; w++) wgs.push_back((int)jobs.size());
          jobs.push_back(J);
        }
        if ((int)wgs.size() > first) B.groups.push_back({q, sq, first, (int)wgs.size() - first});
      }
    B.njobs = (int)jobs.size() - B.job0;
    P.batches.push_back(B);
    last[bat[b].back().mi] = (int)b;
  }
  if (!jobs.empty()) {
    CHK(P.jobs.ensure(sizeof(NormJob) * jobs.size()));
    CHK(P.wg.ensure(sizeof(int) * wgs.size()));
    HIPCHK(hipMemcpyAsync(P.jobs.p, jobs.data(), sizeof(NormJob) * jobs.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(P.wg.p, wgs.data(), sizeof(int) * wgs.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));  // the host vectors go out of scope
  }
  P.corr_off = off;
  P.last_of = last;
  for (int cn = 0; cn < 5; cn++) {
    P.used_cn[cn] = usedCn[cn];
    P.prep_member[cn] = prepM[cn];
  }
  P.ptrs = now;
  P.built = true;
  return 0;
}

// FindAll fleet.go:176-332 on a device-resident parent.
static int fleet_find_dev(Ctx& X, hipStream_t s, cvm_fleet* f, const uint8_t* d_img, int64_t istride, int pw,
                          int ph, int pConst, Res6* d_res, std::vector<int>& flat_members) {
  std::shared_ptr<FleetGeom> g;
  {
    std::lock_guard<std::mutex> lk(f->mu);
    g = f->geom;
  }
  if (!g || g->pw != pw || g->ph != ph || g->pConst != pConst) {
    std::shared_ptr<FleetGeom> ng;
    CHK(fleet_build_geom(X, s, f, pw, ph, pConst, ng));
    std::lock_guard<std::mutex> lk(f->mu);
    f->geom = ng;
    g = ng;
  }
  // Batched normalisations (members taking the wave form): the chunk prefix
  // sums depend on the frame and the channel count only, built once per
  // channel count before the members' tails.
  CHK(plain_norm_plan(X, s, f, *g, pw, ph, d_res));
  const PlainNormPlan& NP = g->norm;
  for (int cn = 1; cn <= 4; cn++) {
    if (!NP.used_cn[cn]) continue;
    const int mi = NP.prep_member[cn];
    Stage S(X, "norm_prep", s);
    launch_norm_prep(d_img, istride, pw, ph, f->step, f->ms[mi]->w, f->ms[mi]->h, cn, g->states + mi,
                     f->dscr[cn].p, s);
  }
  for (size_t bi = 0; bi < f->buckets.size(); bi++) {
    const FleetBucket& b = f->buckets[bi];
    BucketGeom& bg = *g->bg[bi];
    const Plan& p = bg.p;
    PlanTables pt;
    CHK(plan_tables(X, p, bg.dh2, pt));
    int rwF = pw - b.tw + 1, rhF = ph - b.th + 1;
    int ntx = (rwF + p.blockW - 1) / p.blockW, nty = (rhF + p.blockH - 1) / p.blockH, ntiles = ntx * nty;
    TileGrid gF{};
    gF.ntx = ntx;
    gF.nty = nty;
    gF.blockW = p.blockW;
    gF.blockH = p.blockH;
    gF.extW = rwF;
    gF.extH = rhF;
    gF.padW = b.tw - 1;
    gF.padH = b.th - 1;
    gF.lastY0 = ((rhF - 1) / p.blockH) * p.blockH;
    gF.dftW = p.dftW;
    gF.log2W = ilog2(p.dftW);
    gF.hw = p.hw;
    gF.hwPad = bg.hwPad;
    gF.dftH = p.dftH;
    gF.log2H = ilog2(p.dftH);
    gF.dftH2 = bg.dh2;
    gF.log2H2 = bg.dh2 ? ilog2(bg.dh2) : 0;
    int64_t chan = (int64_t)p.dftH * bg.hwPad;
    int64_t ftile = chan * bg.cnMax;
    CHK(f->frame.ensure(sizeof(float2) * (size_t)ftile * ntiles));
    float2* frame = f->frame.as<float2>();
    CallState* fst = g->frame_state + bi;
    {
      // Phase 1 fleet.go:240-263: the parent's spectrum per tile, once
      Stage S(X, "fleet_fwd", s);
      launch_row_fwd(d_img, istride, f->step, bg.cnMax, fst, gF, pt, frame, ftile, chan, ntiles, s);
      launch_cols(COL_FWD, frame, frame, ftile, ftile, chan, bg.cnMax, fst, gF, pt, nullptr, nullptr, 0, 0, 1.0f,
                  ntiles, s);
    }
    // Phase 2 fleet.go:265-328: each member's tail.  Members go in pairs
    // through one column launch (the frame slab read once for both) when the
    // bucket's tiles are full height; each member's row pass and scoring
    // then follow in member order.
    const bool pairs = b.members.size() > 1 && bg.dh2 == 0;
    CHK(f->mspec.ensure(sizeof(float2) * (size_t)ftile * ntiles));
    if (pairs) CHK(f->mspec2.ensure(sizeof(float2) * (size_t)ftile * ntiles));
    float2* mspecs[2] = {f->mspec.as<float2>(), pairs ? f->mspec2.as<float2>() : nullptr};
    int64_t tchan2 = (int64_t)bg.dh2 * bg.hwPad;
    auto grid_of = [&](size_t j) {
      TileGrid gM = gF;
      gM.extW = pw - f->ms[b.members[j]]->w + 1;
      gM.extH = ph - f->ms[b.members[j]]->h + 1;
      return gM;
    };
    auto cols_one = [&](size_t j, float2* mspec) {
      const int mi = b.members[j];
      Stage S(X, "fleet_cols", s);
      launch_cols(COL_MULINV, frame, mspec, ftile, ftile, chan, bg.cns[j], g->states + mi, grid_of(j), pt,
                  bg.ts[j].spec, bg.ts[j].spec2, chan, tchan2, 1.0f, ntiles, s);
    };
    auto tail = [&](size_t j, const float2* mspec) -> int {
      const int mi = b.members[j];
      cvm_matcher* m = f->ms[mi];
      const int cn = bg.cns[j];
      const int rw = pw - m->w + 1, rh = ph - m->h + 1;
      CallState* mst = g->states + mi;
      const TileGrid gM = grid_of(j);
      const bool batched = NP.corr_off[mi] >= 0;
      float* corr;
      if (batched) {
        corr = f->pcorr.as<float>() + NP.corr_off[mi];
      } else {
        CHK(f->corr.ensure(sizeof(float) * (size_t)rw * rh));
        corr = f->corr.as<float>();
      }
      {
        Stage S(X, "fleet_row_inv", s);
        launch_row_inv(mspec, ftile, chan, mst, cn, gM, pt, corr, rw, ntiles, s);
      }
      if (!batched) {  // past the wave form: the int64 normalisation, on its own
        int nparts = norm_partials_count(rw, rh, m->w, m->h, cn);
        CHK(X.parts.ensure(sizeof(Extrema) * (size_t)nparts));
        Stage S(X, "fleet_normalize", s);
        launch_normalize(d_img, istride, pw, ph, f->step, m->w, m->h, rw, rh, cn, mst, corr, nullptr,
                         X.parts.as<Extrema>(), nullptr, s);
        launch_reduce(X.parts.as<Extrema>(), nparts, mst, rw, d_res + mi, nullptr, rh, s);
      }
      if (NP.last_of[mi] >= 0) {  // this member's map completes a batch
        const PlainNormPlan::Batch& B = NP.batches[NP.last_of[mi]];
        Stage S(X, "fleet_normalize", s);
        for (auto& gq : B.groups)
          launch_norm_batch(d_img, istride, pw, ph, f->step, gq[0], gq[1], false, NP.jobs.as<NormJob>(),
                            NP.wg.as<int>() + gq[2], gq[3], s);
        launch_reduce_batch(NP.jobs.as<NormJob>() + B.job0, B.njobs, s);
      }
      return 0;
    };
    for (size_t j = 0; j < b.members.size();) {
      bool paired = false;
      if (pairs && j + 1 < b.members.size()) {
        const TileGrid gA = grid_of(j), gB = grid_of(j + 1);
        const ColPairMember A{g->states + b.members[j], bg.ts[j].spec, mspecs[0], gA.extW, gA.extH};
        const ColPairMember Bm{g->states + b.members[j + 1], bg.ts[j + 1].spec, mspecs[1], gB.extW, gB.extH};
        Stage S(X, "fleet_cols", s);
        paired = launch_cols_pair(frame, ftile, chan, std::max(bg.cns[j], bg.cns[j + 1]), gF, pt, chan, ntiles, A,
                                  Bm, s);
      }
      if (paired) {
        CHK(tail(j, mspecs[0]));
        CHK(tail(j + 1, mspecs[1]));
        j += 2;
      } else {
        cols_one(j, mspecs[0]);
        CHK(tail(j, mspecs[0]));
        j += 1;
      }
    }
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------- direct correlation --
//
// The mixed-precision form: the window x template sum exactly on int8 MFMA
// (launch_direct), every score the FP64 one rounded to float32 (the exact
// mode of the wave-form normalisation).  Not bit-identical to the FFT path
// (whose float32 rounding it does not share); within float32 rounding of
// the FP64 oracle instead.

// Waves per direct workgroup when nothing is measured: CVMATCH_DIRECT_WAVES
// (1, 2 or 4), else 1.
static int direct_waves_default() {
  static int w = [] {
    const char* v = getenv("CVMATCH_DIRECT_WAVES");
    int x = v ? atoi(v) : 1;
    return direct_waves_ok(x) ? x : 1;
  }();
  return w;
}

static int check_direct(int tw, int th, int cnMax, int kstep) {
  if ((int64_t)tw * th * cnMax >= 131072)
    return fail(CVM_E_BAD_ARGS, "cvmatch: template %dx%d (cn=%d) exceeds the direct-correlation bound area*cn < 131072",
                tw, th, cnMax);
  if (tw > 257 || th > 257 || !norm_wave_ok(tw, th, cnMax) || !direct_shape_ok(tw, kstep))
    return fail(CVM_E_BAD_ARGS, "cvmatch: template %dx%d is outside the direct-correlation shapes", tw, th);
  return 0;
}

// Direct correlation + exact normalisation + extrema of one template whose
// B fragments are ready; the call state is final on s.  corr receives the
// int32 C' map (the caller's map, rescored in place, or scratch) and parts
// the extrema partials (norm_partials_max entries).
static int run_direct(Ctx& X, hipStream_t s, const CallState* st, const uint8_t* d_img, int64_t istride, int iw,
                      int ih, int tw, int th, int step, int kstep, int cnMax, const int8_t* bf, void* nscratch,
                      bool prep, float* d_map, float* corr, Extrema* parts, Res6* d_res) {
  int rw = iw - tw + 1, rh = ih - th + 1;
  // waves per workgroup: measured per shape on the first call (the C' map
  // the trial launches write is rewritten below); the allocation-free call
  // takes the default
  int waves = direct_waves_default();
  if (!X.noprof) {
    char what[96];
    snprintf(what, sizeof(what), "direct.waves/k%d/c%d/%dx%d/%dx%d", kstep, direct_cmax(direct_nk(tw, kstep), kstep),
             iw, ih, tw, th);
    waves = tune_choice(what, {1, 2, 4}, "CVMATCH_DIRECT_WAVES", [&](int w) {
      return time_launches(
          s, [&] { return launch_direct(d_img, istride, iw, ih, kstep, bf, tw, th, rw, rh, (int32_t*)corr, w, s); },
          1);
    });
  }
  {
    Stage S(X, "direct", s);
    if (launch_direct(d_img, istride, iw, ih, kstep, bf, tw, th, rw, rh, (int32_t*)corr, waves, s) != 0)
      return fail(CVM_E_BAD_ARGS, "cvmatch: template %dx%d is outside the direct-correlation shapes", tw, th);
  }
  if (prep) {
    Stage S(X, "norm_prep", s);
    launch_norm_prep(d_img, istride, iw, ih, step, tw, th, cnMax, st, nscratch, s);
  }
  const int q = norm_q_for(X, s, d_img, istride, iw, ih, step, tw, th, cnMax, st, corr, parts, nscratch, true);
  {
    Stage S(X, "normalize", s);
    launch_normalize(d_img, istride, iw, ih, step, tw, th, rw, rh, cnMax, st, corr, d_map, parts, nscratch, s, true,
                     q);
  }
  {
    Stage S(X, "reduce", s);
    launch_reduce(parts, norm_partials_count(rw, rh, tw, th, cnMax, q), st, rw, d_res, nullptr, rh, s);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// The scratch of one direct-form matchU8 call (run_match_direct): its call
// state, the template's B fragments, the normalisation's chunk sums, the C'
// map when the caller passes none, the extrema partials and the six-tuple.
struct DirectScratch {
  CallState* st;
  int8_t* bf;
  void* nscratch;
  float* corr;
  Extrema* parts;
  Res6* res;
};

// Byte sizes of DirectScratch's buffers after the state (which is
// sizeof(CallState)) in the order bf, nscratch, corr, parts; 0 = none.
static void direct_scratch_sizes(int iw, int ih, int tw, int th, int cnMax, int kstep, bool want_corr,
                                 size_t out[4]) {
  const int rw = iw - tw + 1, rh = ih - th + 1;
  out[0] = direct_bfrag_bytes(tw, th, kstep);
  out[1] = norm_scratch_bytes(iw, ih, tw, th, cnMax);
  out[2] = want_corr ? sizeof(float) * (size_t)rw * rh : 0;
  out[3] = sizeof(Extrema) * (size_t)norm_partials_max(rw, rh, tw, th, cnMax);
}

// The argument rules of the direct form; sets cnMax and kstep.
static int direct_args(int tw, int th, int cn_mode, int step, int* cnMax, int* kstep) {
  // the staging loads read 1 byte per pixel (gray) or 4 (RGBA): a packed
  // 2- or 3-byte layout would be read past its rows
  if (!((step == 1 && cn_mode == 1) || (step == 4 && (cn_mode == 0 || cn_mode == 3 || cn_mode == 4))))
    return fail(CVM_E_BAD_ARGS, "cvmatch: the direct form takes cn=1 step=1 or cn=0/3/4 step=4 (got cn=%d step=%d)",
                cn_mode, step);
  *cnMax = cn_mode ? cn_mode : 4;
  // with the alpha rule cn is decided on the device: stage all four bytes
  // (the fragments zero the alpha column when cn comes out 3)
  *kstep = direct_kstep(step, *cnMax);
  return check_direct(tw, th, *cnMax, *kstep);
}

// matchU8 in the direct form on device buffers with its scratch given
// (cn_mode 0 = the RGBA alpha rule, decided on the device; the B fragments
// then read cn from the state).  Allocates nothing.
static int direct_pipeline(Ctx& X, hipStream_t s, const uint8_t* d_img, int64_t istride, int iw, int ih,
                           const uint8_t* d_tpl, int64_t tstride, int tw, int th, int cn_mode, int step, int cnMax,
                           int kstep, const DirectScratch& B, float* d_map) {
  {
    Stage S(X, "stats", s);
    launch_state_init(B.st, s);
    launch_call_state(d_img, istride, iw, ih, d_tpl, tstride, tw, th, cnMax, step, cn_mode, B.st, s);
  }
  {
    Stage S(X, "bfrag", s);
    launch_bfrag(d_tpl, tstride, tw, th, kstep, cnMax, B.st, B.bf, s);
  }
  return run_direct(X, s, B.st, d_img, istride, iw, ih, tw, th, step, kstep, cnMax, B.bf, B.nscratch, true, d_map,
                    d_map ? d_map : B.corr, B.parts, B.res);
}

// matchU8 in the direct form on device buffers, scratch from the context.
static int run_match_direct(Ctx& X, hipStream_t s, const uint8_t* d_img, int64_t istride, int iw, int ih,
                            const uint8_t* d_tpl, int64_t tstride, int tw, int th, int cn_mode, int step,
                            float* d_map) {
  int cnMax, kstep;
  CHK(direct_args(tw, th, cn_mode, step, &cnMax, &kstep));
  size_t sz[4];
  direct_scratch_sizes(iw, ih, tw, th, cnMax, kstep, d_map == nullptr, sz);
  CHK(X.dbf.ensure(sz[0]));
  CHK(X.nscratch.ensure(sz[1]));
  if (sz[2]) CHK(X.corr.ensure(sz[2]));
  CHK(X.parts.ensure(sz[3]));
  const DirectScratch B{X.st, X.dbf.as<int8_t>(), X.nscratch.p, X.corr.as<float>(), X.parts.as<Extrema>(),
                        res_out(X)};
  return direct_pipeline(X, s, d_img, istride, iw, ih, d_tpl, tstride, tw, th, cn_mode, step, cnMax, kstep, B, d_map);
}

// B fragments of a matcher for cn channels (cached; cn known on the host).
static int matcher_bfrag(Ctx& X, hipStream_t s, cvm_matcher* m, int cn, const int8_t** out) {
  std::lock_guard<std::mutex> lk(m->mu);
  if (!m->dbf_ok[cn]) {
    const int kstep = direct_kstep(m->step, cn);
    CHK(m->dbf[cn].ensure(direct_bfrag_bytes(m->w, m->h, kstep)));
    launch_bfrag(m->d_pix, (int64_t)m->w * m->step, m->w, m->h, kstep, cn, nullptr, m->dbf[cn].as<int8_t>(), s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));  // built before another stream may read it
    m->dbf_ok[cn] = true;
  }
  *out = m->dbf[cn].as<int8_t>();
  return 0;
}

// Matcher.Find in the direct form.  With a constant-alpha RGBA template the
// channel count depends on the parent's alpha, so its B fragments are built
// for the call from the device state; otherwise they are cached.
static int matcher_find_direct_dev(Ctx& X, hipStream_t s, cvm_matcher* m, const uint8_t* d_img, int64_t istride,
                                   int iw, int ih, Res6* d_res) {
  int cnMax = m->step;
  const int kstep = direct_kstep(m->step, cnMax);
  CHK(check_direct(m->w, m->h, cnMax, kstep));
  CHK(X.nscratch.ensure(norm_scratch_bytes(iw, ih, m->w, m->h, cnMax)));
  matcher_state(X, s, X.st, m, d_img, istride, iw, ih);
  const int8_t* bf = nullptr;
  if (m->step == 4 && m->aConst) {
    CHK(X.dbf.ensure(direct_bfrag_bytes(m->w, m->h, kstep)));
    launch_bfrag(m->d_pix, (int64_t)m->w * m->step, m->w, m->h, kstep, 4, X.st, X.dbf.as<int8_t>(), s);
    bf = X.dbf.as<int8_t>();
  } else {
    CHK(matcher_bfrag(X, s, m, cnMax, &bf));
  }
  const int rw = iw - m->w + 1, rh = ih - m->h + 1;
  CHK(X.corr.ensure(sizeof(float) * (size_t)rw * rh));
  CHK(X.parts.ensure(sizeof(Extrema) * (size_t)norm_partials_max(rw, rh, m->w, m->h, cnMax)));
  return run_direct(X, s, X.st, d_img, istride, iw, ih, m->w, m->h, m->step, kstep, cnMax, bf, X.nscratch.p, true,
                    nullptr, X.corr.as<float>(), X.parts.as<Extrema>(), d_res);
}

// Fleet.FindAll in the direct form: every non-flat member against the frame
// (results independent of the bucketing, each equal to the member's direct
// Find).  The frame's chunk column sums are built once per channel count.
// Fused form: one launch per (kstep, cmax, sq32) group correlates and scores
// every member's blocks (largest member first), one reduces their extrema.
// Unfused form (frames whose layout the fused kernel does not take, or when
// it measures slower): one launch per (kstep, cmax) group writes each
// member's C' map, one per wave-form variant scores them, one reduces.
static int fleet_find_direct_dev(Ctx& X, hipStream_t s, cvm_fleet* f, const uint8_t* d_img, int64_t istride,
                                 int pw, int ph, int pConst, Res6* d_res) {
  size_t n = f->ms.size();
  CHK(f->dstates.ensure(sizeof(CallState) * n));
  CallState* states = f->dstates.as<CallState>();
  auto cn_of = [&](const cvm_matcher* m) { return m->step == 4 ? ((m->aConst && pConst) ? 3 : 4) : 1; };
  for (size_t i = 0; i < n; i++)
    if (f->ms[i]->varSum != 0) {
      cvm_matcher* m = f->ms[i];
      CHK(check_direct(m->w, m->h, cn_of(m), direct_kstep(m->step, cn_of(m))));
    }
  if (f->dpConst != pConst) {
    for (size_t i = 0; i < n; i++) {
      cvm_matcher* m = f->ms[i];
      if (m->varSum == 0) continue;
      launch_state_set(states + i, m->stats, 0, s);
      launch_state_finalize(states + i, m->w, m->h, cn_of(m), s);
    }
    HIPCHK(hipGetLastError());
    f->dpConst = pConst;
  }
  bool used[5] = {false, false, false, false, false};
  std::vector<size_t> live;
  for (size_t i = 0; i < n; i++)
    if (f->ms[i]->varSum != 0) {
      live.push_back(i);
      used[cn_of(f->ms[i])] = true;
    }
  if (live.empty()) return 0;
  for (int cn = 1; cn <= 4; cn++)
    if (used[cn]) CHK(f->dscr[cn].ensure(norm_scratch_bytes(pw, ph, 1, 1, cn)));

  bool moved = false;
  for (int cn = 1; cn <= 4; cn++)
    if (used[cn] && f->bj_ps[cn] != f->dscr[cn].p) moved = true;
  if (moved || f->bj_pw != pw || f->bj_ph != ph || f->bj_pConst != pConst) {
    // job tables for this frame shape
    std::vector<int> order(live.size());
    for (size_t k = 0; k < live.size(); k++) order[k] = (int)k;
    std::vector<int64_t> cost(live.size());
    std::vector<int> ks(live.size()), cms(live.size()), sq32s(live.size());
    for (size_t k = 0; k < live.size(); k++) {
      cvm_matcher* m = f->ms[live[k]];
      int rw = pw - m->w + 1, rh = ph - m->h + 1;
      ks[k] = direct_kstep(m->step, cn_of(m));
      cms[k] = direct_cmax(direct_nk(m->w, ks[k]), ks[k]);
      int q = 0;
      norm_form(m->w, m->h, cn_of(m), &q, &sq32s[k]);
      cost[k] = (int64_t)m->h * direct_nk(m->w, ks[k]) * direct_grid_x(rw) * direct_grid_y(rh);
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    for (int cn = 1; cn <= 4; cn++) {  // the scoring reads the chunk sums
      if (!used[cn]) continue;
      for (size_t i : live)
        if (cn_of(f->ms[i]) == cn) {
          launch_norm_prep(d_img, istride, pw, ph, f->step, f->ms[i]->w, f->ms[i]->h, cn, states + i,
                           f->dscr[cn].p, s);
          break;
        }
    }

    // the chunk sums of a group's channel count (every member of a group has
    // the same kstep, hence the same channel count)
    auto ps_of = [&](int kstep) -> const uint32_t* {
      for (size_t k = 0; k < live.size(); k++)
        if (ks[k] == kstep) return f->dscr[cn_of(f->ms[live[k]])].as<uint32_t>();
      return nullptr;
    };

    // The fused form: every member's blocks correlated and scored in one
    // launch per (kstep, cmax, sq32) group; no C' maps.  Available when the
    // frame's layout allows it for every member.
    bool fusable = true;
    for (size_t k = 0; k < live.size(); k++)
      if (!fused_ok(d_img, istride, pw, ks[k], cms[k], sq32s[k] != 0)) fusable = false;
    // merge: members of up to five chunks per class share one launch per
    // (kstep, sq32) at the largest such class count of the Fleet (their
    // MFMA loops skip the classes they lack); otherwise one launch per
    // (kstep, cmax, sq32).  Wider members keep their own launches, since
    // their scoring strips need eight columns per lane.
    auto build_fused = [&](int merge) -> int {
      std::vector<int> ecm(cms);
      if (merge) {
        int top[5][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
        for (size_t k = 0; k < live.size(); k++)
          if (cms[k] <= 5) top[ks[k]][sq32s[k]] = std::max(top[ks[k]][sq32s[k]], cms[k]);
        for (size_t k = 0; k < live.size(); k++)
          if (cms[k] <= 5) ecm[k] = top[ks[k]][sq32s[k]];
      }
      std::vector<FusedJob> fj;
      std::vector<NormJob> rj(live.size());
      std::vector<int> fwg;
      std::vector<std::array<int, 7>> fgroups;  // (kstep, cmax, sq32, first wg, wg count, max pitch, waves)
      std::vector<size_t> pt_off(live.size());
      size_t pt_total = 0;
      for (size_t k = 0; k < live.size(); k++) {
        cvm_matcher* m = f->ms[live[k]];
        pt_off[k] = pt_total;
        pt_total += (size_t)direct_grid_x(pw - m->w + 1) * direct_grid_y(ph - m->h + 1);
      }
      CHK(f->bj_fparts.ensure(sizeof(Extrema) * pt_total));
      Extrema* parts = f->bj_fparts.as<Extrema>();
      for (int kst : {4, 3, 1})
        for (int cm : {9, 6, 5, 4, 3, 2})
          for (int sq : {1, 0}) {
            int first = (int)fwg.size(), max_pitch = 0;
            for (int k : order) {
              if (ks[k] != kst || ecm[k] != cm || sq32s[k] != sq) continue;
              cvm_matcher* m = f->ms[live[k]];
              int rw = pw - m->w + 1, rh = ph - m->h + 1;
              const int8_t* bf = nullptr;
              CHK(matcher_bfrag(X, s, m, cn_of(m), &bf));
              FusedJob J{};
              J.bf = bf;
              J.st = states + live[k];
              J.parts = parts + pt_off[k];
              J.tw = m->w;
              J.th = m->h;
              J.nk = direct_nk(m->w, kst);
              J.pitch = direct_pitch_of(m->w, kst);
              J.rw = rw;
              J.rh = rh;
              J.gx = direct_grid_x(rw);
              J.wg0 = (int)fwg.size() - first;
              max_pitch = std::max(max_pitch, J.pitch);
              const int nb = J.gx * direct_grid_y(rh);
              for (int b = 0; b < nb; b++) fwg.push_back((int)fj.size());
              fj.push_back(J);
              NormJob R{};
              R.st = J.st;
              R.parts = J.parts;
              R.res = d_res + live[k];
              R.tw = m->w;
              R.th = m->h;
              R.rw = rw;
              R.rh = rh;
              R.gx = J.gx;
              R.gy = direct_grid_y(rh);
              rj[k] = R;
            }
            if ((int)fwg.size() > first)
              fgroups.push_back({kst, cm, sq, first, (int)fwg.size() - first, max_pitch, 1});
          }
      CHK(f->bj_fused.ensure(sizeof(FusedJob) * fj.size()));
      CHK(f->bj_fwg.ensure(sizeof(int) * fwg.size()));
      CHK(f->bj_fred.ensure(sizeof(NormJob) * rj.size()));
      HIPCHK(hipMemcpyAsync(f->bj_fused.p, fj.data(), sizeof(FusedJob) * fj.size(), hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(f->bj_fwg.p, fwg.data(), sizeof(int) * fwg.size(), hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(f->bj_fred.p, rj.data(), sizeof(NormJob) * rj.size(), hipMemcpyHostToDevice, s));
      HIPCHK(hipStreamSynchronize(s));  // the host vectors go out of scope
      // waves per workgroup of each group, measured on this frame
      for (auto& g : fgroups) {
        char what[112];
        snprintf(what, sizeof(what), "fleet.fused.waves/k%d/c%d/s%d/%dx%d/n%d", g[0], g[1], g[2], pw, ph, g[4]);
        g[6] = tune_choice(what, {1, 2, 4}, "CVMATCH_DIRECT_WAVES", [&](int w) {
          return time_launches(
              s,
              [&] {
                return launch_direct_fused(d_img, istride, pw, ph, ps_of(g[0]), g[0], g[1], g[2] != 0,
                                           f->bj_fused.as<FusedJob>(), f->bj_fwg.as<int>() + g[3], g[4], g[5], w,
                                           s);
              },
              1);
        });
      }
      f->bj_fgroups = fgroups;
      f->bj_nfred = (int)rj.size();
      return 0;
    };

    auto build_unfused = [&]() -> int {
      std::vector<DirectJob> dj;
      size_t cm_total = 0;
      std::vector<size_t> cm_off(live.size());
      for (size_t k = 0; k < live.size(); k++) {
        cvm_matcher* m = f->ms[live[k]];
        cm_off[k] = cm_total;
        cm_total += (size_t)(pw - m->w + 1) * (ph - m->h + 1);
      }
      CHK(f->bj_cmaps.ensure(sizeof(int32_t) * cm_total));
      int32_t* cmaps = f->bj_cmaps.as<int32_t>();
      // direct jobs grouped by (kstep, cmax), largest first inside a group
      std::vector<int> dwg;
      std::vector<std::array<int, 6>> dgroups;  // (kstep, cmax, first wg, wg count, max pitch, waves)
      for (int kst : {4, 3, 1})
        for (int cm : {9, 6, 5, 4, 3, 2}) {
          int first = (int)dwg.size(), max_pitch = 0;
          for (int k : order) {
            if (ks[k] != kst || cms[k] != cm) continue;
            cvm_matcher* m = f->ms[live[k]];
            int rw = pw - m->w + 1, rh = ph - m->h + 1;
            const int8_t* bf = nullptr;
            CHK(matcher_bfrag(X, s, m, cn_of(m), &bf));
            DirectJob J{};
            J.bf = bf;
            J.out = cmaps + cm_off[k];
            J.th = m->h;
            J.nk = direct_nk(m->w, kst);
            J.pitch = direct_pitch_of(m->w, kst);
            J.rw = rw;
            J.rh = rh;
            J.gx = direct_grid_x(rw);
            J.wg0 = (int)dwg.size() - first;
            max_pitch = std::max(max_pitch, J.pitch);
            int nb = J.gx * direct_grid_y(rh);
            for (int b = 0; b < nb; b++) dwg.push_back((int)dj.size());
            dj.push_back(J);
          }
          if ((int)dwg.size() > first) dgroups.push_back({kst, cm, first, (int)dwg.size() - first, max_pitch, 1});
        }
      CHK(f->bj_direct.ensure(sizeof(DirectJob) * dj.size()));
      CHK(f->bj_dwg.ensure(sizeof(int) * dwg.size()));
      HIPCHK(hipMemcpyAsync(f->bj_direct.p, dj.data(), sizeof(DirectJob) * dj.size(), hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(f->bj_dwg.p, dwg.data(), sizeof(int) * dwg.size(), hipMemcpyHostToDevice, s));
      HIPCHK(hipStreamSynchronize(s));  // the host vectors go out of scope
      // waves per workgroup of each group, measured on this frame (the C'
      // maps these trial launches write are rewritten by every FindAll)
      for (auto& g : dgroups) {
        char what[96];
        snprintf(what, sizeof(what), "fleet.direct.waves/k%d/c%d/%dx%d/n%d", g[0], g[1], pw, ph, g[3]);
        g[5] = tune_choice(what, {1, 2, 4}, "CVMATCH_DIRECT_WAVES", [&](int w) {
          return time_launches(
              s,
              [&] {
                return launch_direct_batch(d_img, istride, pw, ph, g[0], g[1], f->bj_direct.as<DirectJob>(),
                                           f->bj_dwg.as<int>() + g[2], g[3], g[4], w, s);
              },
              1);
        });
      }

      // Normalisation jobs grouped by wave-form variant (the reduction takes
      // the whole job array; a job's wg0 counts from its group's first entry).
      // qcap 4 gives each member the widest strip its template allows, qcap 2
      // half-width strips; the tuner times both on this frame.
      auto build_norm = [&](int qcap, std::vector<NormJob>& nj, std::vector<int>& nwg,
                            std::vector<std::array<int, 4>>& groups) -> size_t {
        nj.clear();
        nwg.clear();
        groups.clear();
        std::vector<int> qs(live.size()), sqs(live.size());
        std::vector<size_t> pt_off(live.size());
        size_t pt_total = 0;
        for (size_t k = 0; k < live.size(); k++) {
          cvm_matcher* m = f->ms[live[k]];
          norm_form(m->w, m->h, cn_of(m), &qs[k], &sqs[k]);
          qs[k] = std::min(qs[k], qcap);
          pt_off[k] = pt_total;
          pt_total += (size_t)norm_grid_x(pw - m->w + 1, qs[k]) * norm_grid_y(ph - m->h + 1);
        }
        Extrema* parts = f->bj_parts.as<Extrema>();
        for (int q : {4, 2})
          for (int sq : {1, 0}) {
            int first = (int)nwg.size();
            for (int k : order) {
              if (qs[k] != q || sqs[k] != sq) continue;
              cvm_matcher* m = f->ms[live[k]];
              int cn = cn_of(m);
              int rw = pw - m->w + 1, rh = ph - m->h + 1;
              NormJob J{};
              J.st = states + live[k];
              J.ps = f->dscr[cn].as<uint32_t>();
              J.corr = (const float*)(cmaps + cm_off[k]);
              J.parts = parts ? parts + pt_off[k] : nullptr;
              J.res = d_res + live[k];
              J.tw = m->w;
              J.th = m->h;
              J.rw = rw;
              J.rh = rh;
              J.gx = norm_grid_x(rw, q);
              J.gy = norm_grid_y(rh);
              J.wg0 = (int)nwg.size() - first;
              for (int b = 0; b < J.gx * J.gy; b++) nwg.push_back((int)nj.size());
              nj.push_back(J);
            }
            if ((int)nwg.size() > first) groups.push_back({q, sq, first, (int)nwg.size() - first});
          }
        return pt_total;
      };
      std::vector<NormJob> nj;
      std::vector<int> nwg;
      std::vector<std::array<int, 4>> groups;
      size_t pts = std::max(build_norm(4, nj, nwg, groups), build_norm(2, nj, nwg, groups));
      CHK(f->bj_parts.ensure(sizeof(Extrema) * pts));
      auto upload_norm = [&](int qcap) -> int {
        build_norm(qcap, nj, nwg, groups);
        CHK(f->bj_norm.ensure(sizeof(NormJob) * nj.size()));
        CHK(f->bj_nwg.ensure(sizeof(int) * nwg.size()));
        HIPCHK(hipMemcpyAsync(f->bj_norm.p, nj.data(), sizeof(NormJob) * nj.size(), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(f->bj_nwg.p, nwg.data(), sizeof(int) * nwg.size(), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        return 0;
      };
      char what[96];
      snprintf(what, sizeof(what), "fleet.norm.q/%dx%d/n%zu", pw, ph, live.size());
      int rc = 0;
      const int qcap = tune_choice(what, {4, 2}, "CVMATCH_NORM_Q", [&](int q) -> float {
        if (upload_norm(q) != 0) {
          rc = -1;
          return -1.0f;
        }
        return time_launches(
            s,
            [&] {
              for (auto& g : groups)
                launch_norm_batch(d_img, istride, pw, ph, f->step, g[0], g[1], true, f->bj_norm.as<NormJob>(),
                                  f->bj_nwg.as<int>() + g[2], g[3], s);
              return 0;
            },
            1);
      });
      if (rc) return rc;
      CHK(upload_norm(qcap));
      f->bj_dgroups = dgroups;
      f->bj_nnorm = (int)nj.size();
      f->bj_groups = groups;
      return 0;
    };
    auto run_fused = [&]() -> int {
      for (auto& g : f->bj_fgroups)
        if (launch_direct_fused(d_img, istride, pw, ph, ps_of(g[0]), g[0], g[1], g[2] != 0,
                                f->bj_fused.as<FusedJob>(), f->bj_fwg.as<int>() + g[3], g[4], g[5], g[6], s) != 0)
          return -1;
      launch_reduce_batch(f->bj_fred.as<NormJob>(), f->bj_nfred, s);
      return 0;
    };
    auto run_unfused = [&]() -> int {
      for (auto& g : f->bj_dgroups)
        if (launch_direct_batch(d_img, istride, pw, ph, g[0], g[1], f->bj_direct.as<DirectJob>(),
                                f->bj_dwg.as<int>() + g[2], g[3], g[4], g[5], s) != 0)
          return -1;
      for (auto& g : f->bj_groups)
        launch_norm_batch(d_img, istride, pw, ph, f->step, g[0], g[1], true, f->bj_norm.as<NormJob>(),
                          f->bj_nwg.as<int>() + g[2], g[3], s);
      launch_reduce_batch(f->bj_norm.as<NormJob>(), f->bj_nnorm, s);
      return 0;
    };
    // fused or unfused: both give the same bits; the tuner times a whole
    // frame of each (building the unfused form's C' maps only to time it)
    int use_fused = 0;
    bool unfused_built = false;  // the unfused tables of this frame shape
    if (fusable) {
      {
        char what[96];
        snprintf(what, sizeof(what), "fleet.fused.merge/%dx%d/n%zu", pw, ph, live.size());
        int rc = 0;
        const int merge = tune_choice(what, {1, 0}, "CVMATCH_FLEET_MERGE", [&](int mg) -> float {
          if (build_fused(mg) != 0) {
            rc = -1;
            return -1.0f;
          }
          return time_launches(s, run_fused, 1);
        });
        if (rc) return rc;
        CHK(build_fused(merge));
      }
      char what[96];
      snprintf(what, sizeof(what), "fleet.fused/%dx%d/n%zu", pw, ph, live.size());
      int rc = 0;
      use_fused = tune_choice(what, {1, 0}, "CVMATCH_FLEET_FUSED", [&](int fz) -> float {
        if (fz == 0) {
          if (build_unfused() != 0) {
            rc = -1;
            return -1.0f;
          }
          unfused_built = true;
        }
        return time_launches(s, fz ? std::function<int()>(run_fused) : std::function<int()>(run_unfused), 1);
      });
      if (rc) return rc;
    }
    if (use_fused) {
      // the unfused form's per-member maps are not kept
      for (DevBuf* b : {&f->bj_direct, &f->bj_dwg, &f->bj_norm, &f->bj_nwg, &f->bj_cmaps, &f->bj_parts})
        b->release();
      f->bj_dgroups.clear();
      f->bj_groups.clear();
      f->bj_nnorm = 0;
    } else {
      if (!unfused_built) CHK(build_unfused());
      for (DevBuf* b : {&f->bj_fused, &f->bj_fwg, &f->bj_fred, &f->bj_fparts}) b->release();
      f->bj_fgroups.clear();
      f->bj_nfred = 0;
    }
    f->bj_use_fused = use_fused;
    f->bj_pw = pw;
    f->bj_ph = ph;
    f->bj_pConst = pConst;
    for (int cn = 1; cn <= 4; cn++) f->bj_ps[cn] = f->dscr[cn].p;
  }

  for (int cn = 1; cn <= 4; cn++) {
    if (!used[cn]) continue;
    for (size_t i : live)
      if (cn_of(f->ms[i]) == cn) {
        Stage S(X, "norm_prep", s);
        launch_norm_prep(d_img, istride, pw, ph, f->step, f->ms[i]->w, f->ms[i]->h, cn, states + i, f->dscr[cn].p,
                         s);
        break;
      }
  }
  if (f->bj_use_fused) {
    Stage S(X, "fused", s);
    for (auto& g : f->bj_fgroups) {
      const uint32_t* ps = nullptr;  // the chunk sums of the group's channel count
      for (size_t i : live)
        if (direct_kstep(f->ms[i]->step, cn_of(f->ms[i])) == g[0]) {
          ps = f->dscr[cn_of(f->ms[i])].as<uint32_t>();
          break;
        }
      if (launch_direct_fused(d_img, istride, pw, ph, ps, g[0], g[1], g[2] != 0, f->bj_fused.as<FusedJob>(),
                              f->bj_fwg.as<int>() + g[3], g[4], g[5], g[6], s) != 0)
        return fail(CVM_E_BAD_ARGS, "cvmatch: fleet members outside the direct-correlation shapes");
    }
  } else {
    {
      Stage S(X, "direct", s);
      for (auto& g : f->bj_dgroups)
        if (launch_direct_batch(d_img, istride, pw, ph, g[0], g[1], f->bj_direct.as<DirectJob>(),
                                f->bj_dwg.as<int>() + g[2], g[3], g[4], g[5], s) != 0)
          return fail(CVM_E_BAD_ARGS, "cvmatch: fleet members outside the direct-correlation shapes");
    }
    {
      Stage S(X, "normalize", s);
      for (auto& g : f->bj_groups)
        launch_norm_batch(d_img, istride, pw, ph, f->step, g[0], g[1], true, f->bj_norm.as<NormJob>(),
                          f->bj_nwg.as<int>() + g[2], g[3], s);
    }
  }
  {
    Stage S(X, "reduce", s);
    if (f->bj_use_fused)
      launch_reduce_batch(f->bj_fred.as<NormJob>(), f->bj_nfred, s);
    else
      launch_reduce_batch(f->bj_norm.as<NormJob>(), f->bj_nnorm, s);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

}  // namespace cvm

// ====================================================================== ABI

using namespace cvm;

extern "C" {

const char* cvm_last_error_detail(void) { return g_err; }

const char* cvm_version(void) { return "cvmatch-hip 0.1 (gfx950)"; }

int64_t cvm_stats_cap(int32_t cn) { return stats_cap(cn); }

int64_t cvm_plan(int32_t tw, int32_t th, int32_t rw, int32_t rh, int32_t* out) {
  Plan p = make_plan(tw, th, rw, rh);
  out[0] = p.dftW;
  out[1] = p.dftH;
  out[2] = p.blockW;
  out[3] = p.blockH;
  return p.cost;
}

int cvm_match_u8(const uint8_t* img, int64_t istride, int32_t iw, int32_t ih, const uint8_t* tpl, int64_t tstride,
                 int32_t tw, int32_t th, int32_t cn, int32_t step, float* map, cvm_result* out) {
  if (!img || !tpl) return fail(CVM_E_NIL_IMAGE, "cvmatch: nil image");
  CHK(check_args(istride, iw, ih, tstride, tw, th, cn, step));
  return match_host(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, map, out);
}

static int rgba_args(const uint8_t* img, int32_t iw, int32_t ih, const uint8_t* tpl, int32_t tw, int32_t th) {
  if (!img || !tpl) return fail(CVM_E_NIL_IMAGE, "cvmatch: nil image");
  if (iw <= 0 || ih <= 0 || tw <= 0 || th <= 0) return fail(CVM_E_EMPTY_IMAGE, "cvmatch: empty image");
  return 0;
}

int cvm_match(const uint8_t* img, int64_t istride, int32_t iw, int32_t ih, const uint8_t* tpl, int64_t tstride,
              int32_t tw, int32_t th, cvm_result* out) {
  CHK(rgba_args(img, iw, ih, tpl, tw, th));
  // the shape checks of cn=4; the exact-statistics bound follows the alpha
  // rule (alpha_rule_cn)
  CHK(check_shape(istride, iw, ih, tstride, tw, th, 4, 4));
  return match_host(img, istride, iw, ih, tpl, tstride, tw, th, 0, 4, nullptr, out);
}

int cvm_match_map(const uint8_t* img, int64_t istride, int32_t iw, int32_t ih, const uint8_t* tpl, int64_t tstride,
                  int32_t tw, int32_t th, float* map, cvm_result* out) {
  CHK(rgba_args(img, iw, ih, tpl, tw, th));
  CHK(check_shape(istride, iw, ih, tstride, tw, th, 4, 4));
  if (!map) return fail(CVM_E_BAD_ARGS, "cvmatch: MatchMap needs an output map");
  return match_host(img, istride, iw, ih, tpl, tstride, tw, th, 0, 4, map, out);
}

int cvm_match_peaks(const uint8_t* img, int64_t istride, int32_t iw, int32_t ih, const uint8_t* tpl, int64_t tstride,
                    int32_t tw, int32_t th, int32_t cn, int32_t step, float thresh, int32_t rx, int32_t ry,
                    int32_t* xy, float* vals, int32_t cap, int32_t* count, cvm_result* out) {
  CHK(rgba_args(img, iw, ih, tpl, tw, th));
  if (cn == 0) {
    if (step != 4) return fail(CVM_E_BAD_ARGS, "cvmatch: the alpha rule (cn=0) needs step 4");
    CHK(check_shape(istride, iw, ih, tstride, tw, th, 4, 4));
  } else {
    CHK(check_args(istride, iw, ih, tstride, tw, th, cn, step));
  }
  PeakReq pk{thresh, rx, ry, cap, xy, vals, count};
  CHK(check_peaks(iw - tw + 1, ih - th + 1, pk.rx, pk.ry, cap, xy, vals, count));
  return match_host(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, nullptr, out, false, &pk);
}

int cvm_find_peaks_device(const float* d_map, int32_t w, int32_t h, float thresh, int32_t rx, int32_t ry,
                          int32_t* d_xy, float* d_vals, int32_t cap, int32_t* count, void* stream) {
  if (!d_map) return fail(CVM_E_NIL_IMAGE, "cvmatch: nil image");
  CHK(check_peaks(w, h, rx, ry, cap, d_xy, d_vals, count));
  LEASE(X);
  hipStream_t s = stream ? (hipStream_t)stream : X.stream;
  CHK(peaks_dev(X, s, d_map, w, h, thresh, rx, ry, cap, d_xy, d_vals));
  int tot = 0;
  CHK(peaks_total_host(X, s, w, h, &tot));
  X.stage_collect();
  *count = tot;
  return 0;
}

int cvm_match_gray(const uint8_t* img, int64_t istride, int32_t iw, int32_t ih, int32_t istep, const uint8_t* tpl,
                   int64_t tstride, int32_t tw, int32_t th, int32_t tstep, cvm_result* out) {
  CHK(rgba_args(img, iw, ih, tpl, tw, th));
  if ((istep != 1 && istep != 4) || (tstep != 1 && tstep != 4))
    return fail(CVM_E_BAD_ARGS, "cvmatch: gray input step must be 1 or 4");
  if (istride < (int64_t)iw * istep || tstride < (int64_t)tw * tstep)
    return fail(CVM_E_BAD_ARGS, "cvmatch: bad match arguments (%dx%d in %dx%d, cn=%d step=%d)", tw, th, iw, ih, 1, 1);
  CHK(check_args(iw, iw, ih, tw, tw, th, 1, 1));
  LEASE(X);
  hipStream_t s = X.stream;
  size_t ib = (size_t)img_bytes(istride, iw, ih, istep), tb = (size_t)img_bytes(tstride, tw, th, tstep);
  CHK(X.img.ensure(ib));
  CHK(X.tpl.ensure(tb));
  {
    Stage S(X, "h2d", s);
    HIPCHK(xfer_h2d(X.img.p, img, ib, s));
    HIPCHK(xfer_h2d(X.tpl.p, tpl, tb, s));
  }
  const uint8_t* gi = X.img.as<uint8_t>();
  const uint8_t* gt = X.tpl.as<uint8_t>();
  int64_t gis = istride, gts = tstride;
  if (istep == 4 || tstep == 4) {
    Stage S(X, "gray", s);
    if (istep == 4) {
      CHK(X.gray0.ensure((size_t)iw * ih));
      launch_rgba_to_gray(X.img.as<uint8_t>(), istride, iw, ih, X.gray0.as<uint8_t>(), s);
      gi = X.gray0.as<uint8_t>();
      gis = iw;
    }
    if (tstep == 4) {
      CHK(X.gray1.ensure((size_t)tw * th));
      launch_rgba_to_gray(X.tpl.as<uint8_t>(), tstride, tw, th, X.gray1.as<uint8_t>(), s);
      gt = X.gray1.as<uint8_t>();
      gts = tw;
    }
  }
  CHK(run_match(X, s, gi, gis, iw, ih, gt, gts, tw, th, 1, 1, nullptr));
  return finish_result(X, s, out);
}

static int match_device(const uint8_t* d_img, int64_t istride, int32_t iw, int32_t ih, const uint8_t* d_tpl,
                        int64_t tstride, int32_t tw, int32_t th, int32_t cn, int32_t step, float* d_map,
                        cvm_result* out, void* stream, bool direct);

int cvm_match_u8_device(const uint8_t* d_img, int64_t istride, int32_t iw, int32_t ih, const uint8_t* d_tpl,
                        int64_t tstride, int32_t tw, int32_t th, int32_t cn, int32_t step, float* d_map,
                        cvm_result* out, void* stream) {
  return match_device(d_img, istride, iw, ih, d_tpl, tstride, tw, th, cn, step, d_map, out, stream, false);
}

int cvm_match_u8_direct_device(const uint8_t* d_img, int64_t istride, int32_t iw, int32_t ih, const uint8_t* d_tpl,
                               int64_t tstride, int32_t tw, int32_t th, int32_t cn, int32_t step, float* d_map,
                               cvm_result* out, void* stream) {
  return match_device(d_img, istride, iw, ih, d_tpl, tstride, tw, th, cn, step, d_map, out, stream, true);
}

int cvm_match_u8_direct(const uint8_t* img, int64_t istride, int32_t iw, int32_t ih, const uint8_t* tpl,
                        int64_t tstride, int32_t tw, int32_t th, int32_t cn, int32_t step, float* map,
                        cvm_result* out) {
  if (!img || !tpl) return fail(CVM_E_NIL_IMAGE, "cvmatch: nil image");
  if (cn == 0) {
    if (step != 4) return fail(CVM_E_BAD_ARGS, "cvmatch: the alpha rule (cn=0) needs step 4");
    CHK(check_shape(istride, iw, ih, tstride, tw, th, 4, 4));
  } else {
    CHK(check_args(istride, iw, ih, tstride, tw, th, cn, step));
  }
  return match_host(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, map, out, true);
}

}  // extern "C"

static int match_device(const uint8_t* d_img, int64_t istride, int32_t iw, int32_t ih, const uint8_t* d_tpl,
                        int64_t tstride, int32_t tw, int32_t th, int32_t cn, int32_t step, float* d_map,
                        cvm_result* out, void* stream, bool direct) {
  if (!d_img || !d_tpl) return fail(CVM_E_NIL_IMAGE, "cvmatch: nil image");
  if (cn == 0) {
    if (step != 4) return fail(CVM_E_BAD_ARGS, "cvmatch: the alpha rule (cn=0) needs step 4");
    CHK(check_shape(istride, iw, ih, tstride, tw, th, 4, 4));
  } else {
    CHK(check_args(istride, iw, ih, tstride, tw, th, cn, step));
  }
  LEASE(X);
  hipStream_t s = stream ? (hipStream_t)stream : X.stream;
  if (cn == 0) CHK(alpha_rule_cn(X, s, d_img, istride, iw, ih, d_tpl, tstride, tw, th, &cn));
  auto run = [&]() -> int {
    if (direct) return run_match_direct(X, s, d_img, istride, iw, ih, d_tpl, tstride, tw, th, cn, step, d_map);
    return run_match(X, s, d_img, istride, iw, ih, d_tpl, tstride, tw, th, cn, step, d_map);
  };
  const std::vector<uint64_t> key = {u64p(d_img), u64p(d_tpl), u64p(d_map), (uint64_t)istride, (uint64_t)tstride,
                                     (uint64_t)iw, (uint64_t)ih, (uint64_t)tw, (uint64_t)th, (uint64_t)cn,
                                     (uint64_t)step, direct ? 1u : 0u};
  // small parents only: a trial of the eager way costs a large call far more
  // than it can gain
  GraphSlot::Entry* timed = nullptr;
  const bool small = (long long)iw * ih <= compact_max_pixels();
  const auto t0 = std::chrono::steady_clock::now();
  CHK(run_graphed(X.graph, key, s, run, small ? &timed : nullptr));
  CHK(finish_result(X, s, out));
  graph_note(timed, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  return 0;
}

// ---------------------------------------------------- allocation-free call --
//
// cvm_match_u8_direct_device_noalloc: the direct form's device-resident call
// for callers that cannot allocate (code under stream capture, serving loops
// that forbid allocation).  Its scratch -- call state, B fragments, chunk
// sums, the C' map when no map is passed, extrema partials, six-tuple -- is
// carved out of one file-scope device array of kNoAllocBytes; each device
// holds its own copy, whose address is looked up once per device with
// hipGetSymbolAddress (the array's name in host code is only a stand-in).
// Calls on one device are serialised by a lock held until their work has
// completed, and each call rewrites the state it reads (the call state is
// cleared by a kernel), so nothing carries over from the previous call.

constexpr size_t kNoAllocBytes = (size_t)64 << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {
constexpr int kNoAllocDevices = 64;
struct NoAllocDevice {
  std::mutex mu;
  uint8_t* ptr = nullptr;  // this device's copy of scratch
};
NoAllocDevice g_noalloc[kNoAllocDevices];
}  // namespace

static int noalloc_call(const uint8_t* d_img, int64_t istride, int32_t iw, int32_t ih, const uint8_t* d_tpl,
                        int64_t tstride, int32_t tw, int32_t th, int32_t cn, int32_t step, float* d_map,
                        cvm_result* out, void* stream) {
  if (!d_img || !d_tpl) return fail(CVM_E_NIL_IMAGE, "cvmatch: nil image");
  if (!out) return fail(CVM_E_BAD_ARGS, "cvmatch: nil result");
  if (cn == 0) {
    if (step != 4) return fail(CVM_E_BAD_ARGS, "cvmatch: the alpha rule (cn=0) needs step 4");
    CHK(check_shape(istride, iw, ih, tstride, tw, th, 4, 4));
  } else {
    CHK(check_args(istride, iw, ih, tstride, tw, th, cn, step));
  }
  int cnMax, kstep;
  CHK(direct_args(tw, th, cn, step, &cnMax, &kstep));
  size_t sz[4];
  direct_scratch_sizes(iw, ih, tw, th, cnMax, kstep, d_map == nullptr, sz);
  auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
  const size_t off_res = up(sizeof(CallState)), off_bf = off_res + up(sizeof(Res6));
  const size_t off_ns = off_bf + up(sz[0]), off_corr = off_ns + up(sz[1]), off_parts = off_corr + up(sz[2]);
  const size_t total = off_parts + up(sz[3]);
  if (total > kNoAllocBytes)
    return fail(CVM_E_UNSUPPORTED,
                "cvmatch: the allocation-free call needs %zu bytes of scratch for %dx%d in %dx%d; its limit is %zu",
                total, tw, th, iw, ih, kNoAllocBytes);
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices) return fail(CVM_E_UNSUPPORTED, "cvmatch: device ordinal %d", dev);
  NoAllocDevice& D = g_noalloc[dev];
  std::lock_guard<std::mutex> lk(D.mu);
  if (!D.ptr) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(CVM_E_HIP, "cvmatch: no device address for the allocation-free scratch");
    D.ptr = (uint8_t*)p;
  }
  uint8_t* base = D.ptr;
  const DirectScratch B{(CallState*)base,          (int8_t*)(base + off_bf),       base + off_ns,
                        (float*)(base + off_corr), (Extrema*)(base + off_parts), (Res6*)(base + off_res)};
  static Ctx quiet = [] {
    Ctx c;
    c.noprof = true;
    return c;
  }();
  hipStream_t s = (hipStream_t)stream;
  CHK(direct_pipeline(quiet, s, d_img, istride, iw, ih, d_tpl, tstride, tw, th, cn, step, cnMax, kstep, B, d_map));
  HIPCHK(hipMemcpyAsync(out, B.res, sizeof(cvm_result), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

extern "C" {

int cvm_match_u8_direct_device_noalloc(const uint8_t* d_img, int64_t istride, int32_t iw, int32_t ih,
                                       const uint8_t* d_tpl, int64_t tstride, int32_t tw, int32_t th, int32_t cn,
                                       int32_t step, float* d_map, cvm_result* out, void* stream) {
  return noalloc_call(d_img, istride, iw, ih, d_tpl, tstride, tw, th, cn, step, d_map, out, stream);
}

int cvm_alpha_const(const uint8_t* pix, int64_t stride, int32_t w, int32_t h, int32_t* out) {
  if (!pix) return fail(CVM_E_NIL_IMAGE, "cvmatch: nil image");
  if (w <= 0 || h <= 0) return fail(CVM_E_EMPTY_IMAGE, "cvmatch: empty image");
  if (stride < (int64_t)w * 4) return fail(CVM_E_BAD_ARGS, "cvmatch: stride %lld below 4*width", (long long)stride);
  LEASE(X);
  hipStream_t s = X.stream;
  size_t ib = (size_t)img_bytes(stride, w, h, 4);
  CHK(X.img.ensure(ib));
  HIPCHK(xfer_h2d(X.img.p, pix, ib, s));
  launch_state_init(X.st, s);
  launch_alpha_scan(X.img.as<uint8_t>(), stride, w, h, X.st, 0, s);
  CallState hs;
  HIPCHK(hipMemcpyAsync(&hs, X.st, sizeof(CallState), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *out = hs.alpha_bad[0] == 0;
  return 0;
}

int cvm_rgba_to_gray_device(const uint8_t* d_pix, int64_t stride, int32_t w, int32_t h, uint8_t* d_out,
                            void* stream) {
  if (!d_pix || !d_out) return fail(CVM_E_NIL_IMAGE, "cvmatch: nil image");
  if (w <= 0 || h <= 0) return fail(CVM_E_EMPTY_IMAGE, "cvmatch: empty image");
  if (stride < (int64_t)w * 4) return fail(CVM_E_BAD_ARGS, "cvmatch: stride %lld below 4*width", (long long)stride);
  LEASE(X);
  hipStream_t s = stream ? (hipStream_t)stream : X.stream;
  launch_rgba_to_gray(d_pix, stride, w, h, d_out, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int cvm_rgba_to_gray(const uint8_t* pix, int64_t stride, int32_t w, int32_t h, uint8_t* out) {
  if (!pix || !out) return fail(CVM_E_NIL_IMAGE, "cvmatch: nil image");
  if (w <= 0 || h <= 0) return fail(CVM_E_EMPTY_IMAGE, "cvmatch: empty image");
  if (stride < (int64_t)w * 4) return fail(CVM_E_BAD_ARGS, "cvmatch: stride %lld below 4*width", (long long)stride);
  LEASE(X);
  hipStream_t s = X.stream;
  size_t ib = (size_t)img_bytes(stride, w, h, 4);
  CHK(X.img.ensure(ib));
  CHK(X.gray0.ensure((size_t)w * h));
  HIPCHK(xfer_h2d(X.img.p, pix, ib, s));
  launch_rgba_to_gray(X.img.as<uint8_t>(), stride, w, h, X.gray0.as<uint8_t>(), s);
  HIPCHK(xfer_d2h(out, X.gray0.p, (size_t)w * h, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}


// ------------------------------------------------------ Matcher / Fleet --

int cvm_matcher_new(const uint8_t* tpl, int64_t stride, int32_t w, int32_t h, int32_t step, cvm_matcher** out) {
  if (!out) return fail(CVM_E_BAD_ARGS, "cvmatch: null output handle");
  return matcher_create(tpl, stride, w, h, step, out);
}

// img is a host buffer, or with on_device a device buffer (then stream, or
// the library's stream when null, orders the call)
static int matcher_find(cvm_matcher* m, const uint8_t* img, int64_t stride, int32_t w, int32_t h, cvm_result* out,
                        bool direct, bool on_device = false, void* stream = nullptr);

int cvm_matcher_find(cvm_matcher* m, const uint8_t* img, int64_t stride, int32_t w, int32_t h, cvm_result* out) {
  return matcher_find(m, img, stride, w, h, out, false);
}

int cvm_matcher_find_direct(cvm_matcher* m, const uint8_t* img, int64_t stride, int32_t w, int32_t h,
                            cvm_result* out) {
  return matcher_find(m, img, stride, w, h, out, true);
}

int cvm_matcher_find_device(cvm_matcher* m, const uint8_t* d_img, int64_t stride, int32_t w, int32_t h,
                            cvm_result* out, void* stream) {
  return matcher_find(m, d_img, stride, w, h, out, false, true, stream);
}

}  // extern "C"

static int matcher_find(cvm_matcher* m, const uint8_t* img, int64_t stride, int32_t w, int32_t h, cvm_result* out,
                        bool direct, bool on_device, void* stream) {
  if (!m) return fail(CVM_E_BAD_ARGS, "cvmatch: null matcher");
  if (!img) return fail(CVM_E_NIL_IMAGE, "cvmatch: nil image");
  if (w <= 0 || h <= 0) return fail(CVM_E_EMPTY_IMAGE, "cvmatch: empty image");
  if (m->w > w || m->h > h) return fail(CVM_E_BAD_ARGS, "cvmatch: bad match arguments (%dx%d in %dx%d)", m->w, m->h, w, h);
  if (stride < (int64_t)w * m->step) return fail(CVM_E_BAD_ARGS, "cvmatch: stride below width*step");
  if (m->varSum == 0) {  // exactly flat template matcher.go:100-102
    out->min_v = out->max_v = 1.0f;
    out->min_x = out->min_y = out->max_x = out->max_y = 0;
    return 0;
  }
  LEASE(X);
  hipStream_t s = stream ? (hipStream_t)stream : X.stream;
  const uint8_t* d_img = img;
  if (!on_device) {
    size_t ib = (size_t)img_bytes(stride, w, h, m->step);
    CHK(X.img.ensure(ib));
    Stage S(X, "h2d", s);
    HIPCHK(xfer_h2d(X.img.p, img, ib, s));
    d_img = X.img.as<uint8_t>();
  }
  if (direct)
    CHK(matcher_find_direct_dev(X, s, m, d_img, stride, w, h, res_out(X)));
  else
    CHK(matcher_find_dev(X, s, m, d_img, stride, w, h, res_out(X)));
  return finish_result(X, s, out);
}

// ------------------------------------------------------- frame streams --
//
// Find over a stream of host frames with the upload of frame i + 1 under way
// while frame i is matched: every slot owns a device frame and a pinned
// result; uploads go through the staged transfers on the stream's own copy
// stream, the library's stream waits for a slot's upload, matches it and
// copies the six-tuple into the slot's pinned result.  Each frame's result
// is the one cvm_matcher_find(_direct) returns for it.

struct cvm_frames {
  struct Slot {
    cvm::DevBuf img;
    cvm::Res6* res = nullptr;  // pinned
    hipEvent_t up = nullptr, done = nullptr;
    bool busy = false;
    cvm::GraphSlot graph;  // the captured Find of this slot's frame
  };
  struct Entry {
    int64_t tag;
    int slot;  // -1 once the result is in r
    cvm_matcher* m;  // reference held while in flight
    cvm_result r;
  };
  std::mutex mu;
  cvm::Ctx* ctx = nullptr;  // owned: scratch and stream of this stream's frames
  hipStream_t copy = nullptr;
  std::vector<Slot> slots;
  std::deque<Entry> fifo;
};

namespace cvm {

// Moves the oldest in-flight entry's result out of its slot (waiting for it).
static int frames_complete(cvm_frames* q, cvm_frames::Entry& e) {
  if (e.slot < 0) return 0;
  cvm_frames::Slot& sl = q->slots[e.slot];
  HIPCHK(hipEventSynchronize(sl.done));
  memcpy(&e.r, sl.res, sizeof(cvm_result));
  sl.busy = false;
  e.slot = -1;
  matcher_release(e.m);
  e.m = nullptr;
  return 0;
}

}  // namespace cvm

extern "C" {

int cvm_frames_new(int32_t depth, cvm_frames** out) {
  if (!out) return fail(CVM_E_BAD_ARGS, "cvmatch: null output handle");
  if (depth < 2 || depth > 8) return fail(CVM_E_BAD_ARGS, "cvmatch: frame stream depth %d outside [2, 8]", depth);
  {
    Engine& E = engine();
    std::lock_guard<std::mutex> lk(E.mu);
    CHK(E.init());
  }
  std::unique_ptr<cvm_frames> q(new cvm_frames());
  q->ctx = new Ctx();
  HIPCHK(q->ctx->init() == 0 ? hipSuccess : hipErrorOutOfMemory);
  HIPCHK(hipStreamCreateWithFlags(&q->copy, hipStreamNonBlocking));
  q->slots.resize(depth);
  for (auto& sl : q->slots) {
    HIPCHK(hipHostMalloc((void**)&sl.res, sizeof(Res6), hipHostMallocDefault));
    HIPCHK(hipEventCreateWithFlags(&sl.up, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
  }
  *out = q.release();
  return 0;
}

int cvm_frames_find(cvm_frames* q, cvm_matcher* m, const uint8_t* img, int64_t stride, int32_t w, int32_t h,
                    int32_t direct, int64_t tag) {
  if (!q || !m) return fail(CVM_E_BAD_ARGS, "cvmatch: null frame stream or matcher");
  if (!img) return fail(CVM_E_NIL_IMAGE, "cvmatch: nil image");
  if (w <= 0 || h <= 0) return fail(CVM_E_EMPTY_IMAGE, "cvmatch: empty image");
  if (m->w > w || m->h > h) return fail(CVM_E_BAD_ARGS, "cvmatch: bad match arguments (%dx%d in %dx%d)", m->w, m->h, w, h);
  if (stride < (int64_t)w * m->step) return fail(CVM_E_BAD_ARGS, "cvmatch: stride below width*step");
  std::lock_guard<std::mutex> ql(q->mu);
  cvm_frames::Entry e{tag, -1, nullptr, {}};
  if (m->varSum == 0) {  // exactly flat template matcher.go:100-102: no GPU work
    e.r = cvm_result{1.0f, 0, 0, 1.0f, 0, 0};
    q->fifo.push_back(e);
    return 0;
  }
  int si = -1;
  for (int i = 0; i < (int)q->slots.size(); i++)
    if (!q->slots[i].busy) {
      si = i;
      break;
    }
  if (si < 0) {  // every slot in flight: finish the oldest first
    for (auto& o : q->fifo)
      if (o.slot >= 0) {
        si = o.slot;
        CHK(frames_complete(q, o));
        break;
      }
  }
  cvm_frames::Slot& sl = q->slots[si];
  // the stream's own context: its frames stay in flight on it after return
  Ctx& X = *q->ctx;
  hipStream_t s = X.stream;
  size_t ib = (size_t)img_bytes(stride, w, h, m->step);
  CHK(sl.img.ensure(ib));
  {
    Stage S(X, "h2d", q->copy);
    HIPCHK(xfer_h2d(sl.img.p, img, ib, q->copy));
  }
  HIPCHK(hipEventRecord(sl.up, q->copy));
  // a pinned (or small) source is read by the DMA engine after xfer_h2d
  // returns: the caller may reuse its buffer only once the upload is done
  if (xfer_h2d_source_busy(img, ib)) HIPCHK(hipEventSynchronize(sl.up));
  HIPCHK(hipStreamWaitEvent(s, sl.up, 0));
  // the frame's Find, result copy included, replays a captured graph per
  // slot once the same matcher and frame shape repeat
  auto run = [&]() -> int {
    if (direct)
      CHK(matcher_find_direct_dev(X, s, m, sl.img.as<uint8_t>(), stride, w, h, X.res_d));
    else
      CHK(matcher_find_dev(X, s, m, sl.img.as<uint8_t>(), stride, w, h, X.res_d));
    HIPCHK(hipMemcpyAsync(sl.res, X.res_d, sizeof(Res6), hipMemcpyDeviceToHost, s));
    return 0;
  };
  const std::vector<uint64_t> key = {u64p(m), u64p(sl.img.p), (uint64_t)stride, (uint64_t)w, (uint64_t)h,
                                     direct ? 1u : 0u};
  CHK(run_graphed(sl.graph, key, s, run));
  HIPCHK(hipEventRecord(sl.done, s));
  sl.busy = true;
  m->refs.fetch_add(1);
  e.slot = si;
  e.m = m;
  q->fifo.push_back(e);
  return 0;
}

int cvm_frames_result(cvm_frames* q, int64_t* tag, cvm_result* out) {
  if (!q || !out) return fail(CVM_E_BAD_ARGS, "cvmatch: null frame stream or output");
  std::lock_guard<std::mutex> ql(q->mu);
  if (q->fifo.empty()) return fail(CVM_E_BAD_ARGS, "cvmatch: no frame in flight");
  cvm_frames::Entry& e = q->fifo.front();
  CHK(frames_complete(q, e));
  if (tag) *tag = e.tag;
  *out = e.r;
  q->fifo.pop_front();
  bool idle = true;
  for (auto& sl : q->slots) idle = idle && !sl.busy;
  if (idle) q->ctx->stage_collect();  // every stage event of the stream has completed
  return 0;
}

int32_t cvm_frames_pending(cvm_frames* q) {
  if (!q) return 0;
  std::lock_guard<std::mutex> ql(q->mu);
  return (int32_t)q->fifo.size();
}

void cvm_frames_free(cvm_frames* q) {
  if (!q) return;
  {
    std::lock_guard<std::mutex> ql(q->mu);
    for (auto& e : q->fifo) (void)frames_complete(q, e);
    q->fifo.clear();
  }
  if (q->copy) {
    (void)hipStreamSynchronize(q->copy);
    (void)xfer_release_stream(q->copy);  // staging slots last used by its uploads
  }
  if (q->ctx) {
    (void)hipStreamSynchronize(q->ctx->stream);
    q->ctx->stage_collect();
    q->ctx->destroy();
    delete q->ctx;
  }
  for (auto& sl : q->slots) {
    sl.graph.drop();
    sl.img.release();
    if (sl.res) (void)hipHostFree(sl.res);
    if (sl.up) (void)hipEventDestroy(sl.up);
    if (sl.done) (void)hipEventDestroy(sl.done);
  }
  if (q->copy) (void)hipStreamDestroy(q->copy);
  delete q;
}

void cvm_matcher_free(cvm_matcher* m) { matcher_release(m); }

}  // extern "C"

// NewFleet fleet.go:85-100 on a non-empty member list of one step: a private
// member list (each matcher's reference count raised) and its buckets.
static cvm_fleet* fleet_make(const std::vector<cvm_matcher*>& ms) {
  cvm_fleet* f = new cvm_fleet();
  f->step = ms[0]->step;
  for (cvm_matcher* m : ms) {
    m->refs.fetch_add(1);
    f->ms.push_back(m);
  }
  f->buckets = bucketize(f->ms);
  return f;
}

extern "C" {

int cvm_fleet_new(cvm_matcher* const* ms, int32_t n, cvm_fleet** out) {
  if (!out) return fail(CVM_E_BAD_ARGS, "cvmatch: null output handle");
  if (n <= 0 || !ms) return fail(CVM_E_EMPTY_FLEET, "cvmatch: empty fleet");
  for (int i = 0; i < n; i++)
    if (!ms[i]) return fail(CVM_E_BAD_ARGS, "cvmatch: null matcher in fleet");
  for (int i = 0; i < n; i++)
    if (ms[i]->step != ms[0]->step) return fail(CVM_E_MIXED_FLEET, "cvmatch: fleet mixes color and gray matchers");
  *out = fleet_make(std::vector<cvm_matcher*>(ms, ms + n));
  return 0;
}

// img is a host buffer, or with on_device a device buffer (then stream, or
// the library's stream when null, orders the call)
// form: 0 plain, 1 direct, 2 mixed
static int fleet_find_all(cvm_fleet* f, const uint8_t* img, int64_t stride, int32_t w, int32_t h, cvm_result* out,
                          int form, bool on_device = false, void* stream = nullptr);

int cvm_fleet_find_all(cvm_fleet* f, const uint8_t* img, int64_t stride, int32_t w, int32_t h, cvm_result* out) {
  return fleet_find_all(f, img, stride, w, h, out, 0);
}

int cvm_fleet_find_all_direct(cvm_fleet* f, const uint8_t* img, int64_t stride, int32_t w, int32_t h,
                              cvm_result* out) {
  return fleet_find_all(f, img, stride, w, h, out, 1);
}

int cvm_fleet_find_all_device(cvm_fleet* f, const uint8_t* d_img, int64_t stride, int32_t w, int32_t h,
                              cvm_result* out, void* stream) {
  return fleet_find_all(f, d_img, stride, w, h, out, 0, true, stream);
}

int cvm_fleet_find_all_direct_device(cvm_fleet* f, const uint8_t* d_img, int64_t stride, int32_t w, int32_t h,
                                     cvm_result* out, void* stream) {
  return fleet_find_all(f, d_img, stride, w, h, out, 1, true, stream);
}

int cvm_fleet_find_all_mixed(cvm_fleet* f, const uint8_t* img, int64_t stride, int32_t w, int32_t h,
                             cvm_result* out) {
  return fleet_find_all(f, img, stride, w, h, out, 2);
}

int cvm_fleet_find_all_mixed_device(cvm_fleet* f, const uint8_t* d_img, int64_t stride, int32_t w, int32_t h,
                                    cvm_result* out, void* stream) {
  return fleet_find_all(f, d_img, stride, w, h, out, 2, true, stream);
}

int cvm_fleet_mixed_plan(cvm_fleet* f, int32_t* routed, float* gate, int32_t n, int32_t* crossover) {
  if (!f) return fail(CVM_E_BAD_ARGS, "cvmatch: null fleet");
  std::lock_guard<std::mutex> fl(f->call_mu);
  if (!f->mixed) return fail(CVM_E_BAD_ARGS, "cvmatch: no mixed FindAll has run on this fleet");
  if (n != (int32_t)f->ms.size()) return fail(CVM_E_BAD_ARGS, "cvmatch: the fleet has %zu members", f->ms.size());
  const auto& M = *f->mixed;
  for (int i = 0; i < n; i++) {
    if (routed) routed[i] = 0;
    if (gate) gate[i] = M.gate[i];
  }
  if (routed)
    for (int i : M.ridx) routed[i] = 1;
  if (crossover) *crossover = M.crossover;
  return 0;
}

int cvm_tune_report(char* buf, int32_t cap) { return tune_report(buf, cap); }

}  // extern "C"

namespace cvm {

__global__ void k_scatter_res(const Res6* __restrict__ src, const int* __restrict__ idx, int n,
                              Res6* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[idx[i]] = src[i];
}

static void launch_scatter_res(const Res6* src, const int* idx, int n, Res6* dst, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_scatter_res, dim3((n + 255) / 256), dim3(256), 0, s, src, idx, n, dst);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v) {
  if (threadIdx.x == 0) *p = v;
}

// max |a - b| over n floats as the bits of a non-negative float (which order
// like the floats themselves; a NaN difference orders above every number)
__global__ __launch_bounds__(256) void k_max_abs_diff(const float* __restrict__ a, const float* __restrict__ b,
                                                      size_t n, uint32_t* out) {
  uint32_t m = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    m = max(m, __float_as_uint(fabsf(a[i] - b[i])) & 0x7FFFFFFFu);
  for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// The largest |plain score - exact score| over member m's whole map on this
// frame: the plain float32 FFT form (the reference's op sequence) against the
// direct form (the FP64 score rounded once to float32).
static int gate_delta(Ctx& X, hipStream_t s, const uint8_t* d_img, int64_t istride, int pw, int ph, cvm_matcher* m,
                      int cn, float* out) {
  const size_t nmap = (size_t)(pw - m->w + 1) * (size_t)(ph - m->h + 1);
  CHK(X.gate.ensure(sizeof(float) * nmap * 2 + 256));
  float* A = X.gate.as<float>();
  float* B = A + nmap;
  uint32_t* r = (uint32_t*)(B + nmap);
  const int64_t ts = (int64_t)m->w * m->step;
  CHK(run_match(X, s, d_img, istride, pw, ph, m->d_pix, ts, m->w, m->h, cn, m->step, A));
  CHK(run_match_direct(X, s, d_img, istride, pw, ph, m->d_pix, ts, m->w, m->h, cn, m->step, B));
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, r, 0u);
  hipLaunchKernelGGL(k_max_abs_diff, dim3(512), dim3(256), 0, s, A, B, nmap, r);
  HIPCHK(hipGetLastError());
  uint32_t bits = 0;
  HIPCHK(hipMemcpyAsync(&bits, r, sizeof(bits), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  memcpy(out, &bits, sizeof(float));
  return 0;
}

// Score bound of the mixed form: a member leaves the direct form only when
// its FFT scores stay this close to the exact ones on the tuning frame
// (1e-5; CVMATCH_MIXED_GATE sets another bound).
static float mixed_gate() {
  const char* v = getenv("CVMATCH_MIXED_GATE");
  return v && *v ? (float)atof(v) : 1e-5f;
}

}  // namespace cvm

// The mixed form's routing for the frame geometry (pw, ph, pConst), built on
// its first FindAll: the tuner's crossover -- the smallest square template
// side at which the plain FFT call on this frame beats the direct call -- and,
// for each non-flat member at least that wide or tall, the gate: its FFT
// form's largest deviation from the exact scores over its whole map on this
// frame.  Members within mixed_gate() take the FFT form (each its plain
// Find, the path the gate measured); members outside the direct form's
// shapes take it regardless; the rest stay direct (as a direct fleet of their
// own).
static int mixed_plan(Ctx& X, hipStream_t s, cvm_fleet* f, const uint8_t* d_img, int64_t istride, int pw, int ph,
                      int pConst) {
  if (f->mixed && f->mixed->pw == pw && f->mixed->ph == ph && f->mixed->pConst == pConst) return 0;
  f->gmixed.drop();
  if (f->mixed) {
    cvm_fleet_free(f->mixed->dpart);
    f->mixed->idx.release();
    f->mixed.reset();
  }
  const size_t n = f->ms.size();
  const int step = f->step;
  auto M = std::make_unique<cvm_fleet::Mixed>();
  M->pw = pw;
  M->ph = ph;
  M->pConst = pConst;
  M->gate.assign(n, -1.0f);
  const int cnF = step == 1 ? 1 : (pConst ? 3 : 4);
  char what[96];
  snprintf(what, sizeof(what), "fleet.crossover/s%d/cn%d/%dx%d", step, cnF, pw, ph);
  M->crossover = tune_value(what, "CVMATCH_CROSSOVER", 0, [&](std::vector<float>& ms) -> int {
    for (int S : {32, 48, 64, 80, 96, 112, 128}) {
      if (S > pw || S > ph) break;
      const uint8_t* t = d_img + (int64_t)((ph - S) / 2) * istride + (int64_t)((pw - S) / 2) * step;
      const float td = time_launches(
          s, [&] { return run_match_direct(X, s, d_img, istride, pw, ph, t, istride, S, S, cnF, step, nullptr); }, 2);
      const float tf = time_launches(
          s, [&] { return run_match(X, s, d_img, istride, pw, ph, t, istride, S, S, cnF, step, nullptr); }, 2);
      ms.push_back(td);
      ms.push_back(tf);
      if (td >= 0.0f && tf >= 0.0f && tf < td) return S;
    }
    return 0;
  });
  std::vector<cvm_matcher*> dms;
  for (size_t i = 0; i < n; i++) {
    cvm_matcher* m = f->ms[i];
    if (m->varSum == 0) continue;
    const int cn = step == 4 ? ((m->aConst && pConst) ? 3 : 4) : 1;
    bool route = false;
    if (check_direct(m->w, m->h, cn, direct_kstep(step, cn)) != 0) {
      route = true;  // the FFT form is the only one this member has
      M->gate[i] = -2.0f;
    } else if (M->crossover > 0 && std::max(m->w, m->h) >= M->crossover) {
      float d = 0.0f;
      CHK(gate_delta(X, s, d_img, istride, pw, ph, m, cn, &d));
      M->gate[i] = d;
      route = d <= mixed_gate();
    }
    (route ? M->ridx : M->didx).push_back((int)i);
    if (!route) dms.push_back(m);
  }
  if (!dms.empty()) {
    M->dpart = fleet_make(dms);
    CHK(M->dpart->res.ensure(sizeof(Res6) * dms.size()));
  }
  const std::vector<int>& idx = M->didx;
  CHK(M->idx.ensure(sizeof(int) * std::max<size_t>(idx.size(), 1)));
  if (!idx.empty()) {
    HIPCHK(hipMemcpyAsync(M->idx.p, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  f->mixed = std::move(M);
  return 0;
}

static int fleet_find_all(cvm_fleet* f, const uint8_t* img, int64_t stride, int32_t w, int32_t h, cvm_result* out,
                          int form, bool on_device, void* stream) {
  if (!f) return fail(CVM_E_BAD_ARGS, "cvmatch: null fleet");
  if (!img) return fail(CVM_E_NIL_IMAGE, "cvmatch: nil image");
  if (w <= 0 || h <= 0) return fail(CVM_E_EMPTY_IMAGE, "cvmatch: empty image");
  if (stride < (int64_t)w * f->step) return fail(CVM_E_BAD_ARGS, "cvmatch: stride below width*step");
  size_t n = f->ms.size();
  std::vector<int> flat;
  for (size_t i = 0; i < n; i++) {
    cvm_matcher* m = f->ms[i];
    if (m->w > w || m->h > h)
      return fail(CVM_E_BAD_ARGS, "cvmatch: bad match arguments (%dx%d in %dx%d)", m->w, m->h, w, h);
    if (m->varSum == 0) {
      out[i].min_v = out[i].max_v = 1.0f;
      out[i].min_x = out[i].min_y = out[i].max_x = out[i].max_y = 0;
      flat.push_back((int)i);
    }
  }
  if (f->buckets.empty()) return 0;
  // one FindAll of a fleet at a time: its caches and buffers are the fleet's
  std::lock_guard<std::mutex> fl(f->call_mu);
  LEASE(X);
  hipStream_t s = stream ? (hipStream_t)stream : X.stream;
  const uint8_t* d_img = img;
  if (!on_device) {
    size_t ib = (size_t)img_bytes(stride, w, h, f->step);
    CHK(X.img.ensure(ib));
    Stage S(X, "h2d", s);
    HIPCHK(xfer_h2d(X.img.p, img, ib, s));
    d_img = X.img.as<uint8_t>();
  }
  int pConst = 0;
  if (f->step == 4) {
    // the parent's alpha constancy keys the geometry cache (fleet.go:202-212)
    launch_state_init(X.st, s);
    launch_alpha_scan(d_img, stride, w, h, X.st, 0, s);
    CallState hs;
    HIPCHK(hipMemcpyAsync(&hs, X.st, sizeof(CallState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    pConst = hs.alpha_bad[0] == 0;
  }
  CHK(f->res.ensure(sizeof(Res6) * n));
  if (form == 2) CHK(mixed_plan(X, s, f, d_img, stride, w, h, pConst));
  // the whole FindAll is one captured graph from its second sighting on
  // (per frame buffer, shape, parent-alpha key, form, context and stream)
  auto run = [&]() -> int {
    if (form == 1) return fleet_find_direct_dev(X, s, f, d_img, stride, w, h, pConst, f->res.as<Res6>());
    if (form == 0) return fleet_find_dev(X, s, f, d_img, stride, w, h, pConst, f->res.as<Res6>(), flat);
    const cvm_fleet::Mixed& M = *f->mixed;
    const int* idx = M.idx.as<int>();
    if (M.dpart) {
      CHK(fleet_find_direct_dev(X, s, M.dpart, d_img, stride, w, h, pConst, M.dpart->res.as<Res6>()));
      launch_scatter_res(M.dpart->res.as<Res6>(), idx, (int)M.didx.size(), f->res.as<Res6>(), s);
    }
    // routed members: the plain Find each (the form their gate measured)
    for (int i : M.ridx) CHK(matcher_find_dev(X, s, f->ms[i], d_img, stride, w, h, f->res.as<Res6>() + i));
    HIPCHK(hipGetLastError());
    return 0;
  };
  const std::vector<uint64_t> key = {u64p(&X), u64p(d_img), (uint64_t)stride, (uint64_t)w, (uint64_t)h,
                                     (uint64_t)pConst};
  CHK(run_graphed(form == 1 ? f->gdirect : form == 0 ? f->gplain : f->gmixed, key, s, run));
  std::vector<Res6> hr(n);
  HIPCHK(hipMemcpyAsync(hr.data(), f->res.p, sizeof(Res6) * n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  X.stage_collect();
  for (size_t i = 0; i < n; i++)
    if (f->ms[i]->varSum != 0) memcpy(&out[i], &hr[i], sizeof(cvm_result));
  return 0;
}

extern "C" {

void cvm_fleet_free(cvm_fleet* f) {
  if (!f) return;
  {
    // FindAll returns with its work complete; this waits for a call still
    // running in another thread
    std::lock_guard<std::mutex> fl(f->call_mu);
    f->gplain.drop();
    f->gdirect.drop();
    f->gmixed.drop();
    if (f->mixed) {
      cvm_fleet_free(f->mixed->dpart);
      f->mixed->idx.release();
      f->mixed.reset();
    }
    f->geom.reset();
    for (DevBuf* b : {&f->frame, &f->mspec, &f->corr, &f->res, &f->pcorr, &f->pparts, &f->mspec2, &f->dstates,
                      &f->bj_direct, &f->bj_dwg, &f->bj_norm, &f->bj_nwg, &f->bj_cmaps, &f->bj_parts,
                      &f->bj_fused, &f->bj_fwg, &f->bj_fred, &f->bj_fparts})
      b->release();
    for (DevBuf& b : f->dscr) b.release();
  }
  for (auto* m : f->ms) matcher_release(m);
  delete f;
}

// ------------------------------------------------------------- helpers --

int cvm_device_alloc(size_t bytes, void** out) {
  {
    Engine& E = engine();
    std::lock_guard<std::mutex> lk(E.mu);
    CHK(E.init());
  }
  HIPCHK(hipMalloc(out, bytes));
  return 0;
}

int cvm_device_free(void* p) {
  HIPCHK(hipFree(p));
  return 0;
}

int cvm_copy_h2d(void* dst, const void* src, size_t bytes) {
  HIPCHK(xfer_h2d(dst, src, bytes, nullptr));
  HIPCHK(hipStreamSynchronize(nullptr));
  return 0;
}

int cvm_copy_d2h(void* dst, const void* src, size_t bytes) {
  HIPCHK(xfer_d2h(dst, src, bytes, nullptr));
  HIPCHK(hipStreamSynchronize(nullptr));
  return 0;
}

int cvm_host_alloc(size_t bytes, void** out) {
  if (!out) return fail(CVM_E_BAD_ARGS, "cvmatch: nil output pointer");
  HIPCHK(host_alloc(bytes, out));
  return 0;
}

int cvm_host_free(void* p) {
  if (!p) return 0;
  hipError_t e = host_free(p);
  if (e == hipErrorInvalidValue) return fail(CVM_E_BAD_ARGS, "cvmatch: %p is not a cvm_host_alloc block", p);
  HIPCHK(e);
  return 0;
}

int32_t cvm_set_threads(int32_t n) { return set_host_threads(n); }

int32_t cvm_threads(void) { return host_threads(); }

int cvm_selftest_copy(void* dst, const void* src, size_t bytes, int32_t threads) {
  if ((!dst || !src) && bytes) return fail(CVM_E_BAD_ARGS, "cvmatch: nil buffer");
  host_copy(dst, src, bytes, threads);
  return 0;
}

}  // extern "C"

namespace cvm {
// Streaming copy, four 16-byte loads in flight per thread (also the
// reference kernel the HBM traffic counters are calibrated against).
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
  const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
  uint4 v[4];
#pragma unroll
  for (int u = 0; u < 4; u++)
    if (base + 256 * u < n16) v[u] = src[base + 256 * u];
#pragma unroll
  for (int u = 0; u < 4; u++)
    if (base + 256 * u < n16) dst[base + 256 * u] = v[u];
}
__global__ void k_copy1(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace cvm

extern "C" {

int cvm_device_copy(void* dst, const void* src, size_t bytes) {
  LEASE(X);
  hipStream_t s = X.stream;
  bool aligned = (((uintptr_t)dst | (uintptr_t)src) & 15) == 0;
  size_t n16 = aligned ? bytes / 16 : 0;
  if (n16) {
    Stage S(X, "copy", s);
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)((n16 + 1023) / 1024)), dim3(256), 0, s, (const uint4*)src,
                       (uint4*)dst, n16);
  }
  if (bytes > n16 * 16)
    hipLaunchKernelGGL(k_copy1, dim3(64), dim3(256), 0, s, (const uint8_t*)src + n16 * 16, (uint8_t*)dst + n16 * 16,
                       bytes - n16 * 16);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  X.stage_collect();
  return 0;
}

int cvm_synchronize(void) {
  HIPCHK(hipDeviceSynchronize());
  return 0;
}

int cvm_profile_enable(int32_t on) {
  engine().prof.store(on != 0);
  return 0;
}

int cvm_profile_reset(void) {
  Engine& E = engine();
  std::lock_guard<std::mutex> lk(E.pmu);
  E.prof_acc.clear();
  return 0;
}

int cvm_profile_read(char* names, double* total_ms, int64_t* count, int32_t max) {
  Engine& E = engine();
  std::lock_guard<std::mutex> lk(E.pmu);
  int i = 0;
  for (auto& kv : E.prof_acc) {
    if (i >= max) break;
    snprintf(names + 32 * i, 32, "%s", kv.first.c_str());
    total_ms[i] = kv.second.first;
    count[i] = kv.second.second;
    i++;
  }
  return i;
}

}  // extern "C"

// ------------------------------------------------------------ self tests --

namespace cvm {
__global__ void k_selftest_fp(const float* a, const float* b, const long long* ints, int n, float* osq, float* odiv,
                              float* ocvt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  osq[i] = sqrt_rn_f32(a[i]);
  odiv[i] = div_rn_f32(a[i], b[i]);
  ocvt[i] = (float)(double)ints[i];
}
}  // namespace cvm

extern "C" int cvm_selftest_fp(const float* a, const float* b, const int64_t* ints, int32_t n, float* out_sqrt,
                               float* out_div, float* out_cvt) {
  LEASE(X);
  hipStream_t s = X.stream;
  DevBuf da, db, di, o1, o2, o3;
  size_t fb = sizeof(float) * (size_t)n, ibytes = sizeof(int64_t) * (size_t)n;
  CHK(da.ensure(fb));
  CHK(db.ensure(fb));
  CHK(di.ensure(ibytes));
  CHK(o1.ensure(fb));
  CHK(o2.ensure(fb));
  CHK(o3.ensure(fb));
  HIPCHK(hipMemcpyAsync(da.p, a, fb, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(db.p, b, fb, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(di.p, ints, ibytes, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_selftest_fp, dim3((n + 255) / 256), dim3(256), 0, s, da.as<float>(), db.as<float>(),
                     di.as<long long>(), n, o1.as<float>(), o2.as<float>(), o3.as<float>());
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out_sqrt, o1.p, fb, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(out_div, o2.p, fb, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(out_cvt, o3.p, fb, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  da.release();
  db.release();
  di.release();
  o1.release();
  o2.release();
  o3.release();
  return 0;
}

extern "C" int cvm_twiddles(int32_t n, float* tri, int32_t* brev, int32_t* tri_len_out) {
  if (n < 1 || (n & (n - 1)) || n > (1 << 22)) return fail(CVM_E_BAD_ARGS, "cvmatch: twiddle size %d", n);
  LEASE(X);
  const FftTables* t;
  CHK(X.tab(n, &t));
  int len = tri_len(n);
  HIPCHK(hipMemcpyAsync(tri, t->tri, sizeof(float2) * (size_t)len, hipMemcpyDeviceToHost, X.stream));
  HIPCHK(hipMemcpyAsync(brev, t->brev, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, X.stream));
  HIPCHK(hipStreamSynchronize(X.stream));
  *tri_len_out = len;
  return 0;
}
