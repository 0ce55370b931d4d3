// Host <-> device transfers of the host-buffer entry points.
//
// The reference API hands the library pageable host memory (Go image pixels
// in, the MatchMap result slice out).  The driver's pageable path moves such
// buffers well below the PCIe rate (and a freshly allocated destination costs
// a page fault per 4 KiB inside the driver's copy).  Here a pageable buffer
// moves through a ring of pinned staging chunks: host threads copy one chunk
// while the DMA engine moves the previous one, so a transfer runs at the
// pinned-memory rate, and the page faults of a fresh destination are taken by
// several threads at once.  Buffers from cvm_host_alloc are pinned already
// and go straight to the DMA engine.
//
// The copy-thread count mirrors the reference's Threads knob (cvmatch.go:
// 36-63): 0 or less means automatic, anything else is clamped to [1, 16], the
// CVMATCH_THREADS environment variable is read once, and no setting changes a
// single output bit (only the host copies are threaded).
#include <sched.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "cvm_internal.h"

namespace cvm {

// ------------------------------------------------------------- threads --

static std::atomic<int> g_threads_setting{-1};  // -1: environment not read yet

static int auto_threads() {
  cpu_set_t cs;
  int n = 0;
  if (sched_getaffinity(0, sizeof(cs), &cs) == 0) n = CPU_COUNT(&cs);
  if (n <= 0) n = (int)std::thread::hardware_concurrency();
  return n;
}

int set_host_threads(int n) {
  g_threads_setting.store(std::max(n, 0));  // any n <= 0: automatic
  return host_threads();
}

// threads() cvmatch.go:52-63
int host_threads() {
  int n = g_threads_setting.load();
  if (n == -1) {
    const char* v = getenv("CVMATCH_THREADS");
    int e = v ? atoi(v) : 0;
    n = e > 0 ? e : 0;
    int expect = -1;
    g_threads_setting.compare_exchange_strong(expect, n);
    n = g_threads_setting.load();
  }
  if (n <= 0) n = auto_threads();
  return std::min(16, std::max(1, n));
}

// ---------------------------------------------------------- copy pool --

// Persistent workers that split one memcpy into pieces.  A job is tagged with
// a generation in the high half of the piece counter, so a worker that wakes
// late can never claim a piece of a later job with an earlier job's pointers.
class CopyPool {
 public:
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }

  void copy(void* dst, const void* src, size_t n, int threads) {
    constexpr size_t kPiece = 192u << 10;
    int parts = (int)std::min<size_t>((n + kPiece - 1) / kPiece, (size_t)threads * 2);
    if (threads <= 1 || parts <= 1) {
      memcpy(dst, src, n);
      return;
    }
    grow(threads - 1);
    Job j{(uint8_t*)dst, (const uint8_t*)src, n, parts};
    uint32_t g;
    {
      std::lock_guard<std::mutex> lk(m_);
      g = (uint32_t)gen_.load() + 1;
      job_ = j;
      left_.store(parts);
      next_.store((uint64_t)g << 32);
      gen_.store(g);
    }
    cv_.notify_all();
    run(j, g);
    while (left_.load(std::memory_order_acquire) != 0) std::this_thread::yield();
  }

 private:
  struct Job {
    uint8_t* dst;
    const uint8_t* src;
    size_t n;
    int parts;
  };

  static size_t piece_at(const Job& j, int p) { return j.n * (size_t)p / (size_t)j.parts & ~(size_t)63; }

  void run(const Job& j, uint32_t g) {
    uint64_t v = next_.load();
    for (;;) {
      if ((uint32_t)(v >> 32) != g || (uint32_t)v >= (uint32_t)j.parts) return;
      if (!next_.compare_exchange_weak(v, v + 1)) continue;
      int p = (int)(uint32_t)v;
      size_t a = piece_at(j, p), b = p + 1 == j.parts ? j.n : piece_at(j, p + 1);
      memcpy(j.dst + a, j.src + a, b - a);
      left_.fetch_sub(1, std::memory_order_release);
      v = next_.load();
    }
  }

  void grow(int want) {
    std::lock_guard<std::mutex> lk(m_);
    while ((int)th_.size() < want) th_.emplace_back([this] { worker(); });
  }

  void worker() {
    uint32_t seen = 0;
    for (;;) {
      // a transfer issues its chunks at most a chunk's DMA time apart: stay
      // awake for 200 us before sleeping
      auto t0 = std::chrono::steady_clock::now();
      while (gen_.load(std::memory_order_acquire) == seen && !stop_flag() &&
             std::chrono::steady_clock::now() - t0 < std::chrono::microseconds(200))
        std::this_thread::yield();
      Job j;
      uint32_t g;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return stop_ || gen_.load() != seen; });
        if (stop_) return;
        g = gen_.load();
        seen = g;
        j = job_;
      }
      run(j, g);
    }
  }
  bool stop_flag() {
    std::lock_guard<std::mutex> lk(m_);
    return stop_;
  }

  std::mutex m_;
  std::condition_variable cv_;
  std::vector<std::thread> th_;
  Job job_{};
  std::atomic<uint32_t> gen_{0};
  std::atomic<uint64_t> next_{0};
  std::atomic<int> left_{0};
  bool stop_ = false;
};

// ------------------------------------------------------------- staging --

struct Staging {
  static constexpr size_t kChunk = 4u << 20;
  static constexpr int kSlots = 4;
  // below this a transfer goes straight to the driver
  static constexpr size_t kDirect = 1u << 20;
  std::mutex mu;
  uint8_t* pin[kSlots] = {};
  hipEvent_t ev[kSlots] = {};
  bool live[kSlots] = {};
  hipStream_t on[kSlots] = {};  // the stream ev[k] was last recorded on
  CopyPool pool;
  std::map<uintptr_t, size_t> pinned;  // cvm_host_alloc blocks: start -> bytes

  bool unavailable = false;  // the host refused the pinned chunks once: not retried

  hipError_t init() {
    if (pin[0]) return hipSuccess;
    if (unavailable) return hipErrorOutOfMemory;
    for (int k = 0; k < kSlots; k++) {
      hipError_t e = hipHostMalloc((void**)&pin[k], kChunk, hipHostMallocDefault);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[k], hipEventDisableTiming);
      if (e != hipSuccess) {
        for (int i = 0; i <= k; i++) {
          if (pin[i]) (void)hipHostFree(pin[i]);
          if (ev[i]) (void)hipEventDestroy(ev[i]);
          pin[i] = nullptr;
          ev[i] = nullptr;
        }
        unavailable = true;
        (void)hipGetLastError();  // not a launch failure for a later check to report
        return e;
      }
    }
    return hipSuccess;
  }

  bool covers(const void* p, size_t n) {
    uintptr_t a = (uintptr_t)p;
    auto it = pinned.upper_bound(a);
    if (it == pinned.begin()) return false;
    --it;
    return a >= it->first && a + n <= it->first + it->second;
  }

  hipError_t slot_free(int k) {
    if (!live[k]) return hipSuccess;
    hipError_t e = hipEventSynchronize(ev[k]);
    live[k] = false;
    return e;
  }
};

static Staging& staging() {
  static Staging* s = new Staging();
  return *s;
}

#define XCHK(x)                       \
  do {                                \
    hipError_t e_ = (x);              \
    if (e_ != hipSuccess) return e_;  \
  } while (0)

hipError_t host_alloc(size_t bytes, void** out) {
  Staging& S = staging();
  std::lock_guard<std::mutex> lk(S.mu);
  XCHK(hipHostMalloc(out, std::max<size_t>(bytes, 1), hipHostMallocDefault));
  S.pinned[(uintptr_t)*out] = std::max<size_t>(bytes, 1);
  return hipSuccess;
}

hipError_t host_free(void* p) {
  Staging& S = staging();
  std::lock_guard<std::mutex> lk(S.mu);
  auto it = S.pinned.find((uintptr_t)p);
  if (it == S.pinned.end()) return hipErrorInvalidValue;
  S.pinned.erase(it);
  return hipHostFree(p);
}

void host_copy(void* dst, const void* src, size_t n, int threads) {
  Staging& S = staging();
  std::lock_guard<std::mutex> lk(S.mu);
  S.pool.copy(dst, src, n, std::min(16, std::max(1, threads)));
}

// Before a stream is destroyed: every staging slot whose event was last
// recorded on it is waited for and retired, so no later transfer waits on
// an event of a stream that no longer exists.
hipError_t xfer_release_stream(hipStream_t s) {
  Staging& S = staging();
  std::lock_guard<std::mutex> lk(S.mu);
  hipError_t first = hipSuccess;
  for (int k = 0; k < Staging::kSlots; k++)
    if (S.live[k] && S.on[k] == s) {
      const hipError_t e = S.slot_free(k);
      if (first == hipSuccess) first = e;
    }
  return first;
}

// Whether xfer_h2d leaves the copy reading src after it returns: true for
// the direct routes (a pinned block, or a buffer below the staging size).
bool xfer_h2d_source_busy(const void* src, size_t n) {
  Staging& S = staging();
  std::lock_guard<std::mutex> lk(S.mu);
  return n < Staging::kDirect || S.covers(src, n) || S.init() != hipSuccess;
}

// A staged source may be reused on return: the last chunks may still be in
// flight on stream s (from the pinned chunks), which orders them before
// later work on s.  On the direct routes the copy reads src itself.
hipError_t xfer_h2d(void* dst, const void* src, size_t n, hipStream_t s) {
  Staging& S = staging();
  std::lock_guard<std::mutex> lk(S.mu);
  // without pinned staging (the host refused the allocation) the driver's
  // pageable path still moves the buffer, only more slowly
  if (n < Staging::kDirect || S.covers(src, n) || S.init() != hipSuccess)
    return hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s);
  const int threads = host_threads();
  size_t c = 0;
  for (size_t off = 0; off < n; off += Staging::kChunk, c++) {
    int k = (int)(c % Staging::kSlots);
    size_t len = std::min(Staging::kChunk, n - off);
    XCHK(S.slot_free(k));
    S.pool.copy(S.pin[k], (const uint8_t*)src + off, len, threads);
    XCHK(hipMemcpyAsync((uint8_t*)dst + off, S.pin[k], len, hipMemcpyHostToDevice, s));
    XCHK(hipEventRecord(S.ev[k], s));
    S.live[k] = true;
    S.on[k] = s;
  }
  return hipSuccess;
}

// xfer_h2d with the caller told of each chunk: done(bytes) runs once the
// copy of the first `bytes` bytes is enqueued on s (after every staged chunk,
// or once for the direct routes), so the caller can record an event on s
// and start work on those bytes on another stream.  A nonzero return from
// done ends the transfer with an error.
hipError_t xfer_h2d_chunks(void* dst, const void* src, size_t n, hipStream_t s,
                           const std::function<int(size_t)>& done) {
  Staging& S = staging();
  std::lock_guard<std::mutex> lk(S.mu);
  if (n < Staging::kDirect || S.covers(src, n) || S.init() != hipSuccess) {
    XCHK(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s));
    return done(n) ? hipErrorUnknown : hipSuccess;
  }
  const int threads = host_threads();
  const size_t chunk = Staging::kChunk;  // smaller chunks measured slower (per-copy cost)
  size_t c = 0;
  for (size_t off = 0; off < n; off += chunk, c++) {
    int k = (int)(c % Staging::kSlots);
    size_t len = std::min(chunk, n - off);
    XCHK(S.slot_free(k));
    S.pool.copy(S.pin[k], (const uint8_t*)src + off, len, threads);
    XCHK(hipMemcpyAsync((uint8_t*)dst + off, S.pin[k], len, hipMemcpyHostToDevice, s));
    XCHK(hipEventRecord(S.ev[k], s));
    S.live[k] = true;
    S.on[k] = s;
    if (done(off + len)) return hipErrorUnknown;
  }
  return hipSuccess;
}

// Returns with the destination written, except for a pinned destination,
// whose copy is left in flight on stream s like any asynchronous copy.
hipError_t xfer_d2h(void* dst, const void* src, size_t n, hipStream_t s) {
  Staging& S = staging();
  std::lock_guard<std::mutex> lk(S.mu);
  if (S.covers(dst, n)) return hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, s);
  if (n < Staging::kDirect || S.init() != hipSuccess) {
    XCHK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, s));
    return hipStreamSynchronize(s);
  }
  const int threads = host_threads();
  const size_t nc = (n + Staging::kChunk - 1) / Staging::kChunk;
  size_t issued = 0;
  for (size_t c = 0; c < nc; c++) {
    // keep up to kSlots chunks in flight; chunk i reuses the slot of chunk
    // i - kSlots, which the host has already copied out
    for (; issued < nc && issued < c + Staging::kSlots; issued++) {
      int k = (int)(issued % Staging::kSlots);
      size_t off = issued * Staging::kChunk;
      XCHK(S.slot_free(k));
      XCHK(hipMemcpyAsync(S.pin[k], (const uint8_t*)src + off, std::min(Staging::kChunk, n - off),
                          hipMemcpyDeviceToHost, s));
      XCHK(hipEventRecord(S.ev[k], s));
      S.live[k] = true;
      S.on[k] = s;
    }
    int k = (int)(c % Staging::kSlots);
    size_t off = c * Staging::kChunk;
    XCHK(S.slot_free(k));
    S.pool.copy((uint8_t*)dst + off, S.pin[k], std::min(Staging::kChunk, n - off), threads);
  }
  return hipSuccess;
}

}  // namespace cvm
