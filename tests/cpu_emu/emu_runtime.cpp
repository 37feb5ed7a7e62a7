// CPU emulation runtime of the HIP subset libmgpcg uses -- TEST INFRASTRUCTURE ONLY.
//
// Kernel launches run in stream order (at once, or queued when deferred
// streams are on; see below).  Each block's threads are ucontext
// coroutines on one OS thread: a thread runs until it reaches __syncthreads
// (or a shuffle, which is two barriers) or returns; the barrier opens when
// every live thread of the block waits at it.  "Device" memory comes from
// mmap with a PROT_NONE guard page on both sides and the buffer end placed
// against the upper guard, so reading or writing past either end faults.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/mman.h>
#include <ucontext.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <vector>

#undef threadIdx
#undef blockIdx
#undef gridDim
#undef blockDim

namespace hipemu {

dim3 threadIdx, blockIdx, gridDim, blockDim;

namespace {

struct Fiber {
  ucontext_t ctx;
  char* stack = nullptr;
  size_t stack_bytes = 0;
  bool done = false;
  bool waiting = false;
  unsigned tid = 0;
};

const size_t STACK = 256 * 1024;
ucontext_t g_sched;
std::vector<Fiber> g_fibers;
Fiber* g_cur = nullptr;
void (*g_tramp)(void*) = nullptr;
void* g_ctx = nullptr;
std::vector<double> g_xd;
std::vector<float> g_xf;
std::vector<int> g_xi;

void fiber_entry(unsigned idx_lo, unsigned idx_hi) {
  (void)idx_hi;
  Fiber* f = &g_fibers[idx_lo];
  g_tramp(g_ctx);
  f->done = true;
  swapcontext(&f->ctx, &g_sched);
}

// run every live fiber until it waits or finishes; repeat while any waits
void run_block(unsigned nthreads) {
  if (g_fibers.size() < nthreads) g_fibers.resize(nthreads);
  for (unsigned t = 0; t < nthreads; ++t) {
    Fiber& f = g_fibers[t];
    if (!f.stack) {
      f.stack = (char*)mmap(nullptr, STACK, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS,
                            -1, 0);
      if (f.stack == MAP_FAILED) {
        fprintf(stderr, "hipemu: fiber stack mmap failed\n");
        abort();
      }
      f.stack_bytes = STACK;
    }
    f.done = false;
    f.waiting = false;
    f.tid = t;
    getcontext(&f.ctx);
    f.ctx.uc_stack.ss_sp = f.stack;
    f.ctx.uc_stack.ss_size = f.stack_bytes;
    f.ctx.uc_link = &g_sched;
    makecontext(&f.ctx, (void (*)())fiber_entry, 2, t, 0u);
  }
  for (;;) {
    unsigned live = 0, waiting = 0;
    for (unsigned t = 0; t < nthreads; ++t) {
      Fiber& f = g_fibers[t];
      if (f.done) continue;
      f.waiting = false;
      g_cur = &f;
      threadIdx = dim3(t, 0, 0);
      swapcontext(&g_sched, &f.ctx);
      g_cur = nullptr;
      if (!f.done) {
        live++;
        if (f.waiting) waiting++;
      }
    }
    if (live == 0) break;
    if (waiting != live) {
      fprintf(stderr, "hipemu: inconsistent barrier state\n");
      abort();
    }
  }
}

}  // namespace

void syncthreads() {
  Fiber* f = g_cur;
  f->waiting = true;
  swapcontext(&f->ctx, &g_sched);
  threadIdx = dim3(f->tid, 0, 0);
}

template <typename T>
T shfl(std::vector<T>& buf, T v, int mask) {
  const unsigned tid = g_cur->tid;
  if (buf.size() < blockDim.x) buf.resize(blockDim.x);
  buf[tid] = v;
  syncthreads();
  const unsigned src = tid ^ (unsigned)mask;
  T r = (src < blockDim.x) ? buf[src] : v;
  syncthreads();
  return r;
}
template <typename T>
T shfl_src(std::vector<T>& buf, T v, int src, int width) {
  const unsigned tid = g_cur->tid;
  if (buf.size() < blockDim.x) buf.resize(blockDim.x);
  buf[tid] = v;
  syncthreads();
  const unsigned base = tid & ~(unsigned)(width - 1);
  const unsigned from = base + ((unsigned)src & (unsigned)(width - 1));
  T r = from < blockDim.x ? buf[from] : v;
  syncthreads();
  return r;
}
float shfl_idx(float v, int src, int width) { return shfl_src(g_xf, v, src, width); }
double shfl_idx(double v, int src, int width) { return shfl_src(g_xd, v, src, width); }
double shfl_xor_d(double v, int mask) { return shfl(g_xd, v, mask); }
float shfl_xor_f(float v, int mask) { return shfl(g_xf, v, mask); }
int shfl_xor_i(int v, int mask) { return shfl(g_xi, v, mask); }
// lane `lane` of the caller's 64-lane wave
int readlane_i(int v, int lane) { return shfl_src(g_xi, v, lane, 64); }
// bit l of the result: lane l of the caller's 64-lane wave had pred != 0
unsigned long long ballot(int pred) {
  const unsigned tid = g_cur->tid;
  if (g_xi.size() < blockDim.x) g_xi.resize(blockDim.x);
  g_xi[tid] = pred != 0;
  syncthreads();
  const unsigned base = tid & ~63u;
  unsigned long long m = 0;
  for (unsigned l = 0; l < 64 && base + l < blockDim.x; ++l)
    if (g_xi[base + l]) m |= 1ull << l;
  syncthreads();
  return m;
}

void run_grid(dim3 grid, dim3 block, void (*tramp)(void*), void* ctx) {
  gridDim = grid;
  blockDim = block;
  g_tramp = tramp;
  g_ctx = ctx;
  const unsigned nthreads = block.x * block.y * block.z;
  for (unsigned b = 0; b < grid.x; ++b) {
    blockIdx = dim3(b, 0, 0);
    run_block(nthreads);
  }
}

namespace {
std::mutex g_mu;
std::map<void*, std::pair<char*, size_t>> g_allocs;  // user ptr -> (mapping, bytes)
const size_t PAGE = 4096;
}  // namespace

void* guarded_alloc(size_t bytes) {
  size_t body = (bytes + 255) / 256 * 256;
  if (body == 0) body = 256;
  size_t pages = (body + PAGE - 1) / PAGE;
  size_t total = (pages + 2) * PAGE;
  char* m = (char*)mmap(nullptr, total, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1,
                        0);
  if (m == MAP_FAILED) return nullptr;
  mprotect(m, PAGE, PROT_NONE);
  mprotect(m + (pages + 1) * PAGE, PAGE, PROT_NONE);
  char* user = m + (pages + 1) * PAGE - body;  // ends at the upper guard
  memset(user, 0xA5, body);                    // device memory starts undefined
  std::lock_guard<std::mutex> lk(g_mu);
  g_allocs[user] = {m, total};
  return user;
}

void guarded_free(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_allocs.find(p);
  if (it == g_allocs.end()) {
    fprintf(stderr, "hipemu: free of unknown pointer %p\n", p);
    abort();
  }
  munmap(it->second.first, it->second.second);
  g_allocs.erase(it);
}

}  // namespace hipemu

// Streams.  By default every operation runs when it is issued (each stream is
// trivially in order).  With deferred streams on (hipemu_set_deferred), work
// issued to a stream the library created is queued and runs only when the
// host waits (stream/event/device synchronization, blocking copies, frees).
// The scheduler then either always advances the most recently created stream
// whose next operation can run, or takes one operation of each stream in turn:
// orders that are legal but as far from issue order as the events allow, so
// a missing event dependency between streams shows up as a wrong result.  The NULL stream (the caller's stream
// on the CPU) always runs at once; waiting on an event from it first runs the
// queues until the event has passed.
struct ihipEvent_t {
  std::chrono::steady_clock::time_point t;
  unsigned long long recorded = 0;   // records issued
  unsigned long long completed = 0;  // records the stream order has passed
};
struct ihipStream_t {
  struct Op {
    int kind;  // 0 run, 1 signal event, 2 wait for event
    std::function<void()> fn;
    ihipEvent_t* ev;
    unsigned long long gen;
  };
  std::deque<Op> q;
};
struct ihipGraph_t {
  std::vector<std::function<void()>> ops;
};
struct ihipGraphExec_t {
  std::vector<std::function<void()>> ops;
};

namespace hipemu {
namespace {
ihipGraph_t* g_capture = nullptr;  // the graph being captured (one at a time)
hipStream_t g_capture_stream = nullptr;
bool g_defer = false;
int g_defer_mode = 1;  // 1: newest stream first, 2: one operation per stream in turn
size_t g_rr = 0;
std::vector<ihipStream_t*> g_streams;  // in creation order

bool deferred(hipStream_t s) { return g_defer && s != nullptr; }

// Advance the queues until pred() holds (every queue drained when pred is
// empty).  Aborts when nothing can run and pred does not hold.
void run_until(const std::function<bool()>& pred) {
  for (;;) {
    if (pred && pred()) return;
    ihipStream_t* pick = nullptr;
    const size_t ns = g_streams.size();
    for (size_t k = 0; k < ns; ++k) {
      // newest stream first; round robin: the stream after the last one run
      const size_t i = g_defer_mode == 2 ? (g_rr + 1 + k) % ns : ns - 1 - k;
      ihipStream_t* st = g_streams[i];
      if (st->q.empty()) continue;
      const ihipStream_t::Op& op = st->q.front();
      if (op.kind == 2 && op.ev->completed < op.gen) continue;
      pick = st;
      g_rr = i;
      break;
    }
    if (!pick) {
      if (!pred) return;
      fprintf(stderr, "hipemu: deadlock: a wait can never be satisfied\n");
      abort();
    }
    ihipStream_t::Op op = std::move(pick->q.front());
    pick->q.pop_front();
    if (getenv("HIPEMU_TRACE")) {
      int idx = 0;
      for (size_t i = 0; i < g_streams.size(); ++i) if (g_streams[i] == pick) idx = (int)i;
      fprintf(stderr, "TRACE stream %d kind %d\n", idx, op.kind);
    }
    if (op.kind == 0) {
      op.fn();
    } else if (op.kind == 1) {
      op.ev->t = std::chrono::steady_clock::now();
      if (op.ev->completed < op.gen) op.ev->completed = op.gen;
    }
  }
}
void drain_all() { run_until(nullptr); }
}  // namespace
bool capturing() { return g_capture != nullptr; }
void record(std::function<void()> op) { g_capture->ops.push_back(std::move(op)); }
void submit(hipStream_t s, std::function<void()> op) {
  if (deferred(s))
    s->q.push_back({0, std::move(op), nullptr, 0});
  else
    op();
}
}  // namespace hipemu

// operations queued on the library's streams (deferred streams only)
extern "C" long hipemu_queued(void) {
  long n = 0;
  for (ihipStream_t* st : hipemu::g_streams) n += (long)st->q.size();
  return n;
}

// 0: off; 1: run the newest stream with runnable work first; 2: round robin,
// one operation of each stream in turn (the finest interleaving of streams)
extern "C" void hipemu_set_deferred(int mode) {
  hipemu::drain_all();
  hipemu::g_defer = mode != 0;
  hipemu::g_defer_mode = mode;
}

hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  if (hipemu::g_capture) return hipErrorInvalidValue;
  hipemu::g_capture = new ihipGraph_t();
  hipemu::g_capture_stream = s;
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
  if (!hipemu::g_capture || s != hipemu::g_capture_stream) return hipErrorInvalidValue;
  *g = hipemu::g_capture;
  hipemu::g_capture = nullptr;
  hipemu::g_capture_stream = nullptr;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  if (!g) return hipErrorInvalidValue;
  *x = new ihipGraphExec_t{g->ops};
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s) {
  if (hipemu::capturing()) return hipErrorInvalidValue;
  std::vector<std::function<void()>> ops = x->ops;
  hipemu::submit(s, [ops]() {
    for (auto& op : ops) op();
  });
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t x) {
  delete x;
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}

const char* hipGetErrorString(hipError_t e) {
  switch (e) {
    case hipSuccess:
      return "no error";
    case hipErrorInvalidValue:
      return "invalid argument";
    case hipErrorOutOfMemory:
      return "out of memory";
    default:
      return "emulated error";
  }
}
hipError_t hipGetLastError(void) { return hipSuccess; }
unsigned long long wall_clock64() {
  return (unsigned long long)(std::chrono::duration_cast<std::chrono::nanoseconds>(
                                  std::chrono::steady_clock::now().time_since_epoch())
                                  .count() /
                              10);
}
hipError_t hipDeviceGetAttribute(int* value, hipDeviceAttribute_t attr, int device) {
  (void)device;
  if (attr != hipDeviceAttributeWallClockRate) return hipErrorInvalidValue;
  *value = 100000;  // kHz
  return hipSuccess;
}
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipDeviceSynchronize(void) {
  hipemu::drain_all();
  return hipSuccess;
}
hipError_t hipMalloc(void** p, size_t bytes) {
  *p = hipemu::guarded_alloc(bytes);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
  hipemu::drain_all();
  hipemu::guarded_free(p);
  return hipSuccess;
}
hipError_t hipGetSymbolAddress(void** devPtr, const void* symbol) {
  *devPtr = const_cast<void*>(symbol);
  return symbol ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipMemset(void* p, int v, size_t bytes) {
  memset(p, v, bytes);  // NULL stream: not ordered with non-blocking streams
  return hipSuccess;
}
hipError_t hipMemsetAsync(void* p, int v, size_t bytes, hipStream_t s) {
  if (hipemu::capturing()) {
    hipemu::record([p, v, bytes]() { memset(p, v, bytes); });
    return hipSuccess;
  }
  hipemu::submit(s, [p, v, bytes]() { memset(p, v, bytes); });
  return hipSuccess;
}
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind) {
  memmove(dst, src, bytes);  // NULL stream: not ordered with non-blocking streams
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind,
                          hipStream_t s) {
  if (hipemu::capturing()) {
    hipemu::record([dst, src, bytes]() { memmove(dst, src, bytes); });
    return hipSuccess;
  }
  hipemu::submit(s, [dst, src, bytes]() { memmove(dst, src, bytes); });
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) {
  *p = hipemu::guarded_alloc(bytes);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostGetDevicePointer(void** devPtr, void* hostPtr, unsigned) {
  *devPtr = hostPtr;  // one address space
  return hipSuccess;
}
hipError_t hipHostFree(void* p) {
  hipemu::drain_all();
  hipemu::guarded_free(p);
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  *s = new ihipStream_t();
  hipemu::g_streams.push_back(*s);
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  hipemu::run_until([s]() { return s->q.empty(); });
  auto& v = hipemu::g_streams;
  v.erase(std::remove(v.begin(), v.end(), s), v.end());
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
  // the NULL stream's work has run already
  if (s != nullptr) hipemu::run_until([s]() { return s->q.empty(); });
  return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
  const unsigned long long gen = e->recorded;
  if (e->completed >= gen) return hipSuccess;
  if (hipemu::deferred(s))
    s->q.push_back({2, nullptr, e, gen});
  else
    hipemu::run_until([e, gen]() { return e->completed >= gen; });
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = new ihipEvent_t();
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  const unsigned long long gen = ++e->recorded;
  if (hipemu::deferred(s)) {
    s->q.push_back({1, nullptr, e, gen});
  } else {
    e->t = std::chrono::steady_clock::now();
    e->completed = gen;
  }
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
  const unsigned long long gen = e->recorded;
  hipemu::run_until([e, gen]() { return e->completed >= gen; });
  return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  hipemu::drain_all();  // queued signals and waits may still name it
  delete e;
  return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e) {
  return e->completed >= e->recorded ? hipSuccess : hipErrorNotReady;
}
