// CPU emulation runtime of the HIP subset libmgpcg uses -- TEST INFRASTRUCTURE ONLY.
//
// Kernel launches run synchronously.  Each block's threads are ucontext
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

#include <chrono>
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

struct ihipStream_t {
  int dummy;
};
struct ihipEvent_t {
  std::chrono::steady_clock::time_point t;
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
}  // namespace
bool capturing() { return g_capture != nullptr; }
void record(std::function<void()> op) { g_capture->ops.push_back(std::move(op)); }
}  // namespace hipemu

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
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t) {
  if (hipemu::capturing()) return hipErrorInvalidValue;
  for (auto& op : x->ops) op();
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
hipError_t hipDeviceSynchronize(void) { return hipSuccess; }
hipError_t hipMalloc(void** p, size_t bytes) {
  *p = hipemu::guarded_alloc(bytes);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
  hipemu::guarded_free(p);
  return hipSuccess;
}
hipError_t hipMemset(void* p, int v, size_t bytes) {
  memset(p, v, bytes);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void* p, int v, size_t bytes, hipStream_t) {
  if (hipemu::capturing()) {
    hipemu::record([p, v, bytes]() { memset(p, v, bytes); });
    return hipSuccess;
  }
  memset(p, v, bytes);
  return hipSuccess;
}
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind) {
  memmove(dst, src, bytes);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind k,
                          hipStream_t) {
  if (hipemu::capturing()) {
    hipemu::record([dst, src, bytes]() { memmove(dst, src, bytes); });
    return hipSuccess;
  }
  return hipMemcpy(dst, src, bytes, k);
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) {
  *p = hipemu::guarded_alloc(bytes);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void* p) {
  hipemu::guarded_free(p);
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  *s = new ihipStream_t();
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = new ihipEvent_t();
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
  e->t = std::chrono::steady_clock::now();
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
