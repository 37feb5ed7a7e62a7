// CPU emulation of the HIP subset libmgpcg uses -- TEST INFRASTRUCTURE ONLY.
//
// tests/cpu_emu/build_emu.py compiles the library's own HIP sources as host
// C++ with this directory first on the include path, producing
// tests/cpu_emu/libmgpcg_emu.so.  Kernels then run on the CPU: every GPU
// thread is a coroutine, __syncthreads and wave shuffles are emulated, and
// every "device" allocation is surrounded by guard pages, so an out-of-bounds
// access faults (or is reported by AddressSanitizer in the driver build) on
// the CPU, where it can be debugged.  The product never loads this build.
#pragma once
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <cmath>
#include <functional>
#include <memory>
#include <tuple>
#include <utility>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) alignas(n)
// a __device__ variable is a host variable here: its address is its "device" address
#define HIP_SYMBOL(x) ((const void*)&(x))
#define MG_GLOBAL
#define MG_READLANE(v, l) (::hipemu::readlane_i((v), (l)))
#define MG_KEEP(x) ((void)(x))
#define MG_KEEP4(a, b, c, d) ((void)(a), (void)(b), (void)(c), (void)(d))
#define MG_KEEP6(a, b, c, d, e, f) ((void)(a), (void)(b), (void)(c), (void)(d), (void)(e), (void)(f))

typedef enum hipError_t {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorNotReady = 600,
  hipErrorLaunchFailure = 719,
} hipError_t;

typedef enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4,
} hipMemcpyKind;

struct ihipStream_t;
struct ihipEvent_t;
struct ihipGraph_t;
struct ihipGraphExec_t;
struct ihipGraphNode_t;
typedef ihipStream_t* hipStream_t;
typedef ihipEvent_t* hipEvent_t;
typedef ihipGraph_t* hipGraph_t;
typedef ihipGraphExec_t* hipGraphExec_t;
typedef ihipGraphNode_t* hipGraphNode_t;
enum hipStreamCaptureMode {
  hipStreamCaptureModeGlobal = 0,
  hipStreamCaptureModeThreadLocal = 1,
  hipStreamCaptureModeRelaxed = 2,
};

#define hipStreamNonBlocking 1
#define hipEventDisableTiming 2

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace hipemu {
extern dim3 threadIdx, blockIdx, gridDim, blockDim;
void syncthreads();
double shfl_xor_d(double v, int mask);
int shfl_xor_i(int v, int mask);
float shfl_idx(float v, int src, int width);
double shfl_idx(double v, int src, int width);
float shfl_xor_f(float v, int mask);
unsigned long long ballot(int pred);
int readlane_i(int v, int lane);
// run `body(thread)` for every thread of every block of the grid
void run_grid(dim3 grid, dim3 block, void (*tramp)(void*), void* ctx);
void* guarded_alloc(size_t bytes);
void guarded_free(void* p);
// stream capture: while a capture is open, launches and async copies are
// recorded into the graph instead of running
bool capturing();
void record(std::function<void()> op);
// run `op` in stream order on `s` (at once, unless deferred streams are on)
void submit(hipStream_t s, std::function<void()> op);
}  // namespace hipemu

#define threadIdx (::hipemu::threadIdx)
#define blockIdx (::hipemu::blockIdx)
#define gridDim (::hipemu::gridDim)
#define blockDim (::hipemu::blockDim)

inline void __syncthreads() { hipemu::syncthreads(); }
inline double __shfl_xor(double v, int mask, int width = 64) {
  (void)width;
  return hipemu::shfl_xor_d(v, mask);
}
inline float __shfl_xor(float v, int mask, int width = 64) {
  (void)width;
  return hipemu::shfl_xor_f(v, mask);
}
inline unsigned long long __ballot(int pred) { return hipemu::ballot(pred); }
inline int __shfl_xor(int v, int mask, int width = 64) {
  (void)width;
  return hipemu::shfl_xor_i(v, mask);
}
template <typename T>
inline T __shfl(T v, int src, int width = 64) {
  return hipemu::shfl_idx(v, src, width);
}
inline void __threadfence() {}
inline void __threadfence_system() {}
// the emulator runs every lane on one host thread, so a plain update is atomic
inline unsigned atomicAdd(unsigned* p, unsigned v) {
  unsigned o = *p;
  *p = o + v;
  return o;
}
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
  const unsigned long long old = *p;
  *p = old + v;
  return old;
}
inline unsigned atomicMin(unsigned* p, unsigned v) {
  const unsigned old = *p;
  *p = v < old ? v : old;
  return old;
}
inline unsigned atomicMax(unsigned* p, unsigned v) {
  const unsigned old = *p;
  *p = v > old ? v : old;
  return old;
}
inline unsigned long long atomicMin(unsigned long long* p, unsigned long long v) {
  const unsigned long long old = *p;
  *p = v < old ? v : old;
  return old;
}
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) {
  const unsigned long long old = *p;
  *p = v > old ? v : old;
  return old;
}
// the device's constant-rate clock: 100 MHz ticks of the host's steady clock
unsigned long long wall_clock64();
inline unsigned __float_as_uint(float f) {
  unsigned u;
  memcpy(&u, &f, sizeof u);
  return u;
}
inline float __uint_as_float(unsigned u) {
  float f;
  memcpy(&f, &u, sizeof f);
  return f;
}
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline unsigned __umulhi(unsigned a, unsigned b) {
  return (unsigned)(((unsigned long long)a * b) >> 32);
}

// runtime API
const char* hipGetErrorString(hipError_t e);
hipError_t hipGetLastError(void);
hipError_t hipGetDevice(int* d);
enum hipDeviceAttribute_t { hipDeviceAttributeWallClockRate = 1 };
hipError_t hipDeviceGetAttribute(int* value, hipDeviceAttribute_t attr, int device);
hipError_t hipSetDevice(int d);
hipError_t hipDeviceSynchronize(void);
hipError_t hipMalloc(void** p, size_t bytes);
template <typename T>
inline hipError_t hipMalloc(T** p, size_t bytes) {
  return hipMalloc((void**)p, bytes);
}
hipError_t hipFree(void* p);
hipError_t hipMemset(void* p, int v, size_t bytes);
hipError_t hipGetSymbolAddress(void** devPtr, const void* symbol);
hipError_t hipMemsetAsync(void* p, int v, size_t bytes, hipStream_t s);
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind k,
                          hipStream_t s);
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned flags);
hipError_t hipHostFree(void* p);
hipError_t hipHostGetDevicePointer(void** devPtr, void* hostPtr, unsigned flags);
#define hipHostMallocMapped 0x2
#define hipHostMallocCoherent 0x40000000
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, hipGraphNode_t* err_node,
                               char* log, size_t log_size);
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t x);
hipError_t hipGraphDestroy(hipGraph_t g);

// kernel launch: every block's threads as coroutines, in stream order
namespace hipemu {
template <typename F, typename Tup>
struct LaunchCtx {
  F f;
  Tup args;
};
template <typename F, typename Tup>
void tramp(void* p) {
  auto* c = (LaunchCtx<F, Tup>*)p;
  std::apply(c->f, c->args);
}
template <typename... KArgs, typename... Args>
void launch(hipStream_t stream, void (*kernel)(KArgs...), dim3 grid, dim3 block, Args&&... args) {
  typedef std::tuple<typename std::decay<KArgs>::type...> Tup;
  typedef LaunchCtx<void (*)(KArgs...), Tup> Ctx;
  auto ctx = std::make_shared<Ctx>(Ctx{kernel, Tup(std::forward<Args>(args)...)});
  auto run = [ctx, grid, block]() {
    run_grid(grid, block, &tramp<void (*)(KArgs...), Tup>, ctx.get());
  };
  if (capturing())
    record(run);
  else
    submit(stream, run);
}
}  // namespace hipemu

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
  ::hipemu::launch((stream), kernel, dim3(grid), dim3(block), ##__VA_ARGS__)
