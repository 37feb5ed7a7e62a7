// Shared definitions of the MI355X MGPCG library: the colour-split padded
// level layout, fast integer division, the device-resident CG state and the
// error plumbing of the C-ABI.
//
// Level layout ("half-colour" storage)
// ------------------------------------
// A level of shape (n0, n1, n2) (2D grids use the view (1, n0, n1)) is stored
// as two colour arrays, colour c = (i + j + k) & 1 of the original cell index
// (mgpcg.py:107).  Each colour array is C-order over
//     [P0][P1][W],  P0 = n0 + 2 (1 in 2D),  P1 = n1 + 2,  W = (n2 + 3) / 2
// i.e. one ghost layer on every side (ghost cells are SOLID with zero
// coefficients and zero vectors, which realises the "out-of-range neighbour
// contributes 0" convention), and the contiguous axis split by colour: in
// padded row (i', j') colour c holds padded k' = 2 m + q at slot m, where
// q = (c + rp + 1) & 1 and rp = (i + j) & 1 is the original row parity.
// A colour-c cell at flat slot t then finds its other-colour neighbours at
//     t -+ P1*W  (axis 0, 3D),   t -+ W  (axis 1),   t + q - 1, t + q (axis 2)
// so a half-sweep streams each colour array contiguously.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MG_FLUID 0
#define MG_AIR 1
#define MG_SOLID 2

// Global (HBM) address space.  Pointers the kernels read from descriptor
// tables are generic to the compiler, which then emits flat loads; accesses
// through ldg / stg are global loads and stores.  Only for device memory
// from hipMalloc, never for LDS.
#ifndef MG_GLOBAL
#define MG_GLOBAL __attribute__((address_space(1)))
#endif
template <typename T>
__device__ __forceinline__ T ldg(const T* p) {
  return *(const MG_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ void stg(T* p, T v) {
  *(MG_GLOBAL T*)p = v;
}
// Element i (0 <= i, i * sizeof(T) < 2^32) of an array whose base is uniform:
// a 32-bit byte offset from the base lets the load use the scalar-base,
// vector-offset addressing form instead of 64-bit address arithmetic.
template <typename T>
__device__ __forceinline__ T ldgi(const T* base, int i) {
  const uint32_t off = (uint32_t)i * (uint32_t)sizeof(T);
  return *(const MG_GLOBAL T*)((const MG_GLOBAL char*)base + off);
}
template <typename T>
__device__ __forceinline__ void stgi(T* base, int i, T v) {
  const uint32_t off = (uint32_t)i * (uint32_t)sizeof(T);
  *(MG_GLOBAL T*)((MG_GLOBAL char*)base + off) = v;
}
// Makes a loaded value live at this point of the program, so the compiler
// cannot sink its load into a later branch (all of a slot's loads are then in
// flight together).  Emits no instruction.
// Lane `l` (wave-uniform) of a 32-bit value held across the wave.
#ifndef MG_READLANE
#define MG_READLANE(v, l) __builtin_amdgcn_readlane((v), (l))
#endif

#ifndef MG_KEEP
#define MG_KEEP(x) asm volatile("" ::"v"(x))
#define MG_KEEP4(a, b, c, d) asm volatile("" ::"v"(a), "v"(b), "v"(c), "v"(d))
#define MG_KEEP6(a, b, c, d, e, f) \
  asm volatile("" ::"v"(a), "v"(b), "v"(c), "v"(d), "v"(e), "v"(f))
#endif

namespace mg {

// Division by a runtime-invariant divisor (dividend < 2^31), multiply-high form.
struct FastDiv {
  uint32_t d, m, s;
  FastDiv() : d(1), m(0), s(0) {}
  explicit FastDiv(uint32_t div) : d(div), m(0), s(0) {
    if (d > 1) {
      uint32_t l = 0;
      while ((1u << l) < d) ++l;  // l = ceil(log2 d)
      uint32_t p = 31 + l;
      m = (uint32_t)(((1ull << p) + d - 1) / d);
      s = p - 32;
    }
  }
  __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const {
#ifdef __HIP_DEVICE_COMPILE__
    return d == 1 ? n : (__umulhi(n, m) >> s);
#else
    return d == 1 ? n : (uint32_t)(((uint64_t)n * m) >> 32) >> s;
#endif
  }
};

struct Geo {
  int dim;         // 2 or 3
  int n0, n1, n2;  // interior dims of the 3D view (n0 = 1 in 2D)
  int P0, P1, W;   // padded dims of one colour array
  int sx;          // axis-0 stride (P1 * W)
  int size;        // P0 * P1 * W
  int zpad;        // zero slots before and after each z colour array (>= the largest
                   // neighbour offset: sx in 3D, W in 2D; a multiple of 64)
  int ncells;      // n0 * n1 * n2
  FastDiv divW, divP1;
};

inline Geo make_geo(int dim, const int* n) {
  Geo g;
  g.dim = dim;
  if (dim == 2) {
    g.n0 = 1;
    g.n1 = n[0];
    g.n2 = n[1];
    g.P0 = 1;
  } else {
    g.n0 = n[0];
    g.n1 = n[1];
    g.n2 = n[2];
    g.P0 = g.n0 + 2;
  }
  g.P1 = g.n1 + 2;
  g.W = (g.n2 + 3) / 2;
  g.sx = g.P1 * g.W;
  g.size = g.P0 * g.P1 * g.W;
  g.zpad = ((dim == 3 ? g.sx : g.W) + 63) / 64 * 64;
  g.ncells = g.n0 * g.n1 * g.n2;
  g.divW = FastDiv((uint32_t)g.W);
  g.divP1 = FastDiv((uint32_t)g.P1);
  return g;
}

// Decoded slot: padded row (ip, jp), slot m, k' parity q, interior flag and
// original coordinates (i, j, k) of the 3D view.
struct Slot {
  int ip, jp, m, q, kp;
  int i, j, k;
  bool interior;
};

__host__ __device__ __forceinline__ Slot decode(const Geo& g, int t, int colour) {
  Slot s;
  uint32_t row = g.divW.div((uint32_t)t);
  s.m = t - (int)row * g.W;
  uint32_t ipu = g.divP1.div(row);
  s.ip = (int)ipu;
  s.jp = (int)row - s.ip * g.P1;
  // original indices (2D: the view's axis 0 has no ghost layer)
  s.i = (g.dim == 3) ? s.ip - 1 : 0;
  s.j = s.jp - 1;
  int rp = (s.i + s.j) & 1;
  s.q = (colour + rp + 1) & 1;
  s.kp = 2 * s.m + s.q;
  s.k = s.kp - 1;
  s.interior = (s.k >= 0) && (s.k < g.n2) && (s.j >= 0) && (s.j < g.n1) && (s.i >= 0) &&
               (s.i < g.n0);
  return s;
}

// Flat slot of original cell (i, j, k) (3D view) and its colour.
__host__ __device__ __forceinline__ int slot_of(const Geo& g, int i, int j, int k) {
  int ip = (g.dim == 3) ? i + 1 : 0;
  return (ip * g.P1 + (j + 1)) * g.W + ((k + 1) >> 1);
}
__host__ __device__ __forceinline__ int colour_of(int i, int j, int k) { return (i + j + k) & 1; }

// Dot products: every product of two FP32 values is exact in FP64 (for FP64
// operands TwoProd gives the exact product as a pair), and the sum is
// accumulated as a double-double (Knuth's TwoSum), so the rounded FP64 result
// does not depend on the summation order except at an exact rounding tie.
// The CPU oracle uses the same accumulation, so both agree to the last bit
// although they add in different orders.
struct DD {
  double hi, lo;
};
__host__ __device__ __forceinline__ DD dd_zero() { return DD{0.0, 0.0}; }
__host__ __device__ __forceinline__ DD dd_add(DD a, DD b) {
  const double s = a.hi + b.hi;
  const double bp = s - a.hi;
  double e = (a.hi - (s - bp)) + (b.hi - bp);
  e = e + (a.lo + b.lo);
  const double h = s + e;
  const double l = e - (h - s);
  return DD{h, l};
}
__host__ __device__ __forceinline__ DD dd_prod(float a, float b) {
  return DD{(double)a * (double)b, 0.0};  // exact: 24 + 24 bits < 53
}
__host__ __device__ __forceinline__ DD dd_prod(double a, double b) {
  const double p = a * b;
  return DD{p, fma(a, b, -p)};
}
__host__ __device__ __forceinline__ double dd_value(DD a) { return a.hi + a.lo; }

// Device-resident CG state of one solve (all scalars of mgpcg.py:150-209).
struct CGState {
  double r0r0, tol, zr_old, zr_new, pAp, alpha, beta, rr, true_rr;
  double rel_tol, abs_tol, eps;
  float alpha32, beta32;
  int iter, done, status, since_refine;
  int max_iters, refine_every, precision, pending_true;
  int phase;  // 0: before the loop (r = b, first V-cycle, first z.r), 1: in the loop
  int pad[3];
};

enum Precision { PLAIN = 0, MIXED = 1, F64 = 2 };
enum Status { RUNNING = 0, CONVERGED = 1, MAX_ITERS = 2, BREAKDOWN = 3 };

}  // namespace mg

// ---------------------------------------------------------------------------
// error plumbing of the C-ABI
// ---------------------------------------------------------------------------
void mg_set_error(const char* fmt, ...);

#define MG_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t _e = (call);                                                            \
    if (_e != hipSuccess) {                                                            \
      mg_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__,    \
                   __LINE__);                                                          \
      return -2;                                                                       \
    }                                                                                  \
  } while (0)

#define MG_REQUIRE(cond, ...)  \
  do {                         \
    if (!(cond)) {             \
      mg_set_error(__VA_ARGS__); \
      return -1;               \
    }                          \
  } while (0)
