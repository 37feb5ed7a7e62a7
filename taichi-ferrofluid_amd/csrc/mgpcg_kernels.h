// HIP kernels of the MGPCG solver for gfx950 (MI355X), colour-split padded
// layout (see mgpcg_common.h).  Every kernel is memory bound; none uses MFMA.
//
// Batching and persistence: every stage kernel serves a whole batch of
// systems (of possibly different shapes) in one launch.  The host builds,
// per stage, an item table {descriptor, first virtual block}; a launch has a
// fixed number of workgroups that stride over the virtual blocks, each XCD
// taking a contiguous range of them so stencil neighbours share its L2.
// Systems that are done cost no dispatches.
//
// Exactness: floating-point operation order follows mgpcg.py and the library
// is built with -ffp-contract=off, so FP32 values on FLUID cells reproduce a
// CPU restatement bit for bit.  Fusions used here change no FLUID value:
//  * the first half-sweep from z = 0 is recomputed on the fly inside the
//    second one (z0 = r0 / Adiag0);
//  * the prolongation (mgpcg.py:120-123) is folded into the first post-
//    smoothing half-sweep: colour-1 cells read z0 + z_coarse[parent]; the
//    colour-0 FLUID values are rewritten by the next half-sweep without
//    reading their old value, so no prolonged value is ever stored.  Non-FLUID
//    z stays unprolonged; it is inert (every coupling to it is 0).
//  * update_p (mgpcg.py:234-238) is fused into compute_Ap with p double-
//    buffered by iteration parity; reduce(z, r) into the last half-sweep.
//
// Compact couplings: when every nonzero off-diagonal coefficient of a level
// has the same FP32 bit pattern (the pressure-type operators: -scale_A on
// every FLUID-FLUID face, pressure_project.py:55-57), the stencil reads a
// 16-bit per-slot flag word instead of the 2 * dim coefficients and uses
// (bit ? coef : 0).  That value is bit-identical to the stored coefficient,
// and the same multiply and add follow, so the result does not change.  The
// word also carries a 3-bit diagonal code: code n < 7 says Adiag equals -coef
// added n times to 0 in FP32, so the diagonal is rebuilt instead of read;
// code 7 falls back to the stored Adiag.
#pragma once
#include "mgpcg_common.h"

namespace mg {

constexpr int BLOCK = 256;
constexpr int COARSE_THREADS = 1024;
// slots per lane of the half-sweep stages: a virtual block of a half-sweep
// covers BLOCK * SLOT_LANES slots (a build-time choice)
#ifndef MG_SLOT_LANES
#define MG_SLOT_LANES 2
#endif
constexpr int SLOT_LANES = MG_SLOT_LANES;
static_assert(SLOT_LANES >= 1 && SLOT_LANES <= 8, "slots per lane");

// Nonzero-coupling summary of a level, accumulated on the device while the
// level is packed: the smallest and largest FP32 bit pattern of its nonzero
// Ax values (minbits = ~0, maxbits = 0 when there is none).
struct LevelMeta {
  uint32_t minbits, maxbits;
  uint32_t nfluid;  // FLUID cells of the level (the launch probe's byte model counts them)
};

// flag word of a slot: bit a = coupling to the upper neighbour on axis a is
// nonzero (Ax[I][a]), bit NB_DN + a = coupling to the lower one (Ax[I - e_a][a]),
// bits NB_AD.. = the diagonal's code: n < NB_AD_STORED when Adiag equals n
// couplings of the slot's magnitude added one at a time from 0 (diag_of_code),
// NB_AD_STORED when it is read from the Adiag array; bit NB_FL = the slot is
// FLUID (the half-sweeps on uniform-coupling levels read it instead of the
// cell type)
constexpr int NB_DN = 3;
constexpr int NB_AD = 6;
constexpr uint32_t NB_AD_STORED = 7u;
constexpr int NB_FL = 9;

template <typename T>
struct LevelPtrs {
  const int8_t* gt[2];
  const T* Adiag[2];
  const T* Ax[3][2];  // [axis][colour]
  T* r[2];
  T* z[2];
  const uint16_t* nb[2];   // flag word per slot
  const LevelMeta* meta;   // NULL: always read Ax
  T* res[2];               // restriction scratch: the level's residual (not the coarsest)
  // per block of BLOCK slots (slot / BLOCK): 1 when a slot of either colour is
  // FLUID.  A kernel whose work is on FLUID cells only skips the others'
  // blocks ("dry" blocks); what such a block would hold (0 residuals, 0 z on
  // colour 0 before the first half-sweep) is written by the packing kernel or
  // stored without loads.
  const uint8_t* wet;
  // the level's work orders: entry k says which block the k-th virtual block
  // of a system's stage processes and whether it is wet.  [0]: blocks of
  // BLOCK slots, in natural order (written by the packing kernel); [1]:
  // half-sweep virtual blocks of BLOCK * SLOT_LANES slots, the wet ones
  // spread evenly (k_level_order; level 0 on the batched kernels only).
  // NULL: natural order, wet from the block flags.
  const int32_t* ord[2];
};

// The level's single coupling value when its couplings are uniform.
template <typename T>
struct Coupling {
  bool uniform;
  T coef;
  T dg[8];  // diag_of_code(n, coef) for n < 8 (uniform levels; set by with_diags)
};
// The diagonal encoded by code n < NB_AD_STORED on a level whose couplings
// are all `coef`: -coef added n times to 0, in order, as the operator
// assembly accumulates it.  The packing kernel assigns a code only where
// this reproduces the stored Adiag bit for bit.
template <typename T>
__host__ __device__ __forceinline__ T diag_of_code(uint32_t n, T coef) {
  const T m = -coef;
  T a = (T)0;
#pragma unroll
  for (uint32_t i = 1; i < NB_AD_STORED; ++i) {
    const T b = a + m;
    a = i <= n ? b : a;
  }
  return a;
}

// the diagonals of every code, computed once for a coupling (the same
// operations as diag_of_code, on values that are uniform over the launch)
template <typename T>
__host__ __device__ __forceinline__ void with_diags(Coupling<T>& k) {
  T a = (T)0;
  k.dg[0] = a;
#pragma unroll
  for (int i = 1; i < 8; ++i) {
    a = a + (-k.coef);
    k.dg[i] = a;
  }
}

template <typename T>
__device__ __forceinline__ Coupling<T> coupling(const LevelPtrs<T>& P) {
  Coupling<T> k{false, (T)0};
  if (P.meta != nullptr) {
    const uint32_t mn = P.meta->minbits, mx = P.meta->maxbits;
    if (mx == 0u) {
      k.uniform = true;  // no nonzero coupling at all: every flag is 0
    } else if (mn == mx) {
      k.uniform = true;
      k.coef = (T)__uint_as_float(mn);
    }
  }
  return k;
}

// One level of one system.
template <typename T>
struct LevelDesc {
  Geo g, gc;          // this level, next coarser level
  LevelPtrs<T> P, C;  // this level, next coarser level
  CGState* st;
  const LevelMeta* stats;  // the level's metadata, whatever P.meta says
};

// Level 0 of one system and its CG vectors.
template <typename T, typename TX>
struct SysDesc {
  Geo g;
  LevelPtrs<T> P;
  T* b[2];
  TX* x[2];
  T* p[2][2];  // [iteration parity][colour]
  T* Ap[2];
  CGState* st;
  DD* partials;  // one per 256-slot block of level 0 (a workgroup run's total in its first)
  int nparts;
  int coarse_from;  // first level handled by the single-workgroup coarse kernel
  int L;
  int desc0;  // index of this system's level-0 LevelDesc
};

struct Item {
  int desc;    // LevelDesc / SysDesc index
  int block0;  // first virtual block of this item
};

constexpr int SPAN_SLOTS = 256;

struct Stage {
  const Item* items;
  int nitems;
  int nblocks;  // total virtual blocks
  // measurement only: when non-NULL, a level-0 smooth(0) launch adds the
  // level-0 cells of every system it sweeps (done systems are skipped) to
  // swept[0] and their FLUID cells to swept[1], and its span on the device's
  // constant-rate clock: workgroup b folds its start into t0[b % SPAN_SLOTS]
  // (atomic min) and its end into t1[b % SPAN_SLOTS] (atomic max), so few
  // workgroups share an address
  unsigned long long* swept;
  unsigned long long* t0;
  unsigned long long* t1;
};

// ---------------------------------------------------------------------------
// persistent launch helpers
// ---------------------------------------------------------------------------
// Virtual blocks handled by this workgroup.  Workgroups are dispatched
// round-robin over the 8 XCDs, so group x = blockIdx % ng (ng = min(8, grid))
// lives on one XCD; it takes the contiguous range [x*per, (x+1)*per) of
// virtual blocks and its workgroups stride through it.
struct VBIter {
  int vb, end, step;
};
__device__ __forceinline__ VBIter vb_begin(int nblocks) {
  const int G = gridDim.x;
  const int ng = G < 8 ? G : 8;
  const int grp = blockIdx.x % ng;
  const int nx = (G - grp + ng - 1) / ng;  // workgroups in this group
  const int j = blockIdx.x / ng;
  const int per = (nblocks + ng - 1) / ng;
  VBIter it;
  const int start = grp * per;
  it.end = min(nblocks, start + per);
  it.vb = start + j;
  it.step = nx;
  return it;
}

__device__ __forceinline__ int find_item(const Item* items, int n, int vb) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].block0 <= vb) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// wave64 / block reductions (fixed order -> deterministic)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ DD wave_sum_dd(DD v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    DD o;
    o.hi = __shfl_xor(v.hi, off, 64);
    o.lo = __shfl_xor(v.lo, off, 64);
    v = dd_add(v, o);
  }
  return v;
}

// Block total (every thread gets it) of a double-double.
template <int NT>
__device__ __forceinline__ DD block_sum_dd(DD v) {
  __shared__ double smh[NT / 64];
  __shared__ double sml[NT / 64];
  v = wave_sum_dd(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    smh[wid] = v.hi;
    sml[wid] = v.lo;
  }
  __syncthreads();
  DD t = dd_zero();
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) t = dd_add(t, DD{smh[w], sml[w]});
  __syncthreads();
  return t;
}

template <int NT>
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sm[NT / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sm[wid] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) t += sm[w];
  __syncthreads();
  return t;
}

// ---------------------------------------------------------------------------
// stencil helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int row_q(const Geo& g, int t, int colour) {
  uint32_t row = g.divW.div((uint32_t)t);
  uint32_t ip = g.divP1.div(row);
  int jp = (int)row - (int)ip * g.P1;
  int i = (g.dim == 3) ? (int)ip - 1 : 0;
  int j = jp - 1;
  return (colour + ((i + j) & 1) + 1) & 1;
}

// Neighbour slots of a colour-c cell at t (other colour), in the reference's
// accumulation order: axis 0 lower/upper, axis 1, axis 2 (mgpcg.py:98-101).
template <int DIM>
struct Nbr {
  int lo[DIM], hi[DIM];
};
template <int DIM>
__device__ __forceinline__ Nbr<DIM> neighbours(const Geo& g, int t, int q) {
  Nbr<DIM> n;
  int a = 0;
  if (DIM == 3) {
    n.lo[0] = t - g.sx;
    n.hi[0] = t + g.sx;
    a = 1;
  }
  n.lo[a] = t - g.W;
  n.hi[a] = t + g.W;
  n.lo[DIM - 1] = t + q - 1;
  n.hi[DIM - 1] = t + q;
  return n;
}

// neighbour_sum (mgpcg.py:96-102) with the other-colour values supplied by a
// functor zo(slot):  ret = ret + (Ax_lo * z_lo + Ax_own * z_hi)
template <typename T, int DIM, typename ZO>
__device__ __forceinline__ T nsum(const LevelPtrs<T>& P, int c, int t, const Nbr<DIM>& n,
                                  ZO zo) {
  T ret = (T)0;
  const Coupling<T> k = coupling(P);
  if (k.uniform) {
    const uint32_t f = P.nb[c][t];
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      const T alo = ((f >> (NB_DN + a)) & 1u) ? k.coef : (T)0;
      const T ahi = ((f >> a) & 1u) ? k.coef : (T)0;
      T lo = alo * zo(n.lo[a]);
      T hi = ahi * zo(n.hi[a]);
      T s = lo + hi;
      ret = ret + s;
    }
    return ret;
  }
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    T lo = P.Ax[a][c ^ 1][n.lo[a]] * zo(n.lo[a]);
    T hi = P.Ax[a][c][t] * zo(n.hi[a]);
    T s = lo + hi;
    ret = ret + s;
  }
  return ret;
}

// ---------------------------------------------------------------------------
// latency-oriented forms for the batched kernels: every load of a slot is
// issued before its FLUID test (neighbour indices are clamped into the padded
// arrays, so ghost slots stay in bounds; their results are discarded), loads
// and stores go through the global address space, and the work item with
// its per-system fields is looked up once per run of virtual blocks.
// ---------------------------------------------------------------------------
template <int DIM>
__device__ __forceinline__ Nbr<DIM> neighbours_clamped(const Geo& g, int t, int q) {
  Nbr<DIM> n = neighbours<DIM>(g, t, q);
  const int top = g.size - 1;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    n.lo[a] = n.lo[a] < 0 ? 0 : n.lo[a];
    n.hi[a] = n.hi[a] > top ? top : n.hi[a];
  }
  return n;
}

template <typename T>
__device__ __forceinline__ Coupling<T> coupling_g(const LevelPtrs<T>& P) {
  Coupling<T> k{false, (T)0};
  if (P.meta != nullptr) {
    const uint32_t mn = ldg(&P.meta->minbits), mx = ldg(&P.meta->maxbits);
    if (mx == 0u) {
      k.uniform = true;
    } else if (mn == mx) {
      k.uniform = true;
      k.coef = (T)__uint_as_float(mn);
    }
  }
  with_diags(k);
  return k;
}

// the lower / upper coupling of every axis at slot t of colour c
template <typename T, int DIM>
__device__ __forceinline__ void couplings_g(const LevelPtrs<T>& P, const Coupling<T>& k, int c,
                                            int t, const Nbr<DIM>& n, T* al, T* ah) {
  if (k.uniform) {
    const uint32_t f = ldgi(P.nb[c], t);
    MG_KEEP(f);
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      al[a] = ((f >> (NB_DN + a)) & 1u) ? k.coef : (T)0;
      ah[a] = ((f >> a) & 1u) ? k.coef : (T)0;
    }
  } else {
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      al[a] = ldgi(P.Ax[a][c ^ 1], n.lo[a]);
      ah[a] = ldgi(P.Ax[a][c], t);
    }
    if constexpr (DIM == 3)
      MG_KEEP6(al[0], al[1], al[2], ah[0], ah[1], ah[2]);
    else
      MG_KEEP4(al[0], al[1], ah[0], ah[1]);
  }
}

// neighbour_sum (mgpcg.py:96-102) from gathered couplings and values
template <typename T, int DIM>
__device__ __forceinline__ T nsum_vals(const T* al, const T* ah, const T* vl, const T* vh) {
  T ret = (T)0;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    T lo = al[a] * vl[a];
    T hi = ah[a] * vh[a];
    T s = lo + hi;
    ret = ret + s;
  }
  return ret;
}

// whether blocks [b0, b0 + n) of BLOCK slots of a level hold a FLUID slot
// (blocks past the level's end do not count; no flags: every block is wet)
template <typename T>
__device__ __forceinline__ bool wet_blocks(const LevelPtrs<T>& P, const Geo& g, int b0, int n) {
  if (P.wet == nullptr) return true;
  const int nblocks = (g.size + BLOCK - 1) / BLOCK;
  bool w = false;
  for (int i = 0; i < n; ++i)
    if (b0 + i < nblocks) w = w || ldg(P.wet + b0 + i) != 0;
  return w;
}

// Work order entries: the block to process (ORD_BLOCK bits) and whether it
// holds a FLUID slot (ORD_WET).  The packing pass spreads a level's wet blocks
// evenly over the positions of its order, so every workgroup's share of
// positions holds about the same number of wet blocks, whatever the layout of
// the scene's FLUID cells.
constexpr int32_t ORD_WET = (int32_t)0x80000000u;
constexpr int32_t ORD_BLOCK = 0x7fffffff;

// The order entries of a workgroup's next 64 virtual blocks, fetched at once
// (lane l loads the entry of the virtual block l iterations ahead) and read
// back one per iteration from the lane that holds it, instead of one
// dependent load per block.  `iter` counts the workgroup's loop iterations;
// the entries are refetched when the item changes (invalidate) or after 64
// iterations.  Every call is block-uniform.
struct OrdScan {
  int32_t v = 0;
  int base = 0;
  bool valid = false;
};
template <typename T>
__device__ __forceinline__ int32_t ord_here(OrdScan& w, int iter, int vb, int step, int end,
                                           int item_begin, int item_end, const LevelPtrs<T>& P,
                                           const Geo& g, int span) {
  if (!w.valid || iter - w.base >= 64) {
    const int l = threadIdx.x & 63;
    const int v = vb + l * step;
    int32_t e = 0;  // beyond the item: refetched when reached
    if (v < end && v < item_end) {
      const int k = v - item_begin;
      const int32_t* o = P.ord[span == 1 ? 0 : 1];
      e = o != nullptr ? ldg(o + k) : (k | (wet_blocks(P, g, k * span, span) ? ORD_WET : 0));
    }
    w.v = e;
    w.base = iter;
    w.valid = true;
  }
  return MG_READLANE(w.v, iter - w.base);
}

// the item of virtual block vb, kept while the workgroup's next virtual
// blocks fall into it; returns true when it changed
struct ItemCache {
  int begin = 0, end = 0, desc = 0, block0 = 0;
};
__device__ __forceinline__ bool item_at(const Stage& S, int vb, ItemCache& c) {
  if (vb >= c.begin && vb < c.end) return false;
  const int ii = find_item(S.items, S.nitems, vb);
  const Item it = S.items[ii];
  c.desc = it.desc;
  c.block0 = it.block0;
  c.begin = it.block0;
  c.end = ii + 1 < S.nitems ? S.items[ii + 1].block0 : S.nblocks;
  return true;
}

// Waves per SIMD a kernel asks the compiler to fit (its VGPR budget): 8 (64
// VGPRs) lets every workgroup of a full persistent launch (8 per CU) be
// resident at once; FP64 hierarchies would spill there and take the default.
template <typename T>
__host__ __device__ constexpr int full_occupancy() {
  return sizeof(T) == 4 ? 8 : 1;
}

// Dot-product partials of one workgroup over a run of its virtual blocks in
// one system: summed in registers and reduced over the workgroup once per run
// (instead of once per virtual block) into the partial of the run's first
// virtual block; the run's other virtual blocks get a zero partial, so every
// partial of a system is written exactly once per launch.  run_mark and
// run_flush are called by the whole workgroup (block-uniform control flow).
struct RunSum {
  DD acc = dd_zero();
  DD* dst = nullptr;  // receives the run's total
};
__device__ __forceinline__ void stg_dd(DD* p, DD v) {
  stg(&p->hi, v.hi);
  stg(&p->lo, v.lo);
}
__device__ __forceinline__ void run_mark(RunSum& r, DD* partials, int lb) {
  if (r.dst == nullptr) r.dst = partials + lb;
  else if (threadIdx.x == 0) stg_dd(partials + lb, dd_zero());
}
__device__ __forceinline__ void run_flush(RunSum& r) {
  if (r.dst == nullptr) return;
  const DD t = block_sum_dd<BLOCK>(r.acc);
  if (threadIdx.x == 0) stg_dd(r.dst, t);
  r.acc = dd_zero();
  r.dst = nullptr;
}

// Two-phase form of a half-sweep slot: every value a slot needs is loaded
// first (smooth_load), the update is computed and stored afterwards
// (smooth_store).  A lane that loads several slots before storing any keeps
// all their loads in flight at once (a store could alias a later load, so
// the compiler would not reorder them by itself).
template <typename T, int DIM, bool COMPACT>
struct SmoothVals {
  int gt;
  uint32_t f;  // flag word (COMPACT)
  T rv, ad;    // ad: the stored diagonal (not COMPACT)
  T zl[DIM], zh[DIM], al[DIM], ah[DIM];  // al / ah: stored couplings (not COMPACT)
};

// The diagonal of slot t of colour c: decoded from its flag word f on a level
// with uniform couplings (read from Adiag only for the code NB_AD_STORED),
// else the value `stored` loaded with the slot's other operands.
template <typename T, bool COMPACT>
__device__ __forceinline__ T slot_diag(const LevelPtrs<T>& P, const Coupling<T>& k, int c, int t,
                                       uint32_t f, T stored) {
  if constexpr (COMPACT) {
    const uint32_t n = (f >> NB_AD) & 7u;
    if (n == NB_AD_STORED) return ldgi(P.Adiag[c], t);
    // k.dg[n] by a select tree on the code's bits (no per-lane indexing)
    const bool b0 = (n & 1u) != 0u, b1 = (n & 2u) != 0u, b2 = (n & 4u) != 0u;
    const T d01 = b0 ? k.dg[1] : k.dg[0], d23 = b0 ? k.dg[3] : k.dg[2];
    const T d45 = b0 ? k.dg[5] : k.dg[4], d67 = b0 ? k.dg[7] : k.dg[6];
    const T lo = b1 ? d23 : d01, hi = b1 ? d67 : d45;
    return b2 ? hi : lo;
  } else {
    return stored;
  }
}

template <typename T, int DIM, bool COMPACT>
__device__ __forceinline__ void smooth_load(const Geo& g, const LevelPtrs<T>& P, int c, int t,
                                            SmoothVals<T, DIM, COMPACT>& v) {
  const int q = row_q(g, t, c);
  if constexpr (!COMPACT) v.gt = ldgi(P.gt[c], t);
  v.rv = ldgi(P.r[c], t);
  if constexpr (COMPACT) {
    // a z colour array has g.zpad zero slots before and after it, so the
    // unclamped neighbours of any slot stay inside the allocation: indices
    // are taken from the start of the front padding (never negative)
    const Nbr<DIM> n = neighbours<DIM>(g, t, q);
    const T* zb = P.z[c ^ 1] - g.zpad;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      v.zl[a] = ldgi(zb, n.lo[a] + g.zpad);
      v.zh[a] = ldgi(zb, n.hi[a] + g.zpad);
    }
    v.f = ldgi(P.nb[c], t);
    v.gt = ((v.f >> NB_FL) & 1u) ? MG_FLUID : MG_SOLID;  // FLUID or not is all that is read
  } else {
    const Nbr<DIM> n = neighbours_clamped<DIM>(g, t, q);
    v.ad = ldgi(P.Adiag[c], t);
    const T* zo = P.z[c ^ 1];
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      v.zl[a] = ldgi(zo, n.lo[a]);
      v.zh[a] = ldgi(zo, n.hi[a]);
    }
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      v.al[a] = ldgi(P.Ax[a][c ^ 1], n.lo[a]);
      v.ah[a] = ldgi(P.Ax[a][c], t);
    }
  }
}

// keeps a slot's loaded values live here (see MG_KEEP)
template <typename T, int DIM, bool COMPACT>
__device__ __forceinline__ void smooth_keep(const SmoothVals<T, DIM, COMPACT>& v) {
  if constexpr (COMPACT) {
    MG_KEEP(v.f);
  } else {
    MG_KEEP(v.gt);
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      MG_KEEP(v.al[a]);
      MG_KEEP(v.ah[a]);
    }
  }
}

// the new z of a FLUID slot: (r - neighbour_sum) / Adiag (mgpcg.py:104-109)
template <typename T, int DIM, bool COMPACT>
__device__ __forceinline__ T smooth_value(const LevelPtrs<T>& P, const Coupling<T>& k, int c,
                                          int t, const SmoothVals<T, DIM, COMPACT>& v) {
  T al[DIM], ah[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    if constexpr (COMPACT) {
      al[a] = ((v.f >> (NB_DN + a)) & 1u) ? k.coef : (T)0;
      ah[a] = ((v.f >> a) & 1u) ? k.coef : (T)0;
    } else {
      al[a] = v.al[a];
      ah[a] = v.ah[a];
    }
  }
  T sm = nsum_vals<T, DIM>(al, ah, v.zl, v.zh);
  T num = v.rv - sm;
  return num / slot_diag<T, COMPACT>(P, k, c, t, v.f, v.ad);
}

template <typename T, int DIM, bool COMPACT>
__device__ __forceinline__ void smooth_store(const LevelPtrs<T>& P, const Coupling<T>& k, int c,
                                             int t, const SmoothVals<T, DIM, COMPACT>& v) {
  if (v.gt != MG_FLUID) return;
  stgi(P.z[c], t, smooth_value<T, DIM, COMPACT>(P, k, c, t, v));
}

// smooth(l, c) at slot t in the latency-oriented form; returns the slot's z
// after the half-sweep (the new value on a FLUID cell)
template <typename T, int DIM, bool COMPACT>
__device__ __forceinline__ T smooth_slot_g(const Geo& g, const LevelPtrs<T>& P,
                                           const Coupling<T>& k, int c, int t) {
  SmoothVals<T, DIM, COMPACT> v;
  smooth_load<T, DIM, COMPACT>(g, P, c, t, v);
  smooth_keep<T, DIM, COMPACT>(v);
  if (v.gt != MG_FLUID) return (T)0;  // unused: callers mask with the cell type
  const T zn = smooth_value<T, DIM, COMPACT>(P, k, c, t, v);
  stgi(P.z[c], t, zn);
  return zn;
}

// smooth(l, 1) at a colour-1 slot followed by the slot's residual
// r - (Adiag z + neighbour_sum) into res[1] (0 off FLUID), the operation
// order of residual_slot: the colour-0 values do not change during the
// half-sweep, so its neighbour sum is the residual's
template <typename T, int DIM, bool COMPACT>
__device__ __forceinline__ void smooth_store_res(const LevelPtrs<T>& P, const Coupling<T>& k,
                                                 int t, const SmoothVals<T, DIM, COMPACT>& v) {
  T res = (T)0;
  if (v.gt == MG_FLUID) {
    T al[DIM], ah[DIM];
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      if constexpr (COMPACT) {
        al[a] = ((v.f >> (NB_DN + a)) & 1u) ? k.coef : (T)0;
        ah[a] = ((v.f >> a) & 1u) ? k.coef : (T)0;
      } else {
        al[a] = v.al[a];
        ah[a] = v.ah[a];
      }
    }
    const T sm = nsum_vals<T, DIM>(al, ah, v.zl, v.zh);
    const T num = v.rv - sm;
    const T ad = slot_diag<T, COMPACT>(P, k, 1, t, v.f, v.ad);
    const T z = num / ad;
    stgi(P.z[1], t, z);
    T Az = ad * z;
    Az = Az + sm;
    res = v.rv - Az;
  }
  stgi(P.res[1], t, res);
}

// SLOT_LANES slots per lane from t0 (stride BLOCK), all loaded before any is
// stored; RES: colour 1 with its residual (smooth_store_res)
template <typename T, int DIM, bool COMPACT, bool RES = false>
__device__ __forceinline__ void smooth_slots(const Geo& g, const LevelPtrs<T>& P,
                                             const Coupling<T>& k, int c, int t0) {
  SmoothVals<T, DIM, COMPACT> v[SLOT_LANES];
#pragma unroll
  for (int u = 0; u < SLOT_LANES; ++u) {
    const int t = t0 + u * BLOCK;
    smooth_load<T, DIM, COMPACT>(g, P, c, t < g.size ? t : g.size - 1, v[u]);
  }
#pragma unroll
  for (int u = 0; u < SLOT_LANES; ++u) smooth_keep<T, DIM, COMPACT>(v[u]);
#pragma unroll
  for (int u = 0; u < SLOT_LANES; ++u) {
    const int t = t0 + u * BLOCK;
    if (t >= g.size) continue;
    if constexpr (RES) smooth_store_res<T, DIM, COMPACT>(P, k, t, v[u]);
    else smooth_store<T, DIM, COMPACT>(P, k, c, t, v[u]);
  }
}

// Two-phase form of A v at one slot of colour c (both colours of a slot are
// loaded before either result is stored): own value, diagonal, neighbours
// and couplings.
template <typename T, int DIM, bool COMPACT>
struct StencilVals {
  int gt;
  uint32_t f;  // flag word (COMPACT)
  T own, ad;   // ad: the stored diagonal (not COMPACT)
  T vl[DIM], vh[DIM], al[DIM], ah[DIM];
};

// ZPAD: v is a level's z (g.zpad zero slots on both sides of each colour
// array), so on a COMPACT level the neighbours are loaded unclamped
template <typename T, int DIM, bool COMPACT, bool ZPAD = false>
__device__ __forceinline__ void stencil_load(const Geo& g, const LevelPtrs<T>& P, const T* const* v,
                                             int c, int t, StencilVals<T, DIM, COMPACT>& s) {
  constexpr bool UNCLAMPED = COMPACT && ZPAD;
  const int q = row_q(g, t, c);
  const Nbr<DIM> n = UNCLAMPED ? neighbours<DIM>(g, t, q) : neighbours_clamped<DIM>(g, t, q);
  s.gt = ldgi(P.gt[c], t);
  s.own = ldgi(v[c], t);
  if constexpr (!COMPACT) s.ad = ldgi(P.Adiag[c], t);
  const int off = UNCLAMPED ? g.zpad : 0;
  const T* vo = v[c ^ 1] - off;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    s.vl[a] = ldgi(vo, n.lo[a] + off);
    s.vh[a] = ldgi(vo, n.hi[a] + off);
  }
  if constexpr (COMPACT) {
    s.f = ldgi(P.nb[c], t);
  } else {
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      s.al[a] = ldgi(P.Ax[a][c ^ 1], n.lo[a]);
      s.ah[a] = ldgi(P.Ax[a][c], t);
    }
  }
}

template <typename T, int DIM, bool COMPACT>
__device__ __forceinline__ void stencil_keep(const StencilVals<T, DIM, COMPACT>& s) {
  MG_KEEP(s.gt);
  if constexpr (COMPACT) {
    MG_KEEP(s.f);
  } else {
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      MG_KEEP(s.al[a]);
      MG_KEEP(s.ah[a]);
    }
  }
}

// Adiag * own + neighbour_sum at slot t of colour c, in the reference order
// (compute_Ap, mgpcg.py:218-224)
template <typename T, int DIM, bool COMPACT>
__device__ __forceinline__ T stencil_apply(const LevelPtrs<T>& P, const Coupling<T>& k, int c,
                                           int t, const StencilVals<T, DIM, COMPACT>& s) {
  T al[DIM], ah[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    if constexpr (COMPACT) {
      al[a] = ((s.f >> (NB_DN + a)) & 1u) ? k.coef : (T)0;
      ah[a] = ((s.f >> a) & 1u) ? k.coef : (T)0;
    } else {
      al[a] = s.al[a];
      ah[a] = s.ah[a];
    }
  }
  T Av = slot_diag<T, COMPACT>(P, k, c, t, s.f, s.ad) * s.own;
  T ns = nsum_vals<T, DIM>(al, ah, s.vl, s.vh);
  return Av + ns;
}

// prolongate(l); smooth(l, 1) at colour-1 slot t in the latency-oriented form
// (same arithmetic as prolong_smooth1_slot below)
template <typename T, int DIM, bool COMPACT>
__device__ __forceinline__ void prolong_smooth1_slot_g(const Geo& g, const Geo& gc,
                                                       const LevelPtrs<T>& P,
                                                       const Coupling<T>& k, const T* zc0,
                                                       const T* zc1, int t) {
  const Slot c = decode(g, t, 1);
  // z0 has g.zpad zero slots on both sides: unclamped on COMPACT levels
  const Nbr<DIM> n = COMPACT ? neighbours<DIM>(g, t, c.q) : neighbours_clamped<DIM>(g, t, c.q);
  const int zoff = COMPACT ? g.zpad : 0;
  const int8_t gt = ldgi(P.gt[1], t);
  const T rv = ldgi(P.r[1], t);
  T ad = (T)0;
  if constexpr (!COMPACT) ad = ldgi(P.Adiag[1], t);
  const T* z0 = P.z[0];
  // The six colour-0 neighbours have four distinct parents: on every axis the
  // neighbour inside the cell's own 2-block shares the cell's parent p, the
  // other one (lower when the coordinate is even, upper when odd) has the
  // adjacent parent p -+ e_a, of the other colour; a neighbour outside the
  // level (only that second one can be) has no parent and keeps its z.
  // A ghost slot (not interior) decodes to -1 / n coordinates: its indices are
  // clamped into the coarse arrays and its result is discarded.
  const int pi = (DIM == 3 && c.i > 0) ? c.i >> 1 : 0;
  const int pj = c.j > 0 ? c.j >> 1 : 0, pk = c.k > 0 ? c.k >> 1 : 0;
  const int cp = colour_of(pi, pj, pk);
  const int sp = min(slot_of(gc, pi, pj, pk), gc.size - 1);
  const T* zcp = cp ? zc1 : zc0;
  const T* zcq = cp ? zc0 : zc1;
  const T vp = ldgi(zcp, sp);
  T zl[DIM], zh[DIM], vq[DIM], al[DIM], ah[DIM];
  bool up[DIM], inq[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    const int va = a + 3 - DIM;  // view axis of original axis a
    const int x = va == 0 ? c.i : (va == 1 ? c.j : c.k);
    const int nx = va == 0 ? g.n0 : (va == 1 ? g.n1 : g.n2);
    up[a] = (x & 1) != 0;  // the other parent lies above (odd x) or below (even x)
    const int d = up[a] ? 1 : -1;
    inq[a] = x + d >= 0 && x + d < nx;
    int sq;
    if (va == 0) sq = sp + d * gc.sx;
    else if (va == 1) sq = sp + d * gc.W;
    else sq = sp - ((pk + 1) >> 1) + ((pk + d + 1) >> 1);
    sq = (inq[a] && sq >= 0 && sq < gc.size) ? sq : sp;
    vq[a] = ldgi(zcq, sq);
    zl[a] = ldgi(z0 - zoff, n.lo[a] + zoff);
    zh[a] = ldgi(z0 - zoff, n.hi[a] + zoff);
  }
  uint32_t f = 0u;
  if constexpr (COMPACT) {
    f = ldgi(P.nb[1], t);
    MG_KEEP(f);
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      al[a] = ((f >> (NB_DN + a)) & 1u) ? k.coef : (T)0;
      ah[a] = ((f >> a) & 1u) ? k.coef : (T)0;
    }
  } else {
    couplings_g<T, DIM>(P, k, 1, t, n, al, ah);
  }
  if (gt != MG_FLUID) return;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    if (up[a]) {
      zl[a] = zl[a] + vp;
      if (inq[a]) zh[a] = zh[a] + vq[a];
    } else {
      if (inq[a]) zl[a] = zl[a] + vq[a];
      zh[a] = zh[a] + vp;
    }
  }
  T ret = nsum_vals<T, DIM>(al, ah, zl, zh);
  T num = rv - ret;
  stgi(P.z[1], t, num / slot_diag<T, COMPACT>(P, k, 1, t, f, ad));
}

// z.fill(0); smooth(l, 0) on colour 0 at slot t in the latency-oriented form:
// z0 = r0 / Adiag0 on FLUID, else 0
template <typename T, bool COMPACT>
__device__ __forceinline__ void first_slot_g(const LevelPtrs<T>& P, const Coupling<T>& k, int t) {
  const int8_t gt = ldgi(P.gt[0], t);
  const T rv = ldgi(P.r[0], t);
  uint32_t f = 0u;
  T ad = (T)0;
  if constexpr (COMPACT) f = ldgi(P.nb[0], t);
  else ad = ldgi(P.Adiag[0], t);
  T z = (T)0;
  if (gt == MG_FLUID) z = rv / slot_diag<T, COMPACT>(P, k, 0, t, f, ad);
  stgi(P.z[0], t, z);
}

// parent slot in the coarser level of fine-level slot `t` of colour `c`
// (fine original cell (i, j, k) -> (i/2, j/2, k/2)); returns colour in pc.
__device__ __forceinline__ int parent_slot(const Geo& gf, const Geo& gc, int t, int c, int& pc) {
  Slot s = decode(gf, t, c);
  const int pi = s.i >> 1, pj = s.j >> 1, pk = s.k >> 1;
  pc = colour_of(pi, pj, pk);
  return slot_of(gc, pi, pj, pk);
}

// ---------------------------------------------------------------------------
// per-slot stage bodies (shared by the batched kernels and the coarse kernel)
// ---------------------------------------------------------------------------
// The slot bodies take the level's z colour arrays explicitly (z[0], z[1]):
// the batched kernels pass the level's own arrays, the coarse kernel copies
// held in LDS.

// smooth(l, c) (mgpcg.py:104-109) at slot t of colour c.
template <typename T, int DIM>
__device__ __forceinline__ void smooth_slot(const Geo& g, const LevelPtrs<T>& P, T* const* z,
                                            int c, int t) {
  if (P.gt[c][t] != MG_FLUID) return;
  const int q = row_q(g, t, c);
  Nbr<DIM> n = neighbours<DIM>(g, t, q);
  const T* zo = z[c ^ 1];
  T s = nsum<T, DIM>(P, c, t, n, [&](int u) { return zo[u]; });
  T num = P.r[c][t] - s;
  z[c][t] = num / P.Adiag[c][t];
}
template <typename T, int DIM>
__device__ __forceinline__ void smooth_slot(const Geo& g, const LevelPtrs<T>& P, int c, int t) {
  smooth_slot<T, DIM>(g, P, P.z, c, t);
}

// z.fill(0); smooth(l, 0); smooth(l, 1) (mgpcg.py:126-130 / 132-134) at slot t:
// colour 0 from zero is r / Adiag, recomputed for every colour-1 stencil.
template <typename T, int DIM>
__device__ __forceinline__ void smooth_pair_from_zero_slot(const Geo& g, const LevelPtrs<T>& P,
                                                           T* const* z, int t) {
  const bool f0 = P.gt[0][t] == MG_FLUID;
  z[0][t] = f0 ? P.r[0][t] / P.Adiag[0][t] : (T)0;
  if (P.gt[1][t] != MG_FLUID) {
    z[1][t] = (T)0;
    return;
  }
  const int q = row_q(g, t, 1);
  Nbr<DIM> n = neighbours<DIM>(g, t, q);
  auto z0 = [&](int u) -> T {
    return P.gt[0][u] == MG_FLUID ? P.r[0][u] / P.Adiag[0][u] : (T)0;
  };
  T s = nsum<T, DIM>(P, 1, t, n, z0);
  T num = P.r[1][t] - s;
  z[1][t] = num / P.Adiag[1][t];
}
template <typename T, int DIM>
__device__ __forceinline__ void smooth_pair_from_zero_slot(const Geo& g, const LevelPtrs<T>& P,
                                                           int t) {
  smooth_pair_from_zero_slot<T, DIM>(g, P, P.z, t);
}

// prolongate(l); smooth(l, 1) (mgpcg.py:146-148) at colour-1 slot t: the
// colour-0 neighbours read z0 + z_coarse[parent].
template <typename T, int DIM>
__device__ __forceinline__ void prolong_smooth1_slot(const Geo& g, const Geo& gc,
                                                     const LevelPtrs<T>& P, T* const* z,
                                                     const T* zc0, const T* zc1, int t) {
  if (P.gt[1][t] != MG_FLUID) return;
  const Slot c = decode(g, t, 1);  // FLUID cells are interior
  Nbr<DIM> n = neighbours<DIM>(g, t, c.q);
  const T* z0 = z[0];
  // value of the colour-0 neighbour (a, dir) after prolongation; a ghost
  // neighbour (outside the level) has no parent and stays 0
  auto zp = [&](int a, int dir, int u) -> T {
    int i = c.i, j = c.j, k = c.k;
    const int va = a + 3 - DIM;  // view axis of original axis a
    if (va == 0) i += dir;
    else if (va == 1) j += dir;
    else k += dir;
    if (i < 0 || i >= g.n0 || j < 0 || j >= g.n1 || k < 0 || k >= g.n2) return z0[u];
    const int pi = i >> 1, pj = j >> 1, pk = k >> 1;
    const int ps = slot_of(gc, pi, pj, pk);
    const T zcv = colour_of(pi, pj, pk) ? zc1[ps] : zc0[ps];
    return z0[u] + zcv;
  };
  T ret = (T)0;
  const Coupling<T> k = coupling(P);
  const uint32_t f = k.uniform ? (uint32_t)P.nb[1][t] : 0u;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    const T alo = k.uniform ? (((f >> (NB_DN + a)) & 1u) ? k.coef : (T)0) : P.Ax[a][0][n.lo[a]];
    const T ahi = k.uniform ? (((f >> a) & 1u) ? k.coef : (T)0) : P.Ax[a][1][t];
    T lo = alo * zp(a, -1, n.lo[a]);
    T hi = ahi * zp(a, +1, n.hi[a]);
    T sm = lo + hi;
    ret = ret + sm;
  }
  T num = P.r[1][t] - ret;
  z[1][t] = num / P.Adiag[1][t];
}
template <typename T, int DIM>
__device__ __forceinline__ void prolong_smooth1_slot(const Geo& g, const Geo& gc,
                                                     const LevelPtrs<T>& P, const T* zc0,
                                                     const T* zc1, int t) {
  prolong_smooth1_slot<T, DIM>(g, gc, P, P.z, zc0, zc1, t);
}

// prolongate(l) alone (only when no post-smoothing follows): every interior cell.
template <typename T>
__device__ __forceinline__ void prolongate_slot(const Geo& g, const Geo& gc, T* const z[2],
                                                const T* zc0, const T* zc1, int c, int t) {
  Slot s = decode(g, t, c);
  if (!s.interior) return;
  const int pi = s.i >> 1, pj = s.j >> 1, pk = s.k >> 1;
  const int pc = colour_of(pi, pj, pk);
  const int ps = slot_of(gc, pi, pj, pk);
  z[c][t] = z[c][t] + (pc ? zc1[ps] : zc0[ps]);
}

// restrict(l) (mgpcg.py:111-118) for the interior coarse cell at coarse slot tc
// of colour cc: the residuals of the FLUID children, lexicographic order.
template <typename T, int DIM>
__device__ __forceinline__ void restrict_slot(const Geo& gf, const LevelPtrs<T>& F,
                                              T* const* z, const Geo& gc, T* rc0, T* rc1,
                                              int cc, int tc) {
  Slot s = decode(gc, tc, cc);
  if (!s.interior) return;
  T sum = (T)0;
  const int o0max = (DIM == 3) ? 2 : 1;
  for (int o0 = 0; o0 < o0max; ++o0)
    for (int o1 = 0; o1 < 2; ++o1)
      for (int o2 = 0; o2 < 2; ++o2) {
        const int i = (DIM == 3) ? 2 * s.i + o0 : 0;
        const int j = 2 * s.j + o1, k = 2 * s.k + o2;
        const int c = colour_of(i, j, k);
        const int t = slot_of(gf, i, j, k);
        if (F.gt[c][t] != MG_FLUID) continue;
        const int q = (k + 1) & 1;  // padded k' = k + 1 = 2m + q
        Nbr<DIM> n = neighbours<DIM>(gf, t, q);
        const T* zo = z[c ^ 1];
        T Az = F.Adiag[c][t] * z[c][t];
        T ns = nsum<T, DIM>(F, c, t, n, [&](int u) { return zo[u]; });
        Az = Az + ns;
        T res = F.r[c][t] - Az;
        sum = sum + res;
      }
  (cc ? rc1 : rc0)[tc] = sum;
}

// z <- 0 on both colours (a level that gets no smoothing).
template <typename T>
__device__ __forceinline__ void zero_slot(T* const* z, int t) {
  z[0][t] = (T)0;
  z[1][t] = (T)0;
}
template <typename T>
__device__ __forceinline__ void zero_slot(const LevelPtrs<T>& P, int t) {
  zero_slot<T>(P.z, t);
}

// z.fill(0); smooth(l, 0) on colour 0 only: z0 = r0 / Adiag0 (0 off FLUID).
// The colour-1 values left from before are not zeroed: the next half-sweep
// rewrites every colour-1 FLUID value without reading it, and the others are
// inert (zero couplings).
template <typename T>
__device__ __forceinline__ void first_half_sweep_slot(const LevelPtrs<T>& P, int t) {
  P.z[0][t] = P.gt[0][t] == MG_FLUID ? P.r[0][t] / P.Adiag[0][t] : (T)0;
}

// prolongate(l) on colour 0 only (mgpcg.py:120-123): the colour-1 FLUID values
// are rewritten by the smooth(l, 1) that follows without reading them.
template <typename T>
__device__ __forceinline__ void prolongate0_slot(const Geo& g, const Geo& gc, T* z0,
                                                 const T* zc0, const T* zc1, int t) {
  Slot s = decode(g, t, 0);
  if (!s.interior) return;
  const int pi = s.i >> 1, pj = s.j >> 1, pk = s.k >> 1;
  const int ps = slot_of(gc, pi, pj, pk);
  z0[t] = z0[t] + (colour_of(pi, pj, pk) ? zc1[ps] : zc0[ps]);
}

// ---------------------------------------------------------------------------
// batched stage kernels over level descriptors
// ---------------------------------------------------------------------------
enum StageOp {
  ST_SMOOTH0 = 0,       // smooth(l, 0)
  ST_SMOOTH1 = 1,       // smooth(l, 1)
  ST_PAIR_FROM_ZERO = 2,  // fill(0); smooth(l,0); smooth(l,1)
  ST_PROLONG_SMOOTH1 = 3,  // prolongate(l); smooth(l, 1)
  ST_ZERO = 4,
  ST_PROLONGATE = 5,
  ST_FIRST = 6,       // fill(0); smooth(l, 0), colour 0 only
  ST_PROLONGATE0 = 7,  // prolongate(l), colour 0 only
  ST_SMOOTH1_RES = 8,  // smooth(l, 1), and the colour-1 residual into res[1]
};

// L0 only separates instantiations in profiles; the code is the same: 0 the
// coarser levels, 1 level 0, 2 level 0's smooth(0) launches that the launch
// probe records (so a profiler's statistics of that instantiation cover
// exactly the launches the probe timed).
template <typename T, int DIM, int OP, int L0>
__global__ void __launch_bounds__(BLOCK) k_stage(const LevelDesc<T>* __restrict__ descs,
                                                 Stage S) {
  ItemCache ic;
  bool done = true;
  Coupling<T> kc{false, (T)0};
  if constexpr (L0 == 2 && OP == ST_SMOOTH0) {
    if (S.t0 != nullptr && threadIdx.x == 0)
      atomicMin(S.t0 + blockIdx.x % SPAN_SLOTS, (unsigned long long)wall_clock64());
  }
  OrdScan ws;
  int iter = 0;
  for (VBIter it = vb_begin(S.nblocks); it.vb < it.end; it.vb += it.step, ++iter) {
    if (item_at(S, it.vb, ic)) {
      ws.valid = false;
      const LevelDesc<T>& D = descs[ic.desc];
      done = ldg(&D.st->done) != 0;
      kc = coupling_g(D.P);
    }
    if (done) continue;
    const LevelDesc<T>& D = descs[ic.desc];
    if constexpr (L0 == 2 && OP == ST_SMOOTH0) {
      if (S.swept != nullptr && it.vb == ic.block0 && threadIdx.x == 0) {
        atomicAdd(S.swept, (unsigned long long)D.g.ncells);
        atomicAdd(S.swept + 1, (unsigned long long)ldg(&D.stats->nfluid));
      }
    }
    if constexpr (OP == ST_SMOOTH0 || OP == ST_SMOOTH1 || OP == ST_SMOOTH1_RES) {
      constexpr int c = OP == ST_SMOOTH0 ? 0 : 1;
      constexpr bool RES = OP == ST_SMOOTH1_RES;
      const int32_t e = ord_here(ws, iter, it.vb, it.step, it.end, ic.block0, ic.end, D.P, D.g, SLOT_LANES);
      if (e >= 0) continue;  // no FLUID cell
      const int t0 = (e & ORD_BLOCK) * (BLOCK * SLOT_LANES) + threadIdx.x;
      if (kc.uniform) smooth_slots<T, DIM, true, RES>(D.g, D.P, kc, c, t0);
      else smooth_slots<T, DIM, false, RES>(D.g, D.P, kc, c, t0);
      continue;
    }
    int lb = it.vb - ic.block0;
    // level 0's blocks without a FLUID cell keep z = 0 from their packing: no
    // kernel writes them (a FLUID cell's neighbours there have zero couplings)
    if constexpr (OP == ST_PROLONG_SMOOTH1 || (OP == ST_PROLONGATE0 && L0 > 0)) {
      // before any per-lane exit: the scan reads lanes
      const int32_t e = ord_here(ws, iter, it.vb, it.step, it.end, ic.block0, ic.end, D.P, D.g, 1);
      if (e >= 0) continue;
      lb = e & ORD_BLOCK;
    }
    const int t = lb * BLOCK + threadIdx.x;
    if (t >= D.g.size) continue;
    if (OP == ST_FIRST) {
      if (kc.uniform) first_slot_g<T, true>(D.P, kc, t);
      else first_slot_g<T, false>(D.P, kc, t);
    } else if (OP == ST_PROLONGATE0) {
      Slot sl = decode(D.g, t, 0);
      if (!sl.interior) continue;
      const int pi = sl.i >> 1, pj = sl.j >> 1, pk = sl.k >> 1;
      const int ps = slot_of(D.gc, pi, pj, pk);
      const T zc = ldg((colour_of(pi, pj, pk) ? D.C.z[1] : D.C.z[0]) + ps);
      const T zv = ldgi(D.P.z[0], t);
      stgi(D.P.z[0], t, zv + zc);
    } else if (OP == ST_PAIR_FROM_ZERO) {
      smooth_pair_from_zero_slot<T, DIM>(D.g, D.P, t);
    } else if (OP == ST_PROLONG_SMOOTH1) {
      if (kc.uniform)
        prolong_smooth1_slot_g<T, DIM, true>(D.g, D.gc, D.P, kc, D.C.z[0], D.C.z[1], t);
      else
        prolong_smooth1_slot_g<T, DIM, false>(D.g, D.gc, D.P, kc, D.C.z[0], D.C.z[1], t);
    } else if (OP == ST_ZERO) {
      zero_slot<T>(D.P, t);
    } else if (OP == ST_PROLONGATE) {
      prolongate_slot<T>(D.g, D.gc, D.P.z, D.C.z[0], D.C.z[1], 0, t);
      prolongate_slot<T>(D.g, D.gc, D.P.z, D.C.z[0], D.C.z[1], 1, t);
    }
  }
  if constexpr (L0 == 2 && OP == ST_SMOOTH0) {
    if (S.t1 != nullptr) {  // uniform: a kernel argument
      __syncthreads();      // every wave of the workgroup has finished
      if (threadIdx.x == 0)
        atomicMax(S.t1 + blockIdx.x % SPAN_SLOTS, (unsigned long long)wall_clock64());
    }
  }
}

// restrict(l) (mgpcg.py:111-118), row-tiled: a virtual block is one coarse
// row (I, J) of a system.  The workgroup forms the residuals of the 2^(d-1)
// fine rows (2I + o0, 2J + o1) -- consecutive lanes take consecutive k, so
// both colour arrays are read in contiguous runs -- keeps them in LDS, then
// sums every coarse cell's children in lexicographic order (o0, o1, o2)
// (0 for a non-FLUID child, which leaves the sum unchanged).
constexpr int RESTRICT_ROW_MAX = 512;  // longest fine row (n2) of this kernel

template <typename T, int DIM>
__global__ void __launch_bounds__(BLOCK) k_restrict_rows(const LevelDesc<T>* __restrict__ descs,
                                                         Stage S) {
  constexpr int NR = (DIM == 3) ? 4 : 2;  // fine rows per coarse row
  __shared__ T res[NR * RESTRICT_ROW_MAX];
  ItemCache ic;
  bool done = true;
  Coupling<T> kc{false, (T)0};
  for (VBIter it = vb_begin(S.nblocks); it.vb < it.end; it.vb += it.step) {
    if (item_at(S, it.vb, ic)) {
      const LevelDesc<T>& D = descs[ic.desc];
      done = ldg(&D.st->done) != 0;
      kc = coupling_g(D.P);
    }
    if (done) continue;
    const LevelDesc<T>& D = descs[ic.desc];
    const Geo& gf = D.g;
    const Geo& gc = D.gc;
    const int row = it.vb - ic.block0;  // coarse row index I * nc1 + J
    if (row >= gc.n0 * gc.n1) continue;
    const int I = (DIM == 3) ? row / gc.n1 : 0;
    const int J = (DIM == 3) ? row - I * gc.n1 : row;
    const int n2 = gf.n2;
    if (n2 > RESTRICT_ROW_MAX) continue;  // the host never selects this kernel then
    for (int idx = threadIdx.x; idx < NR * n2; idx += BLOCK) {
      const int rr = idx / n2;
      const int k = idx - rr * n2;
      const int i = (DIM == 3) ? 2 * I + (rr >> 1) : 0;
      const int j = 2 * J + (rr & 1);
      const int c = colour_of(i, j, k);
      const int t = slot_of(gf, i, j, k);
      const int q = (k + 1) & 1;  // padded k' = k + 1 = 2m + q
      const Nbr<DIM> n = neighbours<DIM>(gf, t, q);  // an interior cell: in bounds
      const int8_t gt = ldgi(D.P.gt[c], t);
      const T zv = ldgi(D.P.z[c], t);
      const T ad = ldgi(D.P.Adiag[c], t);
      const T rv = ldgi(D.P.r[c], t);
      const T* zo = D.P.z[c ^ 1];
      T zl[DIM], zh[DIM], al[DIM], ah[DIM];
#pragma unroll
      for (int a = 0; a < DIM; ++a) {
        zl[a] = ldgi(zo, n.lo[a]);
        zh[a] = ldgi(zo, n.hi[a]);
      }
      couplings_g<T, DIM>(D.P, kc, c, t, n, al, ah);
      T v = (T)0;
      if (gt == MG_FLUID) {
        T Az = ad * zv;
        T ns = nsum_vals<T, DIM>(al, ah, zl, zh);
        Az = Az + ns;
        v = rv - Az;
      }
      res[idx] = v;
    }
    __syncthreads();
    for (int K = threadIdx.x; K < gc.n2; K += BLOCK) {
      T sum = (T)0;
#pragma unroll
      for (int o = 0; o < NR; ++o) {
        sum = sum + res[o * n2 + 2 * K];
        sum = sum + res[o * n2 + 2 * K + 1];
      }
      const int cc = colour_of(I, J, K);
      stg((cc ? D.C.r[1] : D.C.r[0]) + slot_of(gc, I, J, K), sum);
    }
    __syncthreads();  // res is reused by the next virtual block
  }
}

template <typename T, int DIM, bool COMPACT>
__device__ __forceinline__ void residual_slot(const LevelDesc<T>& D, const Coupling<T>& k,
                                              int t) {
#pragma unroll
  for (int c = 0; c < 2; ++c) {  // one colour after the other
    StencilVals<T, DIM, COMPACT> s;
    stencil_load<T, DIM, COMPACT, true>(D.g, D.P, D.P.z, c, t, s);
    const T rv = ldgi(D.P.r[c], t);
    stencil_keep<T, DIM, COMPACT>(s);
    T v = (T)0;
    if (s.gt == MG_FLUID) v = rv - stencil_apply<T, DIM, COMPACT>(D.P, k, c, t, s);
    stgi(D.P.res[c], t, v);
  }
}

// the colour-0 residual at SLOT_LANES slots per lane from t0 (stride BLOCK),
// all loaded before any is stored
template <typename T, int DIM, bool COMPACT>
__device__ __forceinline__ void residual0_slots(const LevelDesc<T>& D, const Coupling<T>& k,
                                                int t0) {
  StencilVals<T, DIM, COMPACT> s[SLOT_LANES];
  T rv[SLOT_LANES];
#pragma unroll
  for (int u = 0; u < SLOT_LANES; ++u) {
    const int t = t0 + u * BLOCK;
    const int tt = t < D.g.size ? t : D.g.size - 1;
    stencil_load<T, DIM, COMPACT, true>(D.g, D.P, D.P.z, 0, tt, s[u]);
    rv[u] = ldgi(D.P.r[0], tt);
  }
#pragma unroll
  for (int u = 0; u < SLOT_LANES; ++u) stencil_keep<T, DIM, COMPACT>(s[u]);
#pragma unroll
  for (int u = 0; u < SLOT_LANES; ++u) {
    const int t = t0 + u * BLOCK;
    if (t >= D.g.size) continue;
    T v = (T)0;
    if (s[u].gt == MG_FLUID) v = rv[u] - stencil_apply<T, DIM, COMPACT>(D.P, k, 0, t, s[u]);
    stgi(D.P.res[0], t, v);
  }
}

// restrict(l) (mgpcg.py:111-118) in two coalesced passes.  Pass 1: the
// residual r - (Adiag z + neighbour_sum) of every FLUID cell, 0 elsewhere,
// into the level's scratch arrays, both colours of a slot per thread; C0:
// colour 0 only (the last smooth(l, 1) wrote colour 1's, ST_SMOOTH1_RES), on
// a half-sweep stage of BLOCK * SLOT_LANES slots per virtual block.
template <typename T, int DIM, bool C0 = false>
__global__ void __launch_bounds__(BLOCK) k_residual(const LevelDesc<T>* __restrict__ descs,
                                                    Stage S) {
  ItemCache ic;
  bool done = true;
  Coupling<T> kc{false, (T)0};
  OrdScan ws;
  int iter = 0;
  for (VBIter it = vb_begin(S.nblocks); it.vb < it.end; it.vb += it.step, ++iter) {
    if (item_at(S, it.vb, ic)) {
      ws.valid = false;
      const LevelDesc<T>& D = descs[ic.desc];
      done = ldg(&D.st->done) != 0;
      kc = coupling_g(D.P);
    }
    if (done) continue;
    const LevelDesc<T>& D = descs[ic.desc];
    // a block without FLUID cells keeps the 0 residuals the packing wrote
    if constexpr (C0) {
      const int32_t e = ord_here(ws, iter, it.vb, it.step, it.end, ic.block0, ic.end, D.P, D.g, SLOT_LANES);
      if (e >= 0) continue;
      const int t0 = (e & ORD_BLOCK) * (BLOCK * SLOT_LANES) + threadIdx.x;
      if (kc.uniform) residual0_slots<T, DIM, true>(D, kc, t0);
      else residual0_slots<T, DIM, false>(D, kc, t0);
    } else {
      const int32_t e = ord_here(ws, iter, it.vb, it.step, it.end, ic.block0, ic.end, D.P, D.g, 1);
      if (e >= 0) continue;
      const int t = (e & ORD_BLOCK) * BLOCK + threadIdx.x;
      if (t >= D.g.size) continue;
      if (kc.uniform) residual_slot<T, DIM, true>(D, kc, t);
      else residual_slot<T, DIM, false>(D, kc, t);
    }
  }
}

// Pass 2: every interior coarse cell sums its children's residuals in
// lexicographic order (o0, o1, o2) into the coarse r (a 0 residual leaves the
// sum unchanged, as skipping a non-FLUID child does).
template <typename T, int DIM>
__global__ void __launch_bounds__(BLOCK) k_restrict_gather(const LevelDesc<T>* __restrict__ descs,
                                                           Stage S) {
  ItemCache ic;
  bool done = true;
  OrdScan ws;
  int iter = 0;
  for (VBIter it = vb_begin(S.nblocks); it.vb < it.end; it.vb += it.step, ++iter) {
    if (item_at(S, it.vb, ic)) {
      ws.valid = false;
      done = ldg(&descs[ic.desc].st->done) != 0;
    }
    if (done) continue;
    const LevelDesc<T>& D = descs[ic.desc];
    // a coarse cell that is not FLUID never has its r read: a block of such
    // coarse cells is skipped
    const int32_t e = ord_here(ws, iter, it.vb, it.step, it.end, ic.block0, ic.end, D.C, D.gc, 1);
    if (e >= 0) continue;
    const int tc = (e & ORD_BLOCK) * BLOCK + threadIdx.x;
    if (tc >= D.gc.size) continue;
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
      const Slot s = decode(D.gc, tc, cc);
      if (!s.interior) continue;
      T v[8];
      int nch = 0;
#pragma unroll
      for (int o0 = 0; o0 < (DIM == 3 ? 2 : 1); ++o0)
#pragma unroll
        for (int o1 = 0; o1 < 2; ++o1)
#pragma unroll
          for (int o2 = 0; o2 < 2; ++o2) {
            const int i = (DIM == 3) ? 2 * s.i + o0 : 0;
            const int j = 2 * s.j + o1, k = 2 * s.k + o2;
            v[nch++] = ldgi(D.P.res[colour_of(i, j, k)], slot_of(D.g, i, j, k));
          }
      T sum = (T)0;
#pragma unroll
      for (int o = 0; o < (DIM == 3 ? 8 : 4); ++o) sum = sum + v[o];
      stgi(cc ? D.C.r[1] : D.C.r[0], tc, sum);
    }
  }
}

// restrict(l), child-parallel: 16 lanes per coarse slot -- 8 per coarse
// colour, one per child (o0, o1, o2) in lexicographic order.  Every lane
// forms its child's residual (0 when the child is not FLUID, which leaves the
// sum unchanged); the first lane of the eight gathers them with shuffles and
// adds them in the reference order.  Virtual blocks cover 16 coarse slots.
constexpr int RESTRICT_SLOTS_PER_BLOCK = BLOCK / 16;

template <typename T, int DIM>
__device__ __forceinline__ T child_res(const Geo& gf, const LevelPtrs<T>& F, const Coupling<T>& k,
                                       const Slot& s, int ch) {
  const int o0 = (ch >> 2) & 1, o1 = (ch >> 1) & 1, o2 = ch & 1;
  if (DIM == 2 && o0) return (T)0;
  const int i = (DIM == 3) ? 2 * s.i + o0 : 0;
  const int j = 2 * s.j + o1, kk = 2 * s.k + o2;
  const int c = colour_of(i, j, kk);
  const int t = slot_of(gf, i, j, kk);
  const int q = (kk + 1) & 1;  // padded k' = k + 1 = 2m + q
  const Nbr<DIM> n = neighbours<DIM>(gf, t, q);  // a real cell: in bounds
  const int8_t gt = ldgi(F.gt[c], t);
  const T zv = ldgi(F.z[c], t);
  const T ad = ldgi(F.Adiag[c], t);
  const T rv = ldgi(F.r[c], t);
  const T* zo = F.z[c ^ 1];
  T zl[DIM], zh[DIM], al[DIM], ah[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    zl[a] = ldgi(zo, n.lo[a]);
    zh[a] = ldgi(zo, n.hi[a]);
  }
  couplings_g<T, DIM>(F, k, c, t, n, al, ah);
  if (gt != MG_FLUID) return (T)0;
  T Az = ad * zv;
  T ns = nsum_vals<T, DIM>(al, ah, zl, zh);
  Az = Az + ns;
  return rv - Az;
}

template <typename T, int DIM>
__global__ void __launch_bounds__(BLOCK) k_restrict(const LevelDesc<T>* __restrict__ descs,
                                                    Stage S) {
  ItemCache ic;
  bool done = true;
  Coupling<T> kc{false, (T)0};
  for (VBIter it = vb_begin(S.nblocks); it.vb < it.end; it.vb += it.step) {
    if (item_at(S, it.vb, ic)) {
      const LevelDesc<T>& D = descs[ic.desc];
      done = ldg(&D.st->done) != 0;
      kc = coupling_g(D.P);
    }
    if (done) continue;
    const LevelDesc<T>& D = descs[ic.desc];
    const int tc = (it.vb - ic.block0) * RESTRICT_SLOTS_PER_BLOCK + (threadIdx.x >> 4);
    const int cc = (threadIdx.x >> 3) & 1;
    const int ch = threadIdx.x & 7;
    T res = (T)0;
    bool interior = false;
    if (tc < D.gc.size) {
      const Slot s = decode(D.gc, tc, cc);
      interior = s.interior;
      if (interior) res = child_res<T, DIM>(D.g, D.P, kc, s, ch);
    }
    // gather the eight children into lane ch == 0 through LDS (all lanes take part)
    __shared__ T xch[BLOCK];
    xch[threadIdx.x] = res;
    __syncthreads();
    if (ch == 0 && interior) {
      T sum = (T)0;
#pragma unroll
      for (int o = 0; o < 8; ++o) sum = sum + xch[threadIdx.x + o];
      stgi(cc ? D.C.r[1] : D.C.r[0], tc, sum);
    }
    __syncthreads();  // xch is reused by the next virtual block
  }
}

// ---------------------------------------------------------------------------
// coarse levels: one workgroup per system runs levels [from, L) of the V-cycle
// (mgpcg.py:127-148 restricted to those levels), barriers between stages.
// ---------------------------------------------------------------------------
template <typename T, int DIM>
__device__ void coarse_pre(const Geo& g, const LevelPtrs<T>& P, int nu) {
  const int n = g.size;
  if (nu <= 0) {
    for (int t = threadIdx.x; t < n; t += COARSE_THREADS) zero_slot<T>(P, t);
    __syncthreads();
    return;
  }
  for (int t = threadIdx.x; t < n; t += COARSE_THREADS)
    smooth_pair_from_zero_slot<T, DIM>(g, P, t);
  __syncthreads();
  for (int i = 1; i < nu; ++i) {
    for (int t = threadIdx.x; t < n; t += COARSE_THREADS) smooth_slot<T, DIM>(g, P, 0, t);
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += COARSE_THREADS) smooth_slot<T, DIM>(g, P, 1, t);
    __syncthreads();
  }
}

// zc: the next coarser level's z colour arrays
template <typename T, int DIM>
__device__ void coarse_post(const Geo& g, const Geo& gc, const LevelPtrs<T>& P,
                            T* const* zc, int nu) {
  const int n = g.size;
  if (nu <= 0) {
    for (int t = threadIdx.x; t < n; t += COARSE_THREADS) {
      prolongate_slot<T>(g, gc, P.z, zc[0], zc[1], 0, t);
      prolongate_slot<T>(g, gc, P.z, zc[0], zc[1], 1, t);
    }
    __syncthreads();
    return;
  }
  for (int t = threadIdx.x; t < n; t += COARSE_THREADS)
    prolong_smooth1_slot<T, DIM>(g, gc, P, zc[0], zc[1], t);
  __syncthreads();
  for (int t = threadIdx.x; t < n; t += COARSE_THREADS) smooth_slot<T, DIM>(g, P, 0, t);
  __syncthreads();
  for (int i = 1; i < nu; ++i) {
    for (int t = threadIdx.x; t < n; t += COARSE_THREADS) smooth_slot<T, DIM>(g, P, 1, t);
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += COARSE_THREADS) smooth_slot<T, DIM>(g, P, 0, t);
    __syncthreads();
  }
}

// bottom solve (mgpcg.py:136-142): half x (smooth 0, smooth 1) from z = 0,
// then half x (smooth 1, smooth 0).
template <typename T, int DIM>
__device__ void coarse_bottom(const Geo& g, const LevelPtrs<T>& P, int bottom) {
  const int n = g.size;
  const int half = bottom / 2;
  if (half <= 0) {
    for (int t = threadIdx.x; t < n; t += COARSE_THREADS) zero_slot<T>(P, t);
    __syncthreads();
    return;
  }
  for (int t = threadIdx.x; t < n; t += COARSE_THREADS)
    smooth_pair_from_zero_slot<T, DIM>(g, P, t);
  __syncthreads();
  for (int i = 1; i < half; ++i) {
    for (int t = threadIdx.x; t < n; t += COARSE_THREADS) smooth_slot<T, DIM>(g, P, 0, t);
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += COARSE_THREADS) smooth_slot<T, DIM>(g, P, 1, t);
    __syncthreads();
  }
  for (int i = 0; i < half; ++i) {
    for (int t = threadIdx.x; t < n; t += COARSE_THREADS) smooth_slot<T, DIM>(g, P, 1, t);
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += COARSE_THREADS) smooth_slot<T, DIM>(g, P, 0, t);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// The coarse V-cycle on its LDS image.  Every phase gives each lane up to
// CZ_SLOTS slots and gathers the operands of all of them before computing any
// (within a phase no slot reads a value the phase writes), then stores: a lane
// keeps all its slots' LDS reads in flight together instead of one slot's
// chain of reads at a time.  The arithmetic is the global-memory form's,
// operation for operation (smooth_slot, restrict_slot, prolong_smooth1_slot).
// ---------------------------------------------------------------------------
constexpr int CZ_SLOTS = 3;
// slots per lane of a phase that also gathers stored couplings (more
// operands per slot: fewer slots keep them in registers)
template <bool COMPACT>
__host__ __device__ constexpr int cz_slots() {
  return COMPACT ? CZ_SLOTS : 2;
}

// operands of a half-sweep, or of a residual (own = the slot's z), at one
// slot: on a level with uniform couplings (COMPACT) the flag word, else the
// stored couplings (global memory, at clamped indices: a ghost slot's reads
// stay in bounds and are discarded).  Only loads: the couplings are decoded
// when the value is formed, so no read waits for another.
template <typename T, int DIM, bool COMPACT>
struct CzOps {
  int gt;
  uint32_t f;
  T rv, ad, own;
  T al[DIM], ah[DIM], zl[DIM], zh[DIM];
};

template <typename T, int DIM, bool COMPACT>
__device__ __forceinline__ void cz_gather(const Geo& g, const LevelPtrs<T>& P, int c, int t,
                                          bool own, CzOps<T, DIM, COMPACT>& v) {
  const Nbr<DIM> n = neighbours_clamped<DIM>(g, t, row_q(g, t, c));
  v.gt = P.gt[c][t];
  v.rv = P.r[c][t];
  v.ad = P.Adiag[c][t];
  v.own = own ? P.z[c][t] : (T)0;
  const T* zo = P.z[c ^ 1];
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    v.zl[a] = zo[n.lo[a]];
    v.zh[a] = zo[n.hi[a]];
  }
  if constexpr (COMPACT) {
    v.f = P.nb[c][t];
  } else {
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      v.al[a] = P.Ax[a][c ^ 1][n.lo[a]];
      v.ah[a] = P.Ax[a][c][t];
    }
  }
}

// the slot's couplings (decoded from the flag word on COMPACT levels)
template <typename T, int DIM, bool COMPACT>
__device__ __forceinline__ void cz_coefs(const Coupling<T>& k, uint32_t f, const T* sal,
                                         const T* sah, T* al, T* ah) {
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    if constexpr (COMPACT) {
      al[a] = ((f >> (NB_DN + a)) & 1u) ? k.coef : (T)0;
      ah[a] = ((f >> a) & 1u) ? k.coef : (T)0;
    } else {
      al[a] = sal[a];
      ah[a] = sah[a];
    }
  }
}

// smooth(l, c) over the level (mgpcg.py:104-109)
template <typename T, int DIM, bool COMPACT>
__device__ void cz_half_t(const Geo& g, const LevelPtrs<T>& P, const Coupling<T>& k, int c) {
  const int n = g.size;
  constexpr int NS = cz_slots<COMPACT>();
  for (int b = 0; b < n; b += NS * COARSE_THREADS) {
    CzOps<T, DIM, COMPACT> v[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int t = b + s * COARSE_THREADS + (int)threadIdx.x;
      v[s].gt = MG_SOLID;
      if (t < n) cz_gather<T, DIM, COMPACT>(g, P, c, t, false, v[s]);
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int t = b + s * COARSE_THREADS + (int)threadIdx.x;
      if (v[s].gt != MG_FLUID) continue;
      T al[DIM], ah[DIM];
      cz_coefs<T, DIM, COMPACT>(k, v[s].f, v[s].al, v[s].ah, al, ah);
      T sm = nsum_vals<T, DIM>(al, ah, v[s].zl, v[s].zh);
      T num = v[s].rv - sm;
      P.z[c][t] = num / v[s].ad;
    }
  }
  __syncthreads();
}
template <typename T, int DIM>
__device__ void cz_half(const Geo& g, const LevelPtrs<T>& P, const Coupling<T>& k, int c) {
  if (k.uniform) cz_half_t<T, DIM, true>(g, P, k, c);
  else cz_half_t<T, DIM, false>(g, P, k, c);
}

// z.fill(0); smooth(l, 0): z0 = r0 / Adiag0 on FLUID, 0 elsewhere, and z1 = 0
// (the smooth(l, 1) that follows rewrites every colour-1 FLUID value)
template <typename T>
__device__ void cz_first(const Geo& g, const LevelPtrs<T>& P) {
  const int n = g.size;
  for (int b = 0; b < n; b += CZ_SLOTS * COARSE_THREADS) {
    int gt[CZ_SLOTS];
    T rv[CZ_SLOTS], ad[CZ_SLOTS];
#pragma unroll
    for (int s = 0; s < CZ_SLOTS; ++s) {
      const int t = b + s * COARSE_THREADS + (int)threadIdx.x;
      gt[s] = MG_SOLID;
      if (t < n) {
        gt[s] = P.gt[0][t];
        rv[s] = P.r[0][t];
        ad[s] = P.Adiag[0][t];
      }
    }
#pragma unroll
    for (int s = 0; s < CZ_SLOTS; ++s) {
      const int t = b + s * COARSE_THREADS + (int)threadIdx.x;
      if (t >= n) continue;
      P.z[0][t] = gt[s] == MG_FLUID ? rv[s] / ad[s] : (T)0;
      P.z[1][t] = (T)0;
    }
  }
  __syncthreads();
}

template <typename T>
__device__ void cz_zero(const Geo& g, const LevelPtrs<T>& P) {
  for (int t = threadIdx.x; t < g.size; t += COARSE_THREADS) zero_slot<T>(P, t);
  __syncthreads();
}

// restrict(l) (mgpcg.py:111-118) into the coarser level's r: the residual of
// every fine cell (0 off FLUID) into `res` (two colour arrays of the fine
// level's size), then every interior coarse cell sums its eight children in
// lexicographic order (adding a 0 leaves the sum unchanged, as skipping a
// non-FLUID child does: the sum starts at +0 and never becomes -0)
template <typename T, int DIM, bool COMPACT>
__device__ void cz_residuals(const Geo& gf, const LevelPtrs<T>& F, const Coupling<T>& k,
                             T* const* res) {
  const int n = gf.size;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    constexpr int NS = cz_slots<COMPACT>();
    for (int b = 0; b < n; b += NS * COARSE_THREADS) {
    CzOps<T, DIM, COMPACT> v[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const int t = b + s * COARSE_THREADS + (int)threadIdx.x;
        v[s].gt = MG_SOLID;
        if (t < n) cz_gather<T, DIM, COMPACT>(gf, F, c, t, true, v[s]);
      }
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const int t = b + s * COARSE_THREADS + (int)threadIdx.x;
        if (t >= n) continue;
        T rr = (T)0;
        if (v[s].gt == MG_FLUID) {
          T al[DIM], ah[DIM];
          cz_coefs<T, DIM, COMPACT>(k, v[s].f, v[s].al, v[s].ah, al, ah);
          T Az = v[s].ad * v[s].own;
          T ns = nsum_vals<T, DIM>(al, ah, v[s].zl, v[s].zh);
          Az = Az + ns;
          rr = v[s].rv - Az;
        }
        res[c][t] = rr;
      }
    }
  }
  __syncthreads();
}

template <typename T, int DIM>
__device__ void cz_restrict(const Geo& gf, const LevelPtrs<T>& F, const Coupling<T>& k,
                            const Geo& gc, T* rc0, T* rc1, T* const* res) {
  if (k.uniform) cz_residuals<T, DIM, true>(gf, F, k, res);
  else cz_residuals<T, DIM, false>(gf, F, k, res);
  const int nc = gc.size;
  for (int tc = threadIdx.x; tc < nc; tc += COARSE_THREADS) {
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
      const Slot s = decode(gc, tc, cc);
      if (!s.interior) continue;
      T v[8];
      int nch = 0;
#pragma unroll
      for (int o0 = 0; o0 < (DIM == 3 ? 2 : 1); ++o0)
#pragma unroll
        for (int o1 = 0; o1 < 2; ++o1)
#pragma unroll
          for (int o2 = 0; o2 < 2; ++o2) {
            const int i = (DIM == 3) ? 2 * s.i + o0 : 0;
            const int j = 2 * s.j + o1, kk = 2 * s.k + o2;
            v[nch++] = res[colour_of(i, j, kk)][slot_of(gf, i, j, kk)];
          }
      T sum = (T)0;
#pragma unroll
      for (int o = 0; o < (DIM == 3 ? 8 : 4); ++o) sum = sum + v[o];
      (cc ? rc1 : rc0)[tc] = sum;
    }
  }
  __syncthreads();
}

// coarse_pre / coarse_bottom / coarse_post on the LDS image
template <typename T, int DIM>
__device__ void cz_pre(const Geo& g, const LevelPtrs<T>& P, int nu) {
  if (nu <= 0) {
    cz_zero<T>(g, P);
    return;
  }
  const Coupling<T> k = coupling(P);
  cz_first<T>(g, P);
  cz_half<T, DIM>(g, P, k, 1);
  for (int i = 1; i < nu; ++i) {
    cz_half<T, DIM>(g, P, k, 0);
    cz_half<T, DIM>(g, P, k, 1);
  }
}

template <typename T, int DIM>
__device__ void cz_bottom(const Geo& g, const LevelPtrs<T>& P, int bottom) {
  const int half = bottom / 2;
  if (half <= 0) {
    cz_zero<T>(g, P);
    return;
  }
  const Coupling<T> k = coupling(P);
  cz_first<T>(g, P);
  cz_half<T, DIM>(g, P, k, 1);
  for (int i = 1; i < half; ++i) {
    cz_half<T, DIM>(g, P, k, 0);
    cz_half<T, DIM>(g, P, k, 1);
  }
  for (int i = 0; i < half; ++i) {
    cz_half<T, DIM>(g, P, k, 1);
    cz_half<T, DIM>(g, P, k, 0);
  }
}

// prolongate(l) (mgpcg.py:120-123) on colour 0: the smooth(l, 1) that follows
// reads these values, each z0 + z_coarse[parent] formed as the reference
// forms it, and rewrites every colour-1 FLUID value without reading it.
// (Colour-1 cells that are not FLUID keep z unprolongated, as do those of
// the global-memory form: no FLUID result reads them -- their couplings are
// 0 and the children of a cell that is not FLUID are not FLUID.)
template <typename T>
__device__ void cz_prolong0(const Geo& g, const Geo& gc, const LevelPtrs<T>& P, const T* zc0,
                            const T* zc1) {
  const int n = g.size;
  for (int b = 0; b < n; b += CZ_SLOTS * COARSE_THREADS) {
    T zv[CZ_SLOTS], pv[CZ_SLOTS];
    bool in[CZ_SLOTS];
#pragma unroll
    for (int s = 0; s < CZ_SLOTS; ++s) {
      const int t = b + s * COARSE_THREADS + (int)threadIdx.x;
      in[s] = false;
      if (t >= n) continue;
      const Slot sl = decode(g, t, 0);
      if (!sl.interior) continue;
      in[s] = true;
      const int pi = sl.i >> 1, pj = sl.j >> 1, pk = sl.k >> 1;
      const int ps = slot_of(gc, pi, pj, pk);
      pv[s] = colour_of(pi, pj, pk) ? zc1[ps] : zc0[ps];
      zv[s] = P.z[0][t];
    }
#pragma unroll
    for (int s = 0; s < CZ_SLOTS; ++s) {
      const int t = b + s * COARSE_THREADS + (int)threadIdx.x;
      if (in[s]) P.z[0][t] = zv[s] + pv[s];
    }
  }
  __syncthreads();
}

template <typename T, int DIM>
__device__ void cz_post(const Geo& g, const Geo& gc, const LevelPtrs<T>& P, T* const* zc,
                        int nu) {
  if (nu <= 0) {
    coarse_post<T, DIM>(g, gc, P, zc, nu);
    return;
  }
  const Coupling<T> k = coupling(P);
  cz_prolong0<T>(g, gc, P, zc[0], zc[1]);
  cz_half<T, DIM>(g, P, k, 1);
  cz_half<T, DIM>(g, P, k, 0);
  for (int i = 1; i < nu; ++i) {
    cz_half<T, DIM>(g, P, k, 1);
    cz_half<T, DIM>(g, P, k, 0);
  }
}

// LDS image of one system's coarse levels: per level the z, r and Adiag
// colour arrays, the cell types, the coupling flags and the coupling range.
// The stored couplings (Ax) of a level without uniform couplings stay in
// global memory.
constexpr int COARSE_LDS_BYTES = 160 * 1024;
__host__ __device__ constexpr int lds_align(int b) { return (b + 15) & ~15; }
__host__ __device__ constexpr int coarse_image_bytes(int size, int esz) {
  return 6 * lds_align(size * esz) + 2 * lds_align(size) + 2 * lds_align(2 * size) +
         lds_align((int)sizeof(LevelMeta));
}
// the restriction's fine residuals (two colour arrays of the largest level
// that restricts), after the levels' images
__host__ __device__ constexpr int coarse_scratch_bytes(int size, int esz) {
  return 2 * lds_align(size * esz);
}

// One workgroup per item: item.desc = the system's LevelDesc of its first
// coarse level, item.block0 = number of levels from there to the bottom
// (a system's LevelDescs are contiguous).  With USE_LDS every per-slot array
// the V-cycle reads is copied into LDS once (the host checks that the
// levels' coarse_image_bytes and the restriction's coarse_scratch_bytes fit
// COARSE_LDS_BYTES), the V-cycle runs on the copy in the multi-slot phases
// above, and z of every level and the restricted r of the levels below the
// first are copied back at the end.  Without, the slot functions of the
// global-memory form run on the levels in place.
template <typename T, int DIM, bool USE_LDS>
__global__ void __launch_bounds__(COARSE_THREADS) k_coarse(const LevelDesc<T>* __restrict__ descs,
                                                           const Item* __restrict__ items,
                                                           int nitems, int nu, int bottom) {
  if ((int)blockIdx.x >= nitems) return;
  const int d0 = items[blockIdx.x].desc;
  const int nlev = items[blockIdx.x].block0;
  if (descs[d0].st->done) return;
  __shared__ alignas(16) unsigned char img[USE_LDS ? COARSE_LDS_BYTES : 16];
  // the operator and vectors of coarse level k as the V-cycle sees them
  auto lv = [&](int k) -> LevelPtrs<T> {
    LevelPtrs<T> P = descs[d0 + k].P;
    if constexpr (USE_LDS) {
      int off = 0;
      for (int j = 0; j < k; ++j) off += coarse_image_bytes(descs[d0 + j].g.size, (int)sizeof(T));
      const int n = descs[d0 + k].g.size;
      const int vb = lds_align(n * (int)sizeof(T)), bb = lds_align(n), wb = lds_align(2 * n);
      unsigned char* b = img + off;
      P.z[0] = (T*)(b + 0 * vb);
      P.z[1] = (T*)(b + 1 * vb);
      P.r[0] = (T*)(b + 2 * vb);
      P.r[1] = (T*)(b + 3 * vb);
      P.Adiag[0] = (const T*)(b + 4 * vb);
      P.Adiag[1] = (const T*)(b + 5 * vb);
      P.gt[0] = (const int8_t*)(b + 6 * vb);
      P.gt[1] = (const int8_t*)(b + 6 * vb + bb);
      P.nb[0] = (const uint16_t*)(b + 6 * vb + 2 * bb);
      P.nb[1] = (const uint16_t*)(b + 6 * vb + 2 * bb + wb);
      if (P.meta != nullptr) P.meta = (const LevelMeta*)(b + 6 * vb + 2 * bb + 2 * wb);
    }
    return P;
  };
  if constexpr (USE_LDS) {
    for (int k = 0; k < nlev; ++k) {
      const LevelPtrs<T> G = descs[d0 + k].P;
      const LevelPtrs<T> P = lv(k);
      const int n = descs[d0 + k].g.size;
      const bool flags = G.meta != nullptr;
      // all loads of a lane's slots in flight before the first LDS store
      for (int b = 0; b < n; b += CZ_SLOTS * COARSE_THREADS) {
        T rv[CZ_SLOTS][2], ad[CZ_SLOTS][2];
        int8_t gt[CZ_SLOTS][2];
        uint16_t f[CZ_SLOTS][2];
#pragma unroll
        for (int s = 0; s < CZ_SLOTS; ++s) {
          const int t = b + s * COARSE_THREADS + (int)threadIdx.x;
          if (t >= n) continue;
#pragma unroll
          for (int c = 0; c < 2; ++c) {
            rv[s][c] = ldgi(G.r[c], t);
            ad[s][c] = ldgi(G.Adiag[c], t);
            gt[s][c] = ldgi(G.gt[c], t);
            f[s][c] = flags ? ldgi(G.nb[c], t) : (uint16_t)0;
          }
        }
#pragma unroll
        for (int s = 0; s < CZ_SLOTS; ++s) {
          const int t = b + s * COARSE_THREADS + (int)threadIdx.x;
          if (t >= n) continue;
#pragma unroll
          for (int c = 0; c < 2; ++c) {
            P.r[c][t] = rv[s][c];
            const_cast<T*>(P.Adiag[c])[t] = ad[s][c];
            const_cast<int8_t*>(P.gt[c])[t] = gt[s][c];
            if (flags) const_cast<uint16_t*>(P.nb[c])[t] = f[s][c];
          }
        }
      }
      if (flags && threadIdx.x == 0) *const_cast<LevelMeta*>(P.meta) = *G.meta;
    }
    __syncthreads();
  }
  if constexpr (USE_LDS) {
    int off = 0;
    for (int j = 0; j < nlev; ++j) off += coarse_image_bytes(descs[d0 + j].g.size, (int)sizeof(T));
    T* res[2];
    res[0] = (T*)(img + off);
    res[1] = (T*)(img + off + lds_align(descs[d0].g.size * (int)sizeof(T)));
    // the geometry is read once per level into registers: a barrier orders
    // memory, so fields read through the descriptor would be fetched again
    // after every phase
    for (int k = 0; k < nlev - 1; ++k) {
      const Geo g = descs[d0 + k].g, gc = descs[d0 + k].gc;
      const LevelPtrs<T> P = lv(k);
      const LevelPtrs<T> C = lv(k + 1);
      cz_pre<T, DIM>(g, P, nu);
      cz_restrict<T, DIM>(g, P, coupling(P), gc, C.r[0], C.r[1], res);
    }
    {
      const Geo gb = descs[d0 + nlev - 1].g;
      cz_bottom<T, DIM>(gb, lv(nlev - 1), bottom);
    }
    for (int k = nlev - 2; k >= 0; --k) {
      const Geo g = descs[d0 + k].g, gc = descs[d0 + k].gc;
      const LevelPtrs<T> C = lv(k + 1);
      cz_post<T, DIM>(g, gc, lv(k), C.z, nu);
    }
  } else {
    for (int k = 0; k < nlev - 1; ++k) {
      const LevelDesc<T>& D = descs[d0 + k];
      const LevelPtrs<T> P = lv(k);
      const LevelPtrs<T> C = lv(k + 1);
      coarse_pre<T, DIM>(D.g, P, nu);
      for (int tc = threadIdx.x; tc < D.gc.size; tc += COARSE_THREADS) {
        restrict_slot<T, DIM>(D.g, P, P.z, D.gc, C.r[0], C.r[1], 0, tc);
        restrict_slot<T, DIM>(D.g, P, P.z, D.gc, C.r[0], C.r[1], 1, tc);
      }
      __syncthreads();
    }
    coarse_bottom<T, DIM>(descs[d0 + nlev - 1].g, lv(nlev - 1), bottom);
    for (int k = nlev - 2; k >= 0; --k) {
      const LevelDesc<T>& D = descs[d0 + k];
      const LevelPtrs<T> C = lv(k + 1);
      coarse_post<T, DIM>(D.g, D.gc, lv(k), C.z, nu);
    }
  }
  if constexpr (USE_LDS) {
    // the finer level's prolongation reads the first coarse level's z; the
    // other levels get z and r too, as the global-memory form leaves them
    for (int k = 0; k < nlev; ++k) {
      const LevelPtrs<T>& G = descs[d0 + k].P;
      const LevelPtrs<T> P = lv(k);
      const int n = descs[d0 + k].g.size;
      for (int t = threadIdx.x; t < n; t += COARSE_THREADS) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          stgi(G.z[c], t, P.z[c][t]);
          if (k > 0) stgi(G.r[c], t, P.r[c][t]);
        }
      }
    }
  }
}

// The V-cycle's first level-0 half-sweep from z = 0 (z.fill(0); smooth(0, 0),
// mgpcg.py:126-127), z0 = r0 / Adiag0 on FLUID and 0 elsewhere, computed by
// the kernel that has just written r0 (the r of every slot then passes
// through registers once instead of being read back).
template <typename T>
__device__ __forceinline__ void first_z0(const LevelPtrs<T>& P, const Coupling<T>& k, int t,
                                         bool fluid, T r0) {
  T z = (T)0;
  if (fluid) {
    const T ad = k.uniform ? slot_diag<T, true>(P, k, 0, t, ldgi(P.nb[0], t), (T)0)
                           : ldgi(P.Adiag[0], t);
    z = r0 / ad;
  }
  stgi(P.z[0], t, z);
}
// the same with the slot's flag word (uniform couplings) or Adiag (otherwise)
// loaded beforehand, with the kernel's other loads of the slot
template <typename T>
__device__ __forceinline__ void first_z0(const LevelPtrs<T>& P, const Coupling<T>& k, int t,
                                         bool fluid, T r0, uint32_t f, T adl) {
  T z = (T)0;
  if (fluid) z = r0 / (k.uniform ? slot_diag<T, true>(P, k, 0, t, f, (T)0) : adl);
  stgi(P.z[0], t, z);
}

// ---------------------------------------------------------------------------
// CG kernels on level 0 (mgpcg.py:150-238).  Reductions: one double-double
// partial per virtual block (a workgroup run's total in its first virtual
// block, zero in the others); a per-system finalize kernel sums them.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double rnd_plain(const CGState* st, double v) {
  return st->precision == PLAIN ? (double)(float)v : v;
}

// r = b (mgpcg.py:157), x = p = Ap = 0; partial r.r.  Systems in phase 0 only.
template <typename T, typename TX>
__global__ void __launch_bounds__(BLOCK) k_cg_init(const SysDesc<T, TX>* __restrict__ sys,
                                                   Stage S) {
  ItemCache ic;
  RunSum rs;
  Coupling<T> kc{false, (T)0};
  for (VBIter it = vb_begin(S.nblocks); it.vb < it.end; it.vb += it.step) {
    if (item_at(S, it.vb, ic)) {
      run_flush(rs);
      kc = coupling_g(sys[ic.desc].P);
    }
    const SysDesc<T, TX>& D = sys[ic.desc];
    const int lb = it.vb - ic.block0;
    run_mark(rs, D.partials, lb);
    const int t = lb * BLOCK + threadIdx.x;
    if (t < D.g.size) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        T bv = D.b[c][t];
        D.P.r[c][t] = bv;
        D.x[c][t] = (TX)0;
        D.p[0][c][t] = (T)0;
        D.p[1][c][t] = (T)0;
        D.Ap[c][t] = (T)0;
        const bool fl = D.P.gt[c][t] == MG_FLUID;
        if (fl) rs.acc = dd_add(rs.acc, dd_prod(bv, bv));
        if (c == 0) first_z0<T>(D.P, kc, t, fl, bv);
      }
    }
  }
  run_flush(rs);
}

// p = z + beta p (update_p, mgpcg.py:234-238) into the other parity buffer,
// Ap = Adiag p + neighbour_sum(p) (compute_Ap, :218-224), partial p.Ap.
template <typename T, typename TX, int DIM>
__global__ void __launch_bounds__(BLOCK) k_cg_ap(const SysDesc<T, TX>* __restrict__ sys,
                                                 Stage S) {
  ItemCache ic;
  RunSum rs;
  for (VBIter it = vb_begin(S.nblocks); it.vb < it.end; it.vb += it.step) {
    if (item_at(S, it.vb, ic)) run_flush(rs);
    const SysDesc<T, TX>& D = sys[ic.desc];
    const CGState* st = D.st;
    if (st->done) continue;
    const int lb = it.vb - ic.block0;
    run_mark(rs, D.partials, lb);
    const int t = lb * BLOCK + threadIdx.x;
    const int par = st->iter & 1;
    const T* pold[2] = {D.p[par][0], D.p[par][1]};
    T* pnew[2] = {D.p[par ^ 1][0], D.p[par ^ 1][1]};
    const bool plain = st->precision == PLAIN;
    const float b32 = st->beta32;
    const double beta = st->beta;
    auto upd = [&](int c, int u) -> T {
      // update_p on a FLUID cell; other cells keep p (inert)
      if (D.P.gt[c][u] != MG_FLUID) return pold[c][u];
      if (plain) {
        T tt = (T)b32 * pold[c][u];
        return D.P.z[c][u] + tt;
      }
      return (T)((double)D.P.z[c][u] + beta * (double)pold[c][u]);
    };
    if (t < D.g.size) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        if (D.P.gt[c][t] != MG_FLUID) continue;
        const int q = row_q(D.g, t, c);
        Nbr<DIM> n = neighbours<DIM>(D.g, t, q);
        T pv = upd(c, t);
        pnew[c][t] = pv;
        T rr = D.P.Adiag[c][t] * pv;
        T ns = nsum<T, DIM>(D.P, c, t, n, [&](int u) { return upd(c ^ 1, u); });
        T ap = rr + ns;
        D.Ap[c][t] = ap;
        rs.acc = dd_add(rs.acc, dd_prod(pv, ap));
      }
    }
  }
  run_flush(rs);
}

// update_p alone (mgpcg.py:234-238): p_next = z + beta p on FLUID cells.
template <typename T, typename TX>
__global__ void __launch_bounds__(BLOCK) k_cg_update_p(const SysDesc<T, TX>* __restrict__ sys,
                                                       Stage S) {
  ItemCache ic;
  bool done = true, plain = false;
  int par = 0;
  float b32 = 0.0f;
  double beta = 0.0;
  OrdScan ws;
  int iter = 0;
  for (VBIter it = vb_begin(S.nblocks); it.vb < it.end; it.vb += it.step, ++iter) {
    if (item_at(S, it.vb, ic)) {
      ws.valid = false;
      const CGState* st = sys[ic.desc].st;
      done = ldg(&st->done) != 0;
      par = ldg(&st->iter) & 1;
      plain = ldg(&st->precision) == PLAIN;
      b32 = ldg(&st->beta32);
      beta = ldg(&st->beta);
    }
    if (done) continue;
    const SysDesc<T, TX>& D = sys[ic.desc];
    const int32_t e = ord_here(ws, iter, it.vb, it.step, it.end, ic.block0, ic.end, D.P, D.g, 1);
    if (e >= 0) continue;  // p changes on FLUID only
    const int t = (e & ORD_BLOCK) * BLOCK + threadIdx.x;
    if (t >= D.g.size) continue;
    int gts[2];
    T polds[2], zvs[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {  // both colours loaded before either is stored
      gts[c] = ldgi(D.P.gt[c], t);
      polds[c] = ldgi(D.p[par][c], t);
      zvs[c] = ldgi(D.P.z[c], t);
    }
    MG_KEEP(gts[0]);
    MG_KEEP(gts[1]);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int gt = gts[c];
      const T pold = polds[c], zv = zvs[c];
      if (gt != MG_FLUID) continue;
      T pv;
      if (plain) {
        T tt = (T)b32 * pold;
        pv = zv + tt;
      } else {
        pv = (T)((double)zv + beta * (double)pold);
      }
      stgi(D.p[par ^ 1][c], t, pv);
    }
  }
}

// compute_Ap at both colours of slot t, one colour after the other: Ap
// stored, p.Ap added to acc
template <typename T, typename TX, int DIM, bool COMPACT>
__device__ __forceinline__ void ap_slot(const SysDesc<T, TX>& D, const Coupling<T>& k, int par,
                                        int t, DD& acc) {
  // both colours' operands are loaded before either is used (p does not
  // change here), so the two stencils' loads are in flight together
  StencilVals<T, DIM, COMPACT> s[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    stencil_load<T, DIM, COMPACT>(D.g, D.P, D.p[par], c, t, s[c]);
    stencil_keep<T, DIM, COMPACT>(s[c]);
  }
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    if (s[c].gt != MG_FLUID) continue;
    const T ap = stencil_apply<T, DIM, COMPACT>(D.P, k, c, t, s[c]);
    stgi(D.Ap[c], t, ap);
    acc = dd_add(acc, dd_prod(s[c].own, ap));
  }
}

// compute_Ap (mgpcg.py:218-224) on the p written by k_cg_update_p, partial p.Ap.
template <typename T, typename TX, int DIM>
__global__ void __launch_bounds__(BLOCK, full_occupancy<T>()) k_cg_ap_only(const SysDesc<T, TX>* __restrict__ sys,
                                                      Stage S) {
  ItemCache ic;
  RunSum rs;
  bool done = true;
  int par = 0;
  Coupling<T> kc{false, (T)0};
  OrdScan ws;
  int iter = 0;
  for (VBIter it = vb_begin(S.nblocks); it.vb < it.end; it.vb += it.step, ++iter) {
    if (item_at(S, it.vb, ic)) {
      ws.valid = false;
      run_flush(rs);
      const SysDesc<T, TX>& D = sys[ic.desc];
      done = ldg(&D.st->done) != 0;
      par = (ldg(&D.st->iter) & 1) ^ 1;
      kc = coupling_g(D.P);
    }
    if (done) continue;
    const SysDesc<T, TX>& D = sys[ic.desc];
    run_mark(rs, D.partials, it.vb - ic.block0);
    const int32_t e = ord_here(ws, iter, it.vb, it.step, it.end, ic.block0, ic.end, D.P, D.g, 1);
    if (e >= 0) continue;  // Ap and p.Ap on FLUID only
    const int t = (e & ORD_BLOCK) * BLOCK + threadIdx.x;
    if (t < D.g.size) {
      if (kc.uniform) ap_slot<T, TX, DIM, true>(D, kc, par, t, rs.acc);
      else ap_slot<T, TX, DIM, false>(D, kc, par, t, rs.acc);
    }
  }
  run_flush(rs);
}

// update_xr (mgpcg.py:226-232) + partial r.r.  MIXED: x is FP64 and on a
// refinement iteration r is left for k_true_residual.
template <typename T, typename TX>
__global__ void __launch_bounds__(BLOCK, full_occupancy<T>()) k_cg_xr(const SysDesc<T, TX>* __restrict__ sys,
                                                 Stage S) {
  ItemCache ic;
  RunSum rs;
  Coupling<T> kc{false, (T)0};
  bool done = true, refine = false;
  int par = 0, prec = PLAIN;
  float a32 = 0.0f;
  double alpha = 0.0;
  OrdScan ws;
  int iter = 0;
  for (VBIter it = vb_begin(S.nblocks); it.vb < it.end; it.vb += it.step, ++iter) {
    if (item_at(S, it.vb, ic)) {
      ws.valid = false;
      run_flush(rs);
      const CGState* st = sys[ic.desc].st;
      kc = coupling_g(sys[ic.desc].P);
      done = ldg(&st->done) != 0;
      par = (ldg(&st->iter) & 1) ^ 1;  // the p written by k_cg_update_p / k_cg_ap
      prec = ldg(&st->precision);
      const int re = ldg(&st->refine_every);
      refine = prec == MIXED && re > 0 && ldg(&st->since_refine) + 1 >= re;
      a32 = ldg(&st->alpha32);
      alpha = ldg(&st->alpha);
    }
    if (done) continue;
    const SysDesc<T, TX>& D = sys[ic.desc];
    run_mark(rs, D.partials, it.vb - ic.block0);
    const int32_t e = ord_here(ws, iter, it.vb, it.step, it.end, ic.block0, ic.end, D.P, D.g, 1);
    const int t = (e & ORD_BLOCK) * BLOCK + threadIdx.x;
    if (e >= 0) continue;  // x and r change on FLUID only; z0 stays 0 here
    DD& acc = rs.acc;
    if (t < D.g.size) {
      // both colours' operands are loaded before either colour is stored
      int8_t gts[2];
      T pvs[2], apvs[2], r0s[2];
      TX xvs[2];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        gts[c] = ldgi(D.P.gt[c], t);
        pvs[c] = ldgi(D.p[par][c], t);
        apvs[c] = ldgi(D.Ap[c], t);
        xvs[c] = ldgi(D.x[c], t);
        r0s[c] = ldgi(D.P.r[c], t);
      }
      uint32_t f0 = 0u;  // the first half-sweep's diagonal operand (colour 0)
      T ad0 = (T)0;
      if (!refine) {
        if (kc.uniform) f0 = ldgi(D.P.nb[0], t);
        else ad0 = ldgi(D.P.Adiag[0], t);
      }
      MG_KEEP4(pvs[1], apvs[1], r0s[1], (int)gts[1]);
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int8_t gt = gts[c];
        const T pv = pvs[c];
        const T apv = apvs[c];
        const TX xv = xvs[c];
        const T r0 = r0s[c];
        if (gt != MG_FLUID) {
          if (c == 0 && !refine) first_z0<T>(D.P, kc, t, false, (T)0, f0, ad0);
          continue;
        }
        if (prec == PLAIN) {
          T tp = (T)a32 * pv;
          stgi(D.x[c], t, (TX)((T)xv + tp));
          T ta = (T)a32 * apv;
          T rv = r0 - ta;
          stgi(D.P.r[c], t, rv);
          acc = dd_add(acc, dd_prod(rv, rv));
          if (c == 0) first_z0<T>(D.P, kc, t, true, rv, f0, ad0);
        } else if (prec == F64) {
          T tp = (T)(alpha * (double)pv);
          stgi(D.x[c], t, (TX)((double)xv + (double)tp));
          T ta = (T)(alpha * (double)apv);
          T rv = r0 - ta;
          stgi(D.P.r[c], t, rv);
          acc = dd_add(acc, dd_prod(rv, rv));
          if (c == 0) first_z0<T>(D.P, kc, t, true, rv, f0, ad0);
        } else {
          stgi(D.x[c], t, (TX)((double)xv + alpha * (double)pv));
          if (!refine) {
            T rv = (T)((double)r0 - alpha * (double)apv);
            stgi(D.P.r[c], t, rv);
            acc = dd_add(acc, dd_prod(rv, rv));
            if (c == 0) first_z0<T>(D.P, kc, t, true, rv, f0, ad0);
          }
        }
      }
    }
  }
  run_flush(rs);
}

// MIXED: r = b - A x with x in FP64, accumulated in FP64, stored as T; partial r.r.
template <typename T, typename TX, int DIM>
__global__ void __launch_bounds__(BLOCK, full_occupancy<T>()) k_true_residual(const SysDesc<T, TX>* __restrict__ sys,
                                                         Stage S) {
  ItemCache ic;
  RunSum rs;
  bool skip = true;
  Coupling<T> kc{false, (T)0};
  OrdScan ws;
  int iter = 0;
  for (VBIter it = vb_begin(S.nblocks); it.vb < it.end; it.vb += it.step, ++iter) {
    if (item_at(S, it.vb, ic)) {
      ws.valid = false;
      run_flush(rs);
      const SysDesc<T, TX>& D = sys[ic.desc];
      skip = ldg(&D.st->done) != 0 || ldg(&D.st->pending_true) == 0;
      kc = coupling_g(D.P);
    }
    if (skip) continue;
    const SysDesc<T, TX>& D = sys[ic.desc];
    run_mark(rs, D.partials, it.vb - ic.block0);
    const int32_t e = ord_here(ws, iter, it.vb, it.step, it.end, ic.block0, ic.end, D.P, D.g, 1);
    const int t = (e & ORD_BLOCK) * BLOCK + threadIdx.x;
    if (e >= 0) continue;  // r changes on FLUID only; z0 stays 0 here
    DD& acc = rs.acc;
    if (t < D.g.size) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int q = row_q(D.g, t, c);
        const Nbr<DIM> n = neighbours_clamped<DIM>(D.g, t, q);
        const int8_t gt = ldgi(D.P.gt[c], t);
        const TX* xo = D.x[c ^ 1];
        TX xl[DIM], xh[DIM];
        T al[DIM], ah[DIM];
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
          xl[a] = ldgi(xo, n.lo[a]);
          xh[a] = ldgi(xo, n.hi[a]);
        }
        couplings_g<T, DIM>(D.P, kc, c, t, n, al, ah);
        const T ad = ldgi(D.P.Adiag[c], t);
        const TX xv = ldgi(D.x[c], t);
        const T bv = ldgi(D.b[c], t);
        if (gt != MG_FLUID) {
          if (c == 0) stgi(D.P.z[0], t, (T)0);  // the V-cycle's first half-sweep
          continue;
        }
        double ret = 0.0;
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
          double lo = (double)al[a] * (double)xl[a];
          double hi = (double)ah[a] * (double)xh[a];
          double s2 = lo + hi;
          ret = ret + s2;
        }
        double Axv = (double)ad * (double)xv;
        Axv = Axv + ret;
        double rv = (double)bv - Axv;
        T r32 = (T)rv;
        stgi(D.P.r[c], t, r32);
        acc = dd_add(acc, dd_prod(r32, r32));
        if (c == 0) stgi(D.P.z[0], t, r32 / ad);  // the V-cycle's first half-sweep
      }
    }
  }
  run_flush(rs);
}

template <typename T, typename TX, int DIM, bool COMPACT>
__device__ __forceinline__ void smooth0_dot_slot(const SysDesc<T, TX>& D, const Coupling<T>& kc,
                                                 int t, DD& acc) {
  const int8_t g0 = ldgi(D.P.gt[0], t);
  const int8_t g1 = ldgi(D.P.gt[1], t);
  const T r0 = ldgi(D.P.r[0], t);
  const T r1 = ldgi(D.P.r[1], t);
  const T z1 = ldgi(D.P.z[1], t);
  const T z0 = smooth_slot_g<T, DIM, COMPACT>(D.g, D.P, kc, 0, t);
  if (g0 == MG_FLUID) acc = dd_add(acc, dd_prod(z0, r0));
  if (g1 == MG_FLUID) acc = dd_add(acc, dd_prod(z1, r1));
}

// Last post-smoothing half-sweep of level 0, smooth(0, 0) (mgpcg.py:148),
// fused with reduce(z, r) (mgpcg.py:171 / 201) over both colours.
template <typename T, typename TX, int DIM>
__global__ void __launch_bounds__(BLOCK) k_smooth0_dot(const SysDesc<T, TX>* __restrict__ sys,
                                                       Stage S) {
  ItemCache ic;
  RunSum rs;
  bool done = true;
  Coupling<T> kc{false, (T)0};
  OrdScan ws;
  int iter = 0;
  for (VBIter it = vb_begin(S.nblocks); it.vb < it.end; it.vb += it.step, ++iter) {
    if (item_at(S, it.vb, ic)) {
      ws.valid = false;
      run_flush(rs);
      const SysDesc<T, TX>& D = sys[ic.desc];
      done = ldg(&D.st->done) != 0;
      kc = coupling_g(D.P);
    }
    if (done) continue;
    const SysDesc<T, TX>& D = sys[ic.desc];
    run_mark(rs, D.partials, it.vb - ic.block0);
    const int32_t e = ord_here(ws, iter, it.vb, it.step, it.end, ic.block0, ic.end, D.P, D.g, 1);
    if (e >= 0) continue;  // z0 and z.r on FLUID only
    const int t = (e & ORD_BLOCK) * BLOCK + threadIdx.x;
    DD& acc = rs.acc;
    if (t < D.g.size) {
      if (kc.uniform) smooth0_dot_slot<T, TX, DIM, true>(D, kc, t, acc);
      else smooth0_dot_slot<T, TX, DIM, false>(D, kc, t, acc);
    }
  }
  run_flush(rs);
}

// reduce(z, r) alone (used when level 0 is inside the coarse kernel or gets
// no post-smoothing).
template <typename T, typename TX>
__global__ void __launch_bounds__(BLOCK) k_dot_zr(const SysDesc<T, TX>* __restrict__ sys,
                                                  Stage S) {
  ItemCache ic;
  RunSum rs;
  for (VBIter it = vb_begin(S.nblocks); it.vb < it.end; it.vb += it.step) {
    if (item_at(S, it.vb, ic)) run_flush(rs);
    const SysDesc<T, TX>& D = sys[ic.desc];
    if (D.st->done) continue;
    const int lb = it.vb - ic.block0;
    run_mark(rs, D.partials, lb);
    const int t = lb * BLOCK + threadIdx.x;
    DD& acc = rs.acc;
    if (t < D.g.size) {
      if (D.P.gt[0][t] == MG_FLUID) acc = dd_add(acc, dd_prod(D.P.z[0][t], D.P.r[0][t]));
      if (D.P.gt[1][t] == MG_FLUID) acc = dd_add(acc, dd_prod(D.P.z[1][t], D.P.r[1][t]));
    }
  }
  run_flush(rs);
}

// ---------------------------------------------------------------------------
// per-system finalize: one workgroup per system sums its partials in a fixed
// order and performs the scalar step of mgpcg.py:150-209.
// ---------------------------------------------------------------------------
enum FinOp { FIN_INIT = 0, FIN_ALPHA = 1, FIN_RR = 2, FIN_TRUE = 3, FIN_BETA = 4 };

// out (FIN_BETA only, may be NULL): the host's mapped readback slot; item i's
// state after the step goes to out[i], done or not.
template <typename T, typename TX, int OP>
__global__ void __launch_bounds__(BLOCK) k_finalize(const SysDesc<T, TX>* __restrict__ sys,
                                                    const Item* __restrict__ items, int nitems,
                                                    CGState* out) {
  if ((int)blockIdx.x >= nitems) return;
  const SysDesc<T, TX>& D = sys[items[blockIdx.x].desc];
  CGState* st = D.st;
  if (st->done) {
    if (OP == FIN_BETA && out != nullptr && threadIdx.x == 0) {
      out[blockIdx.x] = *st;
      __threadfence_system();
    }
    return;
  }
  if (OP == FIN_TRUE && !st->pending_true) return;
  // four independent accumulators: four partials per thread in flight
  DD v0 = dd_zero(), v1 = dd_zero(), v2 = dd_zero(), v3 = dd_zero();
  int i = threadIdx.x;
  for (; i + 3 * BLOCK < D.nparts; i += 4 * BLOCK) {
    const DD a = D.partials[i], b = D.partials[i + BLOCK];
    const DD c = D.partials[i + 2 * BLOCK], d = D.partials[i + 3 * BLOCK];
    v0 = dd_add(v0, a);
    v1 = dd_add(v1, b);
    v2 = dd_add(v2, c);
    v3 = dd_add(v3, d);
  }
  for (; i < D.nparts; i += BLOCK) v0 = dd_add(v0, D.partials[i]);
  const DD v = dd_add(dd_add(v0, v1), dd_add(v2, v3));
  const double tot = dd_value(block_sum_dd<BLOCK>(v));
  if (threadIdx.x != 0) return;
  if (OP == FIN_INIT) {
    // mgpcg.py:157-164
    st->r0r0 = rnd_plain(st, tot);
    st->rr = st->r0r0;
    st->tol = fmax(st->abs_tol, st->r0r0 * st->rel_tol);
    st->iter = 0;
    st->status = RUNNING;
    st->since_refine = 0;
    st->pending_true = 0;
    st->beta = 0.0;
    st->beta32 = 0.0f;
    st->true_rr = -1.0;
  } else if (OP == FIN_ALPHA) {
    // mgpcg.py:178-181
    st->pAp = rnd_plain(st, tot);
    const double alpha = st->zr_old / (st->pAp + st->eps);
    st->alpha = alpha;
    st->alpha32 = (float)alpha;
    if (!isfinite(alpha)) {
      st->done = 1;
      st->status = BREAKDOWN;
    }
  } else if (OP == FIN_RR) {
    // mgpcg.py:188-194
    if (st->precision == MIXED) {
      const bool refine = st->refine_every > 0 && st->since_refine + 1 >= st->refine_every;
      st->since_refine += 1;
      if (refine) {
        st->pending_true = 1;
      } else {
        st->rr = tot;
        if (tot < st->tol) st->pending_true = 1;  // accept only on the true residual
      }
    } else {
      st->rr = rnd_plain(st, tot);
      if (st->rr < st->tol) {
        st->done = 1;
        st->status = CONVERGED;
      }
    }
  } else if (OP == FIN_TRUE) {
    st->pending_true = 0;
    st->since_refine = 0;
    st->rr = tot;
    st->true_rr = tot;
    if (tot < st->tol) {
      st->done = 1;
      st->status = CONVERGED;
    }
  } else if (OP == FIN_BETA) {
    const double zr = rnd_plain(st, tot);
    if (st->phase == 0) {
      // the dot before the loop (mgpcg.py:171-172)
      st->zr_old = zr;
      st->phase = 1;
      if (st->max_iters == 0) {
        st->done = 1;
        st->status = MAX_ITERS;
      }
    } else {
      // mgpcg.py:201-209
      st->zr_new = zr;
      const double beta = zr / (st->zr_old + st->eps);
      st->beta = beta;
      st->beta32 = (float)beta;
      st->zr_old = zr;
      st->iter += 1;
      if (!isfinite(beta)) {
        st->done = 1;
        st->status = BREAKDOWN;
      } else if (st->max_iters >= 0 && st->iter >= st->max_iters) {
        st->done = 1;
        st->status = MAX_ITERS;
      }
    }
    if (out != nullptr) {
      out[blockIdx.x] = *st;
      __threadfence_system();
    }
  }
}

// ---------------------------------------------------------------------------
// The plain FP32 GPU port of mgpcg.py, structured like the reference: one
// launch per reference kernel on one system, scalars on the host.  It is the
// baseline the batched solver is measured against; each kernel handles both
// colours of one slot per thread.
// ---------------------------------------------------------------------------
enum RefVec { RV_B = 0, RV_X = 1, RV_P = 2, RV_AP = 3, RV_R0 = 4, RV_Z0 = 5 };

template <typename T, typename TX>
__device__ __forceinline__ T* ref_vec(const SysDesc<T, TX>& D, int which, int c) {
  switch (which) {
    case RV_B: return D.b[c];
    case RV_P: return D.p[0][c];
    case RV_AP: return D.Ap[c];
    case RV_R0: return D.P.r[c];
    default: return D.P.z[c];
  }
}

// r[0].copy_from(b) (mgpcg.py:157)
template <typename T, typename TX>
__global__ void __launch_bounds__(BLOCK) k_ref_copy_b(const SysDesc<T, TX>* __restrict__ sys) {
  const SysDesc<T, TX>& D = sys[0];
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= D.g.size) return;
  D.P.r[0][t] = D.b[0][t];
  D.P.r[1][t] = D.b[1][t];
}

// reduce(a, b) (mgpcg.py:211-216): per-block partials of the FLUID-masked dot
template <typename T, typename TX>
__global__ void __launch_bounds__(BLOCK) k_ref_dot(const SysDesc<T, TX>* __restrict__ sys, int a,
                                                   int b) {
  const SysDesc<T, TX>& D = sys[0];
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  DD acc = dd_zero();
  if (t < D.g.size) {
#pragma unroll
    for (int c = 0; c < 2; ++c)
      if (D.P.gt[c][t] == MG_FLUID)
        acc = dd_add(acc, dd_prod(ref_vec(D, a, c)[t], ref_vec(D, b, c)[t]));
  }
  DD s = block_sum_dd<BLOCK>(acc);
  if (threadIdx.x == 0) D.partials[blockIdx.x] = s;
}

// the scalar of the last k_ref_dot, rounded to FP32 as the 0-d f32 field `sum`
template <typename T, typename TX>
__global__ void __launch_bounds__(BLOCK) k_ref_sum(const SysDesc<T, TX>* __restrict__ sys,
                                                   double* out) {
  const SysDesc<T, TX>& D = sys[0];
  DD v = dd_zero();
  for (int i = threadIdx.x; i < D.nparts; i += BLOCK) v = dd_add(v, D.partials[i]);
  const double tot = dd_value(block_sum_dd<BLOCK>(v));
  if (threadIdx.x == 0) *out = (double)(float)tot;
}

// compute_Ap (mgpcg.py:218-224)
template <typename T, typename TX, int DIM>
__global__ void __launch_bounds__(BLOCK) k_ref_compute_Ap(const SysDesc<T, TX>* __restrict__ sys) {
  const SysDesc<T, TX>& D = sys[0];
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= D.g.size) return;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    if (D.P.gt[c][t] != MG_FLUID) continue;
    const int q = row_q(D.g, t, c);
    Nbr<DIM> n = neighbours<DIM>(D.g, t, q);
    const T* po = D.p[0][c ^ 1];
    T r = D.P.Adiag[c][t] * D.p[0][c][t];
    T ns = nsum<T, DIM>(D.P, c, t, n, [&](int u) { return po[u]; });
    D.Ap[c][t] = r + ns;
  }
}

// update_xr (mgpcg.py:226-232) with alpha from the 0-d f32 field
template <typename T, typename TX>
__global__ void __launch_bounds__(BLOCK) k_ref_update_xr(const SysDesc<T, TX>* __restrict__ sys) {
  const SysDesc<T, TX>& D = sys[0];
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= D.g.size) return;
  const T alpha = (T)D.st->alpha32;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    if (D.P.gt[c][t] != MG_FLUID) continue;
    T tp = alpha * D.p[0][c][t];
    D.x[c][t] = (TX)((T)D.x[c][t] + tp);
    T ta = alpha * D.Ap[c][t];
    D.P.r[c][t] = D.P.r[c][t] - ta;
  }
}

// update_p (mgpcg.py:234-238) with beta from the 0-d f32 field
template <typename T, typename TX>
__global__ void __launch_bounds__(BLOCK) k_ref_update_p(const SysDesc<T, TX>* __restrict__ sys) {
  const SysDesc<T, TX>& D = sys[0];
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= D.g.size) return;
  const T beta = (T)D.st->beta32;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    if (D.P.gt[c][t] != MG_FLUID) continue;
    T tt = beta * D.p[0][c][t];
    D.p[0][c][t] = D.P.z[c][t] + tt;
  }
}

// r.fill(0); z.fill(0) of a level (mgpcg.py:126, 132-133)
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_ref_fill(const LevelDesc<T>* __restrict__ descs,
                                                    int desc, int which) {
  const LevelDesc<T>& D = descs[desc];
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= D.g.size) return;
  if (which & 1) {
    D.P.z[0][t] = (T)0;
    D.P.z[1][t] = (T)0;
  }
  if (which & 2) {
    D.P.r[0][t] = (T)0;
    D.P.r[1][t] = (T)0;
  }
}

// ---------------------------------------------------------------------------
// layout conversion: canonical C-order (the strategies' fields) <-> colour split
// ---------------------------------------------------------------------------
__device__ __forceinline__ long canon_index(const Geo& g, int i, int j, int k) {
  return ((long)i * g.n1 + j) * g.n2 + k;
}

// the MAC velocity components of a level-0 grid (component k: res + e_k)
struct MacVel {
  const float* u[3];
};

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t w = (uint32_t)__shfl_xor((int)v, o);
    v = w < v ? w : v;
  }
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t w = (uint32_t)__shfl_xor((int)v, o);
    v = w > v ? w : v;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ bool same_bits(T a, T b) {
  unsigned char x[sizeof(T)], y[sizeof(T)];
  __builtin_memcpy(x, &a, sizeof(T));
  __builtin_memcpy(y, &b, sizeof(T));
  for (unsigned i = 0; i < sizeof(T); ++i)
    if (x[i] != y[i]) return false;
  return true;
}

// One slot of one colour in the packed layout: its cell type, diagonal and
// couplings (upper and, where the lower neighbour exists, lower per axis), as
// the level's canonical arrays hold them; a ghost slot is SOLID with zeros.
template <int DIM>
struct SlotCoefs {
  bool inn;
  bool hdn[DIM];
  int8_t gv;
  float ad, up[DIM], dn[DIM];
};

// Writes slot t of colour c from its coefficients: cell type, diagonal,
// couplings, the flag word (coupling flags and diagonal code), zero r, z and
// residual scratch; folds its nonzero upper couplings into mn / mx, and counts
// a FLUID cell into wet / nfl.  A slot's diagonal code is checked against the
// first nonzero coupling of the slot itself: on a level whose couplings turn
// out uniform that is the level's value, and the kernels read codes only on
// such levels.
template <typename T, int DIM>
__device__ __forceinline__ void pack_slot(const SlotCoefs<DIM>& v, int c, int t,
                                          const LevelPtrs<T>& P, uint32_t& mn, uint32_t& mx,
                                          uint32_t& wet, uint32_t& nfl) {
  int8_t gvc = MG_SOLID;
  T ad = (T)0, a[3] = {(T)0, (T)0, (T)0};
  T cl = (T)0;  // the slot's first nonzero coupling
  bool have = false;
  uint32_t f = 0u;
  if (v.inn) {
    gvc = v.gv;
    ad = (T)v.ad;
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      a[k] = (T)v.up[k];
      const uint32_t ub = __float_as_uint(v.up[k]);
      if (ub != 0u) {
        f |= 1u << k;
        mn = ub < mn ? ub : mn;
        mx = ub > mx ? ub : mx;
        if (!have) cl = (T)v.up[k];
        have = true;
      }
      if (v.hdn[k] && __float_as_uint(v.dn[k]) != 0u) {
        f |= 1u << (NB_DN + k);
        if (!have) cl = (T)v.dn[k];
        have = true;
      }
    }
  }
  ((int8_t*)P.gt[c])[t] = gvc;
  ((T*)P.Adiag[c])[t] = ad;
#pragma unroll
  for (int k = 0; k < DIM; ++k) ((T*)P.Ax[k][c])[t] = a[k];
  uint32_t code = NB_AD_STORED;
  for (uint32_t n = 0; n < NB_AD_STORED; ++n) {
    if (n > 0 && !have) break;
    if (same_bits(diag_of_code<T>(n, cl), ad)) {
      code = n;
      break;
    }
  }
  if (gvc == MG_FLUID) f |= 1u << NB_FL;
  ((uint16_t*)P.nb[c])[t] = (uint16_t)(f | (code << NB_AD));
  P.r[c][t] = (T)0;
  P.z[c][t] = (T)0;
  if (P.res[c] != nullptr) P.res[c][t] = (T)0;
  if (gvc == MG_FLUID) {
    wet = 1u;
    nfl += 1u;
  }
}

// End of a packing workgroup: its block's FLUID flag and natural-order work
// entry, and its share of the level's summary in `part`, a structure of
// arrays over the level's nb workgroups (b: this one): [0, nb) the smallest and
// [nb, 2 nb) the largest nonzero upper-coupling bit pattern, [2 nb, 3 nb) the
// FLUID cells.  k_meta_reduce folds them into the level's LevelMeta.  (One
// device-scope atomic per workgroup on the level's counters instead costs a
// level-0 packing over 100 us: atomics on one address from thousands of
// workgroups serialise.)  Called by the whole workgroup.
__device__ __forceinline__ void pack_block_tail(uint32_t mn, uint32_t mx, uint32_t wet,
                                                uint32_t nfl, uint8_t* wet_out, int32_t* ord0,
                                                uint32_t* part, unsigned b, unsigned nb) {
  __shared__ uint32_t smn[BLOCK / 64], smx[BLOCK / 64], swet[BLOCK / 64], snfl[BLOCK / 64];
  mn = wave_min_u32(mn);
  mx = wave_max_u32(mx);
  wet = wave_max_u32(wet);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) nfl += (uint32_t)__shfl_xor((int)nfl, o);
  if ((threadIdx.x & 63) == 0) {
    smn[threadIdx.x >> 6] = mn;
    smx[threadIdx.x >> 6] = mx;
    swet[threadIdx.x >> 6] = wet;
    snfl[threadIdx.x >> 6] = nfl;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < BLOCK / 64; ++w) {
      mn = smn[w] < mn ? smn[w] : mn;
      mx = smx[w] > mx ? smx[w] : mx;
      wet = swet[w] > wet ? swet[w] : wet;
      nfl += snfl[w];
    }
    part[b] = mn;
    part[nb + b] = mx;
    part[2 * nb + b] = nfl;
    if (wet_out != nullptr) wet_out[b] = (uint8_t)wet;
    if (ord0 != nullptr)  // per-block work order: natural, with the wet bit
      ord0[b] = (int32_t)b | (wet ? ORD_WET : 0);
  }
}

// The packing partials of up to 16 levels folded into their LevelMeta: one
// workgroup per level (blockIdx.x), after the levels' packing kernels.
struct MetaReduceList {
  LevelMeta* meta[16];
  const uint32_t* part[16];
  int nblocks[16];
};
constexpr int META_THREADS = 1024;
__device__ __forceinline__ void meta_reduce_level(const MetaReduceList& ml, int l) {
  __shared__ uint32_t smn[META_THREADS / 64], smx[META_THREADS / 64], snf[META_THREADS / 64];
  const int nb = ml.nblocks[l];
  const uint32_t* p = ml.part[l];
  uint32_t mn = 0xffffffffu, mx = 0u, nf = 0u;
  for (int b = threadIdx.x; b < nb; b += META_THREADS) {
    const uint32_t a = p[b], c = p[nb + b];
    mn = a < mn ? a : mn;
    mx = c > mx ? c : mx;
    nf += p[2 * nb + b];
  }
  mn = wave_min_u32(mn);
  mx = wave_max_u32(mx);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) nf += (uint32_t)__shfl_xor((int)nf, o);
  if ((threadIdx.x & 63) == 0) {
    smn[threadIdx.x >> 6] = mn;
    smx[threadIdx.x >> 6] = mx;
    snf[threadIdx.x >> 6] = nf;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < META_THREADS / 64; ++w) {
      mn = smn[w] < mn ? smn[w] : mn;
      mx = smx[w] > mx ? smx[w] : mx;
      nf += snf[w];
    }
    ml.meta[l]->minbits = mn;
    ml.meta[l]->maxbits = mx;
    ml.meta[l]->nfluid = nf;
  }
}
__global__ void __launch_bounds__(META_THREADS) k_meta_reduce(MetaReduceList ml) {
  meta_reduce_level(ml, blockIdx.x);
}

// Packs one level from the canonical arrays (the strategies' output).
template <typename T, int DIM>
__global__ void __launch_bounds__(BLOCK) k_pack_level(Geo g, const int8_t* __restrict__ gt,
                                                      const float* __restrict__ Adiag,
                                                      const float* __restrict__ Ax,
                                                      LevelPtrs<T> P, uint32_t* part) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  uint32_t mn = 0xffffffffu, mx = 0u, wet = 0u, nfl = 0u;
  if (t < g.size) {
    // every canonical value of both colours is loaded first (ghost slots and
    // missing lower neighbours read cell 0 and discard it)
    SlotCoefs<DIM> v[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const Slot sl = decode(g, t, c);
      v[c].inn = sl.interior;
      const long ci = sl.interior ? canon_index(g, sl.i, sl.j, sl.k) : 0;
      const int x[3] = {sl.i, sl.j, sl.k};
      long wi[DIM];
#pragma unroll
      for (int k = 0; k < DIM; ++k) {
        // lower neighbour on original axis k (view axis k + 3 - DIM)
        const int va = k + 3 - DIM;
        v[c].hdn[k] = sl.interior && x[va] > 0;
        int w[3] = {sl.i, sl.j, sl.k};
        w[va] -= 1;
        wi[k] = v[c].hdn[k] ? canon_index(g, w[0], w[1], w[2]) : 0;
      }
      v[c].gv = gt[ci];
      v[c].ad = Adiag[ci];
#pragma unroll
      for (int k = 0; k < DIM; ++k) {
        v[c].up[k] = Ax[ci * DIM + k];
        v[c].dn[k] = Ax[wi[k] * DIM + k];
      }
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) pack_slot<T, DIM>(v[c], c, t, P, mn, mx, wet, nfl);
  }
  pack_block_tail(mn, mx, wet, nfl, (uint8_t*)P.wet, (int32_t*)P.ord[0], part, blockIdx.x,
                  gridDim.x);
}

// The pressure operator of one level assembled straight into the packed
// layout: per cell, PressureProjectStrategy.build_A_kernel
// (pressure_project.py:45-64) -- for each axis, lower then upper neighbour:
// FLUID adds sA to the diagonal (and sets the upper coupling to -sA), AIR adds
// sA, or sA * c with the ghost-fluid weight c at level 0, SOLID or out of
// range adds nothing -- in the same operation order as the canonical
// assembly, so the packed values equal k_pack_level's of that assembly bit for
// bit.  A cell's lower coupling on axis k is the lower neighbour's upper one:
// -sA when both are FLUID.  gt: the level's canonical cell types; phi: the
// level set (ghost-fluid levels only, else NULL).
__device__ __forceinline__ float gfm_weight(float phi_i, float phi_j, float c_max) {
  const float c = (phi_i - phi_j) / phi_i;
  return phi_i == 0.0f ? c_max : fminf(c, c_max);
}

// One slot (both colours) of the pressure operator packed straight into the
// level's arrays (run by k_pack_pressure_all); folds its coupling extremes and
// FLUID count into mn / mx / wet / nfl.
template <typename T, int DIM>
__device__ __forceinline__ void pack_pressure_slot(const Geo& g, const int8_t* __restrict__ gt,
                                                   const float* __restrict__ phi, float sA,
                                                   float c_max, const LevelPtrs<T>& P, int t,
                                                   uint32_t& mn, uint32_t& mx, uint32_t& wet,
                                                   uint32_t& nfl) {
  SlotCoefs<DIM> v[2];
  int8_t nbt[2][DIM][2];
  float nbp[2][DIM][2], ph[2];
  // every load of both colours first (out-of-range neighbours read the
  // cell itself and count as SOLID)
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const Slot sl = decode(g, t, c);
    v[c].inn = sl.interior;
    const long ci = sl.interior ? canon_index(g, sl.i, sl.j, sl.k) : 0;
    const int x[3] = {sl.i, sl.j, sl.k};
    const int nn[3] = {g.n0, g.n1, g.n2};
    v[c].gv = gt[ci];
    ph[c] = phi ? phi[ci] : 0.0f;
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      const int va = k + 3 - DIM;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        int w[3] = {sl.i, sl.j, sl.k};
        w[va] += h ? 1 : -1;
        const bool in = sl.interior && w[va] >= 0 && w[va] < nn[va];
        const long wi = in ? canon_index(g, w[0], w[1], w[2]) : ci;
        const int8_t tv = gt[wi];
        nbt[c][k][h] = in ? tv : (int8_t)MG_SOLID;
        nbp[c][k][h] = phi ? phi[wi] : 0.0f;
      }
      v[c].hdn[k] = sl.interior && x[va] > 0;
    }
  }
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    float ad = 0.0f, up[3] = {0.0f, 0.0f, 0.0f};
    const bool fl = v[c].inn && v[c].gv == MG_FLUID;
    if (fl) {
#pragma unroll
      for (int k = 0; k < DIM; ++k) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {  // sgn = -1, then +1
          const int nb = nbt[c][k][h];
          if (nb == MG_FLUID) {
            ad = ad + sA;
            if (h) up[k] = -sA;
          } else if (nb == MG_AIR) {
            if (phi != nullptr) {
              const float cw = gfm_weight(ph[c], nbp[c][k][h], c_max);
              ad = ad + sA * cw;
            } else {
              ad = ad + sA;
            }
          }
        }
      }
    }
    v[c].ad = ad;
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      v[c].up[k] = up[k];
      v[c].dn[k] = (fl && nbt[c][k][0] == MG_FLUID) ? -sA : 0.0f;
    }
  }
#pragma unroll
  for (int c = 0; c < 2; ++c) pack_slot<T, DIM>(v[c], c, t, P, mn, mx, wet, nfl);
}

// The pressure right-hand side straight into the packed b (both colours of a
// slot per thread): PressureProjectStrategy.build_b_kernel
// (pressure_project.py:15-39) in the operation order of the canonical
// assembly -- the velocity divergence times sb, SOLID faces removing their
// velocity term, AIR faces adding sA * p0 (times the ghost-fluid weight when
// phi is given) -- then, when add_c, the volume-control constant c added on
// FLUID cells (volume_control.py:33-37).  u[k]: MAC component k, shape
// res + e_k (canonical).  Run by k_pack_pressure_all on level 0's workgroups.
template <typename T, int DIM>
__device__ __forceinline__ void pack_pressure_b_slot(const Geo& g, const int8_t* __restrict__ gt,
                                                     const MacVel& V,
                                                     const float* __restrict__ phi, float sA,
                                                     float sb, float p0, float c_max, int add_c,
                                                     float cadd, T* b0, T* b1, int t) {
  float out[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const Slot sl = decode(g, t, c);
    out[c] = 0.0f;
    if (!sl.interior) continue;
    const long ci = canon_index(g, sl.i, sl.j, sl.k);
    if (gt[ci] != MG_FLUID) continue;
    const int x[3] = {sl.i, sl.j, sl.k};
    const int nn[3] = {g.n0, g.n1, g.n2};
    float ulo[DIM], uhi[DIM];
    int nbt[DIM][2];
    float nbp[DIM][2];
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      const int va = k + 3 - DIM;
      // MAC component k has one more entry on axis k: index I and I + e_k
      int e[3] = {g.n0, g.n1, g.n2};
      e[va] += 1;
      const long base = ((long)(DIM == 3 ? x[0] : 0) * e[1] + x[1]) * e[2] + x[2];
      const long step = va == 2 ? 1 : (va == 1 ? (long)e[2] : (long)e[1] * e[2]);
      ulo[k] = V.u[k][base];
      uhi[k] = V.u[k][base + step];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        int w[3] = {sl.i, sl.j, sl.k};
        w[va] += h ? 1 : -1;
        const bool in = w[va] >= 0 && w[va] < nn[va];
        const long wi = in ? canon_index(g, w[0], w[1], w[2]) : ci;
        const int8_t tv = gt[wi];
        nbt[k][h] = in ? tv : MG_SOLID;
        nbp[k][h] = phi ? phi[wi] : 0.0f;
      }
    }
    const float phi_t = phi ? phi[ci] : 0.0f;
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      float d = ulo[k] - uhi[k];
      acc = acc + d;
    }
    float bv = acc * sb;
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {  // sgn = -1, then +1
        const int nb = nbt[k][h];
        if (nb == MG_SOLID) {
          if (!h) {
            float u = ulo[k] - 0.0f;
            bv = bv - sb * u;
          } else {
            float u = uhi[k] - 0.0f;
            bv = bv + sb * u;
          }
        } else if (nb == MG_AIR) {
          float term;
          if (phi != nullptr) {
            float cw = gfm_weight(phi_t, nbp[k][h], c_max);
            term = (sA * cw) * p0;
          } else {
            term = sA * p0;
          }
          bv = bv + term;
        }
      }
    }
    if (add_c) bv = bv + cadd;
    out[c] = bv;
  }
  b0[t] = (T)out[0];
  b1[t] = (T)out[1];
}

// A whole pressure hierarchy packed in one launch: the workgroups of level l
// are [block0[l], block0[l + 1]); level 0's also write b.  The per-level
// geometry and packed arrays come from the handle's table `tab`, the rest
// (cell types, the level set and the velocity) from the call.
template <typename T>
struct PackLevel {
  Geo g;
  LevelPtrs<T> P;
  uint32_t* part;
};
struct PackCall {
  const int8_t* gt[16];  // canonical cell types of each level
  int block0[17];
  int nl;
  MacVel V;
  const float* phi;  // ghost-fluid weights at level 0 (NULL: none)
  void* b[2];        // level 0's packed b, both colours
  float sA, sb, p0, c_max, cadd;
  int add_c;
};
template <typename T, int DIM>
__global__ void __launch_bounds__(BLOCK) k_pack_pressure_all(const PackLevel<T>* __restrict__ tab,
                                                             PackCall c) {
  int l = 0;
  while ((int)blockIdx.x >= c.block0[l + 1]) ++l;
  const PackLevel<T>& L = tab[l];
  const int b = (int)blockIdx.x - c.block0[l];
  const int t = b * BLOCK + threadIdx.x;
  uint32_t mn = 0xffffffffu, mx = 0u, wet = 0u, nfl = 0u;
  if (t < L.g.size) {
    const float* ph = l == 0 ? c.phi : nullptr;  // ghost-fluid weights at level 0 only
    if (l == 0)
      pack_pressure_b_slot<T, DIM>(L.g, c.gt[0], c.V, c.phi, c.sA, c.sb, c.p0, c.c_max, c.add_c,
                                   c.cadd, (T*)c.b[0], (T*)c.b[1], t);
    pack_pressure_slot<T, DIM>(L.g, c.gt[l], ph, c.sA, c.c_max, L.P, t, mn, mx, wet, nfl);
  }
  pack_block_tail(mn, mx, wet, nfl, (uint8_t*)L.P.wet, (int32_t*)L.P.ord[0], L.part, b,
                  c.block0[l + 1] - c.block0[l]);
}

template <typename S, typename D>
__global__ void __launch_bounds__(BLOCK) k_pack_vec(Geo g, const S* __restrict__ src, D* d0,
                                                    D* d1) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= g.size) return;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    Slot s = decode(g, t, c);
    D v = (D)0;
    if (s.interior) v = (D)src[canon_index(g, s.i, s.j, s.k)];
    (c ? d1 : d0)[t] = v;
  }
}

template <typename S, typename D>
__global__ void __launch_bounds__(BLOCK) k_unpack_vec(Geo g, const S* __restrict__ s0,
                                                      const S* __restrict__ s1, D* dst) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= g.size) return;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    Slot s = decode(g, t, c);
    if (!s.interior) continue;
    dst[canon_index(g, s.i, s.j, s.k)] = (D)(c ? s1 : s0)[t];
  }
}

// The half-sweep work order of one level (LevelPtrs::ord[1]) from its
// per-block FLUID flags: n = ceil(n1 / SLOT_LANES) virtual blocks, one wet
// when one of its blocks is.  With W wet entries, position k is wet when
// floor((k+1) W / n) > floor(k W / n): the W wet positions are spread evenly
// over [0, n).  The m-th wet position (m = floor(k W / n)) gets the m-th wet
// virtual block in ascending order, the d-th dry position (d = k - m) the
// d-th dry one, so the order is a permutation.  (A half-sweep stage strides
// its workgroups over the virtual blocks; in natural order a stride close to
// a multiple of a plane's length gives some workgroups only wet blocks and
// others none.)  One workgroup; `tmp` holds n ints.
constexpr int ORDER_THREADS = 1024;
__device__ __forceinline__ void level_order_body(const uint8_t* __restrict__ wet, int n1,
                                                 int32_t* ord, int32_t* tmp) {
  __shared__ int sc[ORDER_THREADS];
  const int tid = threadIdx.x;
  constexpr int span = SLOT_LANES;
  const int n = (n1 + span - 1) / span;
  auto flag = [&](int v) {
    int f = 0;
    for (int i = 0; i < span; ++i)
      if (v * span + i < n1) f |= wet[v * span + i] != 0;
    return f;
  };
  // each thread takes a contiguous chunk of entries; the wet entries before
  // its chunk come from an inclusive scan over the threads' counts
  const int chunk = (n + ORDER_THREADS - 1) / ORDER_THREADS;
  const int v0 = min(n, tid * chunk), v1 = min(n, v0 + chunk);
  int cnt = 0;
  for (int v = v0; v < v1; ++v) cnt += flag(v);
  sc[tid] = cnt;
  __syncthreads();
  for (int d = 1; d < ORDER_THREADS; d <<= 1) {
    const int add = tid >= d ? sc[tid - d] : 0;
    __syncthreads();
    sc[tid] += add;
    __syncthreads();
  }
  const int W = sc[ORDER_THREADS - 1];
  int wb = sc[tid] - cnt;  // wet entries before the chunk
  // stable partition into tmp: wet entries at [0, W), dry ones at [W, n)
  for (int v = v0; v < v1; ++v) {
    if (flag(v)) tmp[wb++] = v;
    else tmp[W + (v - wb)] = v;
  }
  __threadfence();
  __syncthreads();
  for (int k = tid; k < n; k += ORDER_THREADS) {
    const int m = (int)((long long)k * W / n);
    const int m1 = (int)((long long)(k + 1) * W / n);
    ord[k] = m1 > m ? (tmp[m] | ORD_WET) : tmp[W + (k - m)];
  }
}
__global__ void __launch_bounds__(ORDER_THREADS) k_level_order(const uint8_t* __restrict__ wet,
                                                               int n1, int32_t* ord,
                                                               int32_t* tmp) {
  level_order_body(wet, n1, ord, tmp);
}
static_assert(ORDER_THREADS == META_THREADS, "k_finish_levels runs both bodies");

// init_gridtype (mgpcg.py:46-60) on canonical int8 arrays; bit-exact.
__global__ void __launch_bounds__(BLOCK) k_coarsen(int dim, int f0, int f1, int f2,
                                                   const int8_t* __restrict__ fine,
                                                   int8_t* __restrict__ coarse, int ncoarse) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= ncoarse) return;
  const int c1n = f1 / 2, c2n = f2 / 2;
  const int c2 = t % c2n;
  const int rest = t / c2n;
  const int c1 = rest % c1n;
  const int c0 = rest / c1n;
  int tot_air = 0, tot_fluid = 0;
  const int o0max = (dim == 3) ? 2 : 1;
  for (int o0 = 0; o0 < o0max; ++o0)
    for (int o1 = 0; o1 < 2; ++o1)
      for (int o2 = 0; o2 < 2; ++o2) {
        const int i = (dim == 3) ? 2 * c0 + o0 : 0;
        const long fi = ((long)i * f1 + (2 * c1 + o1)) * f2 + (2 * c2 + o2);
        const int attr = fine[fi];
        if (attr == MG_AIR) tot_air++;
        else if (attr == MG_FLUID) tot_fluid++;
      }
  coarse[t] = tot_air > 0 ? MG_AIR : (tot_fluid > 0 ? MG_FLUID : MG_SOLID);
}

// The end of a packing: workgroup l < ml_levels folds level l's packing
// partials into its LevelMeta (meta_reduce_level); one more workgroup, when
// `ord` is given, writes level 0's half-sweep work order (level_order_body).
__global__ void __launch_bounds__(META_THREADS) k_finish_levels(MetaReduceList ml, int ml_levels,
                                                                const uint8_t* __restrict__ wet,
                                                                int n1, int32_t* ord,
                                                                int32_t* tmp) {
  if ((int)blockIdx.x < ml_levels) meta_reduce_level(ml, blockIdx.x);
  else level_order_body(wet, n1, ord, tmp);
}

}  // namespace mg
