// HIP kernels of the MGPCG solver for gfx950 (MI355X), colour-split padded
// layout (see mgpcg_common.h).  Every kernel here is memory bound; none uses
// MFMA.  Floating-point operation order follows mgpcg.py exactly and the
// library is compiled with -ffp-contract=off so that FP32 results are
// reproducible bit for bit against a CPU restatement with the same order.
#pragma once
#include "mgpcg_common.h"

namespace mg {

constexpr int BLOCK = 256;

template <typename T>
struct LevelPtrs {
  const int8_t* gt[2];
  const T* Adiag[2];
  const T* Ax[3][2];  // [axis][colour]
  T* r[2];
  T* z[2];
};

// blocks are dispatched round-robin over the 8 XCDs; hand each XCD a
// contiguous range of logical blocks so stencil neighbours share its L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  int q = nb >> 3, rem = nb & 7;
  int xcd = b & 7, local = b >> 3;
  return (xcd < rem) ? xcd * (q + 1) + local : rem * (q + 1) + (xcd - rem) * q + local;
}

// ---------------------------------------------------------------------------
// wave64 / block reductions (fixed order -> deterministic)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Returns the block total in every thread.
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sm[BLOCK / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sm[wid] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < BLOCK / 64; ++w) t += sm[w];
  __syncthreads();
  return t;
}

// Writes this block's partial, returns true in the (single) block that
// arrives last; that block then sees every partial.
__device__ __forceinline__ bool publish_partial(double* partials, uint32_t* counter, int slot,
                                                int nparts, double v) {
  __shared__ bool last;
  if (threadIdx.x == 0) {
    partials[slot] = v;
    __threadfence();
    uint32_t prev = atomicAdd(counter, 1u);
    last = (prev == (uint32_t)nparts - 1);
  }
  __syncthreads();
  if (last) __threadfence();
  return last;
}

// Sum of all partials in a fixed order (run by the last block).
__device__ __forceinline__ double gather_partials(const double* partials, int nparts) {
  double v = 0.0;
  for (int i = threadIdx.x; i < nparts; i += BLOCK) v += __builtin_nontemporal_load(&partials[i]);
  return block_sum(v);
}

// ---------------------------------------------------------------------------
// neighbour sum (mgpcg.py:96-102) for the colour-c cell at slot t:
//   ret = ret + (Ax[I-e_a][a] * x[I-e_a] + Ax[I][a] * x[I+e_a]),  a = 0..dim-1
// xo = other-colour vector, Axo/Axc = other/own colour coefficient arrays.
// ---------------------------------------------------------------------------
template <typename T, int DIM, typename V>
__device__ __forceinline__ T nsum(const Geo& g, const T* const Axc[3], const T* const Axo[3],
                                  const V* xo, int t, int q) {
  T ret = (T)0;
  if (DIM == 3) {
    T lo = Axo[0][t - g.sx] * (T)xo[t - g.sx];
    T hi = Axc[0][t] * (T)xo[t + g.sx];
    T s = lo + hi;
    ret = ret + s;
  }
  {
    const int a = (DIM == 3) ? 1 : 0;
    T lo = Axo[a][t - g.W] * (T)xo[t - g.W];
    T hi = Axc[a][t] * (T)xo[t + g.W];
    T s = lo + hi;
    ret = ret + s;
  }
  {
    const int a = DIM - 1;
    const int tl = t + q - 1, th = t + q;
    T lo = Axo[a][tl] * (T)xo[tl];
    T hi = Axc[a][t] * (T)xo[th];
    T s = lo + hi;
    ret = ret + s;
  }
  return ret;
}

template <typename T>
__device__ __forceinline__ void own_other(const LevelPtrs<T>& L, int c, const T* Axc[3],
                                          const T* Axo[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    Axc[a] = L.Ax[a][c];
    Axo[a] = L.Ax[a][c ^ 1];
  }
}

__device__ __forceinline__ int row_parity_q(const Geo& g, int t, int colour) {
  uint32_t row = g.divW.div((uint32_t)t);
  uint32_t ip = g.divP1.div(row);
  int jp = (int)row - (int)ip * g.P1;
  int i = (g.dim == 3) ? (int)ip - 1 : 0;
  int j = jp - 1;
  return (colour + ((i + j) & 1) + 1) & 1;
}

// ---------------------------------------------------------------------------
// smooth(l, phase) (mgpcg.py:104-109): one red-black half-sweep on colour c.
// FIRST: z was just zeroed (mgpcg.py:126 / :132) -> the half-sweep reduces to
// z = (r - 0) / Adiag on colour c and the kernel also writes the zeros of the
// whole level, replacing the separate fill.
// ---------------------------------------------------------------------------
template <typename T, int DIM, bool FIRST>
__global__ void __launch_bounds__(BLOCK) k_smooth(Geo g, LevelPtrs<T> L, int c,
                                                  const int* __restrict__ done) {
  if (done && *done) return;
  const int nb = gridDim.x;
  const int t = xcd_remap(blockIdx.x, nb) * BLOCK + threadIdx.x;
  if (t >= g.size) return;
  if (FIRST) {
    // neighbour_sum over an all-zero z is +-0, so (r - ns) / Adiag equals
    // r / Adiag (up to the sign of a zero result); the Ax reads are skipped.
    T zc = (T)0;
    if (L.gt[c][t] == MG_FLUID) zc = L.r[c][t] / L.Adiag[c][t];
    L.z[c][t] = zc;
    L.z[c ^ 1][t] = (T)0;
    return;
  }
  if (L.gt[c][t] != MG_FLUID) return;
  const T* Axc[3];
  const T* Axo[3];
  own_other(L, c, Axc, Axo);
  const int q = row_parity_q(g, t, c);
  T s = nsum<T, DIM>(g, Axc, Axo, L.z[c ^ 1], t, q);
  T num = L.r[c][t] - s;
  L.z[c][t] = num / L.Adiag[c][t];
}

// z <- 0 on both colours (used only when a level gets no smoothing).
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_zero2(int size, T* a0, T* a1, const int* done) {
  if (done && *done) return;
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= size) return;
  a0[t] = (T)0;
  a1[t] = (T)0;
}

// ---------------------------------------------------------------------------
// restrict(l) (mgpcg.py:111-118), gather form: every interior coarse cell sums
// the residuals r - (Adiag z + neighbour_sum) of its FLUID children in
// lexicographic offset order, starting from 0.
// ---------------------------------------------------------------------------
template <typename T, int DIM>
__device__ __forceinline__ T child_residual(const Geo& gf, const LevelPtrs<T>& F, int i, int j,
                                            int k) {
  const int c = colour_of(i, j, k);
  const int t = slot_of(gf, i, j, k);
  const T* Axc[3];
  const T* Axo[3];
  own_other(F, c, Axc, Axo);
  const int q = ((k + 1) & 1);  // padded k' = k + 1 = 2m + q
  T Az = F.Adiag[c][t] * F.z[c][t];
  T s = nsum<T, DIM>(gf, Axc, Axo, F.z[c ^ 1], t, q);
  Az = Az + s;
  return F.r[c][t] - Az;
}

template <typename T, int DIM>
__global__ void __launch_bounds__(BLOCK) k_restrict(Geo gf, Geo gc, LevelPtrs<T> F, T* rc0,
                                                    T* rc1, const int* __restrict__ done) {
  if (done && *done) return;
  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int cc = lb & 1;
  const int t = (lb >> 1) * BLOCK + threadIdx.x;
  if (t >= gc.size) return;
  Slot s = decode(gc, t, cc);
  if (!s.interior) return;
  T sum = (T)0;
  const int o0max = (DIM == 3) ? 2 : 1;
  for (int o0 = 0; o0 < o0max; ++o0)
    for (int o1 = 0; o1 < 2; ++o1)
      for (int o2 = 0; o2 < 2; ++o2) {
        const int i = (DIM == 3) ? 2 * s.i + o0 : 0;
        const int j = 2 * s.j + o1, k = 2 * s.k + o2;
        const int c = colour_of(i, j, k);
        const int tf = slot_of(gf, i, j, k);
        if (F.gt[c][tf] != MG_FLUID) continue;
        T res = child_residual<T, DIM>(gf, F, i, j, k);
        sum = sum + res;
      }
  (cc ? rc1 : rc0)[t] = sum;
}

// ---------------------------------------------------------------------------
// prolongate(l) (mgpcg.py:120-123): z_l[I] += z_{l+1}[I // 2] on every cell.
// ---------------------------------------------------------------------------
template <typename T, int DIM>
__global__ void __launch_bounds__(BLOCK) k_prolongate(Geo gf, Geo gc, T* zf0, T* zf1,
                                                      const T* __restrict__ zc0,
                                                      const T* __restrict__ zc1,
                                                      const int* __restrict__ done) {
  if (done && *done) return;
  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int c = lb & 1;
  const int t = (lb >> 1) * BLOCK + threadIdx.x;
  if (t >= gf.size) return;
  Slot s = decode(gf, t, c);
  if (!s.interior) return;
  const int pi = s.i >> 1, pj = s.j >> 1, pk = s.k >> 1;
  const int pc = colour_of(pi, pj, pk);
  const int tc = slot_of(gc, pi, pj, pk);
  T* zf = c ? zf1 : zf0;
  const T* zc = pc ? zc1 : zc0;
  zf[t] = zf[t] + zc[tc];
}

// ---------------------------------------------------------------------------
// CG kernels on level 0 (mgpcg.py:150-238) with fused FLUID-masked FP64 dots
// and the scalar updates done on device by the last block.
// ---------------------------------------------------------------------------
template <typename T, typename TX>
struct CGPtrs {
  T* b[2];
  TX* x[2];
  T* p[2];
  T* Ap[2];
};

struct RedPtrs {
  CGState* st;
  double* partials;
};

__device__ __forceinline__ double rnd_plain(const CGState* st, double v) {
  return st->precision == PLAIN ? (double)(float)v : v;
}

// r = b (copy_from, mgpcg.py:157), x = p = Ap = 0, reduce(r, r) -> r0r0, tol.
template <typename T, typename TX>
__global__ void __launch_bounds__(BLOCK) k_cg_init(Geo g, LevelPtrs<T> L0, CGPtrs<T, TX> V,
                                                   RedPtrs R) {
  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int c = lb & 1;
  const int t = (lb >> 1) * BLOCK + threadIdx.x;
  double acc = 0.0;
  if (t < g.size) {
    T bv = V.b[c][t];
    L0.r[c][t] = bv;
    V.x[c][t] = (TX)0;
    V.p[c][t] = (T)0;
    V.Ap[c][t] = (T)0;
    if (L0.gt[c][t] == MG_FLUID) acc = (double)bv * (double)bv;
  }
  double s = block_sum(acc);
  if (publish_partial(R.partials, &R.st->counter, blockIdx.x, nb, s)) {
    double tot = gather_partials(R.partials, nb);
    if (threadIdx.x == 0) {
      CGState* st = R.st;
      st->counter = 0;
      st->r0r0 = rnd_plain(st, tot);
      st->rr = st->r0r0;
      st->tol = fmax(st->abs_tol, st->r0r0 * st->rel_tol);
      st->iter = 0;
      st->done = 0;
      st->status = RUNNING;
      st->since_refine = 0;
      st->pending_true = 0;
      st->beta = 0.0;
      st->beta32 = 0.0f;
      st->true_rr = -1.0;
    }
  }
}

// update_p (mgpcg.py:234-238): p = z + beta p on FLUID cells.
template <typename T, typename TX>
__global__ void __launch_bounds__(BLOCK) k_update_p(Geo g, LevelPtrs<T> L0, CGPtrs<T, TX> V,
                                                    const CGState* __restrict__ st) {
  if (st->done) return;
  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int c = lb & 1;
  const int t = (lb >> 1) * BLOCK + threadIdx.x;
  if (t >= g.size) return;
  if (L0.gt[c][t] != MG_FLUID) return;
  if (st->precision == PLAIN) {
    const float b32 = st->beta32;
    T tt = (T)b32 * V.p[c][t];
    V.p[c][t] = L0.z[c][t] + tt;
  } else {
    const double beta = st->beta;
    V.p[c][t] = (T)((double)L0.z[c][t] + beta * (double)V.p[c][t]);
  }
}

// reduce(z, r) (mgpcg.py:171, 201) -> beta (mgpcg.py:203), iteration count.
// FIRSTDOT: the dot before the loop (mgpcg.py:171-172) only sets old_zTr.
template <typename T, bool FIRSTDOT>
__global__ void __launch_bounds__(BLOCK) k_dot_zr(Geo g, LevelPtrs<T> L0, RedPtrs R) {
  if (R.st->done) return;
  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int c = lb & 1;
  const int t = (lb >> 1) * BLOCK + threadIdx.x;
  double acc = 0.0;
  if (t < g.size && L0.gt[c][t] == MG_FLUID) acc = (double)L0.z[c][t] * (double)L0.r[c][t];
  double s = block_sum(acc);
  if (publish_partial(R.partials, &R.st->counter, blockIdx.x, nb, s)) {
    double tot = gather_partials(R.partials, nb);
    if (threadIdx.x == 0) {
      CGState* st = R.st;
      st->counter = 0;
      double zr = rnd_plain(st, tot);
      if (FIRSTDOT) {
        st->zr_old = zr;
      } else {
        st->zr_new = zr;
        double beta = zr / (st->zr_old + st->eps);
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
    }
  }
}

// compute_Ap (mgpcg.py:218-224) + reduce(p, Ap) -> alpha (mgpcg.py:178-181).
template <typename T, typename TX, int DIM>
__global__ void __launch_bounds__(BLOCK) k_compute_Ap(Geo g, LevelPtrs<T> L0, CGPtrs<T, TX> V,
                                                      RedPtrs R) {
  if (R.st->done) return;
  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int c = lb & 1;
  const int t = (lb >> 1) * BLOCK + threadIdx.x;
  double acc = 0.0;
  if (t < g.size && L0.gt[c][t] == MG_FLUID) {
    const T* Axc[3];
    const T* Axo[3];
    own_other(L0, c, Axc, Axo);
    const int q = row_parity_q(g, t, c);
    T pv = V.p[c][t];
    T rr = L0.Adiag[c][t] * pv;
    T s = nsum<T, DIM>(g, Axc, Axo, V.p[c ^ 1], t, q);
    T ap = rr + s;
    V.Ap[c][t] = ap;
    acc = (double)pv * (double)ap;
  }
  double s = block_sum(acc);
  if (publish_partial(R.partials, &R.st->counter, blockIdx.x, nb, s)) {
    double tot = gather_partials(R.partials, nb);
    if (threadIdx.x == 0) {
      CGState* st = R.st;
      st->counter = 0;
      st->pAp = rnd_plain(st, tot);
      double alpha = st->zr_old / (st->pAp + st->eps);
      st->alpha = alpha;
      st->alpha32 = (float)alpha;
      if (!isfinite(alpha)) {
        st->done = 1;
        st->status = BREAKDOWN;
      }
    }
  }
}

// update_xr (mgpcg.py:226-232) + reduce(r, r) + convergence test (:188-194).
// MIXED: x is FP64; on a refinement iteration r is left for k_true_residual.
template <typename T, typename TX>
__global__ void __launch_bounds__(BLOCK) k_update_xr(Geo g, LevelPtrs<T> L0, CGPtrs<T, TX> V,
                                                     RedPtrs R) {
  const CGState* cst = R.st;
  if (cst->done) return;
  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int c = lb & 1;
  const int t = (lb >> 1) * BLOCK + threadIdx.x;
  const bool mixed = cst->precision != PLAIN;
  const bool refine = cst->precision == MIXED && cst->refine_every > 0 &&
                      cst->since_refine + 1 >= cst->refine_every;
  double acc = 0.0;
  if (t < g.size && L0.gt[c][t] == MG_FLUID) {
    if (!mixed) {
      const float a32 = cst->alpha32;
      T tp = (T)a32 * V.p[c][t];
      V.x[c][t] = (TX)((T)V.x[c][t] + tp);
      T ta = (T)a32 * V.Ap[c][t];
      T rv = L0.r[c][t] - ta;
      L0.r[c][t] = rv;
      acc = (double)rv * (double)rv;
    } else {
      const double alpha = cst->alpha;
      if (cst->precision == F64) {
        // everything FP64: the same op order as the reference in double
        T tp = (T)(alpha * (double)V.p[c][t]);
        V.x[c][t] = (TX)((double)V.x[c][t] + (double)tp);
        T ta = (T)(alpha * (double)V.Ap[c][t]);
        T rv = L0.r[c][t] - ta;
        L0.r[c][t] = rv;
        acc = (double)rv * (double)rv;
      } else {
        V.x[c][t] = (TX)((double)V.x[c][t] + alpha * (double)V.p[c][t]);
        if (!refine) {
          T rv = (T)((double)L0.r[c][t] - alpha * (double)V.Ap[c][t]);
          L0.r[c][t] = rv;
          acc = (double)rv * (double)rv;
        }
      }
    }
  }
  double s = block_sum(acc);
  if (publish_partial(R.partials, &R.st->counter, blockIdx.x, nb, s)) {
    double tot = gather_partials(R.partials, nb);
    if (threadIdx.x == 0) {
      CGState* st = R.st;
      st->counter = 0;
      if (st->precision == MIXED) {
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
    }
  }
}

// MIXED: r = b - A x with x in FP64, accumulated in FP64, stored as T; then
// reduce(r, r) and the convergence test on this true residual.
template <typename T, typename TX, int DIM>
__global__ void __launch_bounds__(BLOCK) k_true_residual(Geo g, LevelPtrs<T> L0,
                                                         CGPtrs<T, TX> V, RedPtrs R) {
  const CGState* cst = R.st;
  if (cst->done || !cst->pending_true) return;
  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int c = lb & 1;
  const int t = (lb >> 1) * BLOCK + threadIdx.x;
  double acc = 0.0;
  if (t < g.size && L0.gt[c][t] == MG_FLUID) {
    const T* Axc[3];
    const T* Axo[3];
    own_other(L0, c, Axc, Axo);
    const int q = row_parity_q(g, t, c);
    const TX* xo = V.x[c ^ 1];
    double ret = 0.0;
    if (DIM == 3) {
      double lo = (double)Axo[0][t - g.sx] * (double)xo[t - g.sx];
      double hi = (double)Axc[0][t] * (double)xo[t + g.sx];
      double s = lo + hi;
      ret = ret + s;
    }
    {
      const int a = (DIM == 3) ? 1 : 0;
      double lo = (double)Axo[a][t - g.W] * (double)xo[t - g.W];
      double hi = (double)Axc[a][t] * (double)xo[t + g.W];
      double s = lo + hi;
      ret = ret + s;
    }
    {
      const int a = DIM - 1;
      const int tl = t + q - 1, th = t + q;
      double lo = (double)Axo[a][tl] * (double)xo[tl];
      double hi = (double)Axc[a][t] * (double)xo[th];
      double s = lo + hi;
      ret = ret + s;
    }
    double Axv = (double)L0.Adiag[c][t] * (double)V.x[c][t];
    Axv = Axv + ret;
    double rv = (double)V.b[c][t] - Axv;
    T r32 = (T)rv;
    L0.r[c][t] = r32;
    acc = (double)r32 * (double)r32;
  }
  double s = block_sum(acc);
  if (publish_partial(R.partials, &R.st->counter, blockIdx.x, nb, s)) {
    double tot = gather_partials(R.partials, nb);
    if (threadIdx.x == 0) {
      CGState* st = R.st;
      st->counter = 0;
      st->pending_true = 0;
      st->since_refine = 0;
      st->rr = tot;
      st->true_rr = tot;
      if (tot < st->tol) {
        st->done = 1;
        st->status = CONVERGED;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// layout conversion: canonical C-order (the strategies' fields) <-> colour split
// ---------------------------------------------------------------------------
__device__ __forceinline__ long canon_index(const Geo& g, int i, int j, int k) {
  return ((long)i * g.n1 + j) * g.n2 + k;
}

// one level: gt (int8), Adiag (f32), Ax (f32, [cells, dim] AoS) -> colour
// arrays of T; ghost slots get SOLID and zeros; r and z are zeroed.
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_pack_level(Geo g, const int8_t* __restrict__ gt,
                                                      const float* __restrict__ Adiag,
                                                      const float* __restrict__ Ax,
                                                      int8_t* gt0, int8_t* gt1, T* Ad0, T* Ad1,
                                                      T* Ax00, T* Ax01, T* Ax10, T* Ax11,
                                                      T* Ax20, T* Ax21, T* r0, T* r1, T* z0,
                                                      T* z1) {
  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int c = lb & 1;
  const int t = (lb >> 1) * BLOCK + threadIdx.x;
  if (t >= g.size) return;
  Slot s = decode(g, t, c);
  int8_t gv = MG_SOLID;
  T ad = (T)0, a0 = (T)0, a1 = (T)0, a2 = (T)0;
  if (s.interior) {
    long ci = canon_index(g, s.i, s.j, s.k);
    gv = gt[ci];
    ad = (T)Adiag[ci];
    a0 = (T)Ax[ci * g.dim + 0];
    a1 = (T)Ax[ci * g.dim + 1];
    if (g.dim == 3) a2 = (T)Ax[ci * g.dim + 2];
  }
  (c ? gt1 : gt0)[t] = gv;
  (c ? Ad1 : Ad0)[t] = ad;
  (c ? Ax01 : Ax00)[t] = a0;
  (c ? Ax11 : Ax10)[t] = a1;
  if (g.dim == 3) (c ? Ax21 : Ax20)[t] = a2;
  (c ? r1 : r0)[t] = (T)0;
  (c ? z1 : z0)[t] = (T)0;
}

// canonical f32 vector -> colour arrays of T (ghosts 0)
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_pack_vec(Geo g, const float* __restrict__ src, T* d0,
                                                    T* d1) {
  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int c = lb & 1;
  const int t = (lb >> 1) * BLOCK + threadIdx.x;
  if (t >= g.size) return;
  Slot s = decode(g, t, c);
  T v = (T)0;
  if (s.interior) v = (T)src[canon_index(g, s.i, s.j, s.k)];
  (c ? d1 : d0)[t] = v;
}

// colour arrays of S -> canonical vector of D (interior cells only)
template <typename S, typename D>
__global__ void __launch_bounds__(BLOCK) k_unpack_vec(Geo g, const S* __restrict__ s0,
                                                      const S* __restrict__ s1, D* dst) {
  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int c = lb & 1;
  const int t = (lb >> 1) * BLOCK + threadIdx.x;
  if (t >= g.size) return;
  Slot s = decode(g, t, c);
  if (!s.interior) return;
  dst[canon_index(g, s.i, s.j, s.k)] = (D)(c ? s1 : s0)[t];
}

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

}  // namespace mg
