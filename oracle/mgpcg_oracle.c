/*
 * CPU restatement of the reference MGPCG Poisson solver (mgpcg.py).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only
 * as the checker (or the timed CPU baseline), never as the product path.
 *
 * Layout is the reference's canonical one: every field is a C-order array
 * over the level's shape res_l = res >> l (last index contiguous); Ax is
 * array-of-structs [*res_l, dim] (mgpcg.py:34-35, a ti.Vector.field).
 *
 * Conventions fixed here (the reference leaves them undefined or
 * order-dependent; the HIP library follows the same ones):
 *  - An out-of-range neighbour contributes Ax_lo = 0, x = 0 to
 *    neighbor_sum (the reference reads out of bounds at coarse-level edges,
 *    mgpcg.py:96-102 with pressure_project.py:55).
 *  - neighbor_sum accumulates  ret = ret + (Ax[I-e_i][i]*x[I-e_i] + Ax[I][i]*x[I+e_i])
 *    for i = 0..dim-1, with no fused multiply-add (mgpcg.py:96-102).
 *  - restrict sums FLUID children in lexicographic offset order starting
 *    from 0 (the reference uses an atomic scatter, mgpcg.py:111-118).
 *  - reduce forms each product exactly in FP64 and accumulates in FP64 over
 *    fixed 4096-cell chunks summed in order (the reference uses FP32 atomics,
 *    mgpcg.py:211-216).  Precision mode 0 (PLAIN) rounds every reduced
 *    scalar, alpha and beta to FP32 exactly where the reference stores them
 *    in 0-d f32 fields (mgpcg.py:41-44, 159, 180-181, 189, 202-203).
 *
 * Precision modes:
 *  0 PLAIN  : FP32 everything, FP32 scalars (the reference's arithmetic).
 *  1 MIXED  : FP32 hierarchy/V-cycle/Laplacian, FP64 x and CG scalars,
 *             true residual r = b - A x recomputed (FP64) every
 *             `refine_every` iterations and before accepting convergence.
 *  2 F64    : everything FP64 (reference solves to 1e-12).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define FLUID 0
#define AIR 1
#define SOLID 2

typedef struct {
  int dim;
  long V[3]; /* view dims; 2D grids use V[0] = 1 */
  long S[3]; /* C-order strides of the view */
  long N;    /* number of cells */
} geo_t;

typedef struct {
  int iters;
  int status; /* 1 converged, 2 max_iters, 3 breakdown */
  double r0r0, rr, true_rr, tol;
} orc_result_t;

static geo_t make_geo(int dim, const int *n) {
  geo_t g;
  g.dim = dim;
  if (dim == 2) {
    g.V[0] = 1;
    g.V[1] = n[0];
    g.V[2] = n[1];
  } else {
    g.V[0] = n[0];
    g.V[1] = n[1];
    g.V[2] = n[2];
  }
  g.S[2] = 1;
  g.S[1] = g.V[2];
  g.S[0] = g.V[1] * g.V[2];
  g.N = g.V[0] * g.V[1] * g.V[2];
  return g;
}

static void level_dims(int dim, const int *res, int l, int *out) {
  for (int a = 0; a < dim; ++a) out[a] = res[a] >> l;
}

/* ---------------------------------------------------------------------
 * init_gridtype (mgpcg.py:46-60): coarse = AIR if any of the 2^d children
 * is AIR, else FLUID if any child is FLUID, else SOLID.
 * ------------------------------------------------------------------- */
int orc_coarsen(int dim, const int *nfine, const int8_t *fine, int8_t *coarse) {
  int nc[3] = {0, 0, 0};
  for (int a = 0; a < dim; ++a) nc[a] = nfine[a] / 2;
  geo_t gf = make_geo(dim, nfine), gc = make_geo(dim, nc);
  for (long c0 = 0; c0 < gc.V[0]; ++c0)
    for (long c1 = 0; c1 < gc.V[1]; ++c1)
      for (long c2 = 0; c2 < gc.V[2]; ++c2) {
        int tot_air = 0, tot_fluid = 0;
        long o0max = (dim == 3) ? 2 : 1;
        for (long o0 = 0; o0 < o0max; ++o0)
          for (long o1 = 0; o1 < 2; ++o1)
            for (long o2 = 0; o2 < 2; ++o2) {
              long f0 = (dim == 3) ? 2 * c0 + o0 : 0;
              int attr = fine[f0 * gf.S[0] + (2 * c1 + o1) * gf.S[1] + (2 * c2 + o2)];
              if (attr == AIR) tot_air++;
              else if (attr == FLUID) tot_fluid++;
            }
        int8_t t = tot_air > 0 ? AIR : (tot_fluid > 0 ? FLUID : SOLID);
        coarse[c0 * gc.S[0] + c1 * gc.S[1] + c2] = t;
      }
  return 0;
}

/* Masked dot: each product exact (FP32 x FP32 fits FP64; TwoProd for FP64),
 * summed as a double-double (TwoSum) over 4096-cell chunks in order, so the
 * rounded result does not depend on the summation order (except at an exact
 * rounding tie) and matches the HIP library's tree reduction. */
#define DOT_CHUNK 4096

typedef struct {
  double hi, lo;
} dd_t;

static inline dd_t dd_add(dd_t a, dd_t b) {
  const double s = a.hi + b.hi;
  const double bp = s - a.hi;
  double e = (a.hi - (s - bp)) + (b.hi - bp);
  e = e + (a.lo + b.lo);
  const double h = s + e;
  const double l = e - (h - s);
  dd_t r = {h, l};
  return r;
}
static inline dd_t dd_prod_f32(float a, float b) {
  dd_t r = {(double)a * (double)b, 0.0};
  return r;
}
static inline dd_t dd_prod_f64(double a, double b) {
  const double p = a * b;
  dd_t r = {p, fma(a, b, -p)};
  return r;
}

#define DEFINE_KERNELS(T, SFX)                                                                  \
  /* neighbor_sum (mgpcg.py:96-102) */                                                          \
  static inline T ns_##SFX(const geo_t *g, const T *Ax, const T *x, long c0, long c1, long c2,  \
                           long idx) {                                                          \
    T ret = (T)0;                                                                               \
    long cc[3] = {c0, c1, c2};                                                                  \
    for (int a = 0; a < g->dim; ++a) {                                                          \
      int va = a + 3 - g->dim;                                                                  \
      long s = g->S[va];                                                                        \
      T lo, hi;                                                                                 \
      if (cc[va] - 1 >= 0) {                                                                    \
        T p = Ax[(idx - s) * g->dim + a] * x[idx - s];                                          \
        lo = p;                                                                                 \
      } else {                                                                                  \
        lo = (T)0 * (T)0;                                                                       \
      }                                                                                         \
      if (cc[va] + 1 < g->V[va]) {                                                              \
        T p = Ax[idx * g->dim + a] * x[idx + s];                                                \
        hi = p;                                                                                 \
      } else {                                                                                  \
        T p = Ax[idx * g->dim + a] * (T)0;                                                      \
        hi = p;                                                                                 \
      }                                                                                         \
      T t = lo + hi;                                                                            \
      ret = ret + t;                                                                            \
    }                                                                                           \
    return ret;                                                                                 \
  }                                                                                             \
                                                                                                \
  /* smooth (mgpcg.py:104-109): one red-black Gauss-Seidel half-sweep */                       \
  void orc_smooth_##SFX(int dim, const int *n, const int8_t *gt, const T *Adiag, const T *Ax,   \
                        const T *r, T *z, int phase) {                                          \
    geo_t g = make_geo(dim, n);                                                                 \
    _Pragma("omp parallel for collapse(2) schedule(static)")                                    \
    for (long c0 = 0; c0 < g.V[0]; ++c0)                                                        \
      for (long c1 = 0; c1 < g.V[1]; ++c1)                                                      \
        for (long c2 = 0; c2 < g.V[2]; ++c2) {                                                  \
          if (((c0 + c1 + c2) & 1) != phase) continue;                                          \
          long idx = c0 * g.S[0] + c1 * g.S[1] + c2;                                            \
          if (gt[idx] != FLUID) continue;                                                       \
          T s = ns_##SFX(&g, Ax, z, c0, c1, c2, idx);                                           \
          T num = r[idx] - s;                                                                   \
          z[idx] = num / Adiag[idx];                                                            \
        }                                                                                       \
  }                                                                                             \
                                                                                                \
  /* restrict (mgpcg.py:111-118), gather form over FLUID children */                           \
  void orc_restrict_##SFX(int dim, const int *n, const int8_t *gt, const T *Adiag, const T *Ax, \
                          const T *r, const T *z, T *rc) {                                      \
    geo_t g = make_geo(dim, n);                                                                 \
    int ncd[3] = {0, 0, 0};                                                                               \
    for (int a = 0; a < dim; ++a) ncd[a] = n[a] / 2;                                            \
    geo_t gc = make_geo(dim, ncd);                                                              \
    long o0max = (dim == 3) ? 2 : 1;                                                            \
    _Pragma("omp parallel for collapse(2) schedule(static)")                                    \
    for (long J0 = 0; J0 < gc.V[0]; ++J0)                                                       \
      for (long J1 = 0; J1 < gc.V[1]; ++J1)                                                     \
        for (long J2 = 0; J2 < gc.V[2]; ++J2) {                                                 \
          T sum = (T)0;                                                                         \
          for (long o0 = 0; o0 < o0max; ++o0)                                                   \
            for (long o1 = 0; o1 < 2; ++o1)                                                     \
              for (long o2 = 0; o2 < 2; ++o2) {                                                 \
                long c0 = (dim == 3) ? 2 * J0 + o0 : 0, c1 = 2 * J1 + o1, c2 = 2 * J2 + o2;      \
                long idx = c0 * g.S[0] + c1 * g.S[1] + c2;                                      \
                if (gt[idx] != FLUID) continue;                                                 \
                T Az = Adiag[idx] * z[idx];                                                     \
                T s = ns_##SFX(&g, Ax, z, c0, c1, c2, idx);                                     \
                Az = Az + s;                                                                    \
                T res = r[idx] - Az;                                                            \
                sum = sum + res;                                                                \
              }                                                                                 \
          rc[J0 * gc.S[0] + J1 * gc.S[1] + J2] = sum;                                           \
        }                                                                                       \
  }                                                                                             \
                                                                                                \
  /* prolongate (mgpcg.py:120-123): injection into every cell */                               \
  void orc_prolongate_##SFX(int dim, const int *n, T *z, const T *zc) {                         \
    geo_t g = make_geo(dim, n);                                                                 \
    int ncd[3] = {0, 0, 0};                                                                               \
    for (int a = 0; a < dim; ++a) ncd[a] = n[a] / 2;                                            \
    geo_t gc = make_geo(dim, ncd);                                                              \
    _Pragma("omp parallel for collapse(2) schedule(static)")                                    \
    for (long c0 = 0; c0 < g.V[0]; ++c0)                                                        \
      for (long c1 = 0; c1 < g.V[1]; ++c1)                                                      \
        for (long c2 = 0; c2 < g.V[2]; ++c2) {                                                  \
          long idx = c0 * g.S[0] + c1 * g.S[1] + c2;                                            \
          long cidx = (c0 / 2) * gc.S[0] + (c1 / 2) * gc.S[1] + (c2 / 2);                      \
          z[idx] = z[idx] + zc[cidx];                                                           \
        }                                                                                       \
  }                                                                                             \
                                                                                                \
  /* compute_Ap (mgpcg.py:218-224) */                                                          \
  void orc_compute_Ap_##SFX(int dim, const int *n, const int8_t *gt, const T *Adiag,            \
                            const T *Ax, const T *p, T *Ap) {                                   \
    geo_t g = make_geo(dim, n);                                                                 \
    _Pragma("omp parallel for collapse(2) schedule(static)")                                    \
    for (long c0 = 0; c0 < g.V[0]; ++c0)                                                        \
      for (long c1 = 0; c1 < g.V[1]; ++c1)                                                      \
        for (long c2 = 0; c2 < g.V[2]; ++c2) {                                                  \
          long idx = c0 * g.S[0] + c1 * g.S[1] + c2;                                            \
          if (gt[idx] != FLUID) continue;                                                       \
          T r = Adiag[idx] * p[idx];                                                            \
          T s = ns_##SFX(&g, Ax, p, c0, c1, c2, idx);                                           \
          Ap[idx] = r + s;                                                                      \
        }                                                                                       \
  }                                                                                             \
                                                                                                \
  /* reduce (mgpcg.py:211-216): FLUID-masked dot, double-double, deterministic */              \
  double orc_dot_##SFX(long N, const int8_t *gt, const T *p, const T *q) {                      \
    long nch = (N + DOT_CHUNK - 1) / DOT_CHUNK;                                                 \
    dd_t *part = (dd_t *)malloc(sizeof(dd_t) * (nch > 0 ? nch : 1));                            \
    _Pragma("omp parallel for schedule(static)")                                                \
    for (long ch = 0; ch < nch; ++ch) {                                                         \
      dd_t acc = {0.0, 0.0};                                                                    \
      long e = (ch + 1) * DOT_CHUNK < N ? (ch + 1) * DOT_CHUNK : N;                             \
      for (long i = ch * DOT_CHUNK; i < e; ++i)                                                 \
        if (gt[i] == FLUID) acc = dd_add(acc, dd_prod_##SFX(p[i], q[i]));                       \
      part[ch] = acc;                                                                           \
    }                                                                                           \
    dd_t s = {0.0, 0.0};                                                                        \
    for (long ch = 0; ch < nch; ++ch) s = dd_add(s, part[ch]);                                  \
    free(part);                                                                                 \
    return s.hi + s.lo;                                                                         \
  }                                                                                             \
                                                                                                \
  /* true residual r = b - A x in FP64 (reported in every mode) */                             \
  double orc_true_residual_##SFX(int dim, const int *n, const int8_t *gt, const T *Adiag,       \
                                 const T *Ax, const T *b, const double *x, double *rout) {      \
    geo_t g = make_geo(dim, n);                                                                 \
    double acc = 0.0;                                                                           \
    for (long c0 = 0; c0 < g.V[0]; ++c0)                                                        \
      for (long c1 = 0; c1 < g.V[1]; ++c1)                                                      \
        for (long c2 = 0; c2 < g.V[2]; ++c2) {                                                  \
          long idx = c0 * g.S[0] + c1 * g.S[1] + c2;                                            \
          if (gt[idx] != FLUID) {                                                               \
            if (rout) rout[idx] = 0.0;                                                          \
            continue;                                                                           \
          }                                                                                     \
          double ret = 0.0;                                                                     \
          long cc[3] = {c0, c1, c2};                                                            \
          for (int a = 0; a < g.dim; ++a) {                                                     \
            int va = a + 3 - g.dim;                                                             \
            long s = g.S[va];                                                                   \
            double lo = 0.0, hi = 0.0;                                                          \
            if (cc[va] - 1 >= 0) lo = (double)Ax[(idx - s) * g.dim + a] * x[idx - s];           \
            if (cc[va] + 1 < g.V[va]) hi = (double)Ax[idx * g.dim + a] * x[idx + s];            \
            double t = lo + hi;                                                                 \
            ret = ret + t;                                                                      \
          }                                                                                     \
          double Axv = (double)Adiag[idx] * x[idx];                                             \
          Axv = Axv + ret;                                                                      \
          double rv = (double)b[idx] - Axv;                                                     \
          if (rout) rout[idx] = rv;                                                             \
          acc += rv * rv;                                                                       \
        }                                                                                       \
    return acc;                                                                                 \
  }                                                                                             \
                                                                                                \
  /* v_cycle (mgpcg.py:125-148) */                                                             \
  void orc_vcycle_##SFX(int dim, const int *res, int L, int nu, int bottom,                     \
                        const int8_t *const *gt, const T *const *Adiag, const T *const *Ax,     \
                        T *const *r, T *const *z) {                                             \
    int n[3] = {0, 0, 0};                                                                       \
    level_dims(dim, res, 0, n);                                                                 \
    geo_t g0 = make_geo(dim, n);                                                                \
    memset(z[0], 0, sizeof(T) * g0.N);                                                          \
    for (int l = 0; l < L - 1; ++l) {                                                           \
      level_dims(dim, res, l, n);                                                               \
      for (int i = 0; i < nu; ++i) {                                                            \
        orc_smooth_##SFX(dim, n, gt[l], Adiag[l], Ax[l], r[l], z[l], 0);                        \
        orc_smooth_##SFX(dim, n, gt[l], Adiag[l], Ax[l], r[l], z[l], 1);                        \
      }                                                                                         \
      int nc[3] = {0, 0, 0};                                                                                \
      level_dims(dim, res, l + 1, nc);                                                          \
      geo_t gc = make_geo(dim, nc);                                                             \
      memset(z[l + 1], 0, sizeof(T) * gc.N);                                                    \
      orc_restrict_##SFX(dim, n, gt[l], Adiag[l], Ax[l], r[l], z[l], r[l + 1]);                 \
    }                                                                                           \
    level_dims(dim, res, L - 1, n);                                                             \
    for (int i = 0; i < bottom / 2; ++i) {                                                      \
      orc_smooth_##SFX(dim, n, gt[L - 1], Adiag[L - 1], Ax[L - 1], r[L - 1], z[L - 1], 0);      \
      orc_smooth_##SFX(dim, n, gt[L - 1], Adiag[L - 1], Ax[L - 1], r[L - 1], z[L - 1], 1);      \
    }                                                                                           \
    for (int i = 0; i < bottom / 2; ++i) {                                                      \
      orc_smooth_##SFX(dim, n, gt[L - 1], Adiag[L - 1], Ax[L - 1], r[L - 1], z[L - 1], 1);      \
      orc_smooth_##SFX(dim, n, gt[L - 1], Adiag[L - 1], Ax[L - 1], r[L - 1], z[L - 1], 0);      \
    }                                                                                           \
    for (int l = L - 2; l >= 0; --l) {                                                          \
      level_dims(dim, res, l, n);                                                               \
      orc_prolongate_##SFX(dim, n, z[l], z[l + 1]);                                             \
      for (int i = 0; i < nu; ++i) {                                                            \
        orc_smooth_##SFX(dim, n, gt[l], Adiag[l], Ax[l], r[l], z[l], 1);                        \
        orc_smooth_##SFX(dim, n, gt[l], Adiag[l], Ax[l], r[l], z[l], 0);                        \
      }                                                                                         \
    }                                                                                           \
  }

DEFINE_KERNELS(float, f32)
DEFINE_KERNELS(double, f64)

/* Allocate the per-level r and z of a hierarchy. */
static void *alloc_levels(int dim, const int *res, int L, size_t esz, void **out) {
  for (int l = 0; l < L; ++l) {
    int n[3] = {0, 0, 0};
    level_dims(dim, res, l, n);
    geo_t g = make_geo(dim, n);
    out[l] = calloc((size_t)g.N, esz);
  }
  return out;
}
static void free_levels(int L, void **p) {
  for (int l = 0; l < L; ++l) free(p[l]);
}

/* ---------------------------------------------------------------------
 * solve (mgpcg.py:150-209), FP32 hierarchy: precision 0 (PLAIN) or 1 (MIXED).
 * x_out receives x as FP64 (in PLAIN mode every value is FP32-representable).
 * rr_hist (optional) receives rTr after every update_xr.
 * ------------------------------------------------------------------- */
int orc_solve_f32(int dim, const int *res, int L, int nu, int bottom, int mixed,
                  const int8_t *const *gt, const float *const *Adiag, const float *const *Ax,
                  const float *b, double *x_out, int max_iters, double rel_tol, double abs_tol,
                  double eps, int refine_every, double *rr_hist, int hist_cap,
                  orc_result_t *out) {
  int n[3] = {0, 0, 0};
  level_dims(dim, res, 0, n);
  geo_t g = make_geo(dim, n);
  long N = g.N;
  void *rl[32], *zl[32];
  if (L < 1 || L > 32) return -1;
  alloc_levels(dim, res, L, sizeof(float), rl);
  alloc_levels(dim, res, L, sizeof(float), zl);
  float **r = (float **)rl, **z = (float **)zl;
  float *p = (float *)calloc(N, sizeof(float));
  float *Ap = (float *)calloc(N, sizeof(float));
  float *x32 = (float *)calloc(N, sizeof(float));
  double *x64 = (double *)calloc(N, sizeof(double));
  double *rtrue = (double *)calloc(N, sizeof(double));
  const int8_t *gt0 = gt[0];

  memcpy(r[0], b, sizeof(float) * N); /* r0 <- b  (x = 0) */
  double r0r0 = orc_dot_f32(N, gt0, r[0], r[0]);
  if (!mixed) r0r0 = (double)(float)r0r0;
  double tol = fmax(abs_tol, r0r0 * rel_tol);

  orc_vcycle_f32(dim, res, L, nu, bottom, gt, Adiag, Ax, r, z);
  /* update_p with beta = 0 on p = 0:  p = z + 0*p */
  for (long i = 0; i < N; ++i)
    if (gt0[i] == FLUID) {
      if (mixed) p[i] = (float)((double)z[0][i] + 0.0 * (double)p[i]);
      else p[i] = z[0][i] + 0.0f * p[i];
    }
  double old_zr = orc_dot_f32(N, gt0, z[0], r[0]);
  if (!mixed) old_zr = (double)(float)old_zr;

  int iter = 0, status = 2, since_refine = 0;
  double rr = r0r0, true_rr = -1.0;
  while (max_iters == -1 || iter < max_iters) {
    orc_compute_Ap_f32(dim, n, gt0, Adiag[0], Ax[0], p, Ap);
    double pAp = orc_dot_f32(N, gt0, p, Ap);
    if (!mixed) pAp = (double)(float)pAp;
    double alpha = old_zr / (pAp + eps);
    if (!isfinite(alpha)) {
      status = 3;
      break;
    }
    since_refine++;
    int refine = mixed && refine_every > 0 && since_refine >= refine_every;
    if (!mixed) {
      float a32 = (float)alpha;
      for (long i = 0; i < N; ++i)
        if (gt0[i] == FLUID) {
          float t = a32 * p[i];
          x32[i] = x32[i] + t;
          float u = a32 * Ap[i];
          r[0][i] = r[0][i] - u;
        }
    } else {
      for (long i = 0; i < N; ++i)
        if (gt0[i] == FLUID) {
          x64[i] = x64[i] + alpha * (double)p[i];
          if (!refine) r[0][i] = (float)((double)r[0][i] - alpha * (double)Ap[i]);
        }
    }
    if (refine) {
      true_rr = orc_true_residual_f32(dim, n, gt0, Adiag[0], Ax[0], b, x64, rtrue);
      for (long i = 0; i < N; ++i)
        if (gt0[i] == FLUID) r[0][i] = (float)rtrue[i];
      since_refine = 0;
    }
    rr = orc_dot_f32(N, gt0, r[0], r[0]);
    if (!mixed) rr = (double)(float)rr;
    if (rr_hist && iter < hist_cap) rr_hist[iter] = rr;
    if (rr < tol) {
      if (mixed && !refine) {
        /* accept only on the true residual */
        true_rr = orc_true_residual_f32(dim, n, gt0, Adiag[0], Ax[0], b, x64, rtrue);
        for (long i = 0; i < N; ++i)
          if (gt0[i] == FLUID) r[0][i] = (float)rtrue[i];
        since_refine = 0;
        rr = orc_dot_f32(N, gt0, r[0], r[0]);
      }
      if (rr < tol) {
        status = 1;
        break;
      }
    }
    orc_vcycle_f32(dim, res, L, nu, bottom, gt, Adiag, Ax, r, z);
    double new_zr = orc_dot_f32(N, gt0, z[0], r[0]);
    if (!mixed) new_zr = (double)(float)new_zr;
    double beta = new_zr / (old_zr + eps);
    if (!mixed) {
      float b32 = (float)beta;
      for (long i = 0; i < N; ++i)
        if (gt0[i] == FLUID) {
          float t = b32 * p[i];
          p[i] = z[0][i] + t;
        }
    } else {
      for (long i = 0; i < N; ++i)
        if (gt0[i] == FLUID) p[i] = (float)((double)z[0][i] + beta * (double)p[i]);
    }
    old_zr = new_zr;
    iter++;
  }
  if (!mixed)
    for (long i = 0; i < N; ++i) x64[i] = (double)x32[i];
  double trr = orc_true_residual_f32(dim, n, gt0, Adiag[0], Ax[0], b, x64, NULL);
  memcpy(x_out, x64, sizeof(double) * N);
  if (out) {
    out->iters = iter;
    out->status = status;
    out->r0r0 = r0r0;
    out->rr = rr;
    out->true_rr = trr;
    out->tol = tol;
  }
  (void)true_rr;
  free(p);
  free(Ap);
  free(x32);
  free(x64);
  free(rtrue);
  free_levels(L, rl);
  free_levels(L, zl);
  return 0;
}

/* solve, everything FP64 (precision 2). */
int orc_solve_f64(int dim, const int *res, int L, int nu, int bottom, const int8_t *const *gt,
                  const double *const *Adiag, const double *const *Ax, const double *b,
                  double *x_out, int max_iters, double rel_tol, double abs_tol, double eps,
                  double *rr_hist, int hist_cap, orc_result_t *out) {
  int n[3] = {0, 0, 0};
  level_dims(dim, res, 0, n);
  geo_t g = make_geo(dim, n);
  long N = g.N;
  void *rl[32], *zl[32];
  if (L < 1 || L > 32) return -1;
  alloc_levels(dim, res, L, sizeof(double), rl);
  alloc_levels(dim, res, L, sizeof(double), zl);
  double **r = (double **)rl, **z = (double **)zl;
  double *p = (double *)calloc(N, sizeof(double));
  double *Ap = (double *)calloc(N, sizeof(double));
  double *x = (double *)calloc(N, sizeof(double));
  const int8_t *gt0 = gt[0];
  memcpy(r[0], b, sizeof(double) * N);
  double r0r0 = orc_dot_f64(N, gt0, r[0], r[0]);
  double tol = fmax(abs_tol, r0r0 * rel_tol);
  orc_vcycle_f64(dim, res, L, nu, bottom, gt, Adiag, Ax, r, z);
  for (long i = 0; i < N; ++i)
    if (gt0[i] == FLUID) p[i] = z[0][i] + 0.0 * p[i];
  double old_zr = orc_dot_f64(N, gt0, z[0], r[0]);
  int iter = 0, status = 2;
  double rr = r0r0;
  while (max_iters == -1 || iter < max_iters) {
    orc_compute_Ap_f64(dim, n, gt0, Adiag[0], Ax[0], p, Ap);
    double pAp = orc_dot_f64(N, gt0, p, Ap);
    double alpha = old_zr / (pAp + eps);
    if (!isfinite(alpha)) {
      status = 3;
      break;
    }
    for (long i = 0; i < N; ++i)
      if (gt0[i] == FLUID) {
        double t = alpha * p[i];
        x[i] = x[i] + t;
        double u = alpha * Ap[i];
        r[0][i] = r[0][i] - u;
      }
    rr = orc_dot_f64(N, gt0, r[0], r[0]);
    if (rr_hist && iter < hist_cap) rr_hist[iter] = rr;
    if (rr < tol) {
      status = 1;
      break;
    }
    orc_vcycle_f64(dim, res, L, nu, bottom, gt, Adiag, Ax, r, z);
    double new_zr = orc_dot_f64(N, gt0, z[0], r[0]);
    double beta = new_zr / (old_zr + eps);
    for (long i = 0; i < N; ++i)
      if (gt0[i] == FLUID) {
        double t = beta * p[i];
        p[i] = z[0][i] + t;
      }
    old_zr = new_zr;
    iter++;
  }
  double trr = orc_true_residual_f64(dim, n, gt0, Adiag[0], Ax[0], b, x, NULL);
  memcpy(x_out, x, sizeof(double) * N);
  if (out) {
    out->iters = iter;
    out->status = status;
    out->r0r0 = r0r0;
    out->rr = rr;
    out->true_rr = trr;
    out->tol = tol;
  }
  free(p);
  free(Ap);
  free(x);
  free_levels(L, rl);
  free_levels(L, zl);
  return 0;
}
