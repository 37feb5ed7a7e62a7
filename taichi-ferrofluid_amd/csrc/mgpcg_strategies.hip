// Operator assembly ("Strategy" build_b / build_A) on the device, canonical
// C-order layout, one thread per cell.  Each kernel restates one reference
// Taichi kernel in FP32 with the same per-cell operation order; Python-scope
// constants become FP32 constants as Taichi bakes them (fluid_simulator.py:16
// leaves default_fp at f32).
//
// Shared conventions (the reference leaves them undefined):
//  * an out-of-range neighbour cell is SOLID (pressure_project.py:55 reads
//    grid_type out of bounds at coarse-level edges);
//  * the ghost-fluid weight c = (phi_i - phi_j) / phi_i becomes c_max when
//    phi_i == 0 (the reference stores -inf there, pressure_project.py:36-37).
#include <hip/hip_runtime.h>
#include <math.h>

#include <cstring>
#include <mutex>

#include "../../include/mgpcg.h"
#include "mgpcg_common.h"

namespace {

constexpr int SB = 256;

struct Shape {
  int dim;
  int n[3];  // C-order dims (n[2] unused in 2D)
  mg::FastDiv dn1, dn2;  // division by n[1] and n[2] (coords below 2^31)
  __host__ __device__ long cells() const {
    return dim == 3 ? (long)n[0] * n[1] * n[2] : (long)n[0] * n[1];
  }
  __host__ __device__ long idx(const int* I) const {
    return dim == 3 ? ((long)I[0] * n[1] + I[1]) * n[2] + I[2] : (long)I[0] * n[1] + I[1];
  }
  __host__ __device__ void coords(long t, int* I) const {
    if (t < 0x80000000L) {  // multiply-high division, no 64-bit divide
      const uint32_t u = (uint32_t)t;
      if (dim == 3) {
        const uint32_t r = dn2.div(u);
        I[2] = (int)(u - r * (uint32_t)n[2]);
        I[0] = (int)dn1.div(r);
        I[1] = (int)(r - (uint32_t)I[0] * (uint32_t)n[1]);
      } else {
        I[0] = (int)dn1.div(u);
        I[1] = (int)(u - (uint32_t)I[0] * (uint32_t)n[1]);
        I[2] = 0;
      }
      return;
    }
    if (dim == 3) {
      I[2] = (int)(t % n[2]);
      long r = t / n[2];
      I[1] = (int)(r % n[1]);
      I[0] = (int)(r / n[1]);
    } else {
      I[1] = (int)(t % n[1]);
      I[0] = (int)(t / n[1]);
      I[2] = 0;
    }
  }
  __host__ __device__ bool inside(const int* I) const {
    for (int a = 0; a < dim; ++a)
      if (I[a] < 0 || I[a] >= n[a]) return false;
    return true;
  }
};

Shape make_shape(int dim, const int32_t* res, int level = 0, int grow_axis = -1) {
  Shape s;
  s.dim = dim;
  for (int a = 0; a < 3; ++a) s.n[a] = 1;
  for (int a = 0; a < dim; ++a) s.n[a] = (res[a] >> level) + (a == grow_axis ? 1 : 0);
  s.dn1 = mg::FastDiv((uint32_t)s.n[1]);
  s.dn2 = mg::FastDiv((uint32_t)s.n[2]);
  return s;
}

// A field together with its shape (MAC components are res + e_k).
struct Field {
  const float* p;
  Shape s;
};

// utils.sample (utils.py:12-32): clamped multilinear interpolation at
// index-space position pos; int() truncates toward zero.
__device__ float sample(const Field& f, const float* pos) {
  int i0[3], i1[3];
  float w[3];
  for (int a = 0; a < f.s.dim; ++a) {
    const int n = f.s.n[a];
    int i = (int)pos[a];
    i = i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
    int ip = i + 1;
    ip = ip < 0 ? 0 : (ip > n - 1 ? n - 1 : ip);
    float s = pos[a] - (float)i;
    s = s < 0.0f ? 0.0f : (s > 1.0f ? 1.0f : s);
    i0[a] = i;
    i1[a] = ip;
    w[a] = s;
  }
  const float* d = f.p;
  if (f.s.dim == 2) {
    const int n1 = f.s.n[1];
    auto at = [&](int i, int j) { return d[(long)i * n1 + j]; };
    const float s = w[0], t = w[1];
    const int i = i0[0], j = i0[1], ip = i1[0], jp = i1[1];
    float a = at(i, j) * (1.0f - s) + at(ip, j) * s;
    float b = at(i, jp) * (1.0f - s) + at(ip, jp) * s;
    return a * (1.0f - t) + b * t;
  }
  const int n1 = f.s.n[1], n2 = f.s.n[2];
  auto at = [&](int i, int j, int k) { return d[((long)i * n1 + j) * n2 + k]; };
  const float s = w[0], t = w[1], u = w[2];
  const int i = i0[0], j = i0[1], k = i0[2], ip = i1[0], jp = i1[1], kp = i1[2];
  float a = (at(i, j, k) * (1.0f - s) + at(ip, j, k) * s) * (1.0f - t) +
            (at(i, jp, k) * (1.0f - s) + at(ip, jp, k) * s) * t;
  float b = (at(i, j, kp) * (1.0f - s) + at(ip, j, kp) * s) * (1.0f - t) +
            (at(i, jp, kp) * (1.0f - s) + at(ip, jp, kp) * s) * t;
  return a * (1.0f - u) + b * u;
}

// LevelSet.theta / delta (level_set.py:24-37), eps = 5 dx (level_set.py:22).
struct LS {
  float eps, neg_eps, two_eps, inv_2pi, pi;
};
LS make_ls(double dx) {
  const double eps = 5.0 * dx;
  LS c;
  c.eps = (float)eps;
  c.neg_eps = (float)(-eps);
  c.two_eps = (float)(2.0 * eps);
  c.inv_2pi = (float)(1.0 / (2.0 * M_PI));
  c.pi = (float)M_PI;
  return c;
}
__device__ float theta(float phi, const LS& c) {
  if (phi <= c.neg_eps) return 0.0f;
  if (phi >= c.eps) return 1.0f;
  float a = 0.5f + phi / c.two_eps;
  float arg = c.pi * phi / c.eps;
  return a + c.inv_2pi * sinf(arg);
}
__device__ float delta(float phi, const LS& c) {
  if (phi <= c.neg_eps || phi >= c.eps) return 0.0f;
  float arg = c.pi * phi / c.eps;
  return (1.0f + cosf(arg)) / c.two_eps;
}

__device__ float gfm_c(float phi_i, float phi_j, float c_max) {
  float c = (phi_i - phi_j) / phi_i;
  return phi_i == 0.0f ? c_max : fminf(c, c_max);
}

__device__ int nb_type(const int8_t* ct, const Shape& s, const int* I, int k, int sgn) {
  int J[3] = {I[0], I[1], I[2]};
  J[k] += sgn;
  if (!s.inside(J)) return MG_SOLID;
  return ct[s.idx(J)];
}

struct Vel {
  Field u[3];
};

// ---------------------------------------------------------------------------
// PressureProjectStrategy.build_b_kernel (pressure_project.py:15-39).
// mode 0: pressure (air term sA*c*p0 with GFM, or sA*p0)
// mode 1: MagneticForceStrategy (magnetic_force.py:17-39): air term
//         sA*(p0 + coef*|H(I + s e_k + 0.5)|^2), no GFM weight.
// ---------------------------------------------------------------------------
struct HField {
  Field h[3];
};

__device__ float hdoth(const HField& H, int dim, const float* pos) {
  // FerrofluidSimulator.H_interp (ferrofluid_simulator.py:66-70)
  float acc = 0.0f;
  for (int a = 0; a < dim; ++a) {
    float q[3];
    for (int b = 0; b < dim; ++b) q[b] = pos[b] - (b == a ? 0.0f : 0.5f);
    float h = sample(H.h[a], q);
    float hh = h * h;
    acc = (a == 0) ? hh : acc + hh;
  }
  return acc;
}

__global__ void k_pressure_b(Shape s, const int8_t* __restrict__ ct, Vel V,
                             const float* __restrict__ phi, float sA, float sb, float p0, int gfm,
                             float c_max, int mode, HField H, float coef, float* __restrict__ b) {
  const long t = (long)blockIdx.x * SB + threadIdx.x;
  if (t >= s.cells()) return;
  int I[3];
  s.coords(t, I);
  if (ct[t] != MG_FLUID) {
    b[t] = 0.0f;
    return;
  }
  float acc = 0.0f;
  for (int k = 0; k < s.dim; ++k) {
    int J[3] = {I[0], I[1], I[2]};
    J[k] += 1;
    float d = V.u[k].p[V.u[k].s.idx(I)] - V.u[k].p[V.u[k].s.idx(J)];
    acc = acc + d;
  }
  float bv = acc * sb;
  for (int k = 0; k < s.dim; ++k) {
    for (int sgn = -1; sgn <= 1; sgn += 2) {
      const int nb = nb_type(ct, s, I, k, sgn);
      if (nb == MG_SOLID) {
        if (sgn < 0) {
          float u = V.u[k].p[V.u[k].s.idx(I)] - 0.0f;
          bv = bv - sb * u;
        } else {
          int J[3] = {I[0], I[1], I[2]};
          J[k] += 1;
          float u = V.u[k].p[V.u[k].s.idx(J)] - 0.0f;
          bv = bv + sb * u;
        }
      } else if (nb == MG_AIR) {
        float term;
        if (mode == 1) {
          float pos[3];
          for (int a = 0; a < s.dim; ++a) pos[a] = (float)(I[a] + (a == k ? sgn : 0)) + 0.5f;
          float hh = hdoth(H, s.dim, pos);
          term = sA * (p0 + coef * hh);
        } else if (gfm) {
          int J[3] = {I[0], I[1], I[2]};
          J[k] += sgn;
          float c = gfm_c(phi[t], phi[s.idx(J)], c_max);
          term = (sA * c) * p0;
        } else {
          term = sA * p0;
        }
        bv = bv + term;
      }
    }
  }
  b[t] = bv;
}

// The 2 * DIM face neighbours of cell I: their types (out of range counts as
// SOLID) and, when `phi` is given, their level-set values (0 out of range,
// where the type is SOLID and phi is not used).  All loads are issued before
// any is used: one round trip per cell instead of one per neighbour.
template <int DIM>
struct Faces {
  int type[DIM][2];  // [axis][0: -e_k, 1: +e_k]
  float phi[DIM][2];
};
template <int DIM>
__device__ __forceinline__ void load_faces(const Shape& s, const int8_t* __restrict__ ct,
                                           const float* __restrict__ phi, const int* I, long t,
                                           Faces<DIM>& f) {
  bool in[DIM][2];
  long at[DIM][2];
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      int J[3] = {I[0], I[1], I[2]};
      J[k] += h ? 1 : -1;
      in[k][h] = s.inside(J);
      at[k][h] = in[k][h] ? s.idx(J) : t;
    }
  }
#pragma unroll
  for (int k = 0; k < DIM; ++k)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int g = ct[at[k][h]];
      f.type[k][h] = in[k][h] ? g : MG_SOLID;
      f.phi[k][h] = phi ? phi[at[k][h]] : 0.0f;
    }
}

// PressureProjectStrategy.build_A_kernel (pressure_project.py:45-64).
template <int DIM>
__global__ void k_pressure_A(Shape s, int level, const int8_t* __restrict__ gt,
                             const float* __restrict__ phi, float sA, int gfm, float c_max,
                             float* __restrict__ Adiag, float* __restrict__ Ax) {
  const long t = (long)blockIdx.x * SB + threadIdx.x;
  if (t >= s.cells()) return;
  int I[3];
  s.coords(t, I);
  float ad = 0.0f, ax[3] = {0.0f, 0.0f, 0.0f};
  if (gt[t] == MG_FLUID) {
    const bool use_phi = gfm && level == 0;
    const float* ph = use_phi ? phi : nullptr;
    Faces<DIM> f;
    load_faces<DIM>(s, gt, ph, I, t, f);
    const float phi_t = use_phi ? phi[t] : 0.0f;
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {  // sgn = -1, then +1
        const int nb = f.type[k][h];
        if (nb == MG_FLUID) {
          ad = ad + sA;
          if (h) ax[k] = -sA;
        } else if (nb == MG_AIR) {
          if (use_phi) {
            float c = gfm_c(phi_t, f.phi[k][h], c_max);
            ad = ad + sA * c;
          } else {
            ad = ad + sA;
          }
        }
      }
    }
  }
  Adiag[t] = ad;
#pragma unroll
  for (int k = 0; k < DIM; ++k) Ax[t * DIM + k] = ax[k];
}

// build_b_kernel of the pressure strategy (mode 0 of k_pressure_b) in the
// same operation order, every load of a FLUID cell issued up front.
template <int DIM>
__global__ void k_pressure_b0(Shape s, const int8_t* __restrict__ ct, Vel V,
                              const float* __restrict__ phi, float sA, float sb, float p0,
                              int gfm, float c_max, float* __restrict__ b) {
  const long t = (long)blockIdx.x * SB + threadIdx.x;
  if (t >= s.cells()) return;
  int I[3];
  s.coords(t, I);
  if (ct[t] != MG_FLUID) {
    b[t] = 0.0f;
    return;
  }
  float ulo[DIM], uhi[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    int J[3] = {I[0], I[1], I[2]};
    J[k] += 1;
    ulo[k] = V.u[k].p[V.u[k].s.idx(I)];
    uhi[k] = V.u[k].p[V.u[k].s.idx(J)];
  }
  Faces<DIM> f;
  load_faces<DIM>(s, ct, gfm ? phi : nullptr, I, t, f);
  const float phi_t = gfm ? phi[t] : 0.0f;
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
      const int nb = f.type[k][h];
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
        if (gfm) {
          float c = gfm_c(phi_t, f.phi[k][h], c_max);
          term = (sA * c) * p0;
        } else {
          term = sA * p0;
        }
        bv = bv + term;
      }
    }
  }
  b[t] = bv;
}

// PressureProjectWithVolumeControlStrategy.calc_volume (volume_control.py:22-31):
// per-block FP64 partial sums of the cell volumes.
__global__ void k_volume(Shape s, Field phi, LS ls, double* __restrict__ partials) {
  const long t = (long)blockIdx.x * SB + threadIdx.x;
  double v = 0.0;
  if (t < s.cells()) {
    int I[3];
    s.coords(t, I);
    float tot = 0.0f;
    const int nco = 1 << s.dim;
    for (int o = 0; o < nco; ++o) {
      float pos[3];
      for (int a = 0; a < s.dim; ++a) {
        const int bit = (o >> (s.dim - 1 - a)) & 1;  // lexicographic offsets
        pos[a] = (float)(I[a] + bit) - 0.5f;
      }
      tot = tot + theta(-sample(phi, pos), ls);
    }
    v = (double)(tot / (float)nco);
  }
  __shared__ double sm[SB];
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int w = SB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sm[0];
}

__global__ void k_sum_partials(const double* __restrict__ partials, int n, double* out) {
  __shared__ double sm[SB];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += SB) v += partials[i];
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int w = SB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sm[0];
}

// add_c (volume_control.py:33-37)
__global__ void k_add_c(long n, const int8_t* __restrict__ ct, float c, float* __restrict__ b) {
  const long t = (long)blockIdx.x * SB + threadIdx.x;
  if (t >= n) return;
  if (ct[t] == MG_FLUID) b[t] = b[t] + c;
}

// PotentialFunctionStrategy.build_b_kernel (potential_function.py:19-34).
// in_domain(I) = all(I >= 0) and all(I + 1 <= res) on the fine res.
__global__ void k_potential_b(Shape s, Field chi, HField Hext, float* __restrict__ b) {
  const long t = (long)blockIdx.x * SB + threadIdx.x;
  if (t >= s.cells()) return;
  int I[3];
  s.coords(t, I);
  bool origin = true;
  for (int a = 0; a < s.dim; ++a) origin = origin && (I[a] == 0);
  float bv = 0.0f;
  if (!origin) {
    for (int k = 0; k < s.dim; ++k) {
      // I + e_k in domain  <=>  I[k] + 2 <= res[k]
      if (I[k] + 2 <= s.n[k]) {
        int U[3] = {I[0], I[1], I[2]};
        U[k] += 1;
        float pos[3];
        for (int a = 0; a < s.dim; ++a) pos[a] = (float)U[a] + (a == k ? 0.0f : 0.5f);
        float term = sample(chi, pos) * Hext.h[k].p[Hext.h[k].s.idx(U)];
        bv = bv + term;
      }
      if (I[k] - 1 >= 0) {
        float pos[3];
        for (int a = 0; a < s.dim; ++a) pos[a] = (float)I[a] + (a == k ? 0.0f : 0.5f);
        float term = sample(chi, pos) * Hext.h[k].p[Hext.h[k].s.idx(I)];
        bv = bv - term;
      }
    }
  }
  b[t] = bv;
}

// PotentialFunctionStrategy.build_A_kernel (potential_function.py:40-55) at
// level l; `fine` is the fine shape used by in_domain at every level.
__global__ void k_potential_A(Shape s, Shape fine, int level, float dxl, Field chi,
                              float* __restrict__ Adiag, float* __restrict__ Ax) {
  const long t = (long)blockIdx.x * SB + threadIdx.x;
  if (t >= s.cells()) return;
  int I[3];
  s.coords(t, I);
  bool origin = true;
  for (int a = 0; a < s.dim; ++a) origin = origin && (I[a] == 0);
  float ad = 0.0f, ax[3] = {0.0f, 0.0f, 0.0f};
  const float scale = (float)(1 << level);
  if (origin) {
    ad = 1.0f / dxl;
  } else {
    for (int k = 0; k < s.dim; ++k) {
      for (int sgn = -1; sgn <= 1; sgn += 2) {
        const int nbk = I[k] + sgn;
        if (nbk >= 0 && nbk + 1 <= fine.n[k]) {
          float pos[3];
          for (int a = 0; a < s.dim; ++a)
            pos[a] = ((float)I[a] + (a == k ? 0.5f * (float)sgn : 0.0f)) * scale;
          float val = (1.0f + sample(chi, pos)) / dxl;
          ad = ad - val;
          if (sgn > 0) ax[k] = val;
        }
      }
    }
  }
  Adiag[t] = ad;
  for (int k = 0; k < s.dim; ++k) Ax[t * s.dim + k] = ax[k];
}

// SurfaceTensionStrategy.build_A_kernel (surface_tension.py:71-89) at level l;
// the bounds I[k] +- 1 test the fine res at every level (surface_tension.py:82,86).
__global__ void k_surface_A(Shape s, Shape fine, int level, int d, Field phi, LS ls,
                            float inv_dt, float scale, float* __restrict__ Adiag,
                            float* __restrict__ Ax) {
  const long t = (long)blockIdx.x * SB + threadIdx.x;
  if (t >= s.cells()) return;
  int I[3];
  s.coords(t, I);
  const float lvl = (float)(1 << level);
  float ad = inv_dt;
  float ax[3] = {0.0f, 0.0f, 0.0f};
  for (int k = 0; k < s.dim; ++k) {
    if (I[k] + 1 < fine.n[k]) {
      float pos[3];
      for (int a = 0; a < s.dim; ++a)
        pos[a] = (((float)I[a] + (a == k ? 0.5f : 0.0f)) + (a == d ? 0.0f : 0.5f)) * lvl;
      float dl = delta(sample(phi, pos), ls);
      ax[k] = (-scale) * dl;
      ad = ad + scale * dl;
    }
    if (I[k] - 1 >= 0) {
      float pos[3];
      for (int a = 0; a < s.dim; ++a)
        pos[a] = (((float)I[a] - (a == k ? 0.5f : 0.0f)) + (a == d ? 0.0f : 0.5f)) * lvl;
      float dl = delta(sample(phi, pos), ls);
      ad = ad + scale * dl;
    }
  }
  Adiag[t] = ad;
  for (int k = 0; k < s.dim; ++k) Ax[t * s.dim + k] = ax[k];
}

// SurfaceTension.calc_n (surface_tension.py:115-127): n as [cells, dim].
__global__ void k_calc_n(Shape s, int d, Field phi, float dx, float* __restrict__ n) {
  const long t = (long)blockIdx.x * SB + threadIdx.x;
  if (t >= s.cells()) return;
  int I[3];
  s.coords(t, I);
  float v[3] = {0.0f, 0.0f, 0.0f};
  for (int k = 0; k < s.dim; ++k) {
    float pp[3], pm[3];
    for (int a = 0; a < s.dim; ++a) {
      const float off = (a == d) ? 0.0f : 0.5f;
      pp[a] = ((float)I[a] + (a == k ? 0.5f : 0.0f)) + off;
      pm[a] = ((float)I[a] - (a == k ? 0.5f : 0.0f)) + off;
    }
    v[k] = (sample(phi, pp) - sample(phi, pm)) / dx;
  }
  float nn = v[0] * v[0];
  for (int k = 1; k < s.dim; ++k) nn = nn + v[k] * v[k];
  float norm = sqrtf(nn);
  if (norm > 0.0f)
    for (int k = 0; k < s.dim; ++k) v[k] = v[k] / norm;
  for (int k = 0; k < s.dim; ++k) n[t * s.dim + k] = v[k];
}

// SurfaceTension.calc_kappa (surface_tension.py:129-136); n component k is a
// strided view of the AoS n, sampled with the same rule.
__device__ float sample_comp(const float* n, const Shape& s, int comp, const float* pos) {
  int i0[3], i1[3];
  float w[3];
  for (int a = 0; a < s.dim; ++a) {
    const int nn = s.n[a];
    int i = (int)pos[a];
    i = i < 0 ? 0 : (i > nn - 1 ? nn - 1 : i);
    int ip = i + 1;
    ip = ip > nn - 1 ? nn - 1 : ip;
    float ww = pos[a] - (float)i;
    ww = ww < 0.0f ? 0.0f : (ww > 1.0f ? 1.0f : ww);
    i0[a] = i;
    i1[a] = ip;
    w[a] = ww;
  }
  auto at = [&](int i, int j, int k) {
    int J[3] = {i, j, k};
    return n[s.idx(J) * s.dim + comp];
  };
  if (s.dim == 2) {
    const float u = w[0], v = w[1];
    float a = at(i0[0], i0[1], 0) * (1.0f - u) + at(i1[0], i0[1], 0) * u;
    float b = at(i0[0], i1[1], 0) * (1.0f - u) + at(i1[0], i1[1], 0) * u;
    return a * (1.0f - v) + b * v;
  }
  const float u = w[0], v = w[1], q = w[2];
  float a = (at(i0[0], i0[1], i0[2]) * (1.0f - u) + at(i1[0], i0[1], i0[2]) * u) * (1.0f - v) +
            (at(i0[0], i1[1], i0[2]) * (1.0f - u) + at(i1[0], i1[1], i0[2]) * u) * v;
  float b = (at(i0[0], i0[1], i1[2]) * (1.0f - u) + at(i1[0], i0[1], i1[2]) * u) * (1.0f - v) +
            (at(i0[0], i1[1], i1[2]) * (1.0f - u) + at(i1[0], i1[1], i1[2]) * u) * v;
  return a * (1.0f - q) + b * q;
}

__global__ void k_calc_kappa(Shape s, const float* __restrict__ n, float dx,
                             float* __restrict__ kappa) {
  const long t = (long)blockIdx.x * SB + threadIdx.x;
  if (t >= s.cells()) return;
  int I[3];
  s.coords(t, I);
  float kv = 0.0f;
  for (int k = 0; k < s.dim; ++k) {
    float pp[3], pm[3];
    for (int a = 0; a < s.dim; ++a) {
      pp[a] = (float)I[a] + (a == k ? 0.5f : 0.0f);
      pm[a] = (float)I[a] - (a == k ? 0.5f : 0.0f);
    }
    float diff = sample_comp(n, s, k, pp) - sample_comp(n, s, k, pm);
    kv = kv + diff / dx;
  }
  kappa[t] = kv;
}

// SurfaceTensionStrategy.build_b_kernel (surface_tension.py:23-66).
__global__ void k_surface_b(Shape s, int d, Field u, Field phi, const float* __restrict__ kappa,
                            const float* __restrict__ n, LS ls, float dt, float dx, float dx2,
                            float sigma, float* __restrict__ b) {
  const long t = (long)blockIdx.x * SB + threadIdx.x;
  if (t >= s.cells()) return;
  int I[3];
  s.coords(t, I);
  const int dim = s.dim;
  float bv = u.p[u.s.idx(I)] / dt;
  float pc[3];
  for (int a = 0; a < dim; ++a) pc[a] = (float)I[a] + (a == d ? 0.0f : 0.5f);
  float scale = sigma * delta(sample(phi, pc), ls);
  float grad[3];
  for (int k = 0; k < dim; ++k) {
    float pp[3], pm[3];
    for (int a = 0; a < dim; ++a) {
      pp[a] = (float)I[a] + (a == k ? 0.5f : 0.0f);
      pm[a] = (float)I[a] - (a == k ? 0.5f : 0.0f);
    }
    grad[k] = (sample(u, pp) - sample(u, pm)) / dx;
  }
  float nv[3];
  for (int a = 0; a < dim; ++a) nv[a] = n[t * dim + a];
  float Du = grad[0] * nv[0];
  for (int a = 1; a < dim; ++a) Du = Du + grad[a] * nv[a];
  float D2[3][3];
  for (int k1 = 0; k1 < dim; ++k1)
    for (int k2 = 0; k2 < dim; ++k2) {
      float p00[3], p10[3], p01[3], p11[3];
      for (int a = 0; a < dim; ++a) {
        const float e1 = (a == k1) ? 0.5f : 0.0f, e2 = (a == k2) ? 0.5f : 0.0f;
        const float base = (float)I[a];
        p00[a] = (base - e1) - e2;
        p10[a] = (base + e1) - e2;
        p01[a] = (base - e1) + e2;
        p11[a] = (base + e1) + e2;
      }
      float v00 = sample(u, p00), v10 = sample(u, p10), v01 = sample(u, p01),
            v11 = sample(u, p11);
      D2[k1][k2] = (((v11 + v00) - v10) - v01) / dx2;
    }
  float row[3];
  for (int j = 0; j < dim; ++j) {
    float acc = nv[0] * D2[0][j];
    for (int i = 1; i < dim; ++i) acc = acc + nv[i] * D2[i][j];
    row[j] = acc;
  }
  float D2n = row[0] * nv[0];
  for (int j = 1; j < dim; ++j) D2n = D2n + row[j] * nv[j];
  const float kap = kappa[t];
  float inner = D2n + kap * Du;
  float tot = kap * nv[d] + dt * inner;
  bv = bv - scale * tot;
  b[t] = bv;
}

inline unsigned nb_for(long n) { return (unsigned)((n + SB - 1) / SB); }

int check_common(int dim, const int32_t* res) {
  MG_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
  MG_REQUIRE(res != nullptr, "res is NULL");
  for (int a = 0; a < dim; ++a) MG_REQUIRE(res[a] >= 1, "res[%d] must be positive", a);
  return 0;
}

Vel make_vel(int dim, const int32_t* res, const float* const* velocity) {
  Vel V;
  memset(&V, 0, sizeof(V));
  for (int k = 0; k < dim; ++k) {
    V.u[k].p = velocity[k];
    V.u[k].s = make_shape(dim, res, 0, k);
  }
  return V;
}

HField make_hfield(int dim, const int32_t* res, const float* const* H) {
  HField F;
  memset(&F, 0, sizeof(F));
  if (!H) return F;
  for (int k = 0; k < dim; ++k) {
    F.h[k].p = H[k];
    F.h[k].s = make_shape(dim, res, 0, k);
  }
  return F;
}

// ---------------------------------------------------------------------------
// The step after the solve: FluidSimulator.apply_pressure
// (fluid_simulator.py:192-214, mode 0) and FerrofluidSimulator.
// apply_magnetic_force (ferrofluid_simulator.py:114-137, mode 1).  For every
// cell I and axis k the face velocity u_k[I] between I - e_k and I is
// updated when either side is FLUID: 0 next to SOLID (out of range counts as
// SOLID), else minus scale times the pressure jump, the AIR side taking p0
// (ghost-fluid weighted in mode 0, plus coef |H|^2 at the AIR cell's centre
// in mode 1).  One thread per cell writes its k faces; the top faces
// u_k[res_k] are not touched, as in the reference.
// ---------------------------------------------------------------------------
struct VelOut {
  float* u[3];
  Shape s[3];
};

__global__ void k_apply_pressure(Shape s, const int8_t* __restrict__ ct,
                                 const float* __restrict__ phi, const float* __restrict__ p,
                                 VelOut V, float scale, float p0, int gfm, float c_max, int mode,
                                 HField H, float coef) {
  const long t = (long)blockIdx.x * SB + threadIdx.x;
  if (t >= s.cells()) return;
  int I[3];
  s.coords(t, I);
  const int ti = ct[t];
  for (int k = 0; k < s.dim; ++k) {
    int J[3] = {I[0], I[1], I[2]};  // I_1 = I - e_k
    J[k] -= 1;
    const bool jin = s.inside(J);
    const long tj = jin ? s.idx(J) : -1;
    const int tj_type = jin ? ct[tj] : MG_SOLID;
    if (!(tj_type == MG_FLUID || ti == MG_FLUID)) continue;
    float* u = V.u[k] + V.s[k].idx(I);
    if (tj_type == MG_SOLID || ti == MG_SOLID) {
      *u = 0.0f;
    } else if (ti == MG_AIR) {  // fluid (I_1) - air (I)
      if (mode == 1) {
        float pos[3];
        for (int a = 0; a < s.dim; ++a) pos[a] = (float)I[a] + 0.5f;
        const float pa = p0 + coef * hdoth(H, s.dim, pos);
        *u = *u - scale * (pa - p[tj]);
      } else if (gfm) {
        const float c = gfm_c(phi[tj], phi[t], c_max);
        *u = *u - scale * (p0 - p[tj]) * c;
      } else {
        *u = *u - scale * (p0 - p[tj]);
      }
    } else if (tj_type == MG_AIR) {  // air (I_1) - fluid (I)
      if (mode == 1) {
        float pos[3];
        for (int a = 0; a < s.dim; ++a) pos[a] = (float)J[a] + 0.5f;
        const float pa = p0 + coef * hdoth(H, s.dim, pos);
        *u = *u - scale * (p[t] - pa);
      } else if (gfm) {
        const float c = gfm_c(phi[t], phi[tj], c_max);
        *u = *u - scale * (p[t] - p0) * c;
      } else {
        *u = *u - scale * (p[t] - p0);
      }
    } else {  // fluid - fluid
      *u = *u - scale * (p[t] - p[tj]);
    }
  }
}

VelOut make_vel_out(int dim, const int32_t* res, float* const* velocity) {
  VelOut V;
  memset(&V, 0, sizeof(V));
  for (int k = 0; k < dim; ++k) {
    V.u[k] = velocity[k];
    V.s[k] = make_shape(dim, res, 0, k);
  }
  return V;
}

}  // namespace

// ===========================================================================
// C-ABI of the strategies (mgpcg.h)
// ===========================================================================
extern "C" int mgpcg_pressure_build_b(void* stream, int dim, const int32_t* res,
                                      const int8_t* cell_type, const float* const* velocity,
                                      const float* phi, double scale_A, double scale_b,
                                      double p0, int ghost_fluid_method, double c_max,
                                      float* b) {
  if (check_common(dim, res)) return -1;
  MG_REQUIRE(cell_type && velocity && b, "NULL argument");
  for (int k = 0; k < dim; ++k) MG_REQUIRE(velocity[k], "velocity[%d] is NULL", k);
  MG_REQUIRE(!ghost_fluid_method || phi, "ghost-fluid method needs phi");
  Shape s = make_shape(dim, res);
  if (dim == 3)
    hipLaunchKernelGGL((k_pressure_b0<3>), dim3(nb_for(s.cells())), dim3(SB), 0,
                       (hipStream_t)stream, s, cell_type, make_vel(dim, res, velocity), phi,
                       (float)scale_A, (float)scale_b, (float)p0, ghost_fluid_method,
                       (float)c_max, b);
  else
    hipLaunchKernelGGL((k_pressure_b0<2>), dim3(nb_for(s.cells())), dim3(SB), 0,
                       (hipStream_t)stream, s, cell_type, make_vel(dim, res, velocity), phi,
                       (float)scale_A, (float)scale_b, (float)p0, ghost_fluid_method,
                       (float)c_max, b);
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_magnetic_build_b(void* stream, int dim, const int32_t* res,
                                      const int8_t* cell_type, const float* const* velocity,
                                      const float* const* H, double scale_A, double scale_b,
                                      double p0, double mu0, double k, float* b) {
  if (check_common(dim, res)) return -1;
  MG_REQUIRE(cell_type && velocity && H && b, "NULL argument");
  for (int a = 0; a < dim; ++a) MG_REQUIRE(velocity[a] && H[a], "NULL component %d", a);
  Shape s = make_shape(dim, res);
  const float coef = (float)(1.0 / 2.0 * k * mu0);
  hipLaunchKernelGGL(k_pressure_b, dim3(nb_for(s.cells())), dim3(SB), 0, (hipStream_t)stream, s,
                     cell_type, make_vel(dim, res, velocity), (const float*)nullptr,
                     (float)scale_A, (float)scale_b, (float)p0, 0, 0.0f, 1,
                     make_hfield(dim, res, H), coef, b);
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_pressure_build_A(void* stream, int dim, const int32_t* res, int level,
                                      const int8_t* grid_type, const float* phi, double scale_A,
                                      int ghost_fluid_method, double c_max, float* Adiag,
                                      float* Ax) {
  if (check_common(dim, res)) return -1;
  MG_REQUIRE(grid_type && Adiag && Ax, "NULL argument");
  MG_REQUIRE(level >= 0, "level must be >= 0");
  MG_REQUIRE(!(ghost_fluid_method && level == 0) || phi, "ghost-fluid method needs phi");
  Shape s = make_shape(dim, res, level);
  if (dim == 3)
    hipLaunchKernelGGL((k_pressure_A<3>), dim3(nb_for(s.cells())), dim3(SB), 0,
                       (hipStream_t)stream, s, level, grid_type, phi, (float)scale_A,
                       ghost_fluid_method, (float)c_max, Adiag, Ax);
  else
    hipLaunchKernelGGL((k_pressure_A<2>), dim3(nb_for(s.cells())), dim3(SB), 0,
                       (hipStream_t)stream, s, level, grid_type, phi, (float)scale_A,
                       ghost_fluid_method, (float)c_max, Adiag, Ax);
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_calc_volume(void* stream, int dim, const int32_t* res, const float* phi,
                                 double dx, double* partials, int partials_cap,
                                 double* out_device) {
  if (check_common(dim, res)) return -1;
  MG_REQUIRE(phi && partials && out_device, "NULL argument");
  Shape s = make_shape(dim, res);
  const unsigned nb = nb_for(s.cells());
  MG_REQUIRE((long)partials_cap >= (long)nb, "partials buffer needs %u entries", nb);
  Field f{phi, s};
  hipLaunchKernelGGL(k_volume, dim3(nb), dim3(SB), 0, (hipStream_t)stream, s, f, make_ls(dx),
                     partials);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(SB), 0, (hipStream_t)stream, partials,
                     (int)nb, out_device);
  MG_HIP(hipGetLastError());
  return 0;
}

// The allocation-free variant of mgpcg_calc_volume keeps its per-block
// partial sums in this file-scope array (one per 256 cells, so grids of up to
// VOLUME_SCRATCH_PARTIALS * 256 cells).  Each GPU holds its own copy until the
// process ends.  Host code reaches it only through the device address that
// hipGetSymbolAddress returns, once per device.  Every entry a call reads was
// written by the same call, so nothing needs clearing; calls that overlap on
// one GPU share the array and race.
constexpr int VOLUME_SCRATCH_PARTIALS = 1 << 18;
static __device__ __align__(256) double scratch[VOLUME_SCRATCH_PARTIALS];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// The current device's address of `scratch`, looked up on the device's first call.
int scratch_ptr(double** out) {
  int dev = 0;
  MG_HIP(hipGetDevice(&dev));
  MG_REQUIRE(dev >= 0 && dev < SCRATCH_MAX_DEVICES, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    MG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MG_REQUIRE(p != nullptr, "no device address for the scratch array");
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

extern "C" int mgpcg_calc_volume_static(void* stream, int dim, const int32_t* res,
                                        const float* phi, double dx, double* out_device) {
  if (check_common(dim, res)) return -1;
  MG_REQUIRE(phi && out_device, "NULL argument");
  Shape s = make_shape(dim, res);
  const unsigned nb = nb_for(s.cells());
  MG_REQUIRE((long)nb <= (long)VOLUME_SCRATCH_PARTIALS,
             "grid of %ld cells exceeds the allocation-free limit of %ld cells", (long)s.cells(),
             (long)VOLUME_SCRATCH_PARTIALS * SB);
  double* partials = nullptr;
  if (scratch_ptr(&partials)) return -2;
  Field f{phi, s};
  hipLaunchKernelGGL(k_volume, dim3(nb), dim3(SB), 0, (hipStream_t)stream, s, f, make_ls(dx),
                     partials);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(SB), 0, (hipStream_t)stream, partials,
                     (int)nb, out_device);
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_add_c(void* stream, int dim, const int32_t* res, const int8_t* cell_type,
                           double c, float* b) {
  if (check_common(dim, res)) return -1;
  MG_REQUIRE(cell_type && b, "NULL argument");
  Shape s = make_shape(dim, res);
  hipLaunchKernelGGL(k_add_c, dim3(nb_for(s.cells())), dim3(SB), 0, (hipStream_t)stream,
                     s.cells(), cell_type, (float)c, b);
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_potential_build_b(void* stream, int dim, const int32_t* res,
                                       const float* chi, const float* const* H_ext, float* b) {
  if (check_common(dim, res)) return -1;
  MG_REQUIRE(chi && H_ext && b, "NULL argument");
  for (int a = 0; a < dim; ++a) MG_REQUIRE(H_ext[a], "H_ext[%d] is NULL", a);
  Shape s = make_shape(dim, res);
  hipLaunchKernelGGL(k_potential_b, dim3(nb_for(s.cells())), dim3(SB), 0, (hipStream_t)stream,
                     s, Field{chi, s}, make_hfield(dim, res, H_ext), b);
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_potential_build_A(void* stream, int dim, const int32_t* res, int level,
                                       double dx, const float* chi, float* Adiag, float* Ax) {
  if (check_common(dim, res)) return -1;
  MG_REQUIRE(chi && Adiag && Ax, "NULL argument");
  Shape s = make_shape(dim, res, level), fine = make_shape(dim, res);
  const float dxl = (float)(dx * (double)(1 << level));
  hipLaunchKernelGGL(k_potential_A, dim3(nb_for(s.cells())), dim3(SB), 0, (hipStream_t)stream,
                     s, fine, level, dxl, Field{chi, fine}, Adiag, Ax);
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_surface_build_A(void* stream, int dim, const int32_t* res, int level, int d,
                                     const float* phi, double dt, double dx, double sigma,
                                     float* Adiag, float* Ax) {
  if (check_common(dim, res)) return -1;
  MG_REQUIRE(phi && Adiag && Ax, "NULL argument");
  MG_REQUIRE(d >= 0 && d < dim, "axis d out of range");
  Shape s = make_shape(dim, res, level), fine = make_shape(dim, res);
  const float dxl = (float)(dx * (double)(1 << level));
  const float scale = (float)(sigma * dt) / (dxl * dxl);
  const float inv_dt = (float)(1.0 / dt);
  hipLaunchKernelGGL(k_surface_A, dim3(nb_for(s.cells())), dim3(SB), 0, (hipStream_t)stream, s,
                     fine, level, d, Field{phi, fine}, make_ls(dx), inv_dt, scale, Adiag, Ax);
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_calc_n(void* stream, int dim, const int32_t* res, int d, const float* phi,
                            double dx, float* n) {
  if (check_common(dim, res)) return -1;
  MG_REQUIRE(phi && n, "NULL argument");
  Shape s = make_shape(dim, res);
  hipLaunchKernelGGL(k_calc_n, dim3(nb_for(s.cells())), dim3(SB), 0, (hipStream_t)stream, s, d,
                     Field{phi, s}, (float)dx, n);
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_calc_kappa(void* stream, int dim, const int32_t* res, const float* n,
                                double dx, float* kappa) {
  if (check_common(dim, res)) return -1;
  MG_REQUIRE(n && kappa, "NULL argument");
  Shape s = make_shape(dim, res);
  hipLaunchKernelGGL(k_calc_kappa, dim3(nb_for(s.cells())), dim3(SB), 0, (hipStream_t)stream, s,
                     n, (float)dx, kappa);
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_surface_build_b(void* stream, int dim, const int32_t* res, int d,
                                     const float* velocity_d, const float* phi,
                                     const float* kappa, const float* n, double dt, double dx,
                                     double sigma, float* b) {
  if (check_common(dim, res)) return -1;
  MG_REQUIRE(velocity_d && phi && kappa && n && b, "NULL argument");
  MG_REQUIRE(d >= 0 && d < dim, "axis d out of range");
  Shape s = make_shape(dim, res);
  Field u{velocity_d, make_shape(dim, res, 0, d)};
  hipLaunchKernelGGL(k_surface_b, dim3(nb_for(s.cells())), dim3(SB), 0, (hipStream_t)stream, s,
                     d, u, Field{phi, s}, kappa, n, make_ls(dx), (float)dt, (float)dx,
                     (float)(dx * dx), (float)sigma, b);
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_apply_pressure(void* stream, int dim, const int32_t* res,
                                    const int8_t* cell_type, const float* phi,
                                    const float* pressure, float* const* velocity, double dt,
                                    double rho, double dx, double p0, int ghost_fluid_method,
                                    double c_max) {
  if (check_common(dim, res)) return -1;
  MG_REQUIRE(cell_type && pressure && velocity, "NULL argument");
  for (int k = 0; k < dim; ++k) MG_REQUIRE(velocity[k], "velocity[%d] is NULL", k);
  MG_REQUIRE(!ghost_fluid_method || phi, "ghost-fluid method needs phi");
  MG_REQUIRE(rho > 0.0 && dx > 0.0, "rho and dx must be positive");
  Shape s = make_shape(dim, res);
  HField none;
  memset(&none, 0, sizeof(none));
  // scale = dt / (rho * dx) with dt an f32 kernel argument (fluid_simulator.py:194)
  const float scale = (float)dt / (float)(rho * dx);
  hipLaunchKernelGGL(k_apply_pressure, dim3(nb_for(s.cells())), dim3(SB), 0, (hipStream_t)stream,
                     s, cell_type, phi, pressure, make_vel_out(dim, res, velocity), scale,
                     (float)p0, ghost_fluid_method, (float)c_max, 0, none, 0.0f);
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_apply_magnetic_force(void* stream, int dim, const int32_t* res,
                                          const int8_t* cell_type, const float* pressure,
                                          const float* const* H, float* const* velocity,
                                          double dt, double rho, double dx, double p0,
                                          double mu0, double k) {
  if (check_common(dim, res)) return -1;
  MG_REQUIRE(cell_type && pressure && H && velocity, "NULL argument");
  for (int a = 0; a < dim; ++a) MG_REQUIRE(velocity[a] && H[a], "NULL component %d", a);
  MG_REQUIRE(rho > 0.0 && dx > 0.0, "rho and dx must be positive");
  Shape s = make_shape(dim, res);
  const float scale = (float)dt / (float)(rho * dx);
  const float coef = (float)(1.0 / 2.0 * k * mu0);
  hipLaunchKernelGGL(k_apply_pressure, dim3(nb_for(s.cells())), dim3(SB), 0, (hipStream_t)stream,
                     s, cell_type, (const float*)nullptr, pressure,
                     make_vel_out(dim, res, velocity), scale, (float)p0, 0, 0.0f, 1,
                     make_hfield(dim, res, H), coef);
  MG_HIP(hipGetLastError());
  return 0;
}
