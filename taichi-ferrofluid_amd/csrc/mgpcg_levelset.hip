// The step before the solve: cell classification and level-set redistancing
// on the device, canonical C-order layout.
//
//   apply_markers   FluidSimulator.apply_markers (fluid_simulator.py:151-159)
//   init_boundary   FluidSimulator.init_boundary (fluid_simulator.py:337-341)
//   redistance      FastSweepingLevelSet.redistance (level_set.py:260-333):
//                   target_surface (level_set.py:66-86), then repeat_times x
//                   six directional sweeps of the Eikonal update
//                   update_from_neighbor (level_set.py:88-112).
//
// Sweeps: one thread owns one grid line along the sweep axis and walks it in
// the sweep direction, so the cells behind it on the line are already updated
// and the ones ahead are not, as in the reference's serial inner loop.  The
// reference runs the lines in parallel and reads the neighbouring lines while
// they change (the result depends on thread timing); here the neighbouring
// lines are read from a copy taken before the sweep, which makes the result
// deterministic.  Each axis is swept over its own extent (the reference bounds
// the axis-0 sweeps by res[1], level_set.py:297-303/323-327, which only
// differs on non-square grids).
#include <hip/hip_runtime.h>
#include <math.h>

#include "../../include/mgpcg.h"
#include "mgpcg_common.h"

namespace {

constexpr int LB = 256;

struct Grid {
  int dim;
  int n[3];
  long st[3];          // C-order strides
  mg::FastDiv dn[3];   // division by n[a] (coords below 2^31)
  __host__ __device__ long cells() const { return (long)n[0] * n[1] * n[2]; }
};

Grid make_grid(int dim, const int32_t* res) {
  Grid g;
  g.dim = dim;
  for (int a = 0; a < 3; ++a) g.n[a] = a < dim ? res[a] : 1;
  g.st[dim - 1] = 1;
  for (int a = dim - 2; a >= 0; --a) g.st[a] = g.st[a + 1] * g.n[a + 1];
  for (int a = dim; a < 3; ++a) g.st[a] = 0;
  for (int a = 0; a < 3; ++a) g.dn[a] = mg::FastDiv((uint32_t)g.n[a]);
  return g;
}

__device__ __forceinline__ void coords(const Grid& g, long t, int* I) {
  if (t < 0x80000000L) {  // multiply-high division, no 64-bit divide
    uint32_t u = (uint32_t)t;
    for (int a = g.dim - 1; a >= 0; --a) {
      const uint32_t q = g.dn[a].div(u);
      I[a] = (int)(u - q * (uint32_t)g.n[a]);
      u = q;
    }
    for (int a = g.dim; a < 3; ++a) I[a] = 0;
    return;
  }
  for (int a = g.dim - 1; a >= 0; --a) {
    I[a] = (int)(t % g.n[a]);
    t /= g.n[a];
  }
  for (int a = g.dim; a < 3; ++a) I[a] = 0;
}

// taichi_glsl's sign: -1, 0 or 1
__device__ __forceinline__ float sgnf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

__global__ void k_apply_markers(Grid g, const float* __restrict__ phi, int8_t* __restrict__ ct) {
  const long t = (long)blockIdx.x * LB + threadIdx.x;
  if (t >= g.cells()) return;
  if (ct[t] != MG_SOLID) ct[t] = phi[t] <= 0.0f ? (int8_t)MG_FLUID : (int8_t)MG_AIR;
}

__global__ void k_init_boundary(Grid g, int8_t* __restrict__ ct) {
  const long t = (long)blockIdx.x * LB + threadIdx.x;
  if (t >= g.cells()) return;
  int I[3];
  coords(g, t, I);
  for (int a = 0; a < g.dim; ++a)
    if (I[a] == 0 || I[a] + 1 == g.n[a]) ct[t] = MG_SOLID;
}

// target_surface (level_set.py:66-86): cells with a sign change to a face
// neighbour take the signed distance estimate sign(phi) * theta * dx of the
// closest crossing and are fixed; the others start at 1e20.
__global__ void k_target_surface(Grid g, const float* __restrict__ phi, float dx,
                                 float* __restrict__ tmp, int8_t* __restrict__ fixed) {
  const long t = (long)blockIdx.x * LB + threadIdx.x;
  if (t >= g.cells()) return;
  int I[3];
  coords(g, t, I);
  const float pi = phi[t];
  bool change = false;
  float est = 1e20f;
  for (int k = 0; k < g.dim; ++k) {
    for (int s = -1; s <= 1; s += 2) {
      const int j = I[k] + s;
      if (j < 0 || j >= g.n[k]) continue;
      const float pj = phi[t + s * g.st[k]];
      if (sgnf(pi) != sgnf(pj)) {
        const float th = pi / (pi - pj);
        const float est0 = sgnf(pi) * th * dx;
        est = fabsf(est0) < fabsf(est) ? est0 : est;
        change = true;
      }
    }
  }
  tmp[t] = change ? est : 1e20f;
  fixed[t] = change ? 1 : 0;
}

struct EikonalConsts {
  float dx, two_dx2, dx2, half, third;
};

// update_from_neighbor (level_set.py:88-112); along axis `ax` the values come
// from the live array (the sweeping thread's own line), across from `snap`.
__device__ float eikonal(const Grid& g, const int* I, long t, int ax, const float* live,
                         const float* snap, const EikonalConsts& c) {
  float nb[3] = {0.0f, 0.0f, 0.0f};
  for (int k = 0; k < g.dim; ++k) {
    const float* f = (k == ax) ? live : snap;
    const long o = g.st[k];
    if (I[k] == 0 || (I[k] < g.n[k] - 1 && fabsf(f[t + o]) < fabsf(f[t - o])))
      nb[k] = fabsf(f[t + o]);
    else
      nb[k] = fabsf(f[t - o]);
  }
  for (int i = 0; i < g.dim - 1; ++i)
    for (int j = 0; j < g.dim - 1 - i; ++j)
      if (nb[j] > nb[j + 1]) {
        const float x = nb[j];
        nb[j] = nb[j + 1];
        nb[j + 1] = x;
      }
  float d = nb[0] + c.dx;
  if (d > nb[1]) {
    const float df = nb[1] - nb[0];
    d = c.half * (nb[0] + nb[1] + sqrtf(c.two_dx2 - df * df));
    if (g.dim == 3 && d > nb[2]) {
      const float s = nb[0] + nb[1] + nb[2];
      const float q = nb[0] * nb[0] + nb[1] * nb[1] + nb[2] * nb[2] - c.dx2;
      d = c.third * (s + sqrtf(fmaxf(0.0f, s * s - 3.0f * q)));
    }
  }
  return d;
}

// One directional sweep: a thread per line along `ax`, dir = +1 / -1.
__global__ void k_sweep(Grid g, int ax, int dir, const float* __restrict__ phi,
                        const float* __restrict__ snap, float* __restrict__ tmp,
                        const int8_t* __restrict__ fixed, EikonalConsts c) {
  const long nlines = g.cells() / g.n[ax];
  const long line = (long)blockIdx.x * LB + threadIdx.x;
  if (line >= nlines) return;
  // the line's coordinates on the other axes (C order over them)
  int I[3] = {0, 0, 0};
  long r = line;
  for (int a = g.dim - 1; a >= 0; --a) {
    if (a == ax) continue;
    I[a] = (int)(r % g.n[a]);
    r /= g.n[a];
  }
  const int n = g.n[ax];
  for (int step = 0; step < n; ++step) {
    I[ax] = dir > 0 ? step : n - 1 - step;
    long t = 0;
    for (int a = 0; a < g.dim; ++a) t += I[a] * g.st[a];
    if (fixed[t]) continue;
    const float d = eikonal(g, I, t, ax, tmp, snap, c);
    if (fabsf(d) < fabsf(tmp[t])) tmp[t] = d * sgnf(phi[t]);
  }
}

__global__ void k_copy(long n, const float* __restrict__ src, float* __restrict__ dst) {
  const long t = (long)blockIdx.x * LB + threadIdx.x;
  if (t < n) dst[t] = src[t];
}

// ---------------------------------------------------------------------------
// FluidSimulator.extrap_velocity (fluid_simulator.py:240-270) for one MAC
// component k (shape res + e_k): mark_valid -- a face is valid when a cell on
// either side is FLUID (out of range is not) -- then `iterations` passes of
// diffuse_quantity, each averaging the valid faces of the 3^dim block around
// every invalid face (offsets in C order, FP32 sum then one division) into it
// and marking it valid; each pass reads the velocities and flags as they were
// before the pass.  Faces outside the component's own shape count as invalid
// (the reference's shared `valid` field is res + 1 in every axis and keeps
// other components' flags there, and it is read out of range at index -1).
// ---------------------------------------------------------------------------
__global__ void k_mark_valid(Grid f, Grid c, int k, const int8_t* __restrict__ ct,
                             int8_t* __restrict__ valid) {
  const long t = (long)blockIdx.x * LB + threadIdx.x;
  if (t >= f.cells()) return;
  int I[3];
  coords(f, t, I);
  auto fluid = [&](const int* J) {
    for (int a = 0; a < c.dim; ++a)
      if (J[a] < 0 || J[a] >= c.n[a]) return false;
    long u = 0;
    for (int a = 0; a < c.dim; ++a) u += J[a] * c.st[a];
    return ct[u] == MG_FLUID;
  };
  int J[3] = {I[0], I[1], I[2]};
  J[k] -= 1;
  valid[t] = (fluid(J) || fluid(I)) ? 1 : 0;
}

__global__ void k_diffuse(Grid f, float* __restrict__ dst, const float* __restrict__ src,
                          int8_t* __restrict__ valid_dst, const int8_t* __restrict__ valid) {
  const long t = (long)blockIdx.x * LB + threadIdx.x;
  if (t >= f.cells()) return;
  if (valid[t] != 0) return;
  int I[3];
  coords(f, t, I);
  float tot = 0.0f;
  int cnt = 0;
  const int nz = f.dim == 3 ? 3 : 1;
  for (int o0 = -1; o0 <= 1; ++o0)
    for (int o1 = -1; o1 <= 1; ++o1)
      for (int oz = 0; oz < nz; ++oz) {
        const int o2 = f.dim == 3 ? oz - 1 : 0;
        const int J[3] = {I[0] + o0, I[1] + o1, I[2] + o2};
        bool in = true;
        for (int a = 0; a < f.dim; ++a) in = in && J[a] >= 0 && J[a] < f.n[a];
        if (!in) continue;
        long u = 0;
        for (int a = 0; a < f.dim; ++a) u += J[a] * f.st[a];
        if (valid[u] == 1) {
          tot = tot + src[u];
          cnt += 1;
        }
      }
  if (cnt > 0) {
    dst[t] = tot / (float)cnt;
    valid_dst[t] = 1;
  }
}

__global__ void k_copy8(long n, const int8_t* __restrict__ src, int8_t* __restrict__ dst) {
  const long t = (long)blockIdx.x * LB + threadIdx.x;
  if (t < n) dst[t] = src[t];
}

inline unsigned nblk(long n) { return (unsigned)((n + LB - 1) / LB); }

int check_grid(int dim, const int32_t* res) {
  MG_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
  MG_REQUIRE(res != nullptr, "res is NULL");
  for (int a = 0; a < dim; ++a) MG_REQUIRE(res[a] >= 2, "res[%d] = %d must be >= 2", a, res[a]);
  return 0;
}

}  // namespace

extern "C" int mgpcg_apply_markers(void* stream, int dim, const int32_t* res, const float* phi,
                                   int8_t* cell_type) {
  if (check_grid(dim, res)) return -1;
  MG_REQUIRE(phi && cell_type, "NULL argument");
  Grid g = make_grid(dim, res);
  hipLaunchKernelGGL(k_apply_markers, dim3(nblk(g.cells())), dim3(LB), 0, (hipStream_t)stream, g,
                     phi, cell_type);
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_init_boundary(void* stream, int dim, const int32_t* res, int8_t* cell_type) {
  if (check_grid(dim, res)) return -1;
  MG_REQUIRE(cell_type, "NULL argument");
  Grid g = make_grid(dim, res);
  hipLaunchKernelGGL(k_init_boundary, dim3(nblk(g.cells())), dim3(LB), 0, (hipStream_t)stream,
                     g, cell_type);
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_redistance(void* stream, int dim, const int32_t* res, double dx,
                                int repeat_times, float* phi, float* phi_temp, float* snap,
                                int8_t* fixed) {
  if (check_grid(dim, res)) return -1;
  MG_REQUIRE(phi && phi_temp && snap && fixed, "NULL argument");
  MG_REQUIRE(dx > 0.0, "dx must be positive");
  MG_REQUIRE(repeat_times >= 0, "repeat_times must be >= 0");
  hipStream_t s = (hipStream_t)stream;
  Grid g = make_grid(dim, res);
  const long n = g.cells();
  // Python-scope constants of update_from_neighbor, baked as FP32
  EikonalConsts c;
  c.dx = (float)dx;
  c.two_dx2 = (float)(2.0 * (dx * dx));
  c.dx2 = (float)(dx * dx);
  c.half = (float)(1.0 / 2.0);
  c.third = (float)(1.0 / 3.0);
  hipLaunchKernelGGL(k_target_surface, dim3(nblk(n)), dim3(LB), 0, s, g, (const float*)phi,
                     (float)dx, phi_temp, fixed);
  for (int rep = 0; rep < repeat_times; ++rep) {
    for (int ax = dim - 1; ax >= 0; --ax) {
      for (int dir = 1; dir >= -1; dir -= 2) {
        hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(LB), 0, s, n, (const float*)phi_temp,
                           snap);
        hipLaunchKernelGGL(k_sweep, dim3(nblk(n / g.n[ax])), dim3(LB), 0, s, g, ax, dir,
                           (const float*)phi, (const float*)snap, phi_temp,
                           (const int8_t*)fixed, c);
      }
    }
  }
  hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(LB), 0, s, n, (const float*)phi_temp, phi);
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_extrap_velocity(void* stream, int dim, const int32_t* res,
                                     const int8_t* cell_type, float* const* velocity,
                                     int iterations, float* backup, int8_t* valid,
                                     int8_t* valid_temp) {
  if (check_grid(dim, res)) return -1;
  MG_REQUIRE(cell_type && velocity && backup && valid && valid_temp, "NULL argument");
  for (int k = 0; k < dim; ++k) MG_REQUIRE(velocity[k], "velocity[%d] is NULL", k);
  MG_REQUIRE(iterations >= 0, "iterations must be >= 0");
  hipStream_t s = (hipStream_t)stream;
  const Grid c = make_grid(dim, res);
  for (int k = 0; k < dim; ++k) {
    int32_t fr[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) fr[a] = res[a] + (a == k ? 1 : 0);
    const Grid f = make_grid(dim, fr);
    const long n = f.cells();
    hipLaunchKernelGGL(k_mark_valid, dim3(nblk(n)), dim3(LB), 0, s, f, c, k, cell_type, valid);
    for (int it = 0; it < iterations; ++it) {
      hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(LB), 0, s, n, (const float*)velocity[k],
                         backup);
      hipLaunchKernelGGL(k_copy8, dim3(nblk(n)), dim3(LB), 0, s, n, (const int8_t*)valid,
                         valid_temp);
      hipLaunchKernelGGL(k_diffuse, dim3(nblk(n)), dim3(LB), 0, s, f, velocity[k],
                         (const float*)backup, valid, (const int8_t*)valid_temp);
    }
  }
  MG_HIP(hipGetLastError());
  return 0;
}

