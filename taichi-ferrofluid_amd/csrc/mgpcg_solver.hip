// Host side of the MI355X MGPCG solver and its C-ABI (include/mgpcg.h).
//
// One handle = one multigrid hierarchy resident in HBM in the colour-split
// padded layout, the level-0 CG vectors, a device-resident CG state and the
// reduction partials.  The solve loop of mgpcg.py:150-209 runs entirely on
// the device: every scalar (alpha, beta, the residual test) is computed by
// the last block of the reduction kernel that produces it, and every kernel
// of an iteration returns at once when the state says the solve is done.
// The host enqueues iterations as replays of a captured hipGraph and reads
// the state back asynchronously, one burst behind, so the GPU never idles on
// a host round trip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/mgpcg.h"
#include "mgpcg_kernels.h"

using namespace mg;

static thread_local char g_err[512] = "";

void mg_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" const char* mgpcg_last_error(void) { return g_err; }
extern "C" int mgpcg_version(void) { return MGPCG_ABI_VERSION; }

namespace {

struct LevelBuf {
  int8_t* gt[2];
  void* Adiag[2];
  void* Ax[3][2];
  void* r[2];
  void* z[2];
};

}  // namespace

struct mgpcg_solver {
  int dim = 3, L = 1, nu = 2, bottom = 10, precision = PLAIN;
  int res[3] = {0, 0, 0};
  size_t esz = 4, xsz = 4;  // hierarchy / x element sizes
  std::vector<Geo> geo;
  std::vector<LevelBuf> lv;
  void* b[2] = {nullptr, nullptr};
  void* x[2] = {nullptr, nullptr};
  void* p[2] = {nullptr, nullptr};
  void* Ap[2] = {nullptr, nullptr};
  CGState* st = nullptr;
  double* partials = nullptr;
  int nparts = 0;
  char* pool = nullptr;
  size_t pool_bytes = 0;
  CGState* host_st = nullptr;  // pinned, 2 slots
  hipStream_t stream = nullptr;  // own stream (capturable)
  hipStream_t user_stream = nullptr;
  hipEvent_t ev_user = nullptr, ev_done = nullptr, ev_slot[2] = {nullptr, nullptr};
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  int use_graph = 1;
  int burst = 4;
  int device = 0;
};

namespace {

inline int nblk(int n) { return (n + BLOCK - 1) / BLOCK; }

template <typename T>
LevelPtrs<T> level_ptrs(const mgpcg_solver* s, int l) {
  LevelPtrs<T> P;
  const LevelBuf& B = s->lv[l];
  for (int c = 0; c < 2; ++c) {
    P.gt[c] = B.gt[c];
    P.Adiag[c] = (const T*)B.Adiag[c];
    P.r[c] = (T*)B.r[c];
    P.z[c] = (T*)B.z[c];
    for (int a = 0; a < 3; ++a) P.Ax[a][c] = (const T*)B.Ax[a][c];
  }
  return P;
}

template <typename T, typename TX>
CGPtrs<T, TX> cg_ptrs(const mgpcg_solver* s) {
  CGPtrs<T, TX> V;
  for (int c = 0; c < 2; ++c) {
    V.b[c] = (T*)s->b[c];
    V.x[c] = (TX*)s->x[c];
    V.p[c] = (T*)s->p[c];
    V.Ap[c] = (T*)s->Ap[c];
  }
  return V;
}

// ---------------------------------------------------------------------------
// The solve schedule, templated on hierarchy type T, x type TX and dimension.
// ---------------------------------------------------------------------------
template <typename T, typename TX, int DIM>
struct Impl {
  static void smooth(mgpcg_solver* s, int l, int phase, bool first, const int* done) {
    const Geo& g = s->geo[l];
    LevelPtrs<T> P = level_ptrs<T>(s, l);
    if (first)
      hipLaunchKernelGGL((k_smooth<T, DIM, true>), dim3(nblk(g.size)), dim3(BLOCK), 0, s->stream,
                         g, P, phase, done);
    else
      hipLaunchKernelGGL((k_smooth<T, DIM, false>), dim3(nblk(g.size)), dim3(BLOCK), 0,
                         s->stream, g, P, phase, done);
  }
  static void zero_z(mgpcg_solver* s, int l, const int* done) {
    const Geo& g = s->geo[l];
    hipLaunchKernelGGL((k_zero2<T>), dim3(nblk(g.size)), dim3(BLOCK), 0, s->stream, g.size,
                       (T*)s->lv[l].z[0], (T*)s->lv[l].z[1], done);
  }
  static void restrict_(mgpcg_solver* s, int l, const int* done) {
    const Geo& gf = s->geo[l];
    const Geo& gc = s->geo[l + 1];
    LevelPtrs<T> P = level_ptrs<T>(s, l);
    hipLaunchKernelGGL((k_restrict<T, DIM>), dim3(2 * nblk(gc.size)), dim3(BLOCK), 0, s->stream,
                       gf, gc, P, (T*)s->lv[l + 1].r[0], (T*)s->lv[l + 1].r[1], done);
  }
  static void prolongate(mgpcg_solver* s, int l, const int* done) {
    const Geo& gf = s->geo[l];
    const Geo& gc = s->geo[l + 1];
    hipLaunchKernelGGL((k_prolongate<T, DIM>), dim3(2 * nblk(gf.size)), dim3(BLOCK), 0,
                       s->stream, gf, gc, (T*)s->lv[l].z[0], (T*)s->lv[l].z[1],
                       (const T*)s->lv[l + 1].z[0], (const T*)s->lv[l + 1].z[1], done);
  }
  // The smoothing block of one level: `pairs` x (smooth(a), smooth(b)).
  // `from_zero`: z of the level is known to be zero before the block.
  static void smooth_block(mgpcg_solver* s, int l, int pairs, int a, int b, bool from_zero,
                           const int* done) {
    if (pairs <= 0) {
      if (from_zero) zero_z(s, l, done);
      return;
    }
    for (int i = 0; i < pairs; ++i) {
      smooth(s, l, a, from_zero && i == 0 && a == 0, done);
      smooth(s, l, b, false, done);
    }
  }
  // v_cycle (mgpcg.py:125-148)
  static void vcycle(mgpcg_solver* s, const int* done) {
    const int L = s->L;
    for (int l = 0; l < L - 1; ++l) {
      smooth_block(s, l, s->nu, 0, 1, true, done);
      restrict_(s, l, done);
    }
    const int half = s->bottom / 2;
    if (half > 0) {
      smooth_block(s, L - 1, half, 0, 1, true, done);
      smooth_block(s, L - 1, half, 1, 0, false, done);
    } else {
      zero_z(s, L - 1, done);
    }
    for (int l = L - 2; l >= 0; --l) {
      prolongate(s, l, done);
      smooth_block(s, l, s->nu, 1, 0, false, done);
    }
  }
  static RedPtrs red(mgpcg_solver* s) { return RedPtrs{s->st, s->partials}; }

  static void pre_loop(mgpcg_solver* s) {
    const Geo& g = s->geo[0];
    const int nb = 2 * nblk(g.size);
    LevelPtrs<T> P = level_ptrs<T>(s, 0);
    CGPtrs<T, TX> V = cg_ptrs<T, TX>(s);
    hipLaunchKernelGGL((k_cg_init<T, TX>), dim3(nb), dim3(BLOCK), 0, s->stream, g, P, V, red(s));
    vcycle(s, nullptr);
    hipLaunchKernelGGL((k_update_p<T, TX>), dim3(nb), dim3(BLOCK), 0, s->stream, g, P, V,
                       (const CGState*)s->st);
    hipLaunchKernelGGL((k_dot_zr<T, true>), dim3(nb), dim3(BLOCK), 0, s->stream, g, P, red(s));
  }
  // One loop body of mgpcg.py:176-209.
  static void iteration(mgpcg_solver* s) {
    const Geo& g = s->geo[0];
    const int nb = 2 * nblk(g.size);
    LevelPtrs<T> P = level_ptrs<T>(s, 0);
    CGPtrs<T, TX> V = cg_ptrs<T, TX>(s);
    const int* done = &s->st->done;
    hipLaunchKernelGGL((k_compute_Ap<T, TX, DIM>), dim3(nb), dim3(BLOCK), 0, s->stream, g, P, V,
                       red(s));
    hipLaunchKernelGGL((k_update_xr<T, TX>), dim3(nb), dim3(BLOCK), 0, s->stream, g, P, V,
                       red(s));
    if (s->precision == MIXED)
      hipLaunchKernelGGL((k_true_residual<T, TX, DIM>), dim3(nb), dim3(BLOCK), 0, s->stream, g,
                         P, V, red(s));
    vcycle(s, done);
    hipLaunchKernelGGL((k_dot_zr<T, false>), dim3(nb), dim3(BLOCK), 0, s->stream, g, P, red(s));
    hipLaunchKernelGGL((k_update_p<T, TX>), dim3(nb), dim3(BLOCK), 0, s->stream, g, P, V,
                       (const CGState*)s->st);
  }
  // The reference's host loop (3 device->host syncs per iteration); the
  // baseline the device-resident loop is measured against.
  static int solve_hostsync(mgpcg_solver* s, int max_iters) {
    const Geo& g = s->geo[0];
    const int nb = 2 * nblk(g.size);
    LevelPtrs<T> P = level_ptrs<T>(s, 0);
    CGPtrs<T, TX> V = cg_ptrs<T, TX>(s);
    CGState* h = s->host_st;
    auto sync_state = [&]() -> int {
      MG_HIP(hipMemcpyAsync(h, s->st, sizeof(CGState), hipMemcpyDeviceToHost, s->stream));
      MG_HIP(hipStreamSynchronize(s->stream));
      return 0;
    };
    hipLaunchKernelGGL((k_cg_init<T, TX>), dim3(nb), dim3(BLOCK), 0, s->stream, g, P, V, red(s));
    if (sync_state()) return -2;
    vcycle(s, nullptr);
    hipLaunchKernelGGL((k_update_p<T, TX>), dim3(nb), dim3(BLOCK), 0, s->stream, g, P, V,
                       (const CGState*)s->st);
    hipLaunchKernelGGL((k_dot_zr<T, true>), dim3(nb), dim3(BLOCK), 0, s->stream, g, P, red(s));
    if (sync_state()) return -2;
    for (int it = 0; max_iters < 0 || it < max_iters + 1; ++it) {
      hipLaunchKernelGGL((k_compute_Ap<T, TX, DIM>), dim3(nb), dim3(BLOCK), 0, s->stream, g, P,
                         V, red(s));
      if (sync_state()) return -2;
      hipLaunchKernelGGL((k_update_xr<T, TX>), dim3(nb), dim3(BLOCK), 0, s->stream, g, P, V,
                         red(s));
      if (s->precision == MIXED)
        hipLaunchKernelGGL((k_true_residual<T, TX, DIM>), dim3(nb), dim3(BLOCK), 0, s->stream,
                           g, P, V, red(s));
      if (sync_state()) return -2;
      if (h->done) break;
      vcycle(s, nullptr);
      hipLaunchKernelGGL((k_dot_zr<T, false>), dim3(nb), dim3(BLOCK), 0, s->stream, g, P,
                         red(s));
      if (sync_state()) return -2;
      if (h->done) break;
      hipLaunchKernelGGL((k_update_p<T, TX>), dim3(nb), dim3(BLOCK), 0, s->stream, g, P, V,
                         (const CGState*)s->st);
    }
    return 0;
  }
};

template <template <typename, typename, int> class F, typename... Args>
auto dispatch(mgpcg_solver* s, Args&&... args) {
  if (s->precision == F64) {
    if (s->dim == 3) return F<double, double, 3>::run(s, args...);
    return F<double, double, 2>::run(s, args...);
  }
  if (s->precision == MIXED) {
    if (s->dim == 3) return F<float, double, 3>::run(s, args...);
    return F<float, double, 2>::run(s, args...);
  }
  if (s->dim == 3) return F<float, float, 3>::run(s, args...);
  return F<float, float, 2>::run(s, args...);
}

template <typename T, typename TX, int DIM>
struct RunPre {
  static int run(mgpcg_solver* s) {
    Impl<T, TX, DIM>::pre_loop(s);
    return 0;
  }
};
template <typename T, typename TX, int DIM>
struct RunIter {
  static int run(mgpcg_solver* s) {
    Impl<T, TX, DIM>::iteration(s);
    return 0;
  }
};
template <typename T, typename TX, int DIM>
struct RunHostSync {
  static int run(mgpcg_solver* s, int max_iters) {
    return Impl<T, TX, DIM>::solve_hostsync(s, max_iters);
  }
};
template <typename T, typename TX, int DIM>
struct RunVcycle {
  static int run(mgpcg_solver* s) {
    Impl<T, TX, DIM>::vcycle(s, nullptr);
    return 0;
  }
};
template <typename T, typename TX, int DIM>
struct RunOp {
  static int run(mgpcg_solver* s, int op, int level, int arg) {
    using I = Impl<T, TX, DIM>;
    switch (op) {
      case MGPCG_OP_SMOOTH:
        I::smooth(s, level, arg & 1, false, nullptr);
        return 0;
      case MGPCG_OP_SMOOTH_FROM_ZERO:
        I::smooth(s, level, 0, true, nullptr);
        return 0;
      case MGPCG_OP_RESTRICT:
        I::restrict_(s, level, nullptr);
        return 0;
      case MGPCG_OP_PROLONGATE:
        I::prolongate(s, level, nullptr);
        return 0;
      case MGPCG_OP_VCYCLE:
        I::vcycle(s, nullptr);
        return 0;
      case MGPCG_OP_COMPUTE_AP: {
        // Ap = A p on level 0, via the CG kernel; the reduction result lands in the state.
        const Geo& g = s->geo[0];
        const int nb = 2 * nblk(g.size);
        hipLaunchKernelGGL((k_compute_Ap<T, TX, DIM>), dim3(nb), dim3(BLOCK), 0, s->stream, g,
                           level_ptrs<T>(s, 0), cg_ptrs<T, TX>(s), I::red(s));
        return 0;
      }
      default:
        mg_set_error("unknown op %d", op);
        return -1;
    }
  }
};

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

int capture_iteration(mgpcg_solver* s) {
  if (s->graph_exec) return 0;
  MG_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  dispatch<RunIter>(s);
  hipGraph_t graph = nullptr;
  MG_HIP(hipStreamEndCapture(s->stream, &graph));
  MG_HIP(hipGraphInstantiate(&s->graph_exec, graph, nullptr, nullptr, 0));
  s->graph = graph;
  return 0;
}

int enqueue_iteration(mgpcg_solver* s) {
  if (s->use_graph) {
    MG_HIP(hipGraphLaunch(s->graph_exec, s->stream));
  } else {
    dispatch<RunIter>(s);
  }
  return 0;
}

// The handle's own stream is non-blocking, so it is always ordered against
// the caller's stream explicitly -- also when that is the legacy NULL stream.
int begin(mgpcg_solver* s) {
  MG_HIP(hipSetDevice(s->device));
  MG_HIP(hipEventRecord(s->ev_user, s->user_stream));
  MG_HIP(hipStreamWaitEvent(s->stream, s->ev_user, 0));
  return 0;
}

int end(mgpcg_solver* s) {
  MG_HIP(hipEventRecord(s->ev_done, s->stream));
  MG_HIP(hipStreamWaitEvent(s->user_stream, s->ev_done, 0));
  return 0;
}

int set_params(mgpcg_solver* s, int max_iters, double rel_tol, double abs_tol, double eps,
               int refine_every) {
  CGState init;
  memset(&init, 0, sizeof(init));
  init.rel_tol = rel_tol;
  init.abs_tol = abs_tol;
  init.eps = eps;
  init.max_iters = max_iters;
  init.refine_every = refine_every;
  init.precision = s->precision;
  // stage through pinned slot 1 (slot 0 may be read by the caller)
  CGState* h = &s->host_st[1];
  MG_HIP(hipStreamSynchronize(s->stream));
  *h = init;
  MG_HIP(hipMemcpyAsync(s->st, h, sizeof(CGState), hipMemcpyHostToDevice, s->stream));
  MG_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

void fill_result(const CGState& h, mgpcg_result_t* out) {
  if (!out) return;
  out->iters = h.iter;
  out->status = h.status;
  out->r0r0 = h.r0r0;
  out->rr = h.rr;
  out->true_rr = h.true_rr;
  out->tol = h.tol;
  out->alpha = h.alpha;
  out->beta = h.beta;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" int mgpcg_create(int dim, const int32_t* res, int n_levels, int nu, int bottom,
                            int precision, mgpcg_solver_t** out) {
  MG_REQUIRE(out != nullptr, "out is NULL");
  *out = nullptr;
  MG_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3 (got %d)", dim);
  MG_REQUIRE(res != nullptr, "res is NULL");
  MG_REQUIRE(n_levels >= 1 && n_levels <= 16, "n_levels must be in [1, 16] (got %d)", n_levels);
  MG_REQUIRE(nu >= 0 && bottom >= 0, "smoothing counts must be >= 0");
  MG_REQUIRE(precision >= 0 && precision <= 2, "precision must be 0, 1 or 2");
  for (int a = 0; a < dim; ++a) {
    MG_REQUIRE(res[a] >= 1, "res[%d] = %d must be positive", a, res[a]);
    MG_REQUIRE(res[a] % (1 << (n_levels - 1)) == 0,
               "res[%d] = %d is not divisible by 2^(n_levels-1) = %d", a, res[a],
               1 << (n_levels - 1));
  }
  long long cells = 1;
  for (int a = 0; a < dim; ++a) cells *= res[a];
  MG_REQUIRE(cells < (1ll << 29), "grid too large for 32-bit slot indexing");

  mgpcg_solver* s = new (std::nothrow) mgpcg_solver();
  MG_REQUIRE(s != nullptr, "out of host memory");
  s->dim = dim;
  s->L = n_levels;
  s->nu = nu;
  s->bottom = bottom;
  s->precision = precision;
  for (int a = 0; a < dim; ++a) s->res[a] = res[a];
  s->esz = precision == F64 ? 8 : 4;
  s->xsz = precision == PLAIN ? 4 : 8;
  hipError_t e = hipGetDevice(&s->device);
  if (e != hipSuccess) {
    mg_set_error("hipGetDevice failed: %s", hipGetErrorString(e));
    delete s;
    return -2;
  }

  // geometry and pool layout
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  struct Offs {
    size_t gt[2], Ad[2], Ax[3][2], r[2], z[2];
  };
  std::vector<Offs> lo;
  for (int l = 0; l < n_levels; ++l) {
    int n[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) n[a] = res[a] >> l;
    s->geo.push_back(make_geo(dim, n));
    const Geo& g = s->geo.back();
    Offs o;
    for (int c = 0; c < 2; ++c) {
      o.gt[c] = take(g.size);
      o.Ad[c] = take(g.size * s->esz);
      for (int a = 0; a < 3; ++a) o.Ax[a][c] = (a < dim) ? take(g.size * s->esz) : (size_t)-1;
      o.r[c] = take(g.size * s->esz);
      o.z[c] = take(g.size * s->esz);
    }
    lo.push_back(o);
  }
  const Geo& g0 = s->geo[0];
  size_t ob[2], ox[2], op[2], oap[2];
  for (int c = 0; c < 2; ++c) {
    ob[c] = take(g0.size * s->esz);
    ox[c] = take(g0.size * s->xsz);
    op[c] = take(g0.size * s->esz);
    oap[c] = take(g0.size * s->esz);
  }
  size_t ost = take(sizeof(CGState));
  s->nparts = 2 * nblk(g0.size);
  size_t opart = take(sizeof(double) * s->nparts);
  s->pool_bytes = off;

  auto fail = [&](hipError_t err, const char* what) {
    mg_set_error("%s failed: %s", what, hipGetErrorString(err));
    mgpcg_destroy(s);
    return -2;
  };
  if ((e = hipMalloc(&s->pool, s->pool_bytes)) != hipSuccess) return fail(e, "hipMalloc(pool)");
  if ((e = hipMemset(s->pool, 0, s->pool_bytes)) != hipSuccess) return fail(e, "hipMemset");
  for (int l = 0; l < n_levels; ++l) {
    LevelBuf B;
    for (int c = 0; c < 2; ++c) {
      B.gt[c] = (int8_t*)(s->pool + lo[l].gt[c]);
      B.Adiag[c] = s->pool + lo[l].Ad[c];
      for (int a = 0; a < 3; ++a)
        B.Ax[a][c] = (a < dim) ? (void*)(s->pool + lo[l].Ax[a][c]) : nullptr;
      B.r[c] = s->pool + lo[l].r[c];
      B.z[c] = s->pool + lo[l].z[c];
    }
    s->lv.push_back(B);
  }
  for (int c = 0; c < 2; ++c) {
    s->b[c] = s->pool + ob[c];
    s->x[c] = s->pool + ox[c];
    s->p[c] = s->pool + op[c];
    s->Ap[c] = s->pool + oap[c];
  }
  s->st = (CGState*)(s->pool + ost);
  s->partials = (double*)(s->pool + opart);
  // ghost-layer invariants: gt of ghosts must be SOLID (memset gave FLUID=0)
  for (int l = 0; l < n_levels; ++l) {
    const Geo& g = s->geo[l];
    if ((e = hipMemset(s->lv[l].gt[0], MG_SOLID, g.size)) != hipSuccess) return fail(e, "memset");
    if ((e = hipMemset(s->lv[l].gt[1], MG_SOLID, g.size)) != hipSuccess) return fail(e, "memset");
  }
  if ((e = hipHostMalloc((void**)&s->host_st, 2 * sizeof(CGState), 0)) != hipSuccess)
    return fail(e, "hipHostMalloc");
  memset(s->host_st, 0, 2 * sizeof(CGState));
  if ((e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) != hipSuccess)
    return fail(e, "hipStreamCreate");
  if ((e = hipEventCreateWithFlags(&s->ev_user, hipEventDisableTiming)) != hipSuccess)
    return fail(e, "hipEventCreate");
  if ((e = hipEventCreateWithFlags(&s->ev_done, hipEventDisableTiming)) != hipSuccess)
    return fail(e, "hipEventCreate");
  for (int i = 0; i < 2; ++i)
    if ((e = hipEventCreateWithFlags(&s->ev_slot[i], hipEventDisableTiming)) != hipSuccess)
      return fail(e, "hipEventCreate");
  if ((e = hipDeviceSynchronize()) != hipSuccess) return fail(e, "hipDeviceSynchronize");
  *out = s;
  return 0;
}

extern "C" int mgpcg_destroy(mgpcg_solver_t* s) {
  if (!s) return 0;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  if (s->ev_user) (void)hipEventDestroy(s->ev_user);
  if (s->ev_done) (void)hipEventDestroy(s->ev_done);
  for (int i = 0; i < 2; ++i)
    if (s->ev_slot[i]) (void)hipEventDestroy(s->ev_slot[i]);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  if (s->host_st) (void)hipHostFree(s->host_st);
  if (s->pool) (void)hipFree(s->pool);
  delete s;
  return 0;
}

extern "C" int mgpcg_set_stream(mgpcg_solver_t* s, void* stream) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  s->user_stream = (hipStream_t)stream;
  return 0;
}

extern "C" int mgpcg_set_option(mgpcg_solver_t* s, int option, int value) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  switch (option) {
    case MGPCG_OPT_USE_GRAPH:
      s->use_graph = value != 0;
      return 0;
    case MGPCG_OPT_BURST:
      MG_REQUIRE(value >= 1 && value <= 64, "burst must be in [1, 64]");
      s->burst = value;
      return 0;
    default:
      mg_set_error("unknown option %d", option);
      return -1;
  }
}

extern "C" int mgpcg_info(const mgpcg_solver_t* s, int level, int32_t* out6, int64_t* bytes) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  MG_REQUIRE(level >= 0 && level < s->L, "level %d out of range", level);
  const Geo& g = s->geo[level];
  if (out6) {
    out6[0] = g.P0;
    out6[1] = g.P1;
    out6[2] = g.W;
    out6[3] = g.size;
    out6[4] = g.ncells;
    out6[5] = s->nparts;
  }
  if (bytes) *bytes = (int64_t)s->pool_bytes;
  return 0;
}

extern "C" int mgpcg_load_level(mgpcg_solver_t* s, int level, const int8_t* grid_type,
                                const float* Adiag, const float* Ax) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  MG_REQUIRE(level >= 0 && level < s->L, "level %d out of range [0, %d)", level, s->L);
  MG_REQUIRE(grid_type && Adiag && Ax, "NULL field pointer");
  if (begin(s)) return -2;
  const Geo& g = s->geo[level];
  const LevelBuf& B = s->lv[level];
  const int nb = 2 * nblk(g.size);
  if (s->precision == F64) {
    typedef double T;
    hipLaunchKernelGGL((k_pack_level<T>), dim3(nb), dim3(BLOCK), 0, s->stream, g, grid_type,
                       Adiag, Ax, B.gt[0], B.gt[1], (T*)B.Adiag[0], (T*)B.Adiag[1],
                       (T*)B.Ax[0][0], (T*)B.Ax[0][1], (T*)B.Ax[1][0], (T*)B.Ax[1][1],
                       (T*)B.Ax[2][0], (T*)B.Ax[2][1], (T*)B.r[0], (T*)B.r[1], (T*)B.z[0],
                       (T*)B.z[1]);
  } else {
    typedef float T;
    hipLaunchKernelGGL((k_pack_level<T>), dim3(nb), dim3(BLOCK), 0, s->stream, g, grid_type,
                       Adiag, Ax, B.gt[0], B.gt[1], (T*)B.Adiag[0], (T*)B.Adiag[1],
                       (T*)B.Ax[0][0], (T*)B.Ax[0][1], (T*)B.Ax[1][0], (T*)B.Ax[1][1],
                       (T*)B.Ax[2][0], (T*)B.Ax[2][1], (T*)B.r[0], (T*)B.r[1], (T*)B.z[0],
                       (T*)B.z[1]);
  }
  MG_HIP(hipGetLastError());
  return end(s);
}

extern "C" int mgpcg_load_rhs(mgpcg_solver_t* s, const float* b) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  MG_REQUIRE(b != nullptr, "b is NULL");
  if (begin(s)) return -2;
  const Geo& g = s->geo[0];
  const int nb = 2 * nblk(g.size);
  if (s->precision == F64)
    hipLaunchKernelGGL((k_pack_vec<double>), dim3(nb), dim3(BLOCK), 0, s->stream, g, b,
                       (double*)s->b[0], (double*)s->b[1]);
  else
    hipLaunchKernelGGL((k_pack_vec<float>), dim3(nb), dim3(BLOCK), 0, s->stream, g, b,
                       (float*)s->b[0], (float*)s->b[1]);
  MG_HIP(hipGetLastError());
  return end(s);
}

extern "C" int mgpcg_solve(mgpcg_solver_t* s, int max_iters, double rel_tol, double abs_tol,
                           double eps, int refine_every, mgpcg_result_t* out) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  if (begin(s)) return -2;
  if (set_params(s, max_iters, rel_tol, abs_tol, eps, refine_every)) return -2;
  if (s->use_graph && capture_iteration(s)) return -2;
  dispatch<RunPre>(s);
  MG_HIP(hipGetLastError());
  // Bursts of iterations; the state of burst k is read while burst k+1 runs.
  const long hard_cap = max_iters >= 0 ? (long)max_iters + 2 * s->burst + 2 : 1000000L;
  long issued = 0;
  int slot = 0;
  auto issue_burst = [&](int sl) -> int {
    for (int i = 0; i < s->burst; ++i)
      if (enqueue_iteration(s)) return -2;
    issued += s->burst;
    MG_HIP(hipMemcpyAsync(&s->host_st[sl], s->st, sizeof(CGState), hipMemcpyDeviceToHost,
                          s->stream));
    MG_HIP(hipEventRecord(s->ev_slot[sl], s->stream));
    return 0;
  };
  if (issue_burst(slot)) return -2;
  CGState final_state;
  for (;;) {
    if (issued < hard_cap && issue_burst(slot ^ 1)) return -2;
    MG_HIP(hipEventSynchronize(s->ev_slot[slot]));
    if (s->host_st[slot].done || issued >= hard_cap) {
      final_state = s->host_st[slot];
      break;
    }
    slot ^= 1;
  }
  MG_HIP(hipStreamSynchronize(s->stream));
  if (!final_state.done) {
    final_state = s->host_st[slot ^ 1];
  }
  fill_result(final_state, out);
  MG_HIP(hipGetLastError());
  return end(s);
}

extern "C" int mgpcg_solve_hostsync(mgpcg_solver_t* s, int max_iters, double rel_tol,
                                    double abs_tol, double eps, int refine_every,
                                    mgpcg_result_t* out) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  if (begin(s)) return -2;
  if (set_params(s, max_iters, rel_tol, abs_tol, eps, refine_every)) return -2;
  int rc = dispatch<RunHostSync>(s, max_iters);
  if (rc) return rc;
  MG_HIP(hipStreamSynchronize(s->stream));
  fill_result(s->host_st[0], out);
  MG_HIP(hipGetLastError());
  return end(s);
}

extern "C" int mgpcg_store_x(mgpcg_solver_t* s, float* x_f32, double* x_f64) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  MG_REQUIRE((x_f32 != nullptr) != (x_f64 != nullptr), "exactly one output pointer required");
  if (begin(s)) return -2;
  const Geo& g = s->geo[0];
  const int nb = 2 * nblk(g.size);
  if (s->xsz == 4) {
    if (x_f32)
      hipLaunchKernelGGL((k_unpack_vec<float, float>), dim3(nb), dim3(BLOCK), 0, s->stream, g,
                         (const float*)s->x[0], (const float*)s->x[1], x_f32);
    else
      hipLaunchKernelGGL((k_unpack_vec<float, double>), dim3(nb), dim3(BLOCK), 0, s->stream, g,
                         (const float*)s->x[0], (const float*)s->x[1], x_f64);
  } else {
    if (x_f32)
      hipLaunchKernelGGL((k_unpack_vec<double, float>), dim3(nb), dim3(BLOCK), 0, s->stream, g,
                         (const double*)s->x[0], (const double*)s->x[1], x_f32);
    else
      hipLaunchKernelGGL((k_unpack_vec<double, double>), dim3(nb), dim3(BLOCK), 0, s->stream,
                         g, (const double*)s->x[0], (const double*)s->x[1], x_f64);
  }
  MG_HIP(hipGetLastError());
  return end(s);
}

// Debug / parity access: move a level vector between canonical and internal layout.
static int vec_slots(mgpcg_solver* s, int field, int level, void** c0, void** c1, size_t* esz) {
  MG_REQUIRE(level >= 0 && level < s->L, "level %d out of range", level);
  *esz = s->esz;
  switch (field) {
    case MGPCG_FIELD_R:
      *c0 = s->lv[level].r[0];
      *c1 = s->lv[level].r[1];
      return 0;
    case MGPCG_FIELD_Z:
      *c0 = s->lv[level].z[0];
      *c1 = s->lv[level].z[1];
      return 0;
    case MGPCG_FIELD_P:
      MG_REQUIRE(level == 0, "p lives on level 0");
      *c0 = s->p[0];
      *c1 = s->p[1];
      return 0;
    case MGPCG_FIELD_AP:
      MG_REQUIRE(level == 0, "Ap lives on level 0");
      *c0 = s->Ap[0];
      *c1 = s->Ap[1];
      return 0;
    case MGPCG_FIELD_X:
      MG_REQUIRE(level == 0, "x lives on level 0");
      *c0 = s->x[0];
      *c1 = s->x[1];
      *esz = s->xsz;
      return 0;
    default:
      mg_set_error("unknown field %d", field);
      return -1;
  }
}

extern "C" int mgpcg_debug_get(mgpcg_solver_t* s, int field, int level, void* canon) {
  MG_REQUIRE(s != nullptr && canon != nullptr, "NULL argument");
  void *c0, *c1;
  size_t esz;
  int rc = vec_slots(s, field, level, &c0, &c1, &esz);
  if (rc) return rc;
  if (begin(s)) return -2;
  const Geo& g = s->geo[level];
  const int nb = 2 * nblk(g.size);
  if (esz == 4)
    hipLaunchKernelGGL((k_unpack_vec<float, float>), dim3(nb), dim3(BLOCK), 0, s->stream, g,
                       (const float*)c0, (const float*)c1, (float*)canon);
  else
    hipLaunchKernelGGL((k_unpack_vec<double, double>), dim3(nb), dim3(BLOCK), 0, s->stream, g,
                       (const double*)c0, (const double*)c1, (double*)canon);
  MG_HIP(hipGetLastError());
  return end(s);
}

template <typename T>
__global__ void k_pack_any(Geo g, const T* __restrict__ src, T* d0, T* d1) {
  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int c = lb & 1;
  const int t = (lb >> 1) * BLOCK + threadIdx.x;
  if (t >= g.size) return;
  Slot sl = decode(g, t, c);
  T v = (T)0;
  if (sl.interior) v = src[canon_index(g, sl.i, sl.j, sl.k)];
  (c ? d1 : d0)[t] = v;
}

extern "C" int mgpcg_debug_set(mgpcg_solver_t* s, int field, int level, const void* canon) {
  MG_REQUIRE(s != nullptr && canon != nullptr, "NULL argument");
  void *c0, *c1;
  size_t esz;
  int rc = vec_slots(s, field, level, &c0, &c1, &esz);
  if (rc) return rc;
  if (begin(s)) return -2;
  const Geo& g = s->geo[level];
  const int nb = 2 * nblk(g.size);
  if (esz == 4)
    hipLaunchKernelGGL((k_pack_any<float>), dim3(nb), dim3(BLOCK), 0, s->stream, g,
                       (const float*)canon, (float*)c0, (float*)c1);
  else
    hipLaunchKernelGGL((k_pack_any<double>), dim3(nb), dim3(BLOCK), 0, s->stream, g,
                       (const double*)canon, (double*)c0, (double*)c1);
  MG_HIP(hipGetLastError());
  return end(s);
}

extern "C" int mgpcg_debug_op(mgpcg_solver_t* s, int op, int level, int arg) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  MG_REQUIRE(level >= 0 && level < s->L, "level %d out of range", level);
  if ((op == MGPCG_OP_RESTRICT || op == MGPCG_OP_PROLONGATE))
    MG_REQUIRE(level < s->L - 1, "no coarser level below %d", level);
  if (begin(s)) return -2;
  if (op == MGPCG_OP_COMPUTE_AP) {
    // the fused dot needs a fresh reduction ticket
    CGState init;
    memset(&init, 0, sizeof(init));
    init.precision = s->precision;
    init.eps = 1e-12;
    CGState* h = &s->host_st[1];
    MG_HIP(hipStreamSynchronize(s->stream));
    *h = init;
    MG_HIP(hipMemcpyAsync(s->st, h, sizeof(CGState), hipMemcpyHostToDevice, s->stream));
  }
  int rc = dispatch<RunOp>(s, op, level, arg);
  if (rc) return rc;
  MG_HIP(hipGetLastError());
  MG_HIP(hipStreamSynchronize(s->stream));
  return end(s);
}

extern "C" int mgpcg_coarsen(void* stream, int dim, const int32_t* fine_res,
                             const int8_t* fine, int8_t* coarse) {
  MG_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
  MG_REQUIRE(fine_res && fine && coarse, "NULL argument");
  int f0 = 1, f1, f2;
  if (dim == 3) {
    f0 = fine_res[0];
    f1 = fine_res[1];
    f2 = fine_res[2];
  } else {
    f1 = fine_res[0];
    f2 = fine_res[1];
  }
  MG_REQUIRE(f1 % 2 == 0 && f2 % 2 == 0 && (dim == 2 || f0 % 2 == 0),
             "fine resolution must be even on every axis");
  const int nc = (dim == 3 ? f0 / 2 : 1) * (f1 / 2) * (f2 / 2);
  if (nc == 0) return 0;
  hipLaunchKernelGGL(k_coarsen, dim3(nblk(nc)), dim3(BLOCK), 0, (hipStream_t)stream, dim, f0, f1,
                     f2, fine, coarse, nc);
  MG_HIP(hipGetLastError());
  return 0;
}
