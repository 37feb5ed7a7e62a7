// Host side of the MI355X MGPCG solver and its C-ABI (include/mgpcg.h).
//
// mgpcg_solver  one system: its multigrid hierarchy resident in HBM in the
//               colour-split padded layout, the level-0 CG vectors, a
//               device-resident CG state and the reduction partials.
// Engine        runs the PCG loop of mgpcg.py:150-209 for a set of systems at
//               once (possibly of different shapes): every stage of an
//               iteration is ONE persistent kernel launch over all member
//               systems; every scalar (alpha, beta, the residual test) is
//               computed on the device by a per-system finalize kernel.
//               The host only enqueues iterations, one ahead of the state it
//               reads back, and retires / admits systems at iteration
//               boundaries -- which is also how the server does continuous
//               batching.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
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
  uint16_t* nb[2];   // flag words (coupling flags, diagonal code)
  uint8_t* wet;      // per block of BLOCK slots: 1 when one of them is FLUID
  int32_t* ord[2];   // work order per block / per half-sweep virtual block
  int32_t* ordtmp;   // scratch of the order kernel (one int per block)
  LevelMeta* meta;   // nonzero-coupling summary, written when the level is loaded
  uint32_t* part;    // the packing kernel's per-block share of meta (3 ints per block)
  void* res[2];      // restriction scratch (NULL on the coarsest level)
};

int g_coarse_max_slots = 4096;  // levels at most this large go to k_coarse
int g_max_grid = 2048;          // workgroups of a persistent stage launch
int g_fuse_first = 0;    // first pre-smoothing pair in one kernel: 1 every level, 2 levels >= 1
int g_fuse_prolong = 1;  // prolongation inside the first post-smoothing half-sweep
int g_fuse_p = 0;        // update_p inside compute_Ap
int g_compact = 1;       // flag-word stencils on levels with uniform couplings
int g_coarse_lds = 1;    // coarse kernel works on an LDS image of its levels when they fit
// restriction kernels: 0 one lane per child, 1 row-tiled with LDS, 2 residual
// pass + child gather (both coalesced)
int g_restrict_mode = 2;
int g_restrict_coarse = -1;  // restriction kernels on levels >= 1 (-1: g_restrict_mode)
int g_probe_every = 4;    // the launch probe times one eligible launch in this many
int g_server_depth = 2;   // server iterations queued on the device ahead of the host
int g_server_streams = 2; // engines (streams) a server spreads its members over
constexpr int MAX_LANES = 4;  // most engines per server (the lane stream pool's size)
int g_graph = 0;          // replay iterations of an unchanged member set from a hipGraph
long g_knob_epoch = 0;    // bumped by every option change (captured graphs go stale)

}  // namespace

struct Engine;

struct mgpcg_solver {
  int dim = 3, L = 1, nu = 2, bottom = 10, precision = PLAIN;
  int res[3] = {0, 0, 0};
  size_t esz = 4, xsz = 4;
  std::vector<Geo> geo;
  std::vector<LevelBuf> lv;
  void* b[2] = {nullptr, nullptr};
  void* x[2] = {nullptr, nullptr};
  void* p[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  void* Ap[2] = {nullptr, nullptr};
  CGState* st = nullptr;
  DD* partials = nullptr;
  int nparts = 0;
  int coarse_from = 0;
  char* pool = nullptr;
  size_t pool_bytes = 0;
  void* packtab = nullptr;  // PackLevel<T>[L] of k_pack_pressure_all (in the pool)
  hipStream_t user_stream = nullptr;
  Engine* engine = nullptr;  // private engine of the single-system API
  int device = 0;
  bool full_coupling = false;  // descriptors written now read every Ax
  // Recorded on an engine's stream when the system retires from it: iterations
  // queued on that stream before the host saw the system finish still carry it
  // (they skip it while its state says done).  A later admission, on any
  // engine, waits for this event before it resets the state.
  hipEvent_t ev_retired = nullptr;
  bool retired_pending = false;
  // The engine that recorded ev_retired; only compared, never followed.  An
  // engine synchronizes its stream before it is destroyed, so a stale match
  // (a new engine at a freed one's address) skips a wait that is complete.
  const Engine* retired_on = nullptr;
  // Recorded on the system's stream when it is submitted to a server: its
  // assembly is complete when this event is.
  hipEvent_t ev_submitted = nullptr;
};

namespace {

inline int nblk(int n) { return (n + BLOCK - 1) / BLOCK; }
// virtual blocks of a half-sweep stage over n slots
inline int nsweep(int n) { return (n + BLOCK * SLOT_LANES - 1) / (BLOCK * SLOT_LANES); }
inline int rblk(int coarse_slots) {
  return (coarse_slots + RESTRICT_SLOTS_PER_BLOCK - 1) / RESTRICT_SLOTS_PER_BLOCK;
}
size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

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
    P.nb[c] = B.nb[c];
    P.res[c] = (T*)B.res[c];
  }
  P.wet = B.wet;
  P.ord[0] = B.ord[0];
  P.ord[1] = l == 0 && s->coarse_from > 0 ? B.ord[1] : nullptr;  // level 0 on the batched kernels
  P.meta = (g_compact && !s->full_coupling) ? B.meta : nullptr;
  return P;
}

template <typename T>
LevelDesc<T> level_desc(const mgpcg_solver* s, int l) {
  LevelDesc<T> d;
  d.g = s->geo[l];
  d.P = level_ptrs<T>(s, l);
  if (l + 1 < s->L) {
    d.gc = s->geo[l + 1];
    d.C = level_ptrs<T>(s, l + 1);
  } else {
    d.gc = s->geo[l];
    d.C = d.P;
  }
  d.st = s->st;
  d.stats = s->lv[l].meta;
  return d;
}

template <typename T, typename TX>
SysDesc<T, TX> sys_desc(const mgpcg_solver* s, int desc0) {
  SysDesc<T, TX> d;
  d.g = s->geo[0];
  d.P = level_ptrs<T>(s, 0);
  for (int c = 0; c < 2; ++c) {
    d.b[c] = (T*)s->b[c];
    d.x[c] = (TX*)s->x[c];
    d.p[0][c] = (T*)s->p[0][c];
    d.p[1][c] = (T*)s->p[1][c];
    d.Ap[c] = (T*)s->Ap[c];
  }
  d.st = s->st;
  d.partials = s->partials;
  d.nparts = s->nparts;
  d.coarse_from = s->coarse_from;
  d.L = s->L;
  d.desc0 = desc0;
  return d;
}

// Host-to-device copies of the engine's tables (descriptors, initial states,
// work items), read by one kernel from their mapped pinned staging: each
// copied as a DMA of its own cost a dispatch with tens of microseconds of
// idle device before it.  One workgroup per copy; sizes in 4-byte words.
struct PullList {
  static const int N = 32;
  const uint32_t* src[N];
  uint32_t* dst[N];
  int words[N];
  int n;
};
__global__ void __launch_bounds__(256) k_pull(PullList pl) {
  const int i = blockIdx.x;
  if (i >= pl.n) return;
  const uint32_t* src = pl.src[i];
  uint32_t* dst = pl.dst[i];
  for (int w = threadIdx.x; w < pl.words[i]; w += 256) dst[w] = src[w];
}

// gather the CG states of the listed systems into one contiguous array
template <typename T, typename TX>
__global__ void k_gather_states(const SysDesc<T, TX>* __restrict__ sys,
                                const Item* __restrict__ items, int n, CGState* out) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  out[i] = *sys[items[i].desc].st;
  __threadfence_system();  // `out` is the host's mapped readback slot
}

}  // namespace

// ===========================================================================
// Engine
// ===========================================================================
struct Member {
  mgpcg_solver* s = nullptr;
  int phase = 0;     // host view: 0 = its first iteration (pre-loop) is next
  int active = 0;
  int64_t tag = 0;
  long started = 0;  // iteration index at admission
  uint32_t gen = 0;  // bumped whenever the slot gets a new system
};

struct Engine {
  int precision = -1, dim = 0, L = 0, nu = 0, bottom = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;  // false: a server lane's stream from the process-wide pool
  hipStream_t user_stream = nullptr;
  hipEvent_t ev_user = nullptr, ev_done = nullptr;
  int device = 0;

  std::vector<Member> mem;  // slot -> member
  int cap = 0;              // table capacity (members)
  char* d_levels = nullptr;  // LevelDesc<T>[cap * L]
  char* d_sys = nullptr;     // SysDesc<T,TX>[cap]
  size_t level_sz = 0, sys_sz = 0;

  // item lists of the current configuration
  std::vector<Item> items;
  struct StageRef {
    int off = 0, n = 0, nblocks = 0;
    int fine_nblocks = 0;  // one-item restriction stages: the fine level's slot blocks
  };
  StageRef st_init, st_cg, fin_init, fin_cg, fin_all, st_dot_fused, st_dot_plain, st_coarse,
      st_gather;
  bool coarse_lds = false;  // every coarse item's levels fit the coarse kernel's LDS image
  int restrict_mode = 0;    // restriction stages of level 0 are laid out for this mode
  int restrict_mode_c = 0;  // ... and those of the levels below it for this one
  std::vector<StageRef> st_slot, st_restrict, st_bottom;  // per level
  // the half-sweep stages: virtual blocks of BLOCK * SLOT_LANES slots
  std::vector<StageRef> st_sweep, st_bottom_sweep;
  int n_items_cap = 0;
  Item* d_items = nullptr;
  static const int RING = 4;
  Item* h_items[RING] = {nullptr, nullptr, nullptr, nullptr};
  Item* hd_items[RING] = {nullptr, nullptr, nullptr, nullptr};  // their device addresses
  hipEvent_t ev_items[RING] = {nullptr, nullptr, nullptr, nullptr};
  int items_slot = 0;
  bool dirty = true;

  // state readback
  CGState* d_states = nullptr;
  // state readbacks in flight (a ring: the server keeps several iterations queued)
  static const int NSLOT = 4;
  CGState* h_states[NSLOT] = {nullptr, nullptr, nullptr, nullptr};
  // the same pinned slots as the device addresses them: the iteration's last
  // finalize (after a graph replay, the gather kernel) writes the states there
  // itself (no copy dispatched after it)
  CGState* hd_states[NSLOT] = {nullptr, nullptr, nullptr, nullptr};
  // Set while an eager iteration is enqueued: the readback slot FIN_BETA's
  // finalize writes (its items are the gather's, in the same order).  NULL
  // during graph capture, whose replays are followed by k_gather_states.
  CGState* fused_gather = nullptr;
  hipEvent_t ev_states[NSLOT] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<std::pair<int, uint32_t>> gathered[NSLOT];  // (member slot, generation)
  long iteration = 0;

  // pinned staging of admissions (per member slot): descriptors, initial state,
  // and the device addresses of the same (mapped) memory
  char* h_desc = nullptr;
  CGState* h_stinit = nullptr;
  char* hd_desc = nullptr;
  CGState* hd_stinit = nullptr;
  // table copies queued for the next k_pull launch
  PullList pulls{};

  // launch-duration probe of the level-0 half-sweeps, timed on the device
  bool probe_on = false;
  int probe_used = 0;  // launches recorded
  int probe_seen = 0;  // eligible launches since the probe started
  unsigned long long* d_swept = nullptr;  // [2]: level-0 cells, FLUID cells swept by probed launches
  // per recorded launch i: workgroup starts folded into SPAN_SLOTS minima at
  // [i * SPAN_SLOTS], ends into maxima at [(PROBE_POOL + i) * SPAN_SLOTS]
  unsigned long long* d_span = nullptr;

  // hipGraph of one iteration of the current member set (g_graph): captured
  // on the second iteration of an item layout, replayed until the layout or
  // an option changes.  Replaced graphs are destroyed once the stream has
  // passed their last launch.
  long epoch = 0;        // bumped by every engine_build_items
  int epoch_eager = 0;   // iterations launched one kernel at a time in this epoch
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  long graph_epoch = -1, graph_knobs = -1;
  long graph_launches = 0, graph_captures = 0;
  struct RetiredGraph {
    hipGraph_t g;
    hipGraphExec_t x;
    hipEvent_t ev;
  };
  std::vector<RetiredGraph> graph_retired;

  ~Engine() { release(); }
  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (RetiredGraph& r : graph_retired) {
      (void)hipGraphExecDestroy(r.x);
      (void)hipGraphDestroy(r.g);
      (void)hipEventDestroy(r.ev);
    }
    graph_retired.clear();
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    if (h_desc) (void)hipHostFree(h_desc);
    if (h_stinit) (void)hipHostFree(h_stinit);
    h_desc = nullptr;
    h_stinit = nullptr;
    if (d_levels) (void)hipFree(d_levels);
    if (d_sys) (void)hipFree(d_sys);
    if (d_items) (void)hipFree(d_items);
    if (d_states) (void)hipFree(d_states);
    for (int i = 0; i < RING; ++i) {
      if (h_items[i]) (void)hipHostFree(h_items[i]);
      if (ev_items[i]) (void)hipEventDestroy(ev_items[i]);
    }
    for (int i = 0; i < NSLOT; ++i) {
      if (h_states[i]) (void)hipHostFree(h_states[i]);
      if (ev_states[i]) (void)hipEventDestroy(ev_states[i]);
    }
    if (d_span) (void)hipFree(d_span);
    if (d_swept) (void)hipFree(d_swept);
    d_swept = nullptr;
    if (ev_user) (void)hipEventDestroy(ev_user);
    if (ev_done) (void)hipEventDestroy(ev_done);
    if (stream && own_stream) (void)hipStreamDestroy(stream);
    d_levels = d_sys = nullptr;
    d_items = nullptr;
    d_states = nullptr;
    stream = nullptr;
  }
};

namespace {

// pooled: a stream the engine uses without owning it (NULL: its own stream)
int engine_init(Engine* e, hipStream_t pooled) {
  MG_HIP(hipGetDevice(&e->device));
  if (pooled) {
    e->stream = pooled;
    e->own_stream = false;
  } else {
    MG_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  }
  MG_HIP(hipEventCreateWithFlags(&e->ev_user, hipEventDisableTiming));
  MG_HIP(hipEventCreateWithFlags(&e->ev_done, hipEventDisableTiming));
  for (int i = 0; i < Engine::RING; ++i)
    MG_HIP(hipEventCreateWithFlags(&e->ev_items[i], hipEventDisableTiming));
  for (int i = 0; i < Engine::NSLOT; ++i)
    MG_HIP(hipEventCreateWithFlags(&e->ev_states[i], hipEventDisableTiming));
  return 0;
}

// ordered after the caller's stream, and the caller's stream after us
// Queue a copy of `bytes` (a multiple of 4) from mapped pinned staging (its
// device address) into a device table; engine_flush_pulls launches the queue.
int engine_flush_pulls(Engine* e) {
  if (e->pulls.n == 0) return 0;
  hipLaunchKernelGGL(k_pull, dim3(e->pulls.n), dim3(256), 0, e->stream, e->pulls);
  e->pulls.n = 0;
  MG_HIP(hipGetLastError());
  return 0;
}
int engine_pull(Engine* e, void* dst, const void* src_dev, size_t bytes) {
  MG_REQUIRE(bytes % 4 == 0, "table copy of %zu bytes is not a whole number of words", bytes);
  if (e->pulls.n == PullList::N && engine_flush_pulls(e)) return -2;
  const int i = e->pulls.n++;
  e->pulls.src[i] = (const uint32_t*)src_dev;
  e->pulls.dst[i] = (uint32_t*)dst;
  e->pulls.words[i] = (int)(bytes / 4);
  return 0;
}

int engine_begin(Engine* e) {
  MG_HIP(hipSetDevice(e->device));
  MG_HIP(hipEventRecord(e->ev_user, e->user_stream));
  MG_HIP(hipStreamWaitEvent(e->stream, e->ev_user, 0));
  return 0;
}
int engine_end(Engine* e) {
  MG_HIP(hipEventRecord(e->ev_done, e->stream));
  MG_HIP(hipStreamWaitEvent(e->user_stream, e->ev_done, 0));
  return 0;
}

size_t level_desc_size(int precision) {
  return precision == F64 ? sizeof(LevelDesc<double>) : sizeof(LevelDesc<float>);
}
size_t sys_desc_size(int precision) {
  if (precision == F64) return sizeof(SysDesc<double, double>);
  if (precision == MIXED) return sizeof(SysDesc<float, double>);
  return sizeof(SysDesc<float, float>);
}

int engine_reserve(Engine* e, int cap) {
  if (cap <= e->cap) return 0;
  int ncap = std::max(cap, std::max(8, 2 * e->cap));
  // queued copies name the old tables and staging: they run first
  if (engine_flush_pulls(e)) return -2;
  MG_HIP(hipStreamSynchronize(e->stream));
  char *nl = nullptr, *ns = nullptr;
  MG_HIP(hipMalloc(&nl, e->level_sz * (size_t)ncap * e->L));
  MG_HIP(hipMalloc(&ns, e->sys_sz * (size_t)ncap));
  if (e->d_levels) {
    MG_HIP(hipMemcpy(nl, e->d_levels, e->level_sz * (size_t)e->cap * e->L,
                     hipMemcpyDeviceToDevice));
    MG_HIP(hipMemcpy(ns, e->d_sys, e->sys_sz * (size_t)e->cap, hipMemcpyDeviceToDevice));
    MG_HIP(hipFree(e->d_levels));
    MG_HIP(hipFree(e->d_sys));
  }
  e->d_levels = nl;
  e->d_sys = ns;
  if (e->d_states) MG_HIP(hipFree(e->d_states));
  MG_HIP(hipMalloc(&e->d_states, sizeof(CGState) * (size_t)ncap));
  // readback slots of iterations still in flight (written by now: the stream
  // was synchronized above) keep their states for the collects to come
  for (int i = 0; i < Engine::NSLOT; ++i) {
    CGState* h = nullptr;
    MG_HIP(hipHostMalloc((void**)&h, sizeof(CGState) * (size_t)ncap,
                         hipHostMallocMapped | hipHostMallocCoherent));
    if (e->h_states[i]) {
      memcpy(h, e->h_states[i], sizeof(CGState) * (size_t)e->cap);
      MG_HIP(hipHostFree(e->h_states[i]));
    }
    e->h_states[i] = h;
    MG_HIP(hipHostGetDevicePointer((void**)&e->hd_states[i], e->h_states[i], 0));
  }
  const size_t dbytes = e->level_sz * e->L + e->sys_sz;
  char* nd = nullptr;
  CGState* nst = nullptr;
  // coherent: the device reads them where they lie, never from a cached copy
  MG_HIP(hipHostMalloc((void**)&nd, dbytes * (size_t)ncap,
                       hipHostMallocMapped | hipHostMallocCoherent));
  MG_HIP(hipHostMalloc((void**)&nst, sizeof(CGState) * (size_t)ncap,
                       hipHostMallocMapped | hipHostMallocCoherent));
  if (e->h_desc) MG_HIP(hipHostFree(e->h_desc));
  if (e->h_stinit) MG_HIP(hipHostFree(e->h_stinit));
  e->h_desc = nd;
  e->h_stinit = nst;
  MG_HIP(hipHostGetDevicePointer((void**)&e->hd_desc, e->h_desc, 0));
  MG_HIP(hipHostGetDevicePointer((void**)&e->hd_stinit, e->h_stinit, 0));
  e->cap = ncap;
  return 0;
}

// Write member slot `m`'s descriptors into the device tables, asynchronously
// and in stream order (kernels still in flight for the slot's previous member
// run first).  The pinned staging slot of `m` is not touched again before
// this copy ran: a slot is reused only after its member retired, which the
// host learns from a readback recorded after this copy.
int engine_write_desc(Engine* e, int m) {
  const mgpcg_solver* s = e->mem[m].s;
  const size_t dbytes = e->level_sz * e->L + e->sys_sz;
  char* buf = e->h_desc + dbytes * (size_t)m;
  for (int l = 0; l < e->L; ++l) {
    if (e->precision == F64) {
      LevelDesc<double> d = level_desc<double>(s, l);
      memcpy(buf + e->level_sz * l, &d, sizeof(d));
    } else {
      LevelDesc<float> d = level_desc<float>(s, l);
      memcpy(buf + e->level_sz * l, &d, sizeof(d));
    }
  }
  char* sd = buf + e->level_sz * e->L;
  const int desc0 = m * e->L;
  if (e->precision == F64) {
    SysDesc<double, double> d = sys_desc<double, double>(s, desc0);
    memcpy(sd, &d, sizeof(d));
  } else if (e->precision == MIXED) {
    SysDesc<float, double> d = sys_desc<float, double>(s, desc0);
    memcpy(sd, &d, sizeof(d));
  } else {
    SysDesc<float, float> d = sys_desc<float, float>(s, desc0);
    memcpy(sd, &d, sizeof(d));
  }
  const char* dbuf = e->hd_desc + (buf - e->h_desc);
  if (engine_pull(e, e->d_levels + e->level_sz * (size_t)m * e->L, dbuf, e->level_sz * e->L))
    return -2;
  if (engine_pull(e, e->d_sys + e->sys_sz * (size_t)m, dbuf + e->level_sz * e->L, e->sys_sz))
    return -2;
  return 0;
}

int engine_configure(Engine* e, const mgpcg_solver* s) {
  if (e->precision >= 0) {
    MG_REQUIRE(s->precision == e->precision && s->dim == e->dim && s->L == e->L &&
                   s->nu == e->nu && s->bottom == e->bottom,
               "all systems of a batch must share precision, dim, n_levels and smoothing counts");
    return 0;
  }
  e->precision = s->precision;
  e->dim = s->dim;
  e->L = s->L;
  e->nu = s->nu;
  e->bottom = s->bottom;
  // table strides are exactly the device-side array strides (descs[i])
  e->level_sz = level_desc_size(e->precision);
  e->sys_sz = sys_desc_size(e->precision);
  e->st_slot.assign(e->L, {});
  e->st_sweep.assign(e->L, {});
  e->st_bottom_sweep.assign(1, {});
  e->st_restrict.assign(e->L, {});
  e->st_bottom.assign(1, {});
  return 0;
}

struct SolveParams {
  int max_iters;
  double rel_tol, abs_tol, eps;
  int refine_every;
};

// Initial CG state of member slot m (parameters, phase 0), written
// asynchronously in stream order through the slot's pinned staging entry.
int engine_reset_state(Engine* e, int m, const SolveParams& prm) {
  mgpcg_solver* s = e->mem[m].s;
  CGState& init = e->h_stinit[m];
  memset(&init, 0, sizeof(init));
  init.rel_tol = prm.rel_tol;
  init.abs_tol = prm.abs_tol;
  init.eps = prm.eps;
  init.max_iters = prm.max_iters;
  init.refine_every = prm.refine_every;
  init.precision = s->precision;
  init.phase = 0;
  init.true_rr = -1.0;
  return engine_pull(e, s->st, e->hd_stinit + m, sizeof(CGState));
}

int engine_add(Engine* e, mgpcg_solver* s, int64_t tag, const SolveParams& prm) {
  if (engine_configure(e, s)) return -1;
  int slot = -1;
  for (int i = 0; i < (int)e->mem.size(); ++i)
    if (!e->mem[i].active && e->mem[i].s == nullptr) {
      slot = i;
      break;
    }
  if (slot < 0) {
    slot = (int)e->mem.size();
    e->mem.push_back(Member());
  }
  if (engine_reserve(e, (int)e->mem.size())) return -2;
  Member& m = e->mem[slot];
  m.s = s;
  m.phase = 0;
  m.active = 1;
  m.tag = tag;
  m.started = e->iteration;
  m.gen += 1;
  // iterations still queued where the system last ran read its state; the
  // reset below must come after them (on the same engine the stream's order
  // already places it there)
  if (s->retired_pending) {
    if (s->retired_on != e) MG_HIP(hipStreamWaitEvent(e->stream, s->ev_retired, 0));
    s->retired_pending = false;
  }
  if (engine_write_desc(e, slot)) return -2;
  if (engine_reset_state(e, slot, prm)) return -2;
  e->dirty = true;
  return slot;
}

// virtual blocks of one system's restriction into the coarse level gc, for the
// kernel the engine's current layout selects
inline int restrict_blocks(int mode, const Geo& gc) {
  if (mode == 1) return gc.n0 * gc.n1;
  if (mode == 2) return nblk(gc.size);
  return rblk(gc.size);
}
inline int restrict_blocks(const Engine* e, const Geo& gc) {
  return restrict_blocks(e->restrict_mode, gc);
}
// the restriction mode of level l (the fine level) in the engine's layout
inline int restrict_mode_of(const Engine* e, int l) {
  return l == 0 ? e->restrict_mode : e->restrict_mode_c;
}

// Rebuild the item lists of every stage from the member set.
int engine_build_items(Engine* e) {
  std::vector<Item>& it = e->items;
  it.clear();
  auto add_stage = [&](Engine::StageRef& ref, auto pred, auto blocks, auto descof) {
    ref.off = (int)it.size();
    int b0 = 0;
    for (int m = 0; m < (int)e->mem.size(); ++m) {
      const Member& M = e->mem[m];
      if (!M.active || !pred(M)) continue;
      Item x;
      x.desc = descof(m);
      x.block0 = b0;
      it.push_back(x);
      b0 += blocks(M);
    }
    ref.n = (int)it.size() - ref.off;
    ref.nblocks = b0;
  };
  const int L = e->L;
  // row-tiled restriction only when every member's fine rows fit its LDS tile
  e->restrict_mode = g_restrict_mode;
  e->restrict_mode_c = g_restrict_coarse >= 0 ? g_restrict_coarse : g_restrict_mode;
  if (e->restrict_mode == 1)
    for (const Member& M : e->mem)
      if (M.active && M.s->geo[0].n2 > RESTRICT_ROW_MAX) e->restrict_mode = 0;
  if (e->restrict_mode_c == 1)
    for (const Member& M : e->mem)
      if (M.active && M.s->L > 1 && M.s->geo[1].n2 > RESTRICT_ROW_MAX) e->restrict_mode_c = 0;
  auto sysidx = [](int m) { return m; };
  auto lvl0 = [](const Member& M) { return nblk(M.s->geo[0].size); };
  auto one = [](const Member&) { return 1; };
  add_stage(e->st_init, [](const Member& M) { return M.phase == 0; }, lvl0, sysidx);
  add_stage(e->fin_init, [](const Member& M) { return M.phase == 0; }, one, sysidx);
  add_stage(e->st_cg, [](const Member& M) { return M.phase == 1; }, lvl0, sysidx);
  add_stage(e->fin_cg, [](const Member& M) { return M.phase == 1; }, one, sysidx);
  add_stage(e->fin_all, [](const Member&) { return true; }, one, sysidx);
  add_stage(e->st_gather, [](const Member&) { return true; }, one, sysidx);
  for (int l = 0; l < L; ++l) {
    add_stage(e->st_slot[l], [l](const Member& M) { return l < M.s->coarse_from; },
              [l](const Member& M) { return nblk(M.s->geo[l].size); },
              [l, L](int m) { return m * L + l; });
    add_stage(e->st_sweep[l], [l](const Member& M) { return l < M.s->coarse_from; },
              [l](const Member& M) { return nsweep(M.s->geo[l].size); },
              [l, L](int m) { return m * L + l; });
    add_stage(e->st_restrict[l],
              [l, L](const Member& M) { return l < M.s->coarse_from && l + 1 < L; },
              [l, L, e](const Member& M) {
                if (l + 1 >= L) return 0;
                return restrict_blocks(restrict_mode_of(e, l), M.s->geo[l + 1]);
              },
              [l, L](int m) { return m * L + l; });
  }
  // members whose bottom level is too large for the coarse kernel
  add_stage(e->st_bottom[0], [L](const Member& M) { return M.s->coarse_from >= L; },
            [L](const Member& M) { return nblk(M.s->geo[L - 1].size); },
            [L](int m) { return m * L + L - 1; });
  add_stage(e->st_bottom_sweep[0], [L](const Member& M) { return M.s->coarse_from >= L; },
            [L](const Member& M) { return nsweep(M.s->geo[L - 1].size); },
            [L](int m) { return m * L + L - 1; });
  // coarse kernel: desc = first coarse level, block0 = number of coarse levels
  e->st_coarse.off = (int)it.size();
  e->coarse_lds = g_coarse_lds && e->precision != F64;
  for (int m = 0; m < (int)e->mem.size(); ++m) {
    const Member& M = e->mem[m];
    if (!M.active || M.s->coarse_from >= L) continue;
    Item x;
    x.desc = m * L + M.s->coarse_from;
    x.block0 = L - M.s->coarse_from;
    it.push_back(x);
    long bytes = coarse_scratch_bytes(M.s->geo[M.s->coarse_from].size, (int)M.s->esz);
    for (int l = M.s->coarse_from; l < L; ++l)
      bytes += coarse_image_bytes(M.s->geo[l].size, (int)M.s->esz);
    if (bytes > COARSE_LDS_BYTES) e->coarse_lds = false;
  }
  e->st_coarse.n = (int)it.size() - e->st_coarse.off;
  e->st_coarse.nblocks = e->st_coarse.n;
  // z.r after the V-cycle: fused into level 0's last half-sweep when level 0
  // is smoothed by the batched kernels, a plain dot otherwise
  const bool fused_ok = e->nu > 0 && e->L > 1;
  add_stage(e->st_dot_fused,
            [fused_ok](const Member& M) { return fused_ok && M.s->coarse_from > 0; }, lvl0,
            sysidx);
  add_stage(e->st_dot_plain,
            [fused_ok](const Member& M) { return !(fused_ok && M.s->coarse_from > 0); }, lvl0,
            sysidx);

  // upload through a pinned ring slot
  const int n = (int)it.size();
  if (n > e->n_items_cap) {
    MG_HIP(hipStreamSynchronize(e->stream));
    int ncap = std::max(n, 2 * e->n_items_cap + 64);
    if (e->d_items) MG_HIP(hipFree(e->d_items));
    MG_HIP(hipMalloc(&e->d_items, sizeof(Item) * ncap));
    for (int i = 0; i < Engine::RING; ++i) {
      if (e->h_items[i]) MG_HIP(hipHostFree(e->h_items[i]));
      MG_HIP(hipHostMalloc((void**)&e->h_items[i], sizeof(Item) * ncap,
                           hipHostMallocMapped | hipHostMallocCoherent));
      MG_HIP(hipHostGetDevicePointer((void**)&e->hd_items[i], e->h_items[i], 0));
    }
    e->n_items_cap = ncap;
  }
  const int sl = e->items_slot;
  e->items_slot = (sl + 1) % Engine::RING;
  MG_HIP(hipEventSynchronize(e->ev_items[sl]));
  if (n > 0) {
    memcpy(e->h_items[sl], it.data(), sizeof(Item) * n);
    if (engine_pull(e, e->d_items, e->hd_items[sl], sizeof(Item) * n)) return -2;
  }
  // every queued table copy (this layout's items, the admissions' descriptors
  // and states) lands before the kernels that read them
  if (engine_flush_pulls(e)) return -2;
  MG_HIP(hipEventRecord(e->ev_items[sl], e->stream));
  e->dirty = false;
  e->epoch++;
  e->epoch_eager = 0;
  return 0;
}

// Probe: launches beyond the pool are not recorded.
const int PROBE_POOL = 4096;
// Returns whether the launch is recorded.  Every g_probe_every-th eligible
// launch is; it counts the cells it sweeps into d_swept (on the device, so
// systems that finished on the device but are not retired yet are not
// counted) and records its span from the first workgroup's start to the last
// workgroup's end on the device clock, i.e. the kernel's execution, without
// the dispatch before it and the completion signal after it.
bool probe_begin(Engine* e) {
  if (e->probe_seen++ % g_probe_every != 0) return false;
  return e->probe_used < PROBE_POOL && e->d_swept != nullptr && e->d_span != nullptr;
}
void probe_end(Engine* e) { e->probe_used++; }

inline Stage stage_of(const Engine* e, const Engine::StageRef& r) {
  Stage s;
  s.items = e->d_items + r.off;
  s.nitems = r.n;
  s.nblocks = r.nblocks;
  s.swept = nullptr;
  s.t0 = nullptr;
  s.t1 = nullptr;
  return s;
}
inline int grid_of(const Engine::StageRef& r) { return std::min(r.nblocks, g_max_grid); }

// ---------------------------------------------------------------------------
// iteration schedule, templated on hierarchy type T, x type TX and dimension
// ---------------------------------------------------------------------------
template <typename T, typename TX, int DIM>
struct Sched {
  static const LevelDesc<T>* levels(const Engine* e) { return (const LevelDesc<T>*)e->d_levels; }
  static const SysDesc<T, TX>* sys(const Engine* e) { return (const SysDesc<T, TX>*)e->d_sys; }

  template <int OP>
  static void stage(Engine* e, const Engine::StageRef& r, bool level0 = false) {
    if (r.nblocks == 0) return;
    const bool probe = level0 && OP == ST_SMOOTH0 && e->probe_on && probe_begin(e);
    if (level0) {
      Stage st = stage_of(e, r);
      if (probe) {
        st.swept = e->d_swept;
        st.t0 = e->d_span + (size_t)e->probe_used * SPAN_SLOTS;
        st.t1 = e->d_span + (size_t)(PROBE_POOL + e->probe_used) * SPAN_SLOTS;
        hipLaunchKernelGGL((k_stage<T, DIM, OP, 2>), dim3(grid_of(r)), dim3(BLOCK), 0,
                           e->stream, levels(e), st);
      } else
        hipLaunchKernelGGL((k_stage<T, DIM, OP, 1>), dim3(grid_of(r)), dim3(BLOCK), 0,
                           e->stream, levels(e), st);
    } else
      hipLaunchKernelGGL((k_stage<T, DIM, OP, 0>), dim3(grid_of(r)), dim3(BLOCK), 0, e->stream,
                         levels(e), stage_of(e, r));
    if (probe) probe_end(e);
  }
  // fine: the level's slot stage (the residual pass), r: the restriction
  // stage; sweep (not NULL): the level's half-sweep stage, the last
  // smooth(l, 1) having written the colour-1 residual (ST_SMOOTH1_RES)
  static void restrict_(Engine* e, int mode, const Engine::StageRef& fine,
                        const Engine::StageRef& r, const Engine::StageRef* sweep = nullptr) {
    if (r.nblocks == 0) return;
    if (mode == 2) {
      if (sweep)
        hipLaunchKernelGGL((k_residual<T, DIM, true>), dim3(grid_of(*sweep)), dim3(BLOCK), 0,
                           e->stream, levels(e), stage_of(e, *sweep));
      else
        hipLaunchKernelGGL((k_residual<T, DIM>), dim3(grid_of(fine)), dim3(BLOCK), 0,
                           e->stream, levels(e), stage_of(e, fine));
      hipLaunchKernelGGL((k_restrict_gather<T, DIM>), dim3(grid_of(r)), dim3(BLOCK), 0,
                         e->stream, levels(e), stage_of(e, r));
    } else if (mode == 1) {
      hipLaunchKernelGGL((k_restrict_rows<T, DIM>), dim3(grid_of(r)), dim3(BLOCK), 0, e->stream,
                         levels(e), stage_of(e, r));
    } else {
      hipLaunchKernelGGL((k_restrict<T, DIM>), dim3(grid_of(r)), dim3(BLOCK), 0, e->stream,
                         levels(e), stage_of(e, r));
    }
  }
  // out: FIN_BETA's readback slot (see Engine::fused_gather)
  template <int OP>
  static void fin(Engine* e, const Engine::StageRef& r, CGState* out = nullptr) {
    if (r.n == 0) return;
    hipLaunchKernelGGL((k_finalize<T, TX, OP>), dim3(r.n), dim3(BLOCK), 0, e->stream, sys(e),
                       (const Item*)(e->d_items + r.off), r.n, out);
  }

  // v_cycle (mgpcg.py:125-148) for every member.  first0: level 0's first
  // half-sweep from z = 0 was done by the kernel that wrote r (k_cg_init,
  // k_cg_xr, k_true_residual).
  static void vcycle(Engine* e, bool first0 = false) {
    const int L = e->L, nu = e->nu;
    for (int l = 0; l + 1 < L; ++l) {
      const Engine::StageRef& S = e->st_slot[l];
      const Engine::StageRef& W = e->st_sweep[l];
      // the last pre-smoothing smooth(l, 1) also forms the colour-1 residual
      // when the restriction is the two-pass one and that half-sweep exists
      const int rm = restrict_mode_of(e, l);
      const bool ff = g_fuse_first == 1 || (g_fuse_first == 2 && l > 0);
      const bool res1 = rm == 2 && nu > 0 && !(ff && nu == 1);
      if (nu > 0) {
        if (ff) {
          stage<ST_PAIR_FROM_ZERO>(e, S, l == 0);
        } else {
          if (!(first0 && l == 0)) stage<ST_FIRST>(e, S, l == 0);
          if (nu == 1 && res1) stage<ST_SMOOTH1_RES>(e, W, l == 0);
          else stage<ST_SMOOTH1>(e, W, l == 0);
        }
        for (int i = 1; i < nu; ++i) {
          stage<ST_SMOOTH0>(e, W, l == 0);
          if (i == nu - 1 && res1) stage<ST_SMOOTH1_RES>(e, W, l == 0);
          else stage<ST_SMOOTH1>(e, W, l == 0);
        }
      } else {
        stage<ST_ZERO>(e, S, l == 0);
      }
      restrict_(e, rm, e->st_slot[l], e->st_restrict[l], res1 ? &W : nullptr);
    }
    // bottom level on the batched kernels (only when it is too big for k_coarse)
    const Engine::StageRef& B = e->st_bottom[0];
    const Engine::StageRef& BW = e->st_bottom_sweep[0];
    if (B.nblocks) {
      const int half = e->bottom / 2;
      if (half > 0) {
        stage<ST_PAIR_FROM_ZERO>(e, B);
        for (int i = 1; i < half; ++i) {
          stage<ST_SMOOTH0>(e, BW);
          stage<ST_SMOOTH1>(e, BW);
        }
        for (int i = 0; i < half; ++i) {
          stage<ST_SMOOTH1>(e, BW);
          stage<ST_SMOOTH0>(e, BW);
        }
      } else {
        stage<ST_ZERO>(e, B);
      }
    }
    if (e->st_coarse.n) {
      bool lds = false;
      if constexpr (sizeof(T) == 4) {
        lds = e->coarse_lds;
        if (lds)
          hipLaunchKernelGGL((k_coarse<T, DIM, true>), dim3(e->st_coarse.n),
                             dim3(COARSE_THREADS), 0, e->stream, levels(e),
                             (const Item*)(e->d_items + e->st_coarse.off), e->st_coarse.n, nu,
                             e->bottom);
      }
      if (!lds)
        hipLaunchKernelGGL((k_coarse<T, DIM, false>), dim3(e->st_coarse.n), dim3(COARSE_THREADS),
                           0, e->stream, levels(e), (const Item*)(e->d_items + e->st_coarse.off),
                           e->st_coarse.n, nu, e->bottom);
    }
    for (int l = L - 2; l >= 0; --l) {
      const Engine::StageRef& S = e->st_slot[l];
      const Engine::StageRef& W = e->st_sweep[l];
      if (nu > 0) {
        if (g_fuse_prolong) {
          stage<ST_PROLONG_SMOOTH1>(e, S, l == 0);
        } else {
          stage<ST_PROLONGATE0>(e, S, l == 0);
          stage<ST_SMOOTH1>(e, W, l == 0);
        }
        for (int i = 0; i < nu; ++i) {
          if (i > 0) stage<ST_SMOOTH1>(e, W, l == 0);
          const bool last = (i == nu - 1);
          if (!(last && l == 0)) stage<ST_SMOOTH0>(e, W, l == 0);
        }
      } else {
        stage<ST_PROLONGATE>(e, S, l == 0);
      }
    }
    // level 0's last half-sweep fused with reduce(z, r), or a plain dot
    if (e->st_dot_fused.nblocks)
      hipLaunchKernelGGL((k_smooth0_dot<T, TX, DIM>), dim3(grid_of(e->st_dot_fused)),
                         dim3(BLOCK), 0, e->stream, sys(e), stage_of(e, e->st_dot_fused));
    if (e->st_dot_plain.nblocks)
      hipLaunchKernelGGL((k_dot_zr<T, TX>), dim3(grid_of(e->st_dot_plain)), dim3(BLOCK), 0,
                         e->stream, sys(e), stage_of(e, e->st_dot_plain));
  }

  // One iteration for every member: the pre-loop of newly admitted systems
  // (mgpcg.py:157-172) and one loop body (:176-209) of the others.
  static void iteration(Engine* e, bool sync_each) {
    auto sync = [&]() {
      if (sync_each) (void)hipStreamSynchronize(e->stream);
    };
    if (e->st_init.nblocks) {
      hipLaunchKernelGGL((k_cg_init<T, TX>), dim3(grid_of(e->st_init)), dim3(BLOCK), 0,
                         e->stream, sys(e), stage_of(e, e->st_init));
      fin<FIN_INIT>(e, e->fin_init);
      sync();
    }
    if (e->st_cg.nblocks) {
      if (g_fuse_p) {
        hipLaunchKernelGGL((k_cg_ap<T, TX, DIM>), dim3(grid_of(e->st_cg)), dim3(BLOCK), 0,
                           e->stream, sys(e), stage_of(e, e->st_cg));
      } else {
        hipLaunchKernelGGL((k_cg_update_p<T, TX>), dim3(grid_of(e->st_cg)), dim3(BLOCK), 0,
                           e->stream, sys(e), stage_of(e, e->st_cg));
        hipLaunchKernelGGL((k_cg_ap_only<T, TX, DIM>), dim3(grid_of(e->st_cg)), dim3(BLOCK), 0,
                           e->stream, sys(e), stage_of(e, e->st_cg));
      }
      fin<FIN_ALPHA>(e, e->fin_cg);
      sync();
      hipLaunchKernelGGL((k_cg_xr<T, TX>), dim3(grid_of(e->st_cg)), dim3(BLOCK), 0, e->stream,
                         sys(e), stage_of(e, e->st_cg));
      fin<FIN_RR>(e, e->fin_cg);
      if (e->precision == MIXED) {
        hipLaunchKernelGGL((k_true_residual<T, TX, DIM>), dim3(grid_of(e->st_cg)), dim3(BLOCK),
                           0, e->stream, sys(e), stage_of(e, e->st_cg));
        fin<FIN_TRUE>(e, e->fin_cg);
      }
      sync();
    }
    vcycle(e, true);
    fin<FIN_BETA>(e, e->fin_all, e->fused_gather);
    sync();
  }

  static void gather(Engine* e, int slot) {
    const Engine::StageRef& g = e->st_gather;
    if (g.n == 0) return;
    hipLaunchKernelGGL((k_gather_states<T, TX>), dim3((g.n + 63) / 64), dim3(64), 0, e->stream,
                       sys(e), (const Item*)(e->d_items + g.off), g.n, e->hd_states[slot]);
  }
};

template <template <typename, typename, int> class F, typename... Args>
auto dispatch(int precision, int dim, Args&&... args) {
  if (precision == F64) {
    if (dim == 3) return F<double, double, 3>::run(args...);
    return F<double, double, 2>::run(args...);
  }
  if (precision == MIXED) {
    if (dim == 3) return F<float, double, 3>::run(args...);
    return F<float, double, 2>::run(args...);
  }
  if (dim == 3) return F<float, float, 3>::run(args...);
  return F<float, float, 2>::run(args...);
}

template <typename T, typename TX, int DIM>
struct RunIteration {
  static int run(Engine* e, bool sync_each) {
    Sched<T, TX, DIM>::iteration(e, sync_each);
    return 0;
  }
};
template <typename T, typename TX, int DIM>
struct RunGather {
  static int run(Engine* e, int slot) {
    Sched<T, TX, DIM>::gather(e, slot);
    return 0;
  }
};
template <typename T, typename TX, int DIM>
struct RunVcycle {
  static int run(Engine* e) {
    Sched<T, TX, DIM>::vcycle(e);
    return 0;
  }
};

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

// Destroy the replaced graphs whose last launch the stream has passed (all of
// them when `all`, after synchronizing).
int engine_graph_reap(Engine* e, bool all) {
  if (all && !e->graph_retired.empty()) MG_HIP(hipStreamSynchronize(e->stream));
  size_t k = 0;
  for (size_t i = 0; i < e->graph_retired.size(); ++i) {
    Engine::RetiredGraph& r = e->graph_retired[i];
    const hipError_t q = all ? hipSuccess : hipEventQuery(r.ev);
    if (q == hipSuccess) {
      MG_HIP(hipGraphExecDestroy(r.x));
      MG_HIP(hipGraphDestroy(r.g));
      MG_HIP(hipEventDestroy(r.ev));
    } else if (q == hipErrorNotReady) {
      e->graph_retired[k++] = r;
    } else {
      MG_HIP(q);
    }
  }
  e->graph_retired.resize(k);
  return 0;
}

// Launch the current iteration from the engine's graph, capturing it first
// when the item layout or an option changed since the last capture.  The
// iteration's kernels take only the engine's tables and item lists as
// arguments and read every per-system scalar (iteration parity, alpha, beta,
// done) from the device, so one capture serves every iteration of a layout.
int engine_graph_launch(Engine* e) {
  if (!(e->graph_exec && e->graph_epoch == e->epoch && e->graph_knobs == g_knob_epoch)) {
    if (e->graph_exec) {
      Engine::RetiredGraph r{e->graph, e->graph_exec, nullptr};
      MG_HIP(hipEventCreateWithFlags(&r.ev, hipEventDisableTiming));
      MG_HIP(hipEventRecord(r.ev, e->stream));
      e->graph_retired.push_back(r);
      e->graph = nullptr;
      e->graph_exec = nullptr;
    }
    MG_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed));
    dispatch<RunIteration>(e->precision, e->dim, e, false);
    hipGraph_t g = nullptr;
    const hipError_t rc = hipStreamEndCapture(e->stream, &g);
    if (rc != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      MG_HIP(rc);
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ri = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ri != hipSuccess) {
      (void)hipGraphDestroy(g);
      MG_HIP(ri);
    }
    e->graph = g;
    e->graph_exec = x;
    e->graph_epoch = e->epoch;
    e->graph_knobs = g_knob_epoch;
    e->graph_captures++;
  }
  MG_HIP(hipGraphLaunch(e->graph_exec, e->stream));
  e->graph_launches++;
  return 0;
}

// Enqueue one iteration plus the readback of the states of its members.  With
// g_graph the iteration comes from a graph once its item layout has run
// eagerly once (a layout with newly admitted members lasts one iteration);
// the launch probe and host-synchronized iterations always launch eagerly.
int engine_enqueue(Engine* e, int slot, bool sync_each) {
  if (e->dirty && engine_build_items(e)) return -2;
  if (!e->graph_retired.empty() && engine_graph_reap(e, false)) return -2;
  if (g_graph && !sync_each && !e->probe_on && e->epoch_eager > 0) {
    if (engine_graph_launch(e)) return -2;
    dispatch<RunGather>(e->precision, e->dim, e, slot);
  } else {
    // the iteration's last finalize writes the readback slot itself
    e->fused_gather = e->hd_states[slot];
    dispatch<RunIteration>(e->precision, e->dim, e, sync_each);
    e->fused_gather = nullptr;
    e->epoch_eager++;
  }
  e->gathered[slot].clear();
  for (int m = 0; m < (int)e->mem.size(); ++m)
    if (e->mem[m].active) e->gathered[slot].push_back({m, e->mem[m].gen});
  MG_HIP(hipEventRecord(e->ev_states[slot], e->stream));
  MG_HIP(hipGetLastError());
  // phase transitions are deterministic: a system's pre-loop takes one iteration
  bool changed = false;
  for (Member& M : e->mem)
    if (M.active && M.phase == 0) {
      M.phase = 1;
      changed = true;
    }
  if (changed) e->dirty = true;
  e->iteration++;
  return 0;
}

struct Finished {
  int member;
  CGState st;
};

// Wait for readback `slot`, retire the members that are done.
int engine_collect(Engine* e, int slot, std::vector<Finished>* done) {
  MG_HIP(hipEventSynchronize(e->ev_states[slot]));
  const auto& g = e->gathered[slot];
  for (int i = 0; i < (int)g.size(); ++i) {
    const CGState& st = e->h_states[slot][i];
    Member& M = e->mem[g[i].first];
    // a slot re-admitted since this readback holds another system now
    if (M.active && M.gen == g[i].second && st.done) {
      M.active = 0;
      e->dirty = true;
      // every iteration enqueued so far may still hold the system
      mgpcg_solver* s = M.s;
      if (!s->ev_retired) MG_HIP(hipEventCreateWithFlags(&s->ev_retired, hipEventDisableTiming));
      MG_HIP(hipEventRecord(s->ev_retired, e->stream));
      s->retired_pending = true;
      s->retired_on = e;
      if (done) done->push_back({g[i].first, st});
    }
  }
  return 0;
}

int engine_active(const Engine* e) {
  int n = 0;
  for (const Member& M : e->mem) n += M.active;
  return n;
}

// Run every member to completion (iterations pipelined one deep).
int engine_run(Engine* e, long host_cap, std::vector<Finished>* done) {
  int slot = 0;
  bool inflight = false;
  long launched = 0;
  while (true) {
    const int nact = engine_active(e);
    if (nact > 0 && launched < host_cap) {
      if (engine_enqueue(e, slot, false)) return -2;
      launched++;
      if (inflight && engine_collect(e, slot ^ 1, done)) return -2;
      inflight = true;
      slot ^= 1;
    } else if (inflight) {
      if (engine_collect(e, slot ^ 1, done)) return -2;
      inflight = false;
    } else {
      break;
    }
  }
  MG_HIP(hipStreamSynchronize(e->stream));
  return 0;
}

// remove retired members from the table (keeps slots of active ones)
void engine_clear_retired(Engine* e) {
  for (Member& M : e->mem)
    if (!M.active) M.s = nullptr;
  e->dirty = true;
}

Engine* new_engine(hipStream_t pooled = nullptr) {
  Engine* e = new (std::nothrow) Engine();
  if (!e) return nullptr;
  if (engine_init(e, pooled)) {
    delete e;
    return nullptr;
  }
  return e;
}

// The private engine of a single system (holds only that system).
int single_engine(mgpcg_solver* s, Engine** out) {
  if (!s->engine) {
    s->engine = new_engine();
    MG_REQUIRE(s->engine != nullptr, "engine creation failed: %s", g_err);
  }
  Engine* e = s->engine;
  e->user_stream = s->user_stream;
  *out = e;
  return 0;
}

int solve_single(mgpcg_solver* s, int max_iters, double rel_tol, double abs_tol, double eps,
                 int refine_every, mgpcg_result_t* out, bool hostsync) {
  Engine* e = nullptr;
  if (single_engine(s, &e)) return -1;
  if (engine_begin(e)) return -2;
  for (Member& M : e->mem) {
    M.active = 0;
    M.s = nullptr;
  }
  e->mem.clear();
  const SolveParams prm{max_iters, rel_tol, abs_tol, eps, refine_every};
  if (engine_add(e, s, 0, prm) < 0) return -2;
  std::vector<Finished> done;
  const long cap = max_iters >= 0 ? (long)max_iters + 4 : 10000000L;
  if (!hostsync) {
    if (engine_run(e, cap, &done)) return -2;
  } else {
    // the reference's host loop: a round trip after every reduction
    for (long it = 0; it < cap && engine_active(e); ++it) {
      if (engine_enqueue(e, 0, true)) return -2;
      if (engine_collect(e, 0, &done)) return -2;
    }
  }
  if (done.empty()) {
    // host cap reached: report the last state
    MG_HIP(hipStreamSynchronize(e->stream));
    CGState st;
    MG_HIP(hipMemcpy(&st, s->st, sizeof(CGState), hipMemcpyDeviceToHost));
    fill_result(st, out);
  } else {
    fill_result(done.back().st, out);
  }
  MG_HIP(hipGetLastError());
  return engine_end(e);
}

// Layout of one system's device pool: every sub-buffer 256-B aligned, in this
// order per level (both colours of gt, Adiag, Ax per axis, r, z with its zero
// pads, flag words, residual scratch), then the level's metadata, block flags,
// work orders and packing partials; after the levels, level 0's b, x, both p, Ap, the CG state
// the reduction partials and the packing table.
struct PoolPlan {
  struct Offs {
    size_t gt[2], Ad[2], Ax[3][2], r[2], z[2], nb[2], res[2], meta, wet, ord[2], ordtmp, part;
  };
  std::vector<Geo> geo;
  std::vector<Offs> lo;
  size_t ob[2], ox[2], op[2][2], oap[2], ost, opart, opack;
  int nparts;
  size_t bytes;
};

void plan_pool(int dim, const int32_t* res, int n_levels, size_t esz, size_t xsz, PoolPlan* P) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  P->geo.clear();
  P->lo.clear();
  for (int l = 0; l < n_levels; ++l) {
    int n[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) n[a] = res[a] >> l;
    P->geo.push_back(make_geo(dim, n));
    const Geo& g = P->geo.back();
    PoolPlan::Offs o;
    for (int c = 0; c < 2; ++c) {
      o.gt[c] = take(g.size);
      o.Ad[c] = take(g.size * esz);
      for (int a = 0; a < 3; ++a) o.Ax[a][c] = (a < dim) ? take(g.size * esz) : (size_t)-1;
      o.r[c] = take(g.size * esz);
      o.z[c] = take((size_t)(g.size + 2 * g.zpad) * esz);  // zero pads both sides
      o.nb[c] = take(g.size * sizeof(uint16_t));
      o.res[c] = (l + 1 < n_levels) ? take(g.size * esz) : (size_t)-1;
    }
    o.meta = take(sizeof(LevelMeta));
    o.wet = take(nblk(g.size));
    o.ord[0] = take(sizeof(int32_t) * nblk(g.size));
    o.ord[1] = take(sizeof(int32_t) * nsweep(g.size));
    o.ordtmp = take(sizeof(int32_t) * nsweep(g.size));
    o.part = take(sizeof(uint32_t) * 3 * nblk(g.size));
    P->lo.push_back(o);
  }
  const Geo& g0 = P->geo[0];
  for (int c = 0; c < 2; ++c) {
    P->ob[c] = take(g0.size * esz);
    P->ox[c] = take(g0.size * xsz);
    P->op[0][c] = take(g0.size * esz);
    P->op[1][c] = take(g0.size * esz);
    P->oap[c] = take(g0.size * esz);
  }
  P->ost = take(sizeof(CGState));
  P->nparts = nblk(g0.size);
  P->opart = take(sizeof(DD) * P->nparts);
  // the fused packing kernel's per-level table (same size for either precision)
  static_assert(sizeof(PackLevel<float>) == sizeof(PackLevel<double>), "");
  P->opack = take(sizeof(PackLevel<double>) * (size_t)n_levels);
  P->bytes = off;
}

int check_shape(int dim, const int32_t* res, int n_levels, int precision) {
  MG_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3 (got %d)", dim);
  MG_REQUIRE(res != nullptr, "res is NULL");
  MG_REQUIRE(n_levels >= 1 && n_levels <= 16, "n_levels must be in [1, 16] (got %d)", n_levels);
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
  return 0;
}

}  // namespace

// ===========================================================================
// C-ABI: systems
// ===========================================================================
extern "C" int mgpcg_create(int dim, const int32_t* res, int n_levels, int nu, int bottom,
                            int precision, mgpcg_solver_t** out) {
  MG_REQUIRE(out != nullptr, "out is NULL");
  *out = nullptr;
  if (check_shape(dim, res, n_levels, precision)) return -1;
  MG_REQUIRE(nu >= 0 && bottom >= 0, "smoothing counts must be >= 0");

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

  PoolPlan P;
  plan_pool(dim, res, n_levels, s->esz, s->xsz, &P);
  s->geo = P.geo;
  s->coarse_from = n_levels;
  for (int l = 0; l < n_levels; ++l)
    if (s->geo[l].size <= g_coarse_max_slots) {
      s->coarse_from = l;
      break;
    }
  const std::vector<PoolPlan::Offs>& lo = P.lo;
  const size_t* ob = P.ob;
  const size_t* ox = P.ox;
  const size_t(*op)[2] = P.op;
  const size_t* oap = P.oap;
  const size_t ost = P.ost;
  const size_t opart = P.opart;
  s->nparts = P.nparts;
  s->pool_bytes = P.bytes;

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
      B.z[c] = s->pool + lo[l].z[c] + (size_t)s->geo[l].zpad * s->esz;
      B.nb[c] = (uint16_t*)(s->pool + lo[l].nb[c]);
      B.res[c] = (l + 1 < n_levels) ? (void*)(s->pool + lo[l].res[c]) : nullptr;
    }
    B.meta = (LevelMeta*)(s->pool + lo[l].meta);
    B.wet = (uint8_t*)(s->pool + lo[l].wet);
    B.ord[0] = (int32_t*)(s->pool + lo[l].ord[0]);
    B.ord[1] = (int32_t*)(s->pool + lo[l].ord[1]);
    B.ordtmp = (int32_t*)(s->pool + lo[l].ordtmp);
    B.part = (uint32_t*)(s->pool + lo[l].part);
    s->lv.push_back(B);
    const Geo& g = s->geo[l];
    // ghost cells are SOLID (the zero fill made them FLUID)
    if ((e = hipMemset(B.gt[0], MG_SOLID, g.size)) != hipSuccess) return fail(e, "memset");
    if ((e = hipMemset(B.gt[1], MG_SOLID, g.size)) != hipSuccess) return fail(e, "memset");
  }
  for (int c = 0; c < 2; ++c) {
    s->b[c] = s->pool + ob[c];
    s->x[c] = s->pool + ox[c];
    s->p[0][c] = s->pool + op[0][c];
    s->p[1][c] = s->pool + op[1][c];
    s->Ap[c] = s->pool + oap[c];
  }
  s->st = (CGState*)(s->pool + ost);
  s->partials = (DD*)(s->pool + opart);
  s->packtab = s->pool + P.opack;
  {
    std::vector<char> tab(sizeof(PackLevel<double>) * (size_t)n_levels);
    for (int l = 0; l < n_levels; ++l) {
      if (precision == F64) {
        const PackLevel<double> pl{s->geo[l], level_ptrs<double>(s, l), s->lv[l].part};
        memcpy(tab.data() + sizeof(pl) * l, &pl, sizeof(pl));
      } else {
        const PackLevel<float> pl{s->geo[l], level_ptrs<float>(s, l), s->lv[l].part};
        memcpy(tab.data() + sizeof(pl) * l, &pl, sizeof(pl));
      }
    }
    if ((e = hipMemcpy(s->packtab, tab.data(), tab.size(), hipMemcpyHostToDevice)) != hipSuccess)
      return fail(e, "hipMemcpy(packing table)");
  }
  if ((e = hipDeviceSynchronize()) != hipSuccess) return fail(e, "hipDeviceSynchronize");
  *out = s;
  return 0;
}

extern "C" int mgpcg_pool_bytes(int dim, const int32_t* res, int n_levels, int precision,
                                int64_t* bytes) {
  MG_REQUIRE(bytes != nullptr, "bytes is NULL");
  if (check_shape(dim, res, n_levels, precision)) return -1;
  PoolPlan P;
  plan_pool(dim, res, n_levels, precision == F64 ? 8 : 4, precision == PLAIN ? 4 : 8, &P);
  *bytes = (int64_t)P.bytes;
  return 0;
}

extern "C" int mgpcg_destroy(mgpcg_solver_t* s) {
  if (!s) return 0;
  if (s->engine) {
    delete s->engine;
    s->engine = nullptr;
  }
  if (s->pool) {
    (void)hipDeviceSynchronize();
    (void)hipFree(s->pool);
  }
  if (s->ev_retired) (void)hipEventDestroy(s->ev_retired);
  if (s->ev_submitted) (void)hipEventDestroy(s->ev_submitted);
  delete s;
  return 0;
}

extern "C" int mgpcg_set_stream(mgpcg_solver_t* s, void* stream) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  s->user_stream = (hipStream_t)stream;
  if (s->engine) s->engine->user_stream = s->user_stream;
  return 0;
}

extern "C" int mgpcg_set_option(mgpcg_solver_t* s, int option, int value) {
  g_knob_epoch++;
  switch (option) {
    case MGPCG_OPT_COARSE_MAX_SLOTS:
      MG_REQUIRE(value >= 0, "coarse slot limit must be >= 0");
      g_coarse_max_slots = value;
      return 0;
    case MGPCG_OPT_MAX_GRID:
      MG_REQUIRE(value >= 8, "grid limit must be >= 8");
      g_max_grid = value;
      return 0;
    case MGPCG_OPT_FUSE_FIRST:
      MG_REQUIRE(value >= 0 && value <= 2, "first-pair fusion must be 0, 1 or 2");
      g_fuse_first = value;
      return 0;
    case MGPCG_OPT_FUSE_PROLONG:
      g_fuse_prolong = value != 0;
      return 0;
    case MGPCG_OPT_FUSE_P:
      g_fuse_p = value != 0;
      return 0;
    case MGPCG_OPT_COMPACT:
      g_compact = value != 0;
      return 0;
    case MGPCG_OPT_COARSE_LDS:
      g_coarse_lds = value != 0;
      return 0;
    case MGPCG_OPT_RESTRICT_ROWS:
      MG_REQUIRE(value >= 0 && value <= 2, "restriction mode must be 0, 1 or 2");
      g_restrict_mode = value;
      return 0;
    case MGPCG_OPT_RESTRICT_COARSE:
      MG_REQUIRE(value >= -1 && value <= 2, "coarse restriction mode must be -1, 0, 1 or 2");
      g_restrict_coarse = value;
      return 0;
    case MGPCG_OPT_PROBE_EVERY:
      MG_REQUIRE(value >= 1, "probe period must be >= 1");
      g_probe_every = value;
      return 0;
    case MGPCG_OPT_SERVER_DEPTH:
      MG_REQUIRE(value >= 1 && value < Engine::NSLOT, "server depth must be in [1, %d]",
                 Engine::NSLOT - 1);
      g_server_depth = value;
      return 0;
    case MGPCG_OPT_SERVER_STREAMS:
      MG_REQUIRE(value >= 1 && value <= MAX_LANES, "server streams must be in [1, %d]",
                 MAX_LANES);
      g_server_streams = value;
      return 0;
    case MGPCG_OPT_GRAPH:
      g_graph = value != 0;
      return 0;
    default:
      (void)s;
      mg_set_error("unknown option %d", option);
      return -1;
  }
}

extern "C" int mgpcg_get_option(const mgpcg_solver_t* s, int option, int* value) {
  (void)s;
  MG_REQUIRE(value != nullptr, "value is NULL");
  switch (option) {
    case MGPCG_OPT_COARSE_MAX_SLOTS: *value = g_coarse_max_slots; return 0;
    case MGPCG_OPT_MAX_GRID: *value = g_max_grid; return 0;
    case MGPCG_OPT_FUSE_FIRST: *value = g_fuse_first; return 0;
    case MGPCG_OPT_FUSE_PROLONG: *value = g_fuse_prolong; return 0;
    case MGPCG_OPT_FUSE_P: *value = g_fuse_p; return 0;
    case MGPCG_OPT_COMPACT: *value = g_compact; return 0;
    case MGPCG_OPT_COARSE_LDS: *value = g_coarse_lds; return 0;
    case MGPCG_OPT_RESTRICT_ROWS: *value = g_restrict_mode; return 0;
    case MGPCG_OPT_RESTRICT_COARSE: *value = g_restrict_coarse; return 0;
    case MGPCG_OPT_PROBE_EVERY: *value = g_probe_every; return 0;
    case MGPCG_OPT_SERVER_DEPTH: *value = g_server_depth; return 0;
    case MGPCG_OPT_SERVER_STREAMS: *value = g_server_streams; return 0;
    case MGPCG_OPT_GRAPH: *value = g_graph; return 0;
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
    out6[5] = s->coarse_from;
  }
  if (bytes) *bytes = (int64_t)s->pool_bytes;
  return 0;
}

namespace {
// short stream-ordered section on the system's caller stream
struct UserOrder {
  hipStream_t st;
  explicit UserOrder(mgpcg_solver* s) : st(s->user_stream) {}
};
}  // namespace

extern "C" int mgpcg_load_level(mgpcg_solver_t* s, int level, const int8_t* grid_type,
                                const float* Adiag, const float* Ax) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  MG_REQUIRE(level >= 0 && level < s->L, "level %d out of range [0, %d)", level, s->L);
  MG_REQUIRE(grid_type && Adiag && Ax, "NULL field pointer");
  MG_HIP(hipSetDevice(s->device));
  const Geo& g = s->geo[level];
  uint32_t* part = s->lv[level].part;
  const dim3 grid(nblk(g.size)), block(BLOCK);
  if (s->precision == F64) {
    if (s->dim == 3)
      hipLaunchKernelGGL((k_pack_level<double, 3>), grid, block, 0, s->user_stream, g,
                         grid_type, Adiag, Ax, level_ptrs<double>(s, level), part);
    else
      hipLaunchKernelGGL((k_pack_level<double, 2>), grid, block, 0, s->user_stream, g,
                         grid_type, Adiag, Ax, level_ptrs<double>(s, level), part);
  } else {
    if (s->dim == 3)
      hipLaunchKernelGGL((k_pack_level<float, 3>), grid, block, 0, s->user_stream, g,
                         grid_type, Adiag, Ax, level_ptrs<float>(s, level), part);
    else
      hipLaunchKernelGGL((k_pack_level<float, 2>), grid, block, 0, s->user_stream, g,
                         grid_type, Adiag, Ax, level_ptrs<float>(s, level), part);
  }
  MetaReduceList ml;
  ml.meta[0] = s->lv[level].meta;
  ml.part[0] = part;
  ml.nblocks[0] = nblk(g.size);
  hipLaunchKernelGGL(k_meta_reduce, dim3(1), dim3(META_THREADS), 0, s->user_stream, ml);
  if (level == 0 && s->coarse_from > 0) {
    const LevelBuf& B = s->lv[level];
    hipLaunchKernelGGL(k_level_order, dim3(1), dim3(ORDER_THREADS), 0, s->user_stream, B.wet,
                       nblk(g.size), B.ord[1], B.ordtmp);
  }
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_level_flags(mgpcg_solver_t* s, int level, uint16_t* flags, uint8_t* wet) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  MG_REQUIRE(level >= 0 && level < s->L, "level %d out of range [0, %d)", level, s->L);
  MG_REQUIRE(flags != nullptr && wet != nullptr, "NULL argument");
  MG_HIP(hipSetDevice(s->device));
  const Geo& g = s->geo[level];
  const LevelBuf& B = s->lv[level];
  for (int c = 0; c < 2; ++c)
    MG_HIP(hipMemcpyAsync(flags + (size_t)c * g.size, B.nb[c], sizeof(uint16_t) * g.size,
                          hipMemcpyDeviceToDevice, s->user_stream));
  MG_HIP(hipMemcpyAsync(wet, B.wet, nblk(g.size), hipMemcpyDeviceToDevice, s->user_stream));
  MG_HIP(hipStreamSynchronize(s->user_stream));
  return 0;
}

extern "C" int mgpcg_level_order(mgpcg_solver_t* s, int level, int32_t* ord1, int32_t* ord2) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  MG_REQUIRE(level >= 0 && level < s->L, "level %d out of range [0, %d)", level, s->L);
  MG_REQUIRE(ord1 != nullptr, "NULL argument");
  MG_HIP(hipSetDevice(s->device));
  const Geo& g = s->geo[level];
  const LevelBuf& B = s->lv[level];
  MG_HIP(hipMemcpyAsync(ord1, B.ord[0], sizeof(int32_t) * nblk(g.size), hipMemcpyDeviceToDevice,
                        s->user_stream));
  if (ord2 != nullptr && level == 0 && s->coarse_from > 0)
    MG_HIP(hipMemcpyAsync(ord2, B.ord[1], sizeof(int32_t) * nsweep(g.size),
                          hipMemcpyDeviceToDevice, s->user_stream));
  MG_HIP(hipStreamSynchronize(s->user_stream));
  return 0;
}

extern "C" int mgpcg_level_coupling(mgpcg_solver_t* s, int level, int* uniform, float* coef) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  MG_REQUIRE(level >= 0 && level < s->L, "level %d out of range [0, %d)", level, s->L);
  MG_HIP(hipSetDevice(s->device));
  MG_HIP(hipStreamSynchronize(s->user_stream));
  LevelMeta m;
  MG_HIP(hipMemcpy(&m, s->lv[level].meta, sizeof(m), hipMemcpyDeviceToHost));
  const bool uni = m.maxbits == 0u || m.minbits == m.maxbits;
  if (uniform) *uniform = uni ? 1 : 0;
  if (coef) {
    uint32_t bits = (uni && m.maxbits != 0u) ? m.minbits : 0u;
    float f;
    memcpy(&f, &bits, sizeof f);
    *coef = f;
  }
  return 0;
}

extern "C" int mgpcg_load_rhs(mgpcg_solver_t* s, const float* b) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  MG_REQUIRE(b != nullptr, "b is NULL");
  MG_HIP(hipSetDevice(s->device));
  const Geo& g = s->geo[0];
  if (s->precision == F64)
    hipLaunchKernelGGL((k_pack_vec<float, double>), dim3(nblk(g.size)), dim3(BLOCK), 0,
                       s->user_stream, g, b, (double*)s->b[0], (double*)s->b[1]);
  else
    hipLaunchKernelGGL((k_pack_vec<float, float>), dim3(nblk(g.size)), dim3(BLOCK), 0,
                       s->user_stream, g, b, (float*)s->b[0], (float*)s->b[1]);
  MG_HIP(hipGetLastError());
  return 0;
}

namespace {
template <typename T>
int assemble_pressure_t(mgpcg_solver* s, int8_t* const* gt, const float* const* velocity,
                        const float* phi, float sA, float sb, float p0, int gfm, float c_max,
                        int add_c, float cadd) {
  hipStream_t st = s->user_stream;
  const Geo& g0 = s->geo[0];
  const int L = s->L;
  // the coarse levels' cell types, level by level (init_gridtype, mgpcg.py:46-60)
  for (int l = 1; l < L; ++l) {
    const Geo& g = s->geo[l];
    const Geo& gf = s->geo[l - 1];
    const int nc = (s->dim == 3 ? g.n0 : 1) * g.n1 * g.n2;
    hipLaunchKernelGGL(k_coarsen, dim3(nblk(nc)), dim3(BLOCK), 0, st, s->dim,
                       s->dim == 3 ? gf.n0 : 1, gf.n1, gf.n2, gt[l - 1], gt[l], nc);
  }
  // b and every level's operator, flags and block records: one launch
  PackCall c;
  c.nl = L;
  c.block0[0] = 0;
  for (int l = 0; l < L; ++l) {
    c.gt[l] = gt[l];
    c.block0[l + 1] = c.block0[l] + nblk(s->geo[l].size);
  }
  for (int k = 0; k < 3; ++k) c.V.u[k] = k < s->dim ? velocity[k] : nullptr;
  c.phi = gfm ? phi : nullptr;
  c.b[0] = s->b[0];
  c.b[1] = s->b[1];
  c.sA = sA;
  c.sb = sb;
  c.p0 = p0;
  c.c_max = c_max;
  c.cadd = cadd;
  c.add_c = add_c;
  const PackLevel<T>* tab = (const PackLevel<T>*)s->packtab;
  if (s->dim == 3)
    hipLaunchKernelGGL((k_pack_pressure_all<T, 3>), dim3(c.block0[L]), dim3(BLOCK), 0, st, tab,
                       c);
  else
    hipLaunchKernelGGL((k_pack_pressure_all<T, 2>), dim3(c.block0[L]), dim3(BLOCK), 0, st, tab,
                       c);
  // every level's record, and level 0's half-sweep order: one launch
  MetaReduceList ml;
  for (int l = 0; l < L; ++l) {
    ml.meta[l] = s->lv[l].meta;
    ml.part[l] = s->lv[l].part;
    ml.nblocks[l] = nblk(s->geo[l].size);
  }
  const LevelBuf& B = s->lv[0];
  const bool order = s->coarse_from > 0;
  hipLaunchKernelGGL(k_finish_levels, dim3(L + (order ? 1 : 0)), dim3(META_THREADS), 0, st, ml,
                     L, B.wet, nblk(g0.size), order ? B.ord[1] : nullptr, B.ordtmp);
  MG_HIP(hipGetLastError());
  return 0;
}
}  // namespace

extern "C" int mgpcg_assemble_pressure(mgpcg_solver_t* s, int8_t* const* grid_type,
                                       const float* const* velocity, const float* phi,
                                       double scale_A, double scale_b, double p0,
                                       int ghost_fluid_method, double c_max, int add_c,
                                       double c) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  MG_REQUIRE(grid_type != nullptr && velocity != nullptr, "NULL argument");
  for (int l = 0; l < s->L; ++l) MG_REQUIRE(grid_type[l] != nullptr, "grid_type[%d] is NULL", l);
  for (int k = 0; k < s->dim; ++k) MG_REQUIRE(velocity[k], "velocity[%d] is NULL", k);
  MG_REQUIRE(!ghost_fluid_method || phi, "ghost-fluid method needs phi");
  MG_HIP(hipSetDevice(s->device));
  const float sA = (float)scale_A, sb = (float)scale_b, fp0 = (float)p0, cm = (float)c_max;
  if (s->precision == F64)
    return assemble_pressure_t<double>(s, grid_type, velocity, phi, sA, sb, fp0,
                                       ghost_fluid_method, cm, add_c, (float)c);
  return assemble_pressure_t<float>(s, grid_type, velocity, phi, sA, sb, fp0, ghost_fluid_method,
                                    cm, add_c, (float)c);
}

extern "C" int mgpcg_solve(mgpcg_solver_t* s, int max_iters, double rel_tol, double abs_tol,
                           double eps, int refine_every, mgpcg_result_t* out) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  return solve_single(s, max_iters, rel_tol, abs_tol, eps, refine_every, out, false);
}

extern "C" int mgpcg_solve_hostsync(mgpcg_solver_t* s, int max_iters, double rel_tol,
                                    double abs_tol, double eps, int refine_every,
                                    mgpcg_result_t* out) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  return solve_single(s, max_iters, rel_tol, abs_tol, eps, refine_every, out, true);
}

extern "C" int mgpcg_store_x(mgpcg_solver_t* s, float* x_f32, double* x_f64) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  MG_REQUIRE((x_f32 != nullptr) != (x_f64 != nullptr), "exactly one output pointer required");
  MG_HIP(hipSetDevice(s->device));
  const Geo& g = s->geo[0];
  const dim3 gr(nblk(g.size)), bl(BLOCK);
  hipStream_t st = s->user_stream;
  if (s->xsz == 4) {
    if (x_f32)
      hipLaunchKernelGGL((k_unpack_vec<float, float>), gr, bl, 0, st, g, (const float*)s->x[0],
                         (const float*)s->x[1], x_f32);
    else
      hipLaunchKernelGGL((k_unpack_vec<float, double>), gr, bl, 0, st, g, (const float*)s->x[0],
                         (const float*)s->x[1], x_f64);
  } else {
    if (x_f32)
      hipLaunchKernelGGL((k_unpack_vec<double, float>), gr, bl, 0, st, g,
                         (const double*)s->x[0], (const double*)s->x[1], x_f32);
    else
      hipLaunchKernelGGL((k_unpack_vec<double, double>), gr, bl, 0, st, g,
                         (const double*)s->x[0], (const double*)s->x[1], x_f64);
  }
  MG_HIP(hipGetLastError());
  return 0;
}

// ===========================================================================
// The plain FP32 GPU port structured like mgpcg.py (the baseline)
// ===========================================================================
namespace {

template <typename T, typename TX, int DIM>
struct RefPath {
  static int run(Engine* e, mgpcg_solver* s, int max_iters, double rel_tol, double abs_tol,
                 double eps, mgpcg_result_t* out) {
    using Sd = Sched<T, TX, DIM>;
    const int L = s->L, nu = s->nu;
    const SysDesc<T, TX>* sysd = Sd::sys(e);
    const LevelDesc<T>* lv = Sd::levels(e);
    const int nb0 = nblk(s->geo[0].size);
    // one item per level: LevelDesc l of member 0
    std::vector<Item> it(L);
    for (int l = 0; l < L; ++l) it[l] = Item{l, 0};
    MG_HIP(hipStreamSynchronize(e->stream));
    MG_HIP(hipMemcpy(e->d_items, it.data(), sizeof(Item) * L, hipMemcpyHostToDevice));
    e->dirty = true;
    double* dsum = (double*)e->d_states;  // scratch for the reduced scalar
    auto lvl = [&](int l, bool coarse) {
      Engine::StageRef r;
      r.off = l;
      r.n = 1;
      r.nblocks = coarse ? restrict_blocks(e, s->geo[l + 1]) : nblk(s->geo[l].size);
      return r;
    };
    auto sweep = [&](int l) {  // a half-sweep stage of level l
      Engine::StageRef r;
      r.off = l;
      r.n = 1;
      r.nblocks = nsweep(s->geo[l].size);
      return r;
    };
    auto fill = [&](int l, int which) {
      hipLaunchKernelGGL((k_ref_fill<T>), dim3(nblk(s->geo[l].size)), dim3(BLOCK), 0, e->stream,
                         lv, l, which);
    };
    // reduce(a, b) followed by the read of self.sum[None] (a host round trip)
    auto reduce = [&](int a, int b, double* v) -> int {
      hipLaunchKernelGGL((k_ref_dot<T, TX>), dim3(nb0), dim3(BLOCK), 0, e->stream, sysd, a, b);
      hipLaunchKernelGGL((k_ref_sum<T, TX>), dim3(1), dim3(BLOCK), 0, e->stream, sysd, dsum);
      MG_HIP(hipMemcpyAsync(v, dsum, sizeof(double), hipMemcpyDeviceToHost, e->stream));
      MG_HIP(hipStreamSynchronize(e->stream));
      return 0;
    };
    auto write_f32 = [&](float* dst, float v) -> int {
      MG_HIP(hipMemcpyAsync(dst, &v, sizeof(float), hipMemcpyHostToDevice, e->stream));
      MG_HIP(hipStreamSynchronize(e->stream));
      return 0;
    };
    auto vcycle = [&]() {
      fill(0, 1);  // z[0].fill(0)
      for (int l = 0; l + 1 < L; ++l) {
        for (int i = 0; i < nu; ++i) {
          Sd::template stage<ST_SMOOTH0>(e, sweep(l));
          Sd::template stage<ST_SMOOTH1>(e, sweep(l));
        }
        fill(l + 1, 3);  // r[l+1].fill(0); z[l+1].fill(0)
        Sd::restrict_(e, e->restrict_mode, lvl(l, false), lvl(l, true));
      }
      const int half = s->bottom / 2;
      for (int i = 0; i < half; ++i) {
        Sd::template stage<ST_SMOOTH0>(e, sweep(L - 1));
        Sd::template stage<ST_SMOOTH1>(e, sweep(L - 1));
      }
      for (int i = 0; i < half; ++i) {
        Sd::template stage<ST_SMOOTH1>(e, sweep(L - 1));
        Sd::template stage<ST_SMOOTH0>(e, sweep(L - 1));
      }
      for (int l = L - 2; l >= 0; --l) {
        Sd::template stage<ST_PROLONGATE>(e, lvl(l, false));
        for (int i = 0; i < nu; ++i) {
          Sd::template stage<ST_SMOOTH1>(e, sweep(l));
          Sd::template stage<ST_SMOOTH0>(e, sweep(l));
        }
      }
    };
    float* alpha32 = &s->st->alpha32;
    float* beta32 = &s->st->beta32;
    if (write_f32(beta32, 0.0f)) return -2;
    hipLaunchKernelGGL((k_ref_copy_b<T, TX>), dim3(nb0), dim3(BLOCK), 0, e->stream, sysd);
    double r0r0 = 0, old_zr = 0, pAp = 0, rr = 0, new_zr = 0;
    if (reduce(RV_R0, RV_R0, &r0r0)) return -2;
    const double tol = std::max(abs_tol, r0r0 * rel_tol);
    vcycle();
    hipLaunchKernelGGL((k_ref_update_p<T, TX>), dim3(nb0), dim3(BLOCK), 0, e->stream, sysd);
    if (reduce(RV_Z0, RV_R0, &old_zr)) return -2;
    int iter = 0, status = MAX_ITERS;
    double alpha = 0, beta = 0;
    while (max_iters == -1 || iter < max_iters) {
      hipLaunchKernelGGL((k_ref_compute_Ap<T, TX, DIM>), dim3(nb0), dim3(BLOCK), 0, e->stream,
                         sysd);
      if (reduce(RV_P, RV_AP, &pAp)) return -2;
      alpha = old_zr / (pAp + eps);
      if (write_f32(alpha32, (float)alpha)) return -2;
      hipLaunchKernelGGL((k_ref_update_xr<T, TX>), dim3(nb0), dim3(BLOCK), 0, e->stream, sysd);
      if (reduce(RV_R0, RV_R0, &rr)) return -2;
      if (rr < tol) {
        status = CONVERGED;
        break;
      }
      vcycle();
      if (reduce(RV_Z0, RV_R0, &new_zr)) return -2;
      beta = new_zr / (old_zr + eps);
      if (write_f32(beta32, (float)beta)) return -2;
      hipLaunchKernelGGL((k_ref_update_p<T, TX>), dim3(nb0), dim3(BLOCK), 0, e->stream, sysd);
      old_zr = new_zr;
      iter++;
    }
    if (out) {
      out->iters = iter;
      out->status = status;
      out->r0r0 = r0r0;
      out->rr = rr;
      out->true_rr = -1.0;
      out->tol = tol;
      out->alpha = alpha;
      out->beta = beta;
    }
    MG_HIP(hipGetLastError());
    return 0;
  }
};

}  // namespace

extern "C" int mgpcg_solve_reference_structured(mgpcg_solver_t* s, int max_iters, double rel_tol,
                                                double abs_tol, double eps,
                                                mgpcg_result_t* out) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  MG_REQUIRE(s->precision == PLAIN, "the reference-structured path is the FP32 baseline");
  Engine* e = nullptr;
  if (single_engine(s, &e)) return -1;
  if (engine_begin(e)) return -2;
  e->mem.clear();
  const SolveParams prm{max_iters, rel_tol, abs_tol, eps, 0};
  // the baseline reads every coefficient, as mgpcg.py does
  s->full_coupling = true;
  const int added = engine_add(e, s, 0, prm);
  s->full_coupling = false;
  if (added < 0) return -2;
  if (engine_build_items(e)) return -2;  // allocates the item table
  int rc;
  if (s->dim == 3) rc = RefPath<float, float, 3>::run(e, s, max_iters, rel_tol, abs_tol, eps, out);
  else rc = RefPath<float, float, 2>::run(e, s, max_iters, rel_tol, abs_tol, eps, out);
  if (rc) return rc;
  MG_HIP(hipStreamSynchronize(e->stream));
  return engine_end(e);
}

// ===========================================================================
// C-ABI: kernel-level access for parity tests
// ===========================================================================
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
    case MGPCG_FIELD_P_NEXT: {
      MG_REQUIRE(level == 0, "p lives on level 0");
      const int par = field == MGPCG_FIELD_P ? 0 : 1;
      *c0 = s->p[par][0];
      *c1 = s->p[par][1];
      return 0;
    }
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
  const Geo& g = s->geo[level];
  if (esz == 4)
    hipLaunchKernelGGL((k_unpack_vec<float, float>), dim3(nblk(g.size)), dim3(BLOCK), 0,
                       s->user_stream, g, (const float*)c0, (const float*)c1, (float*)canon);
  else
    hipLaunchKernelGGL((k_unpack_vec<double, double>), dim3(nblk(g.size)), dim3(BLOCK), 0,
                       s->user_stream, g, (const double*)c0, (const double*)c1, (double*)canon);
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_debug_set(mgpcg_solver_t* s, int field, int level, const void* canon) {
  MG_REQUIRE(s != nullptr && canon != nullptr, "NULL argument");
  void *c0, *c1;
  size_t esz;
  int rc = vec_slots(s, field, level, &c0, &c1, &esz);
  if (rc) return rc;
  const Geo& g = s->geo[level];
  if (esz == 4)
    hipLaunchKernelGGL((k_pack_vec<float, float>), dim3(nblk(g.size)), dim3(BLOCK), 0,
                       s->user_stream, g, (const float*)canon, (float*)c0, (float*)c1);
  else
    hipLaunchKernelGGL((k_pack_vec<double, double>), dim3(nblk(g.size)), dim3(BLOCK), 0,
                       s->user_stream, g, (const double*)canon, (double*)c0, (double*)c1);
  MG_HIP(hipGetLastError());
  return 0;
}

namespace {

// A one-item stage over level `level` of member 0 (item table slot 0).
int prepare_op(Engine* e, const mgpcg_solver* s, int op, int level, Engine::StageRef* r) {
  Item it;
  it.desc = level;  // member 0: LevelDesc index = level
  it.block0 = 0;
  int nb = nblk(s->geo[level].size);
  if (op == MGPCG_OP_SMOOTH) nb = nsweep(s->geo[level].size);
  if (op == MGPCG_OP_RESTRICT) {
    e->restrict_mode = g_restrict_mode;
    if (e->restrict_mode == 1 && s->geo[0].n2 > RESTRICT_ROW_MAX) e->restrict_mode = 0;
    nb = restrict_blocks(e, s->geo[level + 1]);
    r->fine_nblocks = nblk(s->geo[level].size);
  }
  if (op == MGPCG_OP_COMPUTE_AP) it.desc = 0;  // SysDesc of member 0
  MG_HIP(hipStreamSynchronize(e->stream));
  MG_HIP(hipMemcpy(e->d_items, &it, sizeof(Item), hipMemcpyHostToDevice));
  r->off = 0;
  r->n = 1;
  r->nblocks = nb;
  e->dirty = true;  // the item table no longer matches the member set
  return 0;
}

// Launch one stage on the prepared one-item table.  level0: a half-sweep of
// level 0 runs the level-0 instantiation, as in a solve.
template <typename T, typename TX, int DIM>
struct RunOp {
  static int run(Engine* e, int op, const Engine::StageRef& r, int arg, bool level0) {
    using Sd = Sched<T, TX, DIM>;
    const int nb = r.nblocks;
    switch (op) {
      case MGPCG_OP_SMOOTH:
        if (arg & 1) Sd::template stage<ST_SMOOTH1>(e, r, level0);
        else Sd::template stage<ST_SMOOTH0>(e, r, level0);
        return 0;
      case MGPCG_OP_SMOOTH_PAIR_FROM_ZERO:
        Sd::template stage<ST_PAIR_FROM_ZERO>(e, r);
        return 0;
      case MGPCG_OP_PROLONG_SMOOTH1:
        Sd::template stage<ST_PROLONG_SMOOTH1>(e, r);
        return 0;
      case MGPCG_OP_RESTRICT: {
        Engine::StageRef fine = r;
        fine.nblocks = r.fine_nblocks;
        Sd::restrict_(e, e->restrict_mode, fine, r);
        return 0;
      }
      case MGPCG_OP_PROLONGATE:
        Sd::template stage<ST_PROLONGATE>(e, r);
        return 0;
      case MGPCG_OP_COMPUTE_AP:
        hipLaunchKernelGGL((k_cg_ap<T, TX, DIM>), dim3(std::min(nb, g_max_grid)), dim3(BLOCK), 0,
                           e->stream, Sd::sys(e), stage_of(e, r));
        return 0;
      default:
        mg_set_error("unknown op %d", op);
        return -1;
    }
  }
};

}  // namespace

extern "C" int mgpcg_debug_op(mgpcg_solver_t* s, int op, int level, int arg) {
  MG_REQUIRE(s != nullptr, "solver is NULL");
  MG_REQUIRE(level >= 0 && level < s->L, "level %d out of range", level);
  if (op == MGPCG_OP_RESTRICT || op == MGPCG_OP_PROLONGATE || op == MGPCG_OP_PROLONG_SMOOTH1)
    MG_REQUIRE(level < s->L - 1, "no coarser level below %d", level);
  Engine* e = nullptr;
  if (single_engine(s, &e)) return -1;
  if (engine_begin(e)) return -2;
  // a fresh state (not done, iteration 0, beta 0) and a one-member table
  for (Member& M : e->mem) {
    M.active = 0;
    M.s = nullptr;
  }
  e->mem.clear();
  const SolveParams prm{500, 1e-12, 1e-14, 1e-12, 10};
  if (engine_add(e, s, 0, prm) < 0) return -2;
  if (engine_build_items(e)) return -2;
  MG_HIP(hipStreamSynchronize(e->stream));
  if (op == MGPCG_OP_VCYCLE) {
    dispatch<RunVcycle>(e->precision, e->dim, e);
  } else {
    Engine::StageRef r;
    if (prepare_op(e, s, op, level, &r)) return -2;
    int rc = dispatch<RunOp>(e->precision, e->dim, e, op, r, arg, level == 0);
    if (rc) return rc;
  }
  MG_HIP(hipGetLastError());
  MG_HIP(hipStreamSynchronize(e->stream));
  e->dirty = true;  // the item table was overwritten
  return engine_end(e);
}

extern "C" int mgpcg_time_op(mgpcg_solver_t* s, int op, int level, int arg, int reps,
                             double* ms_per_launch) {
  MG_REQUIRE(s != nullptr && ms_per_launch != nullptr, "NULL argument");
  MG_REQUIRE(reps >= 1, "reps must be >= 1");
  MG_REQUIRE(op == MGPCG_OP_SMOOTH || op == MGPCG_OP_RESTRICT || op == MGPCG_OP_VCYCLE ||
                 op == MGPCG_OP_PROLONG_SMOOTH1 || op == MGPCG_OP_SMOOTH_PAIR_FROM_ZERO,
             "op %d cannot be timed", op);
  // warm-up launch (also sets up the one-member table)
  int rc = mgpcg_debug_op(s, op, level, arg);
  if (rc) return rc;
  Engine* e = s->engine;
  hipEvent_t e0, e1;
  MG_HIP(hipEventCreate(&e0));
  MG_HIP(hipEventCreate(&e1));
  Engine::StageRef r;
  if (op != MGPCG_OP_VCYCLE) {
    if (prepare_op(e, s, op, level, &r)) return -2;
  } else if (engine_build_items(e)) {
    return -2;
  }
  MG_HIP(hipStreamSynchronize(e->stream));
  MG_HIP(hipEventRecord(e0, e->stream));
  for (int i = 0; i < reps; ++i) {
    if (op == MGPCG_OP_VCYCLE) {
      dispatch<RunVcycle>(e->precision, e->dim, e);
    } else {
      rc = dispatch<RunOp>(e->precision, e->dim, e, op, r, arg, level == 0);
      if (rc) return rc;
    }
  }
  MG_HIP(hipEventRecord(e1, e->stream));
  MG_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  MG_HIP(hipEventElapsedTime(&ms, e0, e1));
  *ms_per_launch = (double)ms / reps;
  MG_HIP(hipEventDestroy(e0));
  MG_HIP(hipEventDestroy(e1));
  e->dirty = true;
  MG_HIP(hipGetLastError());
  return 0;
}

// ===========================================================================
// C-ABI: batches and the server
// ===========================================================================
struct mgpcg_batch {
  Engine* e = nullptr;
};

extern "C" int mgpcg_batch_create(mgpcg_batch_t** out) {
  MG_REQUIRE(out != nullptr, "out is NULL");
  mgpcg_batch* b = new (std::nothrow) mgpcg_batch();
  MG_REQUIRE(b != nullptr, "out of host memory");
  b->e = new_engine();
  if (!b->e) {
    delete b;
    return -2;
  }
  *out = b;
  return 0;
}

extern "C" int mgpcg_batch_destroy(mgpcg_batch_t* b) {
  if (!b) return 0;
  delete b->e;
  delete b;
  return 0;
}

extern "C" int mgpcg_batch_set_stream(mgpcg_batch_t* b, void* stream) {
  MG_REQUIRE(b != nullptr, "batch is NULL");
  b->e->user_stream = (hipStream_t)stream;
  return 0;
}

extern "C" int mgpcg_batch_solve(mgpcg_batch_t* b, mgpcg_solver_t* const* systems, int n,
                                 int max_iters, double rel_tol, double abs_tol, double eps,
                                 int refine_every, mgpcg_result_t* results) {
  MG_REQUIRE(b != nullptr && systems != nullptr, "NULL argument");
  MG_REQUIRE(n >= 1, "n must be >= 1");
  for (int i = 0; i < n; ++i) MG_REQUIRE(systems[i] != nullptr, "systems[%d] is NULL", i);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j)
      MG_REQUIRE(systems[i] != systems[j], "system %d appears twice in the batch", i);
  Engine* e = b->e;
  if (engine_begin(e)) return -2;
  e->mem.clear();
  e->dirty = true;
  for (int i = 0; i < n; ++i) {
    const SolveParams prm{max_iters, rel_tol, abs_tol, eps, refine_every};
    const int slot = engine_add(e, systems[i], i, prm);
    if (slot < 0) return slot == -1 ? -1 : -2;
  }
  std::vector<Finished> done;
  const long cap = max_iters >= 0 ? (long)max_iters + 4 : 10000000L;
  if (engine_run(e, cap, &done)) return -2;
  if (results) {
    for (const Finished& f : done) fill_result(f.st, &results[e->mem[f.member].tag]);
    // systems the host cap stopped
    for (const Member& M : e->mem)
      if (M.active) {
        CGState st;
        MG_HIP(hipMemcpy(&st, M.s->st, sizeof(CGState), hipMemcpyDeviceToHost));
        fill_result(st, &results[M.tag]);
      }
  }
  e->mem.clear();
  e->dirty = true;
  MG_HIP(hipGetLastError());
  return engine_end(e);
}

// Server lanes run on streams from a per-device pool, created together the
// first time a server on that device is made and kept until the process ends.
// The runtime maps each stream onto one of a few hardware queues when the
// stream is created, and two lanes on one queue run one after the other:
// lanes created afresh for each server were measured sharing one queue in a
// process that had made a server before (two lanes then ran at two thirds of
// one lane's rate), while streams created together went to distinct queues.
// Lane i of every server uses pool stream i, so servers alive at the same time
// share lane streams (correct, but their iterations then take turns).
namespace {
int lane_stream(int lane, hipStream_t* out) {
  static std::mutex mu;
  static std::vector<std::vector<hipStream_t>> pool;  // [device][lane]
  int dev = 0;
  MG_HIP(hipGetDevice(&dev));
  MG_REQUIRE(lane >= 0 && lane < MAX_LANES, "lane %d outside [0, %d)", lane, MAX_LANES);
  std::lock_guard<std::mutex> lock(mu);
  if ((int)pool.size() <= dev) pool.resize(dev + 1);
  if (pool[dev].empty()) {
    std::vector<hipStream_t> v(MAX_LANES, nullptr);
    for (hipStream_t& st : v) {
      const hipError_t rc = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
      if (rc != hipSuccess) {
        for (hipStream_t x : v)
          if (x) (void)hipStreamDestroy(x);
        MG_HIP(rc);
      }
    }
    pool[dev] = v;
  }
  *out = pool[dev][lane];
  return 0;
}
}  // namespace

// The server: continuous batching.  Systems are submitted at any time; each
// call of mgpcg_server_step admits the submitted ones (they join at the next
// iteration boundary), enqueues one iteration for every member and retires
// the ones whose state (read back a few iterations behind) says they are done.
// Members are spread over lanes, each an engine with its own stream, so one
// lane's latency-bound coarse levels and reductions overlap another lane's
// bandwidth-bound level-0 kernels on the device.
struct mgpcg_server {
  struct Lane {
    Engine* e = nullptr;
    std::deque<int> inflight;  // readback slots of the queued iterations, oldest first
    int next_slot = 0;
  };
  std::vector<Lane> lanes;
  std::deque<std::pair<mgpcg_solver*, int64_t>> pending;
  std::vector<std::pair<int64_t, mgpcg_result_t>> finished;
  struct Params {
    int max_iters = 500;
    double rel_tol = 1e-12, abs_tol = 1e-14, eps = 1e-12;
    int refine_every = 10;
  } prm;
  ~mgpcg_server() {
    for (Lane& ln : lanes) delete ln.e;
  }
};

extern "C" int mgpcg_server_create(int max_iters, double rel_tol, double abs_tol, double eps,
                                   int refine_every, mgpcg_server_t** out) {
  MG_REQUIRE(out != nullptr, "out is NULL");
  mgpcg_server* sv = new (std::nothrow) mgpcg_server();
  MG_REQUIRE(sv != nullptr, "out of host memory");
  sv->lanes.resize(g_server_streams);
  for (int i = 0; i < (int)sv->lanes.size(); ++i) {
    mgpcg_server::Lane& ln = sv->lanes[i];
    hipStream_t st = nullptr;
    if (lane_stream(i, &st)) {
      delete sv;
      return -2;
    }
    ln.e = new_engine(st);
    if (!ln.e) {
      delete sv;
      return -2;
    }
  }
  sv->prm.max_iters = max_iters;
  sv->prm.rel_tol = rel_tol;
  sv->prm.abs_tol = abs_tol;
  sv->prm.eps = eps;
  sv->prm.refine_every = refine_every;
  *out = sv;
  return 0;
}

extern "C" int mgpcg_server_destroy(mgpcg_server_t* sv) {
  delete sv;
  return 0;
}

extern "C" int mgpcg_server_set_stream(mgpcg_server_t* sv, void* stream) {
  MG_REQUIRE(sv != nullptr, "server is NULL");
  for (mgpcg_server::Lane& ln : sv->lanes) ln.e->user_stream = (hipStream_t)stream;
  return 0;
}

extern "C" int mgpcg_server_submit(mgpcg_server_t* sv, mgpcg_solver_t* s, int64_t tag) {
  MG_REQUIRE(sv != nullptr && s != nullptr, "NULL argument");
  for (const mgpcg_server::Lane& ln : sv->lanes)
    for (const Member& M : ln.e->mem)
      MG_REQUIRE(!(M.active && M.s == s), "system is already being solved");
  for (const auto& pr : sv->pending)
    MG_REQUIRE(pr.first != s, "system is already submitted");
  MG_HIP(hipSetDevice(s->device));
  if (!s->ev_submitted) MG_HIP(hipEventCreateWithFlags(&s->ev_submitted, hipEventDisableTiming));
  MG_HIP(hipEventRecord(s->ev_submitted, s->user_stream));
  sv->pending.push_back({s, tag});
  return 0;
}

extern "C" int mgpcg_server_step(mgpcg_server_t* sv, int* n_active) {
  MG_REQUIRE(sv != nullptr, "server is NULL");
  const int nl = (int)sv->lanes.size();
  // admit submitted systems at this iteration boundary, each into the lane
  // with the fewest members
  std::vector<int> count(nl);
  int running = 0;
  for (int i = 0; i < nl; ++i) running += (count[i] = engine_active(sv->lanes[i].e));
  // While systems are running, a submitted system joins once its assembly has
  // finished on the device: a lane made to wait for a running assembly would
  // hold up every member meanwhile.  With none running, all join at once (the
  // lane waits for their assemblies, which may run side by side on their own
  // streams).
  const bool wait_all = running == 0;
  std::vector<char> begun(nl, 0);
  decltype(sv->pending) later;
  while (!sv->pending.empty()) {
    auto pr = sv->pending.front();
    sv->pending.pop_front();
    if (!wait_all) {
      const hipError_t q = hipEventQuery(pr.first->ev_submitted);
      if (q == hipErrorNotReady) {
        later.push_back(pr);
        continue;
      }
      MG_HIP(q);
    }
    // the lane with the fewest members; among those, the one the system last
    // ran on (its queued iterations there need no wait across lanes)
    int li = 0;
    for (int i = 1; i < nl; ++i)
      if (count[i] < count[li]) li = i;
    for (int i = 0; i < nl; ++i)
      if (count[i] == count[li] && pr.first->retired_pending &&
          pr.first->retired_on == sv->lanes[i].e)
        li = i;
    Engine* e = sv->lanes[li].e;
    // the lane also follows the server's stream (the caller's readbacks of the
    // systems' previous results come before their new iterations)
    if (!begun[li]) {
      if (engine_begin(e)) return -2;
      begun[li] = 1;
    }
    MG_HIP(hipStreamWaitEvent(e->stream, pr.first->ev_submitted, 0));
    // free member slots of retired systems are reused
    for (Member& M : e->mem)
      if (!M.active) M.s = nullptr;
    const SolveParams prm{sv->prm.max_iters, sv->prm.rel_tol, sv->prm.abs_tol, sv->prm.eps,
                          sv->prm.refine_every};
    if (engine_add(e, pr.first, pr.second, prm) < 0) return -2;
    ++count[li];
  }
  sv->pending = std::move(later);
  // every lane's iteration is enqueued before the host waits for any readback
  for (mgpcg_server::Lane& ln : sv->lanes) {
    if (engine_active(ln.e) == 0) continue;
    const int slot = ln.next_slot;
    ln.next_slot = (ln.next_slot + 1) % Engine::NSLOT;
    if (engine_enqueue(ln.e, slot, false)) return -2;
    ln.inflight.push_back(slot);
  }
  // Up to g_server_depth iterations per lane stay queued on the device while
  // the host returns to its caller (admissions, assembly); a member's
  // completion is seen that many iterations late, and its later iterations
  // skip it.
  int active = 0;
  for (mgpcg_server::Lane& ln : sv->lanes) {
    Engine* e = ln.e;
    std::vector<Finished> done;
    const size_t keep = engine_active(e) > 0 ? (size_t)g_server_depth : 0;
    while (ln.inflight.size() > keep) {
      if (engine_collect(e, ln.inflight.front(), &done)) return -2;
      ln.inflight.pop_front();
    }
    for (const Finished& f : done) {
      mgpcg_result_t r;
      fill_result(f.st, &r);
      sv->finished.push_back({e->mem[f.member].tag, r});
    }
    active += engine_active(e);
  }
  if (n_active) *n_active = active;
  // no ordering of the caller's stream after the engines: results are released
  // only after the host waited for the readback recorded behind their last
  // kernels, so the caller may read x at once
  MG_HIP(hipGetLastError());
  return 0;
}

extern "C" int mgpcg_server_poll(mgpcg_server_t* sv, int64_t* tags, mgpcg_result_t* results,
                                 int cap, int* n_out) {
  MG_REQUIRE(sv != nullptr && n_out != nullptr, "NULL argument");
  int n = std::min(cap, (int)sv->finished.size());
  for (int i = 0; i < n; ++i) {
    if (tags) tags[i] = sv->finished[i].first;
    if (results) results[i] = sv->finished[i].second;
  }
  sv->finished.erase(sv->finished.begin(), sv->finished.begin() + n);
  *n_out = n;
  return 0;
}

extern "C" int mgpcg_server_pending(const mgpcg_server_t* sv, int* n_pending, int* n_active,
                                    int* inflight) {
  MG_REQUIRE(sv != nullptr, "server is NULL");
  int act = 0, infl = 0;
  for (const mgpcg_server::Lane& ln : sv->lanes) {
    act += engine_active(ln.e);
    infl |= ln.inflight.empty() ? 0 : 1;
  }
  if (n_pending) *n_pending = (int)sv->pending.size();
  if (n_active) *n_active = act;
  if (inflight) *inflight = infl;
  return 0;
}

namespace {
// enable: start a fresh measurement; disable: stop and report it
int engine_probe(Engine* e, int enable, int* n, double* total_ms, int64_t* cells,
                 int64_t* fluid) {
  if (enable) {
    if (e->d_swept == nullptr) MG_HIP(hipMalloc(&e->d_swept, 2 * sizeof(unsigned long long)));
    const size_t half = (size_t)PROBE_POOL * SPAN_SLOTS;
    if (e->d_span == nullptr)
      MG_HIP(hipMalloc(&e->d_span, 2 * half * sizeof(unsigned long long)));
    MG_HIP(hipMemsetAsync(e->d_swept, 0, 2 * sizeof(unsigned long long), e->stream));
    // starts to the largest value (atomic min), ends to 0 (atomic max)
    MG_HIP(hipMemsetAsync(e->d_span, 0xff, half * sizeof(unsigned long long), e->stream));
    MG_HIP(hipMemsetAsync(e->d_span + half, 0, half * sizeof(unsigned long long), e->stream));
    e->probe_on = true;
    e->probe_used = 0;
    e->probe_seen = 0;
    return 0;
  }
  e->probe_on = false;
  MG_HIP(hipStreamSynchronize(e->stream));
  double tot = 0.0;
  if (e->probe_used > 0) {
    const size_t n = (size_t)e->probe_used * SPAN_SLOTS;
    std::vector<unsigned long long> t0(n), t1(n);
    MG_HIP(hipMemcpy(t0.data(), e->d_span, n * sizeof(unsigned long long),
                     hipMemcpyDeviceToHost));
    MG_HIP(hipMemcpy(t1.data(), e->d_span + (size_t)PROBE_POOL * SPAN_SLOTS,
                     n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    int dev = 0, khz = 0;
    MG_HIP(hipGetDevice(&dev));
    MG_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    MG_REQUIRE(khz > 0, "device clock rate unknown");
    for (int i = 0; i < e->probe_used; ++i) {
      unsigned long long a = ~0ull, b = 0ull;
      for (int k = 0; k < SPAN_SLOTS; ++k) {
        a = std::min(a, t0[(size_t)i * SPAN_SLOTS + k]);
        b = std::max(b, t1[(size_t)i * SPAN_SLOTS + k]);
      }
      if (b > a) tot += (double)(b - a) / (double)khz;  // ms
    }
  }
  unsigned long long c[2] = {0, 0};
  if (e->d_swept) MG_HIP(hipMemcpy(c, e->d_swept, sizeof(c), hipMemcpyDeviceToHost));
  if (n) *n = e->probe_used;
  if (total_ms) *total_ms = tot;
  if (cells) *cells = (int64_t)c[0];
  if (fluid) *fluid = (int64_t)c[1];
  return 0;
}
}  // namespace

extern "C" int mgpcg_server_probe(mgpcg_server_t* sv, int enable, int* n_launches,
                                  double* total_ms, int64_t* total_cells,
                                  int64_t* total_fluid) {
  MG_REQUIRE(sv != nullptr, "server is NULL");
  int n = 0;
  double ms = 0.0;
  int64_t cells = 0, fluid = 0;
  for (mgpcg_server::Lane& ln : sv->lanes) {
    int n1 = 0;
    double ms1 = 0.0;
    int64_t c1 = 0, f1 = 0;
    if (engine_probe(ln.e, enable, &n1, &ms1, &c1, &f1)) return -2;
    n += n1;
    ms += ms1;
    cells += c1;
    fluid += f1;
  }
  if (!enable) {
    if (n_launches) *n_launches = n;
    if (total_ms) *total_ms = ms;
    if (total_cells) *total_cells = cells;
    if (total_fluid) *total_fluid = fluid;
  }
  return 0;
}

extern "C" int mgpcg_graph_counts(const mgpcg_solver_t* s, const mgpcg_batch_t* b,
                                  const mgpcg_server_t* sv, int64_t* captures,
                                  int64_t* launches) {
  MG_REQUIRE((s != nullptr) + (b != nullptr) + (sv != nullptr) == 1,
             "exactly one of solver, batch and server must be given");
  int64_t c = 0, n = 0;
  auto add = [&](const Engine* e) {
    if (!e) return;
    c += e->graph_captures;
    n += e->graph_launches;
  };
  if (s) add(s->engine);
  if (b) add(b->e);
  if (sv)
    for (const mgpcg_server::Lane& ln : sv->lanes) add(ln.e);
  if (captures) *captures = c;
  if (launches) *launches = n;
  return 0;
}

extern "C" int mgpcg_batch_probe(mgpcg_batch_t* b, int enable, int* n_launches,
                                 double* total_ms, int64_t* total_cells, int64_t* total_fluid) {
  MG_REQUIRE(b != nullptr, "batch is NULL");
  return engine_probe(b->e, enable, n_launches, total_ms, total_cells, total_fluid);
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
