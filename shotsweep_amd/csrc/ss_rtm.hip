// One native call for the device part of Sim.gradient / Sim.images
// (openfwi/physics/gpu.py:152-210): the forward with its tape, the muted
// residual, the exact adjoint over the whole Laplacian history or over
// checkpoint segments with recompute (gpu.py:169-184), the padding fold with
// the source-amplitude term and the image epilogue.  The caller hands over
// coefficients, data and outputs as device pointers plus one workspace
// (ss_workspace_bytes); every intermediate lives in that workspace, so the
// library allocates no device memory.  The two stages can run in separate
// calls on the same workspace (forward with tape, later the adjoint from the
// tape with any data-space seed), which is what an autograd backward needs.
//
// Checkpoint segments (the reference's scheme): the adjoint walks the
// segments last first; segment k-1 is recomputed from its checkpoint on a
// library-owned side stream while segment k's adjoint runs on the caller's
// stream (two segment history buffers alternate), so the recompute fills CUs
// the adjoint leaves idle.  Each launch computes exactly what it computes in
// sequence, so results do not depend on the overlap.  Small batches (at most a
// quarter as many propagations as CUs), and planes too large for the LDS, run
// their segments on the tiled sweep kernels, every propagation split into tiles
// over all CUs.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "ss_kernels.h"

namespace ss {

namespace {

constexpr int64_t kAlign = 256;

template <typename T>
__global__ void zero_kernel(T *__restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = T(0);
}

// dst[i * dstride + k] = src[i * sstride + k] (0 when src is NULL) for i < n, k < len
template <typename T>
__global__ void copy_rows_kernel(T *__restrict__ dst, int64_t dstride, const T *__restrict__ src, int64_t sstride,
                                 int64_t n, int64_t len) {
  const int64_t total = n * len;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / len, k = e - i * len;
    dst[i * dstride + k] = src ? src[i * sstride + k] : T(0);
  }
}

int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

// Fill kernels rather than hipMemsetAsync, so that a caller may capture the call into a graph.
template <typename T>
int zero(T *p, int64_t n, hipStream_t st) {
  if (n <= 0) return SS_OK;
  zero_kernel<T><<<grid_for(n), 256, 0, st>>>(p, n);
  return hip_status(hipGetLastError(), "zero fill");
}

template <typename T>
int copy_rows(T *dst, int64_t dstride, const T *src, int64_t sstride, int64_t n, int64_t len, hipStream_t st) {
  if (n <= 0 || len <= 0) return SS_OK;
  copy_rows_kernel<T><<<grid_for(n * len), 256, 0, st>>>(dst, dstride, src, sstride, n, len);
  return hip_status(hipGetLastError(), "copy");
}

#define SS_TRY(expr)              \
  do {                            \
    int _rc = (expr);             \
    if (_rc != SS_OK) return _rc; \
  } while (0)

enum Impl { kStepwise = 1, kResident = 2, kTiled = 3 };

int device_cus() {
  static int cus = -1;
  if (cus < 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      v = 256;
    cus = v;
  }
  return cus;
}

// Time-loop implementation of a call.  Whole history: the resident kernels when the
// plane fits the LDS (FP32), else the stepwise ones.  Checkpoint segments: also the
// tiled kernels, chosen by 'auto' for batches of at most a quarter as many
// propagations as the GPU has CUs (one workgroup per propagation would leave most
// CUs idle) and for planes that do not fit the LDS.
template <typename T>
int choose_impl(const Geom &g, int segment, int req, Impl *out) {
  constexpr bool f32 = std::is_same<T, float>::value;
  const bool res_ok = f32 && resident_fits(g);
  const bool tiled_ok = f32 && segment < g.nt && sweep_fits(g);
  switch (req) {
    case SS_IMPL_STEPWISE:
      *out = kStepwise;
      return SS_OK;
    case SS_IMPL_RESIDENT:
      if (!res_ok) {
        set_error("resident kernels need FP32 and the padded plane in LDS");
        return SS_EUNSUP;
      }
      *out = kResident;
      return SS_OK;
    case SS_IMPL_TILED:
      if (!tiled_ok) {
        set_error("tiled segments need FP32, checkpoint segments and H*W <= 131072, pad <= 44, nt*H*W < 2^29");
        return SS_EUNSUP;
      }
      *out = kTiled;
      return SS_OK;
    case SS_IMPL_AUTO:
    case SS_IMPL_UNTILED:
      // resident segments cost about one resident round per segment pair whatever the batch (the
      // recompute runs on CUs the adjoint leaves idle); tiled segments scale with the batch and
      // win below about a quarter of the CUs (80 propagations on 256 CUs: 12.6 vs 13.4 ms per
      // gradient).  Planes that do not fit the LDS take the tiled kernels at any batch size.
      if (req == SS_IMPL_AUTO && tiled_ok && (!res_ok || 4 * (int64_t)g.nb * g.ns <= device_cus()))
        *out = kTiled;
      else
        *out = res_ok ? kResident : kStepwise;
      return SS_OK;
    default:
      set_error("unknown implementation selector");
      return SS_EINVAL;
  }
}

// Workspace regions (elements of T, each start aligned to 256 bytes).  The state
// block is shared: resident/stepwise segments use p0, p1, m1, m2 and two recompute
// states; tiled segments use two double-buffered sweep states of four planes each.
template <typename T>
struct Layout {
  int64_t bytes = 0;
  float *c2a = nullptr;      // (nb, H, W) c2 * A, FP32 only
  T *sim = nullptr;          // (nprop, nt, nx) gathers
  T *res = nullptr;          // (nprop, nt, nx) residual or seed
  T *sg = nullptr;           // (nprop) source-amplitude term
  T *gacc = nullptr;         // (nprop, H, W) imaging sum
  T *state = nullptr;        // 8 planes of (nprop, H, W)
  T *hist = nullptr;         // whole history (nprop, nt, H, W)
  T *ckpt = nullptr;         // (nprop, nck, 2, H, W)
  T *seg[2] = {nullptr, nullptr};  // (nprop, segment, H, W) each
  // tiled segments: coefficient, illumination and source-term copies at fixed
  // workspace addresses (the sweep's graphs are keyed by buffer addresses)
  float *tc2 = nullptr, *tamp = nullptr, *tsd = nullptr, *till = nullptr, *tsg = nullptr;

  T *plane(int k, int64_t nprop, int64_t HW) const { return state + k * nprop * HW; }
};

template <typename T>
void make_layout(const Geom &g, int segment, char *base, Layout<T> *L) {
  const int64_t nprop = (int64_t)g.nb * g.ns, HW = (int64_t)g.H * g.W, traces = (int64_t)g.nt * g.nx;
  const int64_t nzx = (int64_t)g.nz * g.nx;
  int64_t off = 0;
  auto take = [&](int64_t n, int64_t esize) -> char * {
    off = (off + kAlign - 1) / kAlign * kAlign;
    char *p = base ? base + off : nullptr;
    off += n * esize;
    return p;
  };
  const int64_t E = sizeof(T);
  if (std::is_same<T, float>::value) L->c2a = reinterpret_cast<float *>(take(g.nb * HW, 4));
  L->sim = reinterpret_cast<T *>(take(nprop * traces, E));
  L->res = reinterpret_cast<T *>(take(nprop * traces, E));
  L->sg = reinterpret_cast<T *>(take(nprop, E));
  L->gacc = reinterpret_cast<T *>(take(nprop * HW, E));
  L->state = reinterpret_cast<T *>(take(8 * nprop * HW, E));
  if (segment >= g.nt) {
    L->hist = reinterpret_cast<T *>(take(nprop * g.nt * HW, E));
  } else {
    const int64_t nck = (g.nt + segment - 1) / segment;
    L->ckpt = reinterpret_cast<T *>(take(nprop * nck * 2 * HW, E));
    L->seg[0] = reinterpret_cast<T *>(take(nprop * segment * HW, E));
    L->seg[1] = reinterpret_cast<T *>(take(nprop * segment * HW, E));
    if (std::is_same<T, float>::value) {
      L->tc2 = reinterpret_cast<float *>(take(g.nb * HW, 4));
      L->tamp = reinterpret_cast<float *>(take(nprop, 4));
      L->tsd = reinterpret_cast<float *>(take(nprop, 4));
      L->till = reinterpret_cast<float *>(take(nprop * nzx, 4));
      L->tsg = reinterpret_cast<float *>(take(nprop, 4));
    }
  }
  L->bytes = (off + kAlign - 1) / kAlign * kAlign;
}

// ------------------------------------------------------------------ resources
// Per thread and device: the side stream of the segment recomputes and its
// events (created on first use, kept until the process ends), and the timing
// events of the last call.
struct Resources {
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, ready[2] = {nullptr, nullptr}, freed[2] = {nullptr, nullptr};
};
thread_local std::map<int, Resources> t_res;

int resources(Resources **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_status(e, "device");
  Resources &r = t_res[dev];
  if (!r.side) {
    e = hipStreamCreateWithFlags(&r.side, hipStreamNonBlocking);  // default priority
    if (e != hipSuccess) {
      r.side = nullptr;
      return hip_status(e, "side stream");
    }
    hipEvent_t *evs[] = {&r.fork, &r.join, &r.ready[0], &r.ready[1], &r.freed[0], &r.freed[1]};
    for (hipEvent_t *ev : evs) {
      e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
      if (e != hipSuccess) return hip_status(e, "side stream event");
    }
  }
  *out = &r;
  return SS_OK;
}

struct Timing {
  bool on = false;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<std::pair<int, size_t>> marks;  // (0 forward / 1 adjoint, index of its start event)
};
thread_local Timing t_timing;

// Start (and later end) an event interval around one time-loop launch on stream st.
int mark(int kind, hipStream_t st, size_t *idx) {
  Timing &T = t_timing;
  if (!T.on) return SS_OK;
  while (T.pool.size() < T.used + 2) {
    hipEvent_t e;
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) return hip_status(r, "timing event");
    T.pool.push_back(e);
  }
  *idx = T.used;
  T.used += 2;
  T.marks.push_back({kind, *idx});
  return hip_status(hipEventRecord(T.pool[*idx], st), "timing record");
}
int mark_end(size_t idx, hipStream_t st) {
  if (!t_timing.on) return SS_OK;
  return hip_status(hipEventRecord(t_timing.pool[idx + 1], st), "timing record");
}

// ------------------------------------------------------------- time loops
template <typename T>
struct Coef {
  const T *c2, *A, *B, *w, *amp, *sd, *v;
  const float *c2a;  // c2 * A for the resident adjoint, or NULL
};

template <typename T>
int fwd_range(Impl impl, const Geom &g, int t0, int t1, const Coef<T> &c, T *p0, T *p1, T *gathers, T *hist, T *ckpt,
              int every, hipStream_t st) {
  size_t m = 0;
  SS_TRY(mark(0, st, &m));
  int rc;
  if constexpr (std::is_same<T, float>::value) {
    if (impl == kResident)
      // with the c2 * A adjoint the history's rim is never read
      rc = forward_resident(g, t0, t1, c.c2, c.A, c.B, c.w, c.amp, p0, p1, gathers, hist, ckpt, every, st,
                            c.c2a == nullptr);
    else
      rc = forward_stepwise<T>(g, t0, t1, c.c2, c.A, c.B, c.w, c.amp, p0, p1, gathers, hist, ckpt, every, st);
  } else {
    rc = forward_stepwise<T>(g, t0, t1, c.c2, c.A, c.B, c.w, c.amp, p0, p1, gathers, hist, ckpt, every, st);
  }
  SS_TRY(rc);
  return mark_end(m, st);
}

template <typename T>
int adj_range(Impl impl, const Geom &g, int t0, int t1, const Coef<T> &c, const T *res, const T *hist, T *m1, T *m2,
              T *gacc, T *illum, T *sg, hipStream_t st) {
  size_t m = 0;
  SS_TRY(mark(1, st, &m));
  int rc;
  if constexpr (std::is_same<T, float>::value) {
    if (impl == kResident)
      rc = adjoint_resident(g, t0, t1, c.c2, c.A, c.B, c.w, c.sd, res, hist, m1, m2, gacc, illum, sg, st, c.c2a);
    else
      rc = adjoint_stepwise<T>(g, t0, t1, c.c2, c.A, c.B, c.w, c.sd, res, hist, m1, m2, gacc, illum, sg, st);
  } else {
    rc = adjoint_stepwise<T>(g, t0, t1, c.c2, c.A, c.B, c.w, c.sd, res, hist, m1, m2, gacc, illum, sg, st);
  }
  SS_TRY(rc);
  return mark_end(m, st);
}

ss_sweep_group tiled_group(const ss_geom &gg, const Layout<float> &L, const float *A, const float *B, const float *w,
                           int dir, int t0, int t1, int resume, int phase, float *state, float *hist) {
  ss_sweep_group q;
  std::memset(&q, 0, sizeof q);
  q.geom = gg;
  q.c2 = L.tc2;
  q.A = A;
  q.B = B;
  q.w = w;
  q.amp = dir ? L.tsd : L.tamp;
  q.state = state;
  q.hist = hist;
  if (dir) {
    q.res = L.res;
    q.gacc = L.gacc;
    q.illum = L.till;
    q.src_grad = L.tsg;
  } else {
    q.gathers = hist ? nullptr : L.sim;  // the checkpointed forward records gathers; recomputes do not
  }
  q.t0 = t0;
  q.t1 = t1;
  q.resume = resume;
  q.phase = phase;
  return q;
}

int tiled_run(const ss_sweep_group &q, int dir, hipStream_t st, int *launches, bool use_graph) {
  const int n = sweep_launches(q, dir, st);
  if (n < 0) return n;
  *launches = n;
  size_t m = 0;
  SS_TRY(mark(dir, st, &m));
  SS_TRY(sweep_run(1, &q, dir, use_graph ? 1 : 0, st));
  return mark_end(m, st);
}

// --------------------------------------------------------------- the driver
// noalloc: run everything on `st` with direct launches: no side stream, no events,
// no captured graphs and no autotuner trials (untuned grids take the tiled kernels'
// cost-model tiling; any tiling gives the same results).
template <typename T>
int rtm(const ss_geom *gg, const ss_rtm_args *a, void *ws, int64_t ws_bytes, hipStream_t st, bool noalloc) {
  const bool own_streams = !noalloc, use_graph = !noalloc;
  Geom g;
  SS_TRY(check_geom(gg, &g));
  if (!a) {
    set_error("rtm: args is NULL");
    return SS_EINVAL;
  }
  const int stages = a->stages;
  if (stages < 1 || stages > (SS_RTM_FORWARD | SS_RTM_ADJOINT) || a->segment < 1) {
    set_error("rtm: stages must be SS_RTM_FORWARD, SS_RTM_ADJOINT or both, and segment >= 1");
    return SS_EINVAL;
  }
  const bool fwd = stages & SS_RTM_FORWARD, adj = stages & SS_RTM_ADJOINT;
  const T *c2 = static_cast<const T *>(a->c2), *A = static_cast<const T *>(a->A), *B = static_cast<const T *>(a->B);
  const T *w = static_cast<const T *>(a->w), *amp = static_cast<const T *>(a->amp);
  const T *sd = static_cast<const T *>(a->sd), *v = static_cast<const T *>(a->v);
  if (!c2 || !A || !B || !w || !amp || !ws) {
    set_error("rtm: c2, A, B, w, amp and the workspace are required");
    return SS_EINVAL;
  }
  if (adj) {
    if (!sd || !v || !a->grad || !a->illum || (!a->obs && !a->seed)) {
      set_error("rtm adjoint: sd, v, grad, illum and either obs or seed are required");
      return SS_EINVAL;
    }
    if (!a->seed && (a->obs_repeat < 1 || g.nb % a->obs_repeat != 0)) {
      set_error("rtm: obs_repeat must be >= 1 and divide nb");
      return SS_EINVAL;
    }
    if (!a->seed && (!(a->dt > 0) || !(a->dx > 0) || !(a->ramp > 0))) {
      set_error("rtm: dt, dx and the mute ramp must be positive");
      return SS_EINVAL;
    }
    if (a->img && !(a->eps > 0)) {
      set_error("illumination regularization eps must be positive");
      return SS_EINVAL;
    }
  }
  const int segment = std::min(a->segment, g.nt);
  const bool full = segment >= g.nt;
  Impl impl;
  SS_TRY(choose_impl<T>(g, a->segment, a->impl, &impl));
  if (full && impl == kTiled) impl = kResident;  // not reached: tiled needs segments
  Layout<T> L;
  make_layout<T>(g, a->segment, static_cast<char *>(ws), &L);
  if (ws_bytes < L.bytes) {
    set_error("rtm: workspace of " + std::to_string(ws_bytes) + " bytes; this call needs " + std::to_string(L.bytes));
    return SS_EINVAL;
  }
  if ((reinterpret_cast<uintptr_t>(ws) & (kAlign - 1)) != 0) {
    set_error("rtm: the workspace must be 256-byte aligned");
    return SS_EINVAL;
  }
  const int64_t nprop = (int64_t)g.nb * g.ns, HW = (int64_t)g.H * g.W, traces = (int64_t)g.nt * g.nx;
  const int64_t nzx = (int64_t)g.nz * g.nx;
  const int nck = full ? 1 : (g.nt + segment - 1) / segment;
  Coef<T> c{c2, A, B, w, amp, sd, v, nullptr};
  if constexpr (std::is_same<T, float>::value) {
    if (impl == kResident && resident_uses_c2a(g)) c.c2a = L.c2a;
  }

  // ---------------------------------------------------------------- forward
  if (fwd) {
    if (impl == kTiled) {
      if constexpr (std::is_same<T, float>::value) {
        SS_TRY(copy_rows<float>(L.tc2, 0, c2, 0, 1, g.nb * HW, st));
        SS_TRY(copy_rows<float>(L.tamp, 0, amp, 0, 1, nprop, st));
        float *fst = L.state;  // 2 buffers x (p0, p1)
        // the autotuner's trial launches overwrite the state: tune before the first segment
        if (!noalloc && !sweep_tuned(0, g.H, g.W, g.pad, nprop)) {
          ss_sweep_group q = tiled_group(*gg, L, A, B, w, 0, 0, segment, 0, 0, fst, L.seg[0]);
          SS_TRY(sweep_autotune(1, &q, 0, 0, st));
        }
        int phase = 0;
        for (int k = 0; k < nck; ++k) {
          const int t0 = k * segment, t1 = std::min(g.nt, t0 + segment);
          // checkpoint k = the state before step t0: zero for k = 0, else the sweep state
          for (int pl = 0; pl < 2; ++pl)
            SS_TRY(copy_rows<float>(L.ckpt + (int64_t)(2 * k + pl) * HW, (int64_t)nck * 2 * HW,
                                    k ? fst + (2 * phase + pl) * nprop * HW : nullptr, HW, nprop, HW, st));
          int n = 0;
          SS_TRY(tiled_run(tiled_group(*gg, L, A, B, w, 0, t0, t1, k > 0, phase, fst, nullptr), 0, st, &n, use_graph));
          phase = (phase + n) & 1;
        }
      }
    } else {
      T *p0 = L.plane(0, nprop, HW), *p1 = L.plane(1, nprop, HW);
      SS_TRY(zero(p0, 2 * nprop * HW, st));  // p0 and p1 are adjacent planes
      if (full)
        SS_TRY(fwd_range<T>(impl, g, 0, g.nt, c, p0, p1, L.sim, L.hist, nullptr, 1, st));
      else
        SS_TRY(fwd_range<T>(impl, g, 0, g.nt, c, p0, p1, L.sim, nullptr, L.ckpt, segment, st));
    }
    if (a->gathers) SS_TRY(copy_rows<T>(static_cast<T *>(a->gathers), 0, L.sim, 0, 1, nprop * traces, st));
  }
  if (!adj) return SS_OK;

  // ------------------------------------------------------ residual or seed
  if (a->seed) {
    SS_TRY(copy_rows<T>(L.res, 0, static_cast<const T *>(a->seed), 0, 1, nprop * traces, st));
  } else {
    SS_TRY(residual<T>(g, L.sim, static_cast<const T *>(a->obs), a->obs_repeat, static_cast<const T *>(a->v_mute),
                       a->dt, a->dx, a->t_mute, a->ramp, L.res, 0, st));
  }
  T *illum = static_cast<T *>(a->illum);
  if constexpr (std::is_same<T, float>::value) {
    if (c.c2a) SS_TRY(c2a_plane(g.nb, g.H, g.W, c2, A, L.c2a, st));
  }

  // ---------------------------------------------------------------- adjoint
  const bool overlap = own_streams && !t_timing.on && !full;
  Resources *R = nullptr;
  if (overlap) SS_TRY(resources(&R));
  hipStream_t side = overlap ? R->side : st;
  auto span = [&](int k, int *t0, int *t1) {
    *t0 = k * segment;
    *t1 = std::min(g.nt, *t0 + segment);
  };
  if (full) {
    T *m1 = L.plane(2, nprop, HW);
    SS_TRY(zero(m1, 2 * nprop * HW, st));  // m1, m2
    SS_TRY(zero(L.gacc, nprop * HW, st));
    SS_TRY(zero(illum, nprop * nzx, st));
    SS_TRY(zero(L.sg, nprop, st));
    SS_TRY(adj_range<T>(impl, g, 0, g.nt, c, L.res, L.hist, m1, m1 + nprop * HW, L.gacc, illum, L.sg, st));
  } else if (impl != kTiled) {
    T *m1 = L.plane(2, nprop, HW), *m2 = L.plane(3, nprop, HW);
    SS_TRY(zero(m1, 2 * nprop * HW, st));
    SS_TRY(zero(L.gacc, nprop * HW, st));
    SS_TRY(zero(illum, nprop * nzx, st));
    SS_TRY(zero(L.sg, nprop, st));
    if (overlap) {
      SS_TRY(hip_status(hipEventRecord(R->fork, st), "fork"));
      SS_TRY(hip_status(hipStreamWaitEvent(side, R->fork, 0), "fork wait"));
    }
    bool freed_pending[2] = {false, false};
    auto recompute = [&](int k) -> int {
      const int i = k & 1;
      int t0, t1;
      span(k, &t0, &t1);
      if (overlap && freed_pending[i]) SS_TRY(hip_status(hipStreamWaitEvent(side, R->freed[i], 0), "freed wait"));
      T *rp0 = L.plane(4 + 2 * i, nprop, HW), *rp1 = L.plane(5 + 2 * i, nprop, HW);
      SS_TRY(copy_rows<T>(rp0, HW, L.ckpt + (int64_t)(2 * k) * HW, (int64_t)nck * 2 * HW, nprop, HW, side));
      SS_TRY(copy_rows<T>(rp1, HW, L.ckpt + (int64_t)(2 * k + 1) * HW, (int64_t)nck * 2 * HW, nprop, HW, side));
      SS_TRY(fwd_range<T>(impl, g, t0, t1, c, rp0, rp1, nullptr, L.seg[i], nullptr, 1, side));
      if (overlap) SS_TRY(hip_status(hipEventRecord(R->ready[i], side), "ready"));
      return SS_OK;
    };
    SS_TRY(recompute(nck - 1));
    for (int k = nck - 1; k >= 0; --k) {
      if (k > 0) SS_TRY(recompute(k - 1));
      const int i = k & 1;
      if (overlap) SS_TRY(hip_status(hipStreamWaitEvent(st, R->ready[i], 0), "ready wait"));
      int t0, t1;
      span(k, &t0, &t1);
      SS_TRY(adj_range<T>(impl, g, t0, t1, c, L.res, L.seg[i], m1, m2, L.gacc, illum, L.sg, st));
      if (overlap) {
        SS_TRY(hip_status(hipEventRecord(R->freed[i], st), "freed"));
        freed_pending[i] = true;
      }
    }
    if (overlap) {
      SS_TRY(hip_status(hipEventRecord(R->join, side), "join"));
      SS_TRY(hip_status(hipStreamWaitEvent(st, R->join, 0), "join wait"));
    }
  } else {
    if constexpr (std::is_same<T, float>::value) {
      SS_TRY(copy_rows<float>(L.tc2, 0, c2, 0, 1, g.nb * HW, st));
      SS_TRY(copy_rows<float>(L.tsd, 0, sd, 0, 1, nprop, st));
      float *fst = L.plane(0, nprop, HW), *ast = L.plane(4, nprop, HW);
      if (!noalloc && !sweep_tuned(0, g.H, g.W, g.pad, nprop)) {
        int t0, t1;
        span(nck - 1, &t0, &t1);
        ss_sweep_group q = tiled_group(*gg, L, A, B, w, 0, t0, t1, 0, 0, fst, L.seg[0]);
        SS_TRY(sweep_autotune(1, &q, 0, 0, st));
      }
      if (overlap) {
        SS_TRY(hip_status(hipEventRecord(R->fork, st), "fork"));
        SS_TRY(hip_status(hipStreamWaitEvent(side, R->fork, 0), "fork wait"));
      }
      bool freed_pending[2] = {false, false};
      auto recompute = [&](int k) -> int {
        const int i = k & 1;
        int t0, t1;
        span(k, &t0, &t1);
        if (overlap && freed_pending[i]) SS_TRY(hip_status(hipStreamWaitEvent(side, R->freed[i], 0), "freed wait"));
        for (int pl = 0; pl < 2; ++pl)
          SS_TRY(copy_rows<float>(fst + pl * nprop * HW, HW, L.ckpt + (int64_t)(2 * k + pl) * HW,
                                  (int64_t)nck * 2 * HW, nprop, HW, side));
        int n = 0;
        SS_TRY(tiled_run(tiled_group(*gg, L, A, B, w, 0, t0, t1, 1, 0, fst, L.seg[i]), 0, side, &n, use_graph));
        if (overlap) SS_TRY(hip_status(hipEventRecord(R->ready[i], side), "ready"));
        return SS_OK;
      };
      SS_TRY(recompute(nck - 1));
      if (!noalloc && !sweep_tuned(1, g.H, g.W, g.pad, nprop)) {
        // the adjoint trials read the last segment's history; the first real adjoint
        // call (resume 0) then overwrites the outputs the trials left
        if (overlap) SS_TRY(hip_status(hipStreamWaitEvent(st, R->ready[(nck - 1) & 1], 0), "ready wait"));
        int t0, t1;
        span(nck - 1, &t0, &t1);
        ss_sweep_group q = tiled_group(*gg, L, A, B, w, 1, t0, t1, 0, 0, ast, L.seg[(nck - 1) & 1]);
        SS_TRY(sweep_autotune(1, &q, 1, 0, st));
      }
      int phase = 0;
      for (int k = nck - 1; k >= 0; --k) {
        if (k > 0) SS_TRY(recompute(k - 1));
        const int i = k & 1;
        if (overlap) SS_TRY(hip_status(hipStreamWaitEvent(st, R->ready[i], 0), "ready wait"));
        int t0, t1;
        span(k, &t0, &t1);
        int n = 0;
        SS_TRY(tiled_run(tiled_group(*gg, L, A, B, w, 1, t0, t1, k < nck - 1, phase, ast, L.seg[i]), 1, st, &n, use_graph));
        phase = (phase + n) & 1;
        if (overlap) {
          SS_TRY(hip_status(hipEventRecord(R->freed[i], st), "freed"));
          freed_pending[i] = true;
        }
      }
      if (overlap) {
        SS_TRY(hip_status(hipEventRecord(R->join, side), "join"));
        SS_TRY(hip_status(hipStreamWaitEvent(st, R->join, 0), "join wait"));
      }
      SS_TRY(copy_rows<float>(illum, 0, L.till, 0, 1, nprop * nzx, st));
      SS_TRY(copy_rows<float>(L.sg, 0, L.tsg, 0, 1, nprop, st));
    }
  }

  // ------------------------------------------------------------- epilogues
  SS_TRY(gradient_epilogue<T>(g, L.gacc, v, L.sg, (T)a->k, static_cast<T *>(a->grad), st));
  if (a->img)
    SS_TRY(image_epilogue<T>(g, static_cast<const T *>(a->grad), illum, (T)a->eps, a->laplacian,
                             static_cast<T *>(a->img), st));
  return SS_OK;
}

template <typename T>
int rtm_entry(const ss_geom *g, const ss_rtm_args *a, void *ws, int64_t ws_bytes, void *stream) {
  t_timing.on = a && a->timing;
  t_timing.used = 0;
  t_timing.marks.clear();
  return rtm<T>(g, a, ws, ws_bytes, (hipStream_t)stream, false);
}

// ------------------------------------------------- no-allocation entry point
// A fixed device array serves as the workspace of ss_rtm_noalloc_f32, for
// callers that cannot allocate device memory (code under graph capture, tight
// serving loops).  Each GPU the library runs on holds its own copy until the
// process ends.  Its host-side name is no device address: the address of the
// current device's copy comes from hipGetSymbolAddress, once per device, and
// only that address reaches kernels.
constexpr int64_t kScratchBytes = 256ll << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
std::map<int, void *> g_scratch;  // device -> address of its copy of scratch

int scratch_address(void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_status(e, "rtm_noalloc device");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  auto it = g_scratch.find(dev);
  if (it == g_scratch.end()) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_status(e, "rtm_noalloc scratch address");
    if (!p) {
      set_error("rtm_noalloc: no device address for the scratch array");
      return SS_EHIP;
    }
    it = g_scratch.emplace(dev, p).first;
  }
  *out = it->second;
  return SS_OK;
}

int rtm_noalloc(const ss_geom *g, const ss_rtm_args *a, void *stream) {
  if (!g || !a) {
    set_error("rtm_noalloc: geometry and args are required");
    return SS_EINVAL;
  }
  const int64_t need = rtm_workspace_bytes(g, a->segment, 4);
  if (need < 0) return SS_EINVAL;
  if (need > kScratchBytes) {
    set_error("rtm_noalloc: this call needs " + std::to_string(need) + " bytes of workspace; the fixed scratch "
              "holds " + std::to_string(kScratchBytes) + " (256 MiB): use fewer rows or a shorter segment, or "
              "ss_rtm_f32 with a workspace");
    return SS_EINVAL;
  }
  void *ws = nullptr;
  SS_TRY(scratch_address(&ws));
  t_timing.on = false;
  t_timing.used = 0;
  t_timing.marks.clear();
  return rtm<float>(g, a, ws, kScratchBytes, (hipStream_t)stream, true);
}

}  // namespace

int64_t rtm_workspace_bytes(const ss_geom *gg, int segment, int elem_size) {
  Geom g;
  if (check_geom(gg, &g) != SS_OK) return -1;
  if (segment < 1 || (elem_size != 4 && elem_size != 8)) {
    set_error("workspace_bytes: segment must be positive and elem_size 4 or 8");
    return -1;
  }
  if (elem_size == 4) {
    Layout<float> L;
    make_layout<float>(g, segment, nullptr, &L);
    return L.bytes;
  }
  Layout<double> L;
  make_layout<double>(g, segment, nullptr, &L);
  return L.bytes;
}

int rtm_timing(int *nforward, double *forward_ms, int *nadjoint, double *adjoint_ms) {
  Timing &T = t_timing;
  int n[2] = {0, 0};
  double ms[2] = {0.0, 0.0};
  for (auto &m : T.marks) {
    hipError_t e = hipEventSynchronize(T.pool[m.second + 1]);
    if (e != hipSuccess) return hip_status(e, "timing sync");
    float x = 0.0f;
    e = hipEventElapsedTime(&x, T.pool[m.second], T.pool[m.second + 1]);
    if (e != hipSuccess) return hip_status(e, "timing elapsed");
    n[m.first] += 1;
    ms[m.first] += x;
  }
  if (nforward) *nforward = n[0];
  if (forward_ms) *forward_ms = ms[0];
  if (nadjoint) *nadjoint = n[1];
  if (adjoint_ms) *adjoint_ms = ms[1];
  return SS_OK;
}

}  // namespace ss

extern "C" {

int ss_rtm_f32(const ss_geom *g, const ss_rtm_args *a, void *workspace, int64_t workspace_bytes, void *stream) {
  try {
    return ss::rtm_entry<float>(g, a, workspace, workspace_bytes, stream);
  } catch (const std::exception &e) {
    ss::set_error(std::string("rtm: ") + e.what());
    return SS_EHIP;
  }
}

int ss_rtm_f64(const ss_geom *g, const ss_rtm_args *a, void *workspace, int64_t workspace_bytes, void *stream) {
  try {
    return ss::rtm_entry<double>(g, a, workspace, workspace_bytes, stream);
  } catch (const std::exception &e) {
    ss::set_error(std::string("rtm: ") + e.what());
    return SS_EHIP;
  }
}

int ss_rtm_noalloc_f32(const ss_geom *g, const ss_rtm_args *a, void *stream) {
  try {
    return ss::rtm_noalloc(g, a, stream);
  } catch (const std::exception &e) {
    ss::set_error(std::string("rtm_noalloc: ") + e.what());
    return SS_EHIP;
  }
}

int ss_rtm_timing(int *nforward, double *forward_ms, int *nadjoint, double *adjoint_ms) {
  return ss::rtm_timing(nforward, forward_ms, nadjoint, adjoint_ms);
}

}  // extern "C"
