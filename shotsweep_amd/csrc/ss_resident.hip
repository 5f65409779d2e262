// Resident FP32 kernels: one workgroup per propagation holds the whole padded
// plane in LDS for an entire step range; per-cell state lives in registers.
//
// Forward (gpu.py:111-115, 126-135): LDS holds p(t); each thread keeps p(t-1)
// and c2 of its cells in registers.  Per step: read the nine stencil taps from
// LDS, update, then (after a barrier) publish p(t+1).  The Laplacian of every
// step can be streamed to HBM for the adjoint (gpu.py:172-175).
//
// Adjoint (gpu.py:177-184): LDS holds q = (c2 A m1)/12 (zero on the two-cell
// border), the illumination accumulator of the model cells and the sponge
// table; each thread keeps m1, m2, the imaging accumulator and c2*A of its
// cells in registers and reads the forward Laplacian from HBM.
//
// Cell ownership: cell i of the H*W plane belongs to thread i % 1024 as its
// slot j = i / 1024, so every LDS and HBM access of a wavefront touches 64
// consecutive cells.  Sponge coefficients A, B (gpu.py:85-90) depend only on
// the distance (dz, dx) into the sponge, so they are read from an LDS table
// of (pad+1)^2 entries instead of two full planes.
#include <algorithm>
#include <cfloat>

#include "ss_kernels.h"

namespace ss {

namespace {

constexpr int kThreads = 1024;
constexpr int kMaxSlots = 24;

struct CellInfo {
  // bits 0-15: sponge table index dz*(pad+1)+dx for sponge cells, and for
  // model cells (where dz = dx = 0, table entry 0) the illumination index
  // (z-pad)*nx + (x-pad); bit 16: Laplacian interior; bit 17: model cell.
  static constexpr uint32_t kIdxMask = 0xffff, kInterior = 1u << 16, kModel = 1u << 17;
};

__device__ __forceinline__ uint32_t table_index(uint32_t info) {
  return (info & CellInfo::kModel) ? 0u : (info & CellInfo::kIdxMask);
}

__device__ __forceinline__ int sponge_dist(int i, int n, int pad) {
  int d = pad - i;
  int e = i - (n - 1 - pad);
  d = d > e ? d : e;
  return d > 0 ? d : 0;
}

__device__ __forceinline__ uint32_t cell_info(int i, const Geom &g) {
  const int z = i / g.W, x = i - z * g.W;
  uint32_t v;
  if (z >= g.pad && z < g.pad + g.nz && x >= g.pad && x < g.pad + g.nx)
    v = CellInfo::kModel | (uint32_t)((z - g.pad) * g.nx + (x - g.pad));
  else
    v = (uint32_t)(sponge_dist(z, g.H, g.pad) * (g.pad + 1) + sponge_dist(x, g.W, g.pad));
  if (z >= 2 && z < g.H - 2 && x >= 2 && x < g.W - 2) v |= CellInfo::kInterior;
  return v;
}

// Fill the sponge table from the full A, B planes: entry (dz, dx) is read at
// the top-left sponge cell (pad-dz, pad-dx).  mode 0 stores (A, B), mode 1
// stores (A, A*B).
__device__ void fill_table(float2 *tab, const float *A, const float *B, const Geom &g, int mode) {
  const int n = (g.pad + 1) * (g.pad + 1);
  for (int k = threadIdx.x; k < n; k += kThreads) {
    const int dz = k / (g.pad + 1), dx = k - dz * (g.pad + 1);
    const int c = (g.pad - dz) * g.W + (g.pad - dx);
    const float a = A[c], b = B[c];
    tab[k] = make_float2(a, mode ? a * b : b);
  }
}

// LDS layout: [guard | plane H*W | guard | (pad) | sponge table | illumination]
__host__ __device__ __forceinline__ int table_offset(const Geom &g) {
  const int n = g.H * g.W + 2 * (2 * g.W + 2);
  return (n + 1) & ~1;  // float2 alignment
}

size_t fwd_lds_bytes(const Geom &g) {
  return (size_t)table_offset(g) * sizeof(float) + (size_t)(g.pad + 1) * (g.pad + 1) * sizeof(float2);
}

size_t adj_lds_bytes(const Geom &g) { return fwd_lds_bytes(g) + (size_t)g.nz * g.nx * sizeof(float); }

template <int K>
__global__ __launch_bounds__(kThreads) void fwd_resident_kernel(Geom g, int t0, int t1, const float *__restrict__ c2,
                                                               const float *__restrict__ A,
                                                               const float *__restrict__ B,
                                                               const float *__restrict__ w,
                                                               const float *__restrict__ amp, float *__restrict__ p0g,
                                                               float *__restrict__ p1g, float *__restrict__ gathers,
                                                               float *__restrict__ hist, float *__restrict__ ckpt,
                                                               int ckpt_every) {
  extern __shared__ float lds[];
  const int W = g.W, HW = g.H * g.W;
  const int guard = 2 * W + 2;
  float *P = lds + guard;
  float2 *tab = reinterpret_cast<float2 *>(lds + table_offset(g));
  const int64_t prop = blockIdx.x;
  const int b = (int)(prop / g.ns), s = (int)(prop - (int64_t)b * g.ns);
  const int t = threadIdx.x;

  for (int k = t; k < guard; k += kThreads) {
    lds[k] = 0.0f;
    P[HW + k] = 0.0f;
  }
  fill_table(tab, A, B, g, 0);

  float pn[K], cc[K];
  uint32_t info[K];
  const float *p0p = p0g + prop * HW;
  const float *p1p = p1g + prop * HW;
  const float *c2b = c2 + (int64_t)b * HW;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int i = t + j * kThreads;
    if (i < HW) {
      pn[j] = p0p[i];
      P[i] = p1p[i];
      cc[j] = c2b[i];
      info[j] = cell_info(i, g);
    } else {
      pn[j] = 0.0f;
      cc[j] = 0.0f;
      info[j] = 0;
    }
  }
  const int r = g.pad + g.row;
  const int rec0 = r * W + g.pad, rec1 = rec0 + g.nx;
  const int srci = r * W + g.pad + g.src[s];
  const float ampv = amp[prop];
  const int nslots = t1 - t0;
  const int64_t nck = ckpt ? (nslots + ckpt_every - 1) / ckpt_every : 0;
  float *gat = gathers ? gathers + prop * g.nt * g.nx - rec0 : nullptr;
  __syncthreads();

  for (int it = t0; it < t1; ++it) {
    const int slot = it - t0;
    float *hs = hist ? hist + (prop * nslots + slot) * HW : nullptr;
    float *ck = (ckpt && slot % ckpt_every == 0) ? ckpt + (prop * nck + slot / ckpt_every) * 2 * HW : nullptr;
    const float wsrc = w[it] * ampv;  // gpu.py:114 w[it] * amp
    float *gt = gat ? gat + (int64_t)it * g.nx : nullptr;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int i = t + j * kThreads;
      if (i < HW) {
        const float c = P[i];
        if (ck) {
          ck[i] = pn[j];
          ck[HW + i] = c;
        }
        const float sum = lap_sum(P[i - 2 * W], P[i - W], c, P[i + W], P[i + 2 * W], P[i - 2], P[i - 1], P[i + 1],
                                  P[i + 2]);
        const float L = (info[j] & CellInfo::kInterior) ? div12(sum) : 0.0f;
        const float2 ab = tab[table_index(info[j])];
        float p2 = leapfrog(c, pn[j], ab.y, cc[j], L, ab.x);
        if (i == srci) p2 = p2 + wsrc;
        if (hs) hs[i] = L;
        if (gt && i >= rec0 && i < rec1) gt[i] = p2;
        pn[j] = p2;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int i = t + j * kThreads;
      if (i < HW) {
        const float c = P[i];
        P[i] = pn[j];
        pn[j] = c;
      }
    }
    __syncthreads();
  }
  float *p0o = p0g + prop * HW;
  float *p1o = p1g + prop * HW;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int i = t + j * kThreads;
    if (i < HW) {
      p0o[i] = pn[j];
      p1o[i] = P[i];
    }
  }
}

template <int K>
__global__ __launch_bounds__(kThreads) void adj_resident_kernel(Geom g, int t0, int t1, const float *__restrict__ c2,
                                                               const float *__restrict__ A,
                                                               const float *__restrict__ B,
                                                               const float *__restrict__ w,
                                                               const float *__restrict__ sd,
                                                               const float *__restrict__ res,
                                                               const float *__restrict__ hist, float *__restrict__ m1g,
                                                               float *__restrict__ m2g, float *__restrict__ gaccg,
                                                               float *__restrict__ illumg,
                                                               float *__restrict__ src_grad) {
  extern __shared__ float lds[];
  const int W = g.W, HW = g.H * g.W;
  const int guard = 2 * W + 2;
  float *Q = lds + guard;
  float2 *tab = reinterpret_cast<float2 *>(lds + table_offset(g));
  float *il = reinterpret_cast<float *>(tab + (g.pad + 1) * (g.pad + 1));
  const int64_t prop = blockIdx.x;
  const int b = (int)(prop / g.ns), s = (int)(prop - (int64_t)b * g.ns);
  const int t = threadIdx.x;
  const int nzx = g.nz * g.nx;

  for (int k = t; k < guard; k += kThreads) {
    lds[k] = 0.0f;
    Q[HW + k] = 0.0f;
  }
  fill_table(tab, A, B, g, 1);
  for (int k = t; k < nzx; k += kThreads) il[k] = illumg[prop * nzx + k];
  __syncthreads();  // table ready

  float m1[K], m2[K], ga[K], ca[K];
  uint32_t info[K];
  const float *c2b = c2 + (int64_t)b * HW;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int i = t + j * kThreads;
    if (i < HW) {
      info[j] = cell_info(i, g);
      const float a = tab[table_index(info[j])].x;
      ca[j] = c2b[i] * a;  // c2 * A (gpu.py:179)
      m1[j] = m1g[prop * HW + i];
      m2[j] = m2g[prop * HW + i];
      ga[j] = gaccg[prop * HW + i];
      Q[i] = (info[j] & CellInfo::kInterior) ? div12(ca[j] * m1[j]) : 0.0f;
    } else {
      info[j] = 0;
      ca[j] = m1[j] = m2[j] = ga[j] = 0.0f;
    }
  }
  const int r = g.pad + g.row;
  const int rec0 = r * W + g.pad, rec1 = rec0 + g.nx;
  const int srci = r * W + g.pad + g.src[s];
  const float sdv = sd[prop];
  float sg = 0.0f;
  const bool own_src = (srci % kThreads) == t;
  if (own_src) sg = src_grad[prop];
  const int nslots = t1 - t0;
  const float *rs = res + prop * g.nt * g.nx - rec0;
  __syncthreads();

  for (int it = t1 - 1; it >= t0; --it) {
    const int slot = it - t0;
    const float *hs = hist + (prop * nslots + slot) * HW;
    const float *rt = rs + (int64_t)it * g.nx;
    const float wt = w[it];
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int i = t + j * kThreads;
      if (i < HW) {
        const float L = hs[i];
        const float lt = lapt_sum(Q[i], Q[i + 2 * W], Q[i + W], Q[i - W], Q[i - 2 * W], Q[i + 2], Q[i + 1],
                                  Q[i - 1], Q[i - 2]);
        const float2 aab = tab[table_index(info[j])];
        float mu = (2.0f * aab.x) * m1[j];  // gpu.py:179, (2A) m1 + lapT - (AB) m2
        mu = mu + lt;
        mu = mu - aab.y * m2[j];
        if (i >= rec0 && i < rec1) mu = mu + rt[i];  // gpu.py:180
        ga[j] = ga[j] + (aab.x * mu) * L;             // gpu.py:181
        if (i == srci) sg = sg + (mu * wt) * sdv;     // gpu.py:182
        if (info[j] & CellInfo::kModel) {             // gpu.py:183
          float &v = il[info[j] & CellInfo::kIdxMask];
          v = v + L * L;
        }
        m2[j] = m1[j];
        m1[j] = mu;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int i = t + j * kThreads;
      if (i < HW) Q[i] = (info[j] & CellInfo::kInterior) ? div12(ca[j] * m1[j]) : 0.0f;
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int i = t + j * kThreads;
    if (i < HW) {
      m1g[prop * HW + i] = m1[j];
      m2g[prop * HW + i] = m2[j];
      gaccg[prop * HW + i] = ga[j];
    }
  }
  if (own_src) src_grad[prop] = sg;
  for (int k = t; k < nzx; k += kThreads) illumg[prop * nzx + k] = il[k];
}

int device_lds_limit() {
  static int limit = -1;
  if (limit < 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      v = 65536;
    limit = v;
  }
  return limit;
}

int slots_for(const Geom &g) { return (g.H * g.W + kThreads - 1) / kThreads; }

template <typename Fn>
int dispatch_slots(int k, Fn &&fn) {
  // instantiated slot counts; the smallest one that covers the plane is used
  if (k <= 8) return fn(std::integral_constant<int, 8>());
  if (k <= 12) return fn(std::integral_constant<int, 12>());
  if (k <= 16) return fn(std::integral_constant<int, 16>());
  if (k <= 20) return fn(std::integral_constant<int, 20>());
  if (k <= 22) return fn(std::integral_constant<int, 22>());
  return fn(std::integral_constant<int, kMaxSlots>());
}

}  // namespace

int64_t resident_max_cells() { return (int64_t)kMaxSlots * kThreads; }

bool resident_fits(const Geom &g) {
  if (slots_for(g) > kMaxSlots) return false;
  if ((g.pad + 1) * (g.pad + 1) > (int)CellInfo::kIdxMask || g.nz * g.nx > (int)CellInfo::kIdxMask) return false;
  return adj_lds_bytes(g) <= (size_t)device_lds_limit();
}

int forward_resident(const Geom &g, int t0, int t1, const float *c2, const float *A, const float *B, const float *w,
                     const float *amp, float *p0, float *p1, float *gathers, float *hist, float *ckpt,
                     int ckpt_every, hipStream_t st) {
  const size_t lds = fwd_lds_bytes(g);
  return dispatch_slots(slots_for(g), [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    auto kern = fwd_resident_kernel<K>;
    hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return hip_status(e, "forward_resident attribute");
    kern<<<dim3((unsigned)((int64_t)g.nb * g.ns)), dim3(kThreads), lds, st>>>(g, t0, t1, c2, A, B, w, amp, p0, p1,
                                                                              gathers, hist, ckpt, ckpt_every);
    return hip_status(hipGetLastError(), "forward_resident launch");
  });
}

int adjoint_resident(const Geom &g, int t0, int t1, const float *c2, const float *A, const float *B, const float *w,
                     const float *sd, const float *res, const float *hist, float *m1, float *m2, float *gacc,
                     float *illum, float *src_grad, hipStream_t st) {
  const size_t lds = adj_lds_bytes(g);
  return dispatch_slots(slots_for(g), [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    auto kern = adj_resident_kernel<K>;
    hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return hip_status(e, "adjoint_resident attribute");
    kern<<<dim3((unsigned)((int64_t)g.nb * g.ns)), dim3(kThreads), lds, st>>>(g, t0, t1, c2, A, B, w, sd, res, hist,
                                                                              m1, m2, gacc, illum, src_grad);
    return hip_status(hipGetLastError(), "adjoint_resident launch");
  });
}

}  // namespace ss
