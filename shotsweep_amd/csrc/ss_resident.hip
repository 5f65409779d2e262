// Resident FP32 kernels: one workgroup per propagation holds the whole padded
// plane in LDS for an entire step range; per-cell state lives in registers.
//
// Forward (gpu.py:111-115, 126-135): LDS holds p(t); each thread keeps p(t-1)
// and c2 of its cells in registers.  A step reads the nine stencil taps of
// every owned cell from LDS and computes p(t+1); after a barrier each thread
// publishes its p(t+1), the owner of the source cell adds the source term
// (gpu.py:114, after damping), and after the next barrier 70 threads record
// the receiver row (gpu.py:134).  The Laplacian of every step is streamed to
// HBM for the adjoint (gpu.py:172-175).
//
// Adjoint (gpu.py:177-184): LDS holds q = (c2 A m1)/12 (zero on the two-cell
// border), the illumination accumulator of the model cells and the sponge
// table; each thread keeps m1, m2, the imaging accumulator and c2*A of its
// cells in registers and streams the forward Laplacian from HBM.
//
// Cell ownership: cell i of the H*W plane belongs to thread i % 1024 as its
// slot j = i / 1024, so every LDS and HBM access of a wavefront touches 64
// consecutive cells.  Sponge coefficients A, B (gpu.py:85-90) depend only on
// the distance (dz, dx) into the sponge, so they come from an LDS table of
// (pad+1)^2 entries; each cell carries a 16-bit descriptor (table index or,
// for model cells, the illumination index, plus two flags), two per register.
//
// The 150 x 150 plane (70 x 70 model, pad 40) is compiled with its shape as
// constants, so stencil offsets fold into instruction immediates; other shapes
// use the same code with runtime dimensions.
#include <algorithm>

#include "ss_kernels.h"

namespace ss {

namespace {

constexpr int kThreads = 1024;
constexpr int kMaxSlots = 24;
constexpr uint32_t kBufWord3 = 0x00020000;  // raw buffer descriptor word 3 for the gfx9 family

// 16-bit cell descriptor
constexpr uint32_t kIdxMask = 0x1fff;  // sponge table index, or illumination index for model cells
constexpr uint32_t kInterior = 1u << 13;
constexpr uint32_t kModel = 1u << 14;

// Keep the scheduler from hoisting the LDS taps of many cells at once (which
// exhausts the registers that hold per-cell state): cells are scheduled in
// groups of kGroup.
constexpr int kGroup = 4;
__device__ __forceinline__ void fence_cells(int j) {
  if ((j + 1) % kGroup == 0) __builtin_amdgcn_sched_barrier(0);
}

__device__ __forceinline__ int sponge_dist(int i, int n, int pad) {
  int d = pad - i;
  int e = i - (n - 1 - pad);
  d = d > e ? d : e;
  return d > 0 ? d : 0;
}

__device__ __forceinline__ uint32_t cell_desc(int i, int H, int W, const Geom &g) {
  const int z = i / W, x = i - z * W;
  uint32_t v;
  if (z >= g.pad && z < g.pad + g.nz && x >= g.pad && x < g.pad + g.nx)
    v = kModel | (uint32_t)((z - g.pad) * g.nx + (x - g.pad));
  else
    v = (uint32_t)(sponge_dist(z, H, g.pad) * (g.pad + 1) + sponge_dist(x, W, g.pad));
  if (z >= 2 && z < H - 2 && x >= 2 && x < W - 2) v |= kInterior;
  return v;
}

__device__ __forceinline__ uint32_t desc_of(const uint32_t *pk, int j) {
  return (j & 1) ? (pk[j >> 1] >> 16) : (pk[j >> 1] & 0xffffu);
}

// Sponge coefficients of a cell: model cells have A = B = 1 (table entry 0).
__device__ __forceinline__ float2 sponge(const float2 *tab, uint32_t d) {
  return tab[(d & kModel) ? 0u : (d & kIdxMask)];
}

// Table (dz, dx) -> (A, B) (mode 0) or (A, A*B) (mode 1), read from the full
// planes at the top-left sponge cell (pad-dz, pad-dx).
__device__ void fill_table(float2 *tab, const float *A, const float *B, int W, int pad, int mode) {
  const int n = (pad + 1) * (pad + 1);
  for (int k = threadIdx.x; k < n; k += kThreads) {
    const int dz = k / (pad + 1), dx = k - dz * (pad + 1);
    const int c = (pad - dz) * W + (pad - dx);
    const float a = A[c], b = B[c];
    tab[k] = make_float2(a, mode ? a * b : b);
  }
}

// LDS layout: [guard | plane H*W | guard | (align) | sponge table | illumination]
__host__ __device__ __forceinline__ int table_offset(int H, int W) {
  const int n = H * W + 2 * (2 * W + 2);
  return (n + 1) & ~1;
}

size_t fwd_lds_bytes(const Geom &g) {
  return (size_t)table_offset(g.H, g.W) * sizeof(float) + (size_t)(g.pad + 1) * (g.pad + 1) * sizeof(float2);
}

size_t adj_lds_bytes(const Geom &g) { return fwd_lds_bytes(g) + (size_t)g.nz * g.nx * sizeof(float); }

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const float *base, int64_t nfloats) {
  const int64_t bytes = nfloats * 4;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base), (short)0,
                                           (int)(bytes > 0x7fffffff ? 0x7fffffff : bytes), (int)kBufWord3);
}

// Float views of the 32-bit raw buffer intrinsics (which move untyped bits).
__device__ __forceinline__ void buf_store(float v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
}
__device__ __forceinline__ float buf_load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

template <int K, int HC, int WC>
__global__ __launch_bounds__(kThreads) void fwd_resident_kernel(Geom g, int t0, int t1, const float *__restrict__ c2,
                                                               const float *__restrict__ A,
                                                               const float *__restrict__ B,
                                                               const float *__restrict__ w,
                                                               const float *__restrict__ amp, float *__restrict__ p0g,
                                                               float *__restrict__ p1g, float *__restrict__ gathers,
                                                               float *__restrict__ hist, float *__restrict__ ckpt,
                                                               int ckpt_every) {
  extern __shared__ float lds[];
  const int H = HC ? HC : g.H, W = WC ? WC : g.W;
  const int HW = H * W;
  const int guard = 2 * W + 2;
  float *P = lds + guard;
  float2 *tab = reinterpret_cast<float2 *>(lds + table_offset(H, W));
  const int64_t prop = blockIdx.x;
  const int b = (int)(prop / g.ns), s = (int)(prop - (int64_t)b * g.ns);
  const int t = threadIdx.x;

  for (int k = t; k < guard; k += kThreads) {
    lds[k] = 0.0f;
    P[HW + k] = 0.0f;
  }
  fill_table(tab, A, B, W, g.pad, 0);

  float pn[K], cc[K];
  uint32_t pk[(K + 1) / 2];
  const float *p0p = p0g + prop * HW;
  const float *p1p = p1g + prop * HW;
  const float *c2b = c2 + (int64_t)b * HW;
#pragma unroll
  for (int j = 0; j < (K + 1) / 2; ++j) pk[j] = 0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int i = t + j * kThreads;
    pn[j] = 0.0f;
    cc[j] = 0.0f;
    if (i < HW) {
      pn[j] = p0p[i];
      P[i] = p1p[i];
      cc[j] = c2b[i];
      pk[j >> 1] |= cell_desc(i, H, W, g) << (16 * (j & 1));
    }
  }
  const int rec0 = (g.pad + g.row) * W + g.pad;
  const int srci = rec0 + g.src[s];
  const bool src_owner = (srci % kThreads) == t;
  const float ampv = amp[prop];
  const int nslots = t1 - t0;
  const int nck = ckpt ? (nslots + ckpt_every - 1) / ckpt_every : 0;
  const bool rec_thread = gathers && t < g.nx;
  // Per-propagation buffer resources: every global access inside the time
  // loop is a buffer instruction with a per-thread byte offset and a uniform
  // (scalar) offset, so no per-cell 64-bit address stays live across steps.
  const __amdgpu_buffer_rsrc_t hrs = make_rsrc(hist ? hist + prop * nslots * HW : p0g, (int64_t)nslots * HW);
  const __amdgpu_buffer_rsrc_t grs =
      make_rsrc(gathers ? gathers + prop * g.nt * g.nx : p0g, (int64_t)g.nt * g.nx);
  const __amdgpu_buffer_rsrc_t crs = make_rsrc(ckpt ? ckpt + prop * nck * 2 * HW : p0g, nck * 2 * HW);
  __syncthreads();

  for (int it = t0; it < t1; ++it) {
    const int slot = it - t0;
    if (rec_thread && it > t0)  // p(it) after injection is the sample of step it-1
      buf_store(P[rec0 + t], grs, t * 4, (it - 1) * g.nx * 4);
    if (ckpt && slot % ckpt_every == 0) {
      const int base = (slot / ckpt_every) * 2 * HW;
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const int i = t + j * kThreads;
        if (i < HW) {
          buf_store(pn[j], crs, t * 4, (base + j * kThreads) * 4);
          buf_store(P[i], crs, t * 4, (base + HW + j * kThreads) * 4);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int i = t + j * kThreads;
      if (i < HW) {
        const uint32_t d = desc_of(pk, j);
        const float c = P[i];
        const float sum = lap_sum(P[i - 2 * W], P[i - W], c, P[i + W], P[i + 2 * W], P[i - 2], P[i - 1], P[i + 1],
                                  P[i + 2]);
        const float L = (d & kInterior) ? div12(sum) : 0.0f;
        const float2 ab = sponge(tab, d);
        pn[j] = leapfrog(c, pn[j], ab.y, cc[j], L, ab.x);
        if (hist) buf_store(L, hrs, t * 4, (slot * HW + j * kThreads) * 4);
      }
      fence_cells(j);
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
    if (src_owner) P[srci] = P[srci] + w[it] * ampv;  // gpu.py:114, added after damping
    __syncthreads();
  }
  if (rec_thread) buf_store(P[rec0 + t], grs, t * 4, (t1 - 1) * g.nx * 4);
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

template <int K, int HC, int WC>
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
  const int H = HC ? HC : g.H, W = WC ? WC : g.W;
  const int HW = H * W;
  const int guard = 2 * W + 2;
  float *Q = lds + guard;
  float2 *tab = reinterpret_cast<float2 *>(lds + table_offset(H, W));
  float *il = reinterpret_cast<float *>(tab + (g.pad + 1) * (g.pad + 1));
  const int64_t prop = blockIdx.x;
  const int b = (int)(prop / g.ns), s = (int)(prop - (int64_t)b * g.ns);
  const int t = threadIdx.x;
  const int nzx = g.nz * g.nx;

  for (int k = t; k < guard; k += kThreads) {
    lds[k] = 0.0f;
    Q[HW + k] = 0.0f;
  }
  fill_table(tab, A, B, W, g.pad, 1);
  for (int k = t; k < nzx; k += kThreads) il[k] = illumg[prop * nzx + k];
  __syncthreads();  // table ready

  float m1[K], m2[K], ga[K], ca[K];
  uint32_t pk[(K + 1) / 2];
  const float *c2b = c2 + (int64_t)b * HW;
#pragma unroll
  for (int j = 0; j < (K + 1) / 2; ++j) pk[j] = 0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int i = t + j * kThreads;
    ca[j] = m1[j] = m2[j] = ga[j] = 0.0f;
    if (i < HW) {
      const uint32_t d = cell_desc(i, H, W, g);
      pk[j >> 1] |= d << (16 * (j & 1));
      ca[j] = c2b[i] * sponge(tab, d).x;  // c2 * A (gpu.py:179)
      m1[j] = m1g[prop * HW + i];
      m2[j] = m2g[prop * HW + i];
      ga[j] = gaccg[prop * HW + i];
      Q[i] = (d & kInterior) ? div12(ca[j] * m1[j]) : 0.0f;
    }
  }
  const int rec0 = (g.pad + g.row) * W + g.pad, rec1 = rec0 + g.nx;
  const int srci = rec0 + g.src[s];
  const float sdv = sd[prop];
  const bool own_src = (srci % kThreads) == t;
  float sg = own_src ? src_grad[prop] : 0.0f;
  const int nslots = t1 - t0;
  const __amdgpu_buffer_rsrc_t rrs = make_rsrc(res + prop * g.nt * g.nx, (int64_t)g.nt * g.nx);
  const __amdgpu_buffer_rsrc_t hrs = make_rsrc(hist + prop * nslots * HW, (int64_t)nslots * HW);
  __syncthreads();

  for (int it = t1 - 1; it >= t0; --it) {
    const int slot = it - t0;
    const float wt = w[it];
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int i = t + j * kThreads;
      if (i < HW) {
        const float L = buf_load(hrs, t * 4, (slot * HW + j * kThreads) * 4);
        const uint32_t d = desc_of(pk, j);
        const float lt = lapt_sum(Q[i], Q[i + 2 * W], Q[i + W], Q[i - W], Q[i - 2 * W], Q[i + 2], Q[i + 1],
                                  Q[i - 1], Q[i - 2]);
        const float2 aab = sponge(tab, d);
        // gpu.py:179: (2A) m1 + lapT - (AB) m2, AB = A*B rounded (table mode 1)
        float mu = (2.0f * aab.x) * m1[j];
        mu = mu + lt;
        mu = mu - aab.y * m2[j];
        if (j * kThreads < rec1 && (j + 1) * kThreads > rec0) {  // slot holds receiver cells
          if (i >= rec0 && i < rec1)                             // gpu.py:180
            mu = mu + buf_load(rrs, (i - rec0) * 4, it * g.nx * 4);
        }
        ga[j] = ga[j] + (aab.x * mu) * L;  // gpu.py:181
        if (j == srci / kThreads) {
          if (i == srci) sg = sg + (mu * wt) * sdv;  // gpu.py:182
        }
        if (d & kModel) {  // gpu.py:183
          float &v = il[d & kIdxMask];
          v = v + L * L;
        }
        m2[j] = m1[j];
        m1[j] = mu;
      }
      fence_cells(j);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int i = t + j * kThreads;
      if (i < HW) Q[i] = (desc_of(pk, j) & kInterior) ? div12(ca[j] * m1[j]) : 0.0f;
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

// Shapes compiled with constant dimensions: (slots, H, W).  Everything else
// runs the runtime-shape instantiation with the smallest slot count >= needed.
template <typename Fn>
int dispatch(const Geom &g, Fn &&fn) {
  using std::integral_constant;
  if (g.H == 150 && g.W == 150)
    return fn(integral_constant<int, 22>(), integral_constant<int, 150>(), integral_constant<int, 150>());
  const int k = slots_for(g);
  if (k <= 4) return fn(integral_constant<int, 4>(), integral_constant<int, 0>(), integral_constant<int, 0>());
  if (k <= 8) return fn(integral_constant<int, 8>(), integral_constant<int, 0>(), integral_constant<int, 0>());
  if (k <= 16) return fn(integral_constant<int, 16>(), integral_constant<int, 0>(), integral_constant<int, 0>());
  return fn(integral_constant<int, kMaxSlots>(), integral_constant<int, 0>(), integral_constant<int, 0>());
}

}  // namespace

int64_t resident_max_cells() { return (int64_t)kMaxSlots * kThreads; }

bool resident_fits(const Geom &g) {
  if (slots_for(g) > kMaxSlots) return false;
  if ((g.pad + 1) * (g.pad + 1) > (int)kIdxMask + 1 || g.nz * g.nx > (int)kIdxMask + 1) return false;
  if ((int64_t)g.nt * g.H * g.W > 0x1fffffff) return false;  // history slots addressed with 32-bit byte offsets
  return adj_lds_bytes(g) <= (size_t)device_lds_limit();
}

int forward_resident(const Geom &g, int t0, int t1, const float *c2, const float *A, const float *B, const float *w,
                     const float *amp, float *p0, float *p1, float *gathers, float *hist, float *ckpt,
                     int ckpt_every, hipStream_t st) {
  const size_t lds = fwd_lds_bytes(g);
  return dispatch(g, [&](auto kc, auto hc, auto wc) -> int {
    auto kern = fwd_resident_kernel<decltype(kc)::value, decltype(hc)::value, decltype(wc)::value>;
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
  return dispatch(g, [&](auto kc, auto hc, auto wc) -> int {
    auto kern = adj_resident_kernel<decltype(kc)::value, decltype(hc)::value, decltype(wc)::value>;
    hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return hip_status(e, "adjoint_resident attribute");
    kern<<<dim3((unsigned)((int64_t)g.nb * g.ns)), dim3(kThreads), lds, st>>>(g, t0, t1, c2, A, B, w, sd, res, hist,
                                                                              m1, m2, gacc, illum, src_grad);
    return hip_status(hipGetLastError(), "adjoint_resident launch");
  });
}

}  // namespace ss
