// Stepwise kernels: one launch per time step, wavefields in HBM.
//
// These run any grid size in FP32 or FP64 (the FP64 instantiation drives the
// finite-difference and dot-product checks).  One thread owns one padded
// cell of one propagation; grid = (ceil(H*W/256), ns, nb).  The reference
// equivalents are the per-step PyTorch op chains in openfwi/physics/gpu.py.
#include "ss_kernels.h"

namespace ss {

namespace {

constexpr int kBlock = 256;

template <typename T>
__device__ __forceinline__ T lap_cell(const T *__restrict__ p, int i, int W) {
  return lap_sum(p[i - 2 * W], p[i - W], p[i], p[i + W], p[i + 2 * W], p[i - 2], p[i - 1], p[i + 1], p[i + 2]) /
         T(12);
}

__device__ __forceinline__ bool interior(int z, int x, const Geom &g) {
  return z >= 2 && z < g.H - 2 && x >= 2 && x < g.W - 2;
}

// One forward step (gpu.py:111-115, 133-134).  p0 holds the field at t-1 on
// entry and receives the field at t+1 (each cell reads and writes only its
// own p0 entry, so the update is in place).
template <typename T>
__global__ __launch_bounds__(kBlock) void fwd_step_kernel(Geom g, int it, int slot, int nslots,
                                                          const T *__restrict__ c2, const T *__restrict__ A,
                                                          const T *__restrict__ B, const T *__restrict__ w,
                                                          const T *__restrict__ amp, const T *__restrict__ p1,
                                                          T *__restrict__ p0, T *__restrict__ gathers,
                                                          T *__restrict__ hist) {
  const int HW = g.H * g.W;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= HW) return;
  const int s = blockIdx.y, b = blockIdx.z;
  const int64_t prop = plane_index(b, s, g.ns);
  const T *q = p1 + prop * HW;
  const int z = i / g.W, x = i - z * g.W;
  const T L = interior(z, x, g) ? lap_cell(q, i, g.W) : T(0);
  T p2 = leapfrog(q[i], p0[prop * HW + i], B[i], c2[(int64_t)b * HW + i], L, A[i]);
  const int r = g.pad + g.row;
  if (z == r && x == g.pad + g.src[s]) p2 = p2 + w[it] * amp[prop];
  p0[prop * HW + i] = p2;
  if (hist) hist[(prop * nslots + slot) * HW + i] = L;
  if (gathers && z == r && x >= g.pad && x < g.pad + g.nx)
    gathers[(prop * g.nt + it) * g.nx + (x - g.pad)] = p2;
}

// q = (c2 A m1) / 12 on the interior, 0 on the two-cell border (gpu.py:179, 43).
template <typename T>
__device__ __forceinline__ T q_cell(const T *__restrict__ m1, const T *__restrict__ c2b, const T *__restrict__ A,
                                    int z, int x, const Geom &g) {
  if (!interior(z, x, g)) return T(0);
  const int j = z * g.W + x;
  return ((c2b[j] * A[j]) * m1[j]) / T(12);
}

// One adjoint step (gpu.py:179-184).  m2 holds mu(t+2) on entry and receives
// mu(t) (own-cell, in place).
template <typename T>
__global__ __launch_bounds__(kBlock) void adj_step_kernel(Geom g, int it, int slot, int nslots,
                                                          const T *__restrict__ c2, const T *__restrict__ A,
                                                          const T *__restrict__ B, const T *__restrict__ w,
                                                          const T *__restrict__ sd, const T *__restrict__ res,
                                                          const T *__restrict__ hist, const T *__restrict__ m1,
                                                          T *__restrict__ m2, T *__restrict__ gacc,
                                                          T *__restrict__ illum, T *__restrict__ src_grad) {
  const int HW = g.H * g.W;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= HW) return;
  const int s = blockIdx.y, b = blockIdx.z;
  const int64_t prop = plane_index(b, s, g.ns);
  const T *mp = m1 + prop * HW;
  const T *c2b = c2 + (int64_t)b * HW;
  const int z = i / g.W, x = i - z * g.W;
  const T lt = lapt_sum(q_cell(mp, c2b, A, z, x, g), q_cell(mp, c2b, A, z + 2, x, g),
                        q_cell(mp, c2b, A, z + 1, x, g), q_cell(mp, c2b, A, z - 1, x, g),
                        q_cell(mp, c2b, A, z - 2, x, g), q_cell(mp, c2b, A, z, x + 2, g),
                        q_cell(mp, c2b, A, z, x + 1, g), q_cell(mp, c2b, A, z, x - 1, g),
                        q_cell(mp, c2b, A, z, x - 2, g));
  T mu = adjoint_mu(A[i], B[i], mp[i], lt, m2[prop * HW + i]);
  const int r = g.pad + g.row;
  if (z == r && x >= g.pad && x < g.pad + g.nx) mu = mu + res[(prop * g.nt + it) * g.nx + (x - g.pad)];
  const T L = hist[(prop * nslots + slot) * HW + i];
  gacc[prop * HW + i] = gacc[prop * HW + i] + (A[i] * mu) * L;
  if (z == r && x == g.pad + g.src[s]) src_grad[prop] = src_grad[prop] + (mu * w[it]) * sd[prop];
  if (z >= g.pad && z < g.pad + g.nz && x >= g.pad && x < g.pad + g.nx) {
    T *il = illum + (prop * g.nz + (z - g.pad)) * g.nx + (x - g.pad);
    *il = *il + L * L;
  }
  m2[prop * HW + i] = mu;
}

// Tangent-linear step: the background field advances as in fwd_step_kernel
// and the perturbation q obeys the derivative of gpu.py:113-114 with respect
// to c2 and amp.  p0/q0 are updated in place.
template <typename T>
__global__ __launch_bounds__(kBlock) void born_step_kernel(Geom g, int it, const T *__restrict__ c2,
                                                           const T *__restrict__ dc2, const T *__restrict__ A,
                                                           const T *__restrict__ B, const T *__restrict__ w,
                                                           const T *__restrict__ amp, const T *__restrict__ damp,
                                                           const T *__restrict__ p1, T *__restrict__ p0,
                                                           const T *__restrict__ q1, T *__restrict__ q0,
                                                           T *__restrict__ born) {
  const int HW = g.H * g.W;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= HW) return;
  const int s = blockIdx.y, b = blockIdx.z;
  const int64_t prop = plane_index(b, s, g.ns);
  const int z = i / g.W, x = i - z * g.W;
  const bool in = interior(z, x, g);
  const T *pp = p1 + prop * HW;
  const T *qq = q1 + prop * HW;
  const T L = in ? lap_cell(pp, i, g.W) : T(0);
  const T Lq = in ? lap_cell(qq, i, g.W) : T(0);
  const int64_t cb = (int64_t)b * HW + i;
  T p2 = leapfrog(pp[i], p0[prop * HW + i], B[i], c2[cb], L, A[i]);
  T q2 = (((T(2) * qq[i] - B[i] * q0[prop * HW + i]) + c2[cb] * Lq) + dc2[cb] * L) * A[i];
  const int r = g.pad + g.row;
  if (z == r && x == g.pad + g.src[s]) {
    p2 = p2 + w[it] * amp[prop];
    q2 = q2 + w[it] * damp[prop];
  }
  p0[prop * HW + i] = p2;
  q0[prop * HW + i] = q2;
  if (born && z == r && x >= g.pad && x < g.pad + g.nx) born[(prop * g.nt + it) * g.nx + (x - g.pad)] = q2;
}

// Store the state (a, b) of every propagation into checkpoint slot k of a
// (nprop, nck, 2, H*W) buffer (gpu.py:131-132 clones p0 and p1).
template <typename T>
__global__ void ckpt_kernel(const T *__restrict__ a, const T *__restrict__ b, T *__restrict__ ckpt, int64_t nprop,
                            int HW, int64_t nck, int64_t k) {
  const int64_t n = nprop * HW;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t prop = i / HW, c = i - prop * HW;
    T *dst = ckpt + ((prop * nck + k) * 2) * HW + c;
    dst[0] = a[i];
    dst[HW] = b[i];
  }
}

// After an odd number of ping-pong steps the state sits in swapped buffers;
// swap it back in one elementwise pass.
template <typename T>
__global__ void swap_kernel(T *__restrict__ a, T *__restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    T t = a[i];
    a[i] = b[i];
    b[i] = t;
  }
}

template <typename T>
void launch_swap(T *a, T *b, int64_t n, hipStream_t st) {
  int64_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 65536) blocks = 65536;
  swap_kernel<T><<<dim3((unsigned)blocks), dim3(kBlock), 0, st>>>(a, b, n);
}

}  // namespace

template <typename T>
int forward_stepwise(const Geom &g, int t0, int t1, const T *c2, const T *A, const T *B, const T *w, const T *amp,
                     T *p0, T *p1, T *gathers, T *hist, T *ckpt, int ckpt_every, hipStream_t st) {
  const int HW = g.H * g.W;
  const int64_t nprop = (int64_t)g.nb * g.ns;
  const int64_t state = nprop * HW;
  const dim3 grid((HW + kBlock - 1) / kBlock, g.ns, g.nb);
  const int nslots = t1 - t0;
  T *a = p0, *bb = p1;  // a: field at t-1, bb: field at t
  for (int it = t0; it < t1; ++it) {
    if (ckpt && (it - t0) % ckpt_every == 0) {
      const int64_t nck = (nslots + ckpt_every - 1) / ckpt_every;
      int64_t blocks = (state + kBlock - 1) / kBlock;
      if (blocks > 65536) blocks = 65536;
      ckpt_kernel<T><<<dim3((unsigned)blocks), dim3(kBlock), 0, st>>>(a, bb, ckpt, nprop, HW, nck,
                                                                      (it - t0) / ckpt_every);
    }
    fwd_step_kernel<T><<<grid, dim3(kBlock), 0, st>>>(g, it, it - t0, nslots, c2, A, B, w, amp, bb, a, gathers,
                                                       hist);
    T *t = a;
    a = bb;
    bb = t;
  }
  if (a != p0) launch_swap(p0, p1, state, st);
  return hip_status(hipGetLastError(), "forward_stepwise launch");
}

template <typename T>
int adjoint_stepwise(const Geom &g, int t0, int t1, const T *c2, const T *A, const T *B, const T *w, const T *sd,
                     const T *res, const T *hist, T *m1, T *m2, T *gacc, T *illum, T *src_grad, hipStream_t st) {
  const int HW = g.H * g.W;
  const int64_t state = (int64_t)g.nb * g.ns * HW;
  const dim3 grid((HW + kBlock - 1) / kBlock, g.ns, g.nb);
  const int nslots = t1 - t0;
  T *a = m1, *bb = m2;  // a: mu(t+1), bb: mu(t+2)
  for (int it = t1 - 1; it >= t0; --it) {
    adj_step_kernel<T><<<grid, dim3(kBlock), 0, st>>>(g, it, it - t0, nslots, c2, A, B, w, sd, res, hist, a, bb,
                                                       gacc, illum, src_grad);
    T *t = a;
    a = bb;
    bb = t;
  }
  if (a != m1) launch_swap(m1, m2, state, st);
  return hip_status(hipGetLastError(), "adjoint_stepwise launch");
}

template <typename T>
int born_stepwise(const Geom &g, int t0, int t1, const T *c2, const T *dc2, const T *A, const T *B, const T *w,
                  const T *amp, const T *damp, T *p0, T *p1, T *q0, T *q1, T *born, hipStream_t st) {
  const int HW = g.H * g.W;
  const int64_t state = (int64_t)g.nb * g.ns * HW;
  const dim3 grid((HW + kBlock - 1) / kBlock, g.ns, g.nb);
  T *a = p0, *bb = p1, *qa = q0, *qb = q1;
  for (int it = t0; it < t1; ++it) {
    born_step_kernel<T><<<grid, dim3(kBlock), 0, st>>>(g, it, c2, dc2, A, B, w, amp, damp, bb, a, qb, qa, born);
    T *t = a;
    a = bb;
    bb = t;
    t = qa;
    qa = qb;
    qb = t;
  }
  if (a != p0) {
    launch_swap(p0, p1, state, st);
    launch_swap(q0, q1, state, st);
  }
  return hip_status(hipGetLastError(), "born_stepwise launch");
}

#define SS_INSTANTIATE(T)                                                                                         \
  template int forward_stepwise<T>(const Geom &, int, int, const T *, const T *, const T *, const T *, const T *, \
                                   T *, T *, T *, T *, T *, int, hipStream_t);                                    \
  template int adjoint_stepwise<T>(const Geom &, int, int, const T *, const T *, const T *, const T *, const T *, \
                                   const T *, const T *, T *, T *, T *, T *, T *, hipStream_t);                   \
  template int born_stepwise<T>(const Geom &, int, int, const T *, const T *, const T *, const T *, const T *,    \
                                const T *, const T *, T *, T *, T *, T *, T *, hipStream_t);
SS_INSTANTIATE(float)
SS_INSTANTIATE(double)
#undef SS_INSTANTIATE

}  // namespace ss
