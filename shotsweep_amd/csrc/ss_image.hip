// Residual/mute, gradient fold and image epilogue kernels.
//
//   residual_kernel   gpu.py:142-150 (mute weights) and 158-160 (muted residual)
//   gradient_kernel   gpu.py:185-187 and fold_edge_padding gpu.py:56-67
//   image_kernel      gpu.py:206-209 (illumination divide, Laplacian filter)
#include <cmath>
#include <limits>

#include "ss_kernels.h"

namespace ss {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ float cos_rn(float x) { return (float)cos((double)x); }
__device__ __forceinline__ double cos_rn(double x) { return cos(x); }

// Mute weight of sample (it, receiver x) for shot s and mute velocity vm,
// evaluated in the working type in the reference's order:
//   t = it*dt ; t0 = h/vm + t_mute ; ramp = clamp((t - t0)/0.03, 0, 1)
//   w = 0.5 - 0.5*cos(pi*ramp)
template <typename T>
__device__ __forceinline__ T mute_weight(int it, T h, T vm, T dt, T t_mute, T ramp_len) {
  const T t = T(it) * dt;
  const T t0 = h / vm + t_mute;
  T ramp = (t - t0) / ramp_len;
  // Outside the ramp the formula gives exactly 0 (cos 0 = 1) and 1 (the cosine
  // of pi rounded to T rounds to -1): skip the cosine there.  A NaN ramp fails
  // both tests and stays NaN, as in the reference's clamp and cosine.
  if (ramp <= T(0)) return T(0);
  if (ramp >= T(1)) return T(1);
  return T(0.5) - T(0.5) * cos_rn(T(M_PI) * ramp);
}

// grid: (ceil(nt*nx/256), ns, nb)
template <typename T>
__global__ __launch_bounds__(kBlock) void residual_kernel(Geom g, const T *__restrict__ sim,
                                                          const T *__restrict__ obs, int obs_repeat,
                                                          const T *__restrict__ v_mute, T dt, double dx, T t_mute,
                                                          T ramp_len, T *__restrict__ res, int weights_only) {
  const int n = g.nt * g.nx;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int s = blockIdx.y, b = blockIdx.z;
  const int it = i / g.nx, x = i - it * g.nx;
  const int64_t o = plane_index(b, s, g.ns) * n + i;
  T w = T(1);
  if (v_mute) {
    // h = |x*dx - src_x| in float64, rounded once (gpu.py:146)
    const T h = (T)fabs((double)x * dx - (double)g.src[s] * dx);
    w = mute_weight<T>(it, h, v_mute[b], dt, t_mute, ramp_len);
  }
  if (weights_only) {
    res[o] = w;
    return;
  }
  const int64_t oo = plane_index(b / obs_repeat, s, g.ns) * n + i;
  T r = sim[o] - obs[oo];
  if (v_mute) r = r * w;
  res[o] = r;
}

// Scaled imaging sum at padded cell (z, x): gpu.py:185, g * 2 vp (dt/dx)^2 with
// vp the edge-padded velocity.
template <typename T>
__device__ __forceinline__ T grad_scaled(const Geom &g, const T *ga, const T *vb, T k, int z, int x) {
  int zi = z - g.pad, xi = x - g.pad;
  zi = zi < 0 ? 0 : (zi >= g.nz ? g.nz - 1 : zi);
  xi = xi < 0 ? 0 : (xi >= g.nx ? g.nx - 1 : xi);
  return ((ga[z * g.W + x] * T(2)) * vb[zi * g.nx + xi]) * k;
}

// Model pixel (i, j) of the fold: its own value plus the padding sums it
// receives, added in the reference's order (top, bottom, left, right strips,
// then the four corner blocks; fold_edge_padding gpu.py:56-67), plus the
// source-amplitude term (gpu.py:187).
template <typename T>
__device__ __forceinline__ T fold_pixel(const Geom &g, int s, T own, int i, int j, T top, T bottom, T left, T right,
                                        const T *corner, T sgrad) {
  const int nz = g.nz, nx = g.nx;
  T o = own;
  if (i == 0) o = o + top;
  if (i == nz - 1) o = o + bottom;
  if (j == 0) o = o + left;
  if (j == nx - 1) o = o + right;
  for (int zc = 0; zc < 2; ++zc)
    for (int xc = 0; xc < 2; ++xc)
      if ((zc == 0 ? i == 0 : i == nz - 1) && (xc == 0 ? j == 0 : j == nx - 1)) o = o + corner[2 * zc + xc];
  if (i == g.row && j == g.src[s]) o = o + sgrad;
  return o;
}

// One block per propagation.  The 2 (nz + nx) padding strips (one sum of pad
// values per edge pixel) and the four pad x pad corner blocks (block
// reductions) are summed in parallel into LDS, then every model pixel is
// written in one coalesced pass.  The sums inside a strip or corner block run
// in another order than the reference's torch.sum (gradients agree to
// rounding).  LDS: 2 (nz + nx) + 4 + 4 (kBlock / 64) values.
template <typename T>
__global__ __launch_bounds__(kBlock) void gradient_kernel(Geom g, const T *__restrict__ gacc,
                                                          const T *__restrict__ v, const T *__restrict__ src_grad,
                                                          T k, T *__restrict__ grad) {
  extern __shared__ __align__(8) unsigned char grad_lds[];
  T *sums = reinterpret_cast<T *>(grad_lds);
  constexpr int nw = kBlock / 64;
  const int s = blockIdx.x, b = blockIdx.y, t = threadIdx.x;
  const int64_t prop = plane_index(b, s, g.ns);
  const int p = g.pad, nz = g.nz, nx = g.nx;
  const T *ga = gacc + prop * g.H * g.W;
  const T *vb = v + (int64_t)b * nz * nx;
  const int nstrip = 2 * (nx + nz);
  T *corner = sums + nstrip, *part = corner + 4;
  for (int q = t; q < nstrip; q += kBlock) {
    T acc = T(0);
    if (q < nx) {
      for (int z = 0; z < p; ++z) acc = acc + grad_scaled(g, ga, vb, k, z, p + q);
    } else if (q < 2 * nx) {
      for (int z = p + nz; z < g.H; ++z) acc = acc + grad_scaled(g, ga, vb, k, z, p + q - nx);
    } else if (q < 2 * nx + nz) {
      for (int x = 0; x < p; ++x) acc = acc + grad_scaled(g, ga, vb, k, p + q - 2 * nx, x);
    } else {
      for (int x = p + nx; x < g.W; ++x) acc = acc + grad_scaled(g, ga, vb, k, p + q - 2 * nx - nz, x);
    }
    sums[q] = acc;
  }
  T cs[4] = {T(0), T(0), T(0), T(0)};  // top-left, top-right, bottom-left, bottom-right
  for (int e = t; e < p * p; e += kBlock) {
    const int zz = e / p, xx = e - zz * p;
    cs[0] = cs[0] + grad_scaled(g, ga, vb, k, zz, xx);
    cs[1] = cs[1] + grad_scaled(g, ga, vb, k, zz, p + nx + xx);
    cs[2] = cs[2] + grad_scaled(g, ga, vb, k, p + nz + zz, xx);
    cs[3] = cs[3] + grad_scaled(g, ga, vb, k, p + nz + zz, p + nx + xx);
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    for (int off = 32; off > 0; off >>= 1) cs[c] = cs[c] + __shfl_xor(cs[c], off, 64);
    if ((t & 63) == 0) part[c * nw + (t >> 6)] = cs[c];
  }
  __syncthreads();
  if (t < 4) {
    T acc = part[t * nw];
    for (int w = 1; w < nw; ++w) acc = acc + part[t * nw + w];
    corner[t] = acc;
  }
  __syncthreads();
  T *out = grad + prop * nz * nx;
  const T sgrad = src_grad[prop];
  for (int c = t; c < nz * nx; c += kBlock) {
    const int i = c / nx, j = c - i * nx;
    out[c] = fold_pixel(g, s, grad_scaled(g, ga, vb, k, p + i, p + j), i, j, sums[j], sums[nx + j],
                        sums[2 * nx + i], sums[2 * nx + nz + i], corner, sgrad);
  }
}

// The same fold with every sum taken by the thread of its pixel, for grids
// whose strip sums do not fit in the LDS of one block.
template <typename T>
__global__ __launch_bounds__(kBlock) void gradient_serial_kernel(Geom g, const T *__restrict__ gacc,
                                                                 const T *__restrict__ v,
                                                                 const T *__restrict__ src_grad, T k,
                                                                 T *__restrict__ grad) {
  const int s = blockIdx.x, b = blockIdx.y;
  const int64_t prop = plane_index(b, s, g.ns);
  const int p = g.pad, nz = g.nz, nx = g.nx;
  const T *ga = gacc + prop * g.H * g.W;
  const T *vb = v + (int64_t)b * nz * nx;
  T *out = grad + prop * nz * nx;
  for (int c = threadIdx.x; c < nz * nx; c += kBlock) {
    const int i = c / nx, j = c - i * nx;
    T top = T(0), bottom = T(0), left = T(0), right = T(0), corner[4] = {T(0), T(0), T(0), T(0)};
    if (i == 0)
      for (int z = 0; z < p; ++z) top = top + grad_scaled(g, ga, vb, k, z, p + j);
    if (i == nz - 1)
      for (int z = p + nz; z < g.H; ++z) bottom = bottom + grad_scaled(g, ga, vb, k, z, p + j);
    if (j == 0)
      for (int x = 0; x < p; ++x) left = left + grad_scaled(g, ga, vb, k, p + i, x);
    if (j == nx - 1)
      for (int x = p + nx; x < g.W; ++x) right = right + grad_scaled(g, ga, vb, k, p + i, x);
    for (int zc = 0; zc < 2; ++zc)
      for (int xc = 0; xc < 2; ++xc)
        if ((zc == 0 ? i == 0 : i == nz - 1) && (xc == 0 ? j == 0 : j == nx - 1))
          for (int z = zc == 0 ? 0 : p + nz; z < (zc == 0 ? p : g.H); ++z)
            for (int x = xc == 0 ? 0 : p + nx; x < (xc == 0 ? p : g.W); ++x)
              corner[2 * zc + xc] = corner[2 * zc + xc] + grad_scaled(g, ga, vb, k, z, x);
    out[c] = fold_pixel(g, s, grad_scaled(g, ga, vb, k, p + i, p + j), i, j, top, bottom, left, right, corner,
                        src_grad[prop]);
  }
}

template <typename T>
__device__ __forceinline__ T block_max(T v, T *scratch) {
  // wavefront reduction then one value per wavefront through LDS
  for (int off = 32; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    T m = scratch[0];
    for (int k = 1; k < kBlock / 64; ++k) m = scratch[k] > m ? scratch[k] : m;
    scratch[0] = m;
  }
  __syncthreads();
  return scratch[0];
}

// One block per propagation.  img0 = -grad / max(illum + eps*max(illum), tiny)
// is recomputed at each replicate-clamped tap of the Laplacian filter.
template <typename T>
__global__ __launch_bounds__(kBlock) void image_kernel(Geom g, const T *__restrict__ grad,
                                                       const T *__restrict__ illum, T eps, int laplacian,
                                                       T *__restrict__ img) {
  __shared__ T scratch[kBlock / 64];
  const int s = blockIdx.x, b = blockIdx.y;
  const int nz = g.nz, nx = g.nx;
  const int64_t off = plane_index(b, s, g.ns) * nz * nx;
  const T *gr = grad + off;
  const T *il = illum + off;
  T m = -std::numeric_limits<T>::infinity();
  for (int c = threadIdx.x; c < nz * nx; c += kBlock) m = il[c] > m ? il[c] : m;
  m = block_max(m, scratch);
  const T e = eps * m;
  const T tiny = std::numeric_limits<T>::min();
  auto img0 = [&](int i, int j) -> T {
    i = i < 0 ? 0 : (i >= nz ? nz - 1 : i);
    j = j < 0 ? 0 : (j >= nx ? nx - 1 : j);
    const int c = i * nx + j;
    T den = il[c] + e;
    den = den < tiny ? tiny : den;
    return (-gr[c]) / den;
  };
  for (int c = threadIdx.x; c < nz * nx; c += kBlock) {
    const int i = c / nx, j = c - i * nx;
    T v;
    if (laplacian) {
      v = -(lap_sum(img0(i - 2, j), img0(i - 1, j), img0(i, j), img0(i + 1, j), img0(i + 2, j), img0(i, j - 2),
                    img0(i, j - 1), img0(i, j + 1), img0(i, j + 2)) /
            T(12));
    } else {
      v = img0(i, j);
    }
    img[off + c] = v;
  }
}

}  // namespace

template <typename T>
int residual(const Geom &g, const T *sim, const T *obs, int obs_repeat, const T *v_mute, double dt, double dx,
             double t_mute, double ramp, T *res, int weights_only, hipStream_t st) {
  const int n = g.nt * g.nx;
  const dim3 grid((n + kBlock - 1) / kBlock, g.ns, g.nb);
  residual_kernel<T><<<grid, dim3(kBlock), 0, st>>>(g, sim, obs, obs_repeat, v_mute, (T)dt, dx, (T)t_mute,
                                                     (T)ramp, res, weights_only);
  return hip_status(hipGetLastError(), "residual launch");
}

template <typename T>
int gradient_epilogue(const Geom &g, const T *gacc, const T *v, const T *src_grad, T k, T *grad, hipStream_t st) {
  const size_t lds = (size_t)(2 * (g.nz + g.nx) + 4 + 4 * (kBlock / 64)) * sizeof(T);
  if (lds <= 65536)
    gradient_kernel<T><<<dim3(g.ns, g.nb), dim3(kBlock), lds, st>>>(g, gacc, v, src_grad, k, grad);
  else
    gradient_serial_kernel<T><<<dim3(g.ns, g.nb), dim3(kBlock), 0, st>>>(g, gacc, v, src_grad, k, grad);
  return hip_status(hipGetLastError(), "gradient epilogue launch");
}

template <typename T>
int image_epilogue(const Geom &g, const T *grad, const T *illum, T eps, int laplacian, T *img, hipStream_t st) {
  image_kernel<T><<<dim3(g.ns, g.nb), dim3(kBlock), 0, st>>>(g, grad, illum, eps, laplacian, img);
  return hip_status(hipGetLastError(), "image epilogue launch");
}

#define SS_INSTANTIATE(T)                                                                                       \
  template int residual<T>(const Geom &, const T *, const T *, int, const T *, double, double, double, double, \
                           T *, int, hipStream_t);                                                              \
  template int gradient_epilogue<T>(const Geom &, const T *, const T *, const T *, T, T *, hipStream_t);        \
  template int image_epilogue<T>(const Geom &, const T *, const T *, T, int, T *, hipStream_t);
SS_INSTANTIATE(float)
SS_INSTANTIATE(double)
#undef SS_INSTANTIATE

}  // namespace ss
