// Standalone operators of the reference module's free functions:
//   lap_kernel          gpu.py:33-37   lap
//   lap_adjoint_kernel  gpu.py:40-53   lap_adjoint
//   fold_kernel         gpu.py:56-67   fold_edge_padding
// They act on a batch of n planes (any leading shape flattened by the caller).
#include "ss_kernels.h"

namespace ss {

namespace {

constexpr int kBlock = 256;

template <typename T>
__global__ __launch_bounds__(kBlock) void lap_kernel(int64_t n, int H, int W, const T *__restrict__ in,
                                                     T *__restrict__ out) {
  const int64_t HW = (int64_t)H * W;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n * HW; k += (int64_t)gridDim.x * kBlock) {
    const int64_t c = k % HW;
    const int z = (int)(c / W), x = (int)(c - (int64_t)z * W);
    T v = T(0);
    if (z >= 2 && z < H - 2 && x >= 2 && x < W - 2) {
      const T *p = in + k;
      v = lap_sum(p[-2 * W], p[-W], p[0], p[W], p[2 * W], p[-2], p[-1], p[1], p[2]) / T(12);
    }
    out[k] = v;
  }
}

template <typename T>
__device__ __forceinline__ T qtap(const T *__restrict__ plane, int z, int x, int H, int W) {
  if (z < 2 || z >= H - 2 || x < 2 || x >= W - 2) return T(0);
  return plane[(int64_t)z * W + x] / T(12);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void lap_adjoint_kernel(int64_t n, int H, int W, const T *__restrict__ in,
                                                             T *__restrict__ out) {
  const int64_t HW = (int64_t)H * W;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n * HW; k += (int64_t)gridDim.x * kBlock) {
    const int64_t plane = k / HW, c = k - plane * HW;
    const int z = (int)(c / W), x = (int)(c - (int64_t)z * W);
    const T *p = in + plane * HW;
    out[k] = lapt_sum(qtap(p, z, x, H, W), qtap(p, z + 2, x, H, W), qtap(p, z + 1, x, H, W),
                      qtap(p, z - 1, x, H, W), qtap(p, z - 2, x, H, W), qtap(p, z, x + 2, H, W),
                      qtap(p, z, x + 1, H, W), qtap(p, z, x - 1, H, W), qtap(p, z, x - 2, H, W));
  }
}

// out (n, nz, nx) = transpose of edge padding by `pad` applied to g (n, H, W).
template <typename T>
__global__ __launch_bounds__(kBlock) void fold_kernel(int64_t n, int nz, int nx, int pad, const T *__restrict__ g,
                                                      T *__restrict__ out) {
  const int H = nz + 2 * pad, W = nx + 2 * pad;
  const int64_t m = (int64_t)nz * nx;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n * m; k += (int64_t)gridDim.x * kBlock) {
    const int64_t plane = k / m, c = k - plane * m;
    const int i = (int)(c / nx), j = (int)(c - (int64_t)i * nx);
    const T *gp = g + plane * H * W;
    auto at = [&](int z, int x) { return gp[(int64_t)z * W + x]; };
    T o = at(pad + i, pad + j);
    T acc;
    if (i == 0) {
      acc = T(0);
      for (int z = 0; z < pad; ++z) acc = acc + at(z, pad + j);
      o = o + acc;
    }
    if (i == nz - 1) {
      acc = T(0);
      for (int z = pad + nz; z < H; ++z) acc = acc + at(z, pad + j);
      o = o + acc;
    }
    if (j == 0) {
      acc = T(0);
      for (int x = 0; x < pad; ++x) acc = acc + at(pad + i, x);
      o = o + acc;
    }
    if (j == nx - 1) {
      acc = T(0);
      for (int x = pad + nx; x < W; ++x) acc = acc + at(pad + i, x);
      o = o + acc;
    }
    for (int zc = 0; zc < 2; ++zc) {
      for (int xc = 0; xc < 2; ++xc) {
        if (!((zc == 0 ? i == 0 : i == nz - 1) && (xc == 0 ? j == 0 : j == nx - 1))) continue;
        const int z0 = zc == 0 ? 0 : pad + nz, z1 = zc == 0 ? pad : H;
        const int x0 = xc == 0 ? 0 : pad + nx, x1 = xc == 0 ? pad : W;
        acc = T(0);
        for (int z = z0; z < z1; ++z)
          for (int x = x0; x < x1; ++x) acc = acc + at(z, x);
        o = o + acc;
      }
    }
    out[k] = o;
  }
}

unsigned grid_for(int64_t total) {
  int64_t b = (total + kBlock - 1) / kBlock;
  return (unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

}  // namespace

template <typename T>
int lap_op(int64_t n, int H, int W, const T *in, T *out, int adjoint, hipStream_t st) {
  const int64_t total = n * H * W;
  if (adjoint)
    lap_adjoint_kernel<T><<<grid_for(total), kBlock, 0, st>>>(n, H, W, in, out);
  else
    lap_kernel<T><<<grid_for(total), kBlock, 0, st>>>(n, H, W, in, out);
  return hip_status(hipGetLastError(), "lap launch");
}

template <typename T>
int fold_op(int64_t n, int nz, int nx, int pad, const T *g, T *out, hipStream_t st) {
  fold_kernel<T><<<grid_for(n * nz * nx), kBlock, 0, st>>>(n, nz, nx, pad, g, out);
  return hip_status(hipGetLastError(), "fold launch");
}

template int lap_op<float>(int64_t, int, int, const float *, float *, int, hipStream_t);
template int lap_op<double>(int64_t, int, int, const double *, double *, int, hipStream_t);
template int fold_op<float>(int64_t, int, int, int, const float *, float *, hipStream_t);
template int fold_op<double>(int64_t, int, int, int, const double *, double *, hipStream_t);

}  // namespace ss
