// OpenFWI scores of velocity maps (metrics.py:14-73): MAE and RMSE on the
// [-1, 1] normalisation, SSIM on [0, 1] with an 11 x 11 Gaussian window
// (sigma 1.5, weights formed on the host), zero padding, population moments,
// c1 = 0.01^2, c2 = 0.03^2, averaged over the pixels.  Arithmetic in double,
// as the reference's float64 NumPy/PyTorch path.  One workgroup per map: both
// normalised maps are staged in LDS, each thread scores a strided set of
// pixels, and a tree reduction gives the three sums.
#include "ss_kernels.h"

namespace ss {

namespace {

constexpr int kScoreThreads = 256;
constexpr int kWin = 11;

__device__ __forceinline__ double block_sum(double v, double *red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = kScoreThreads / 2; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

template <typename T>
__global__ __launch_bounds__(kScoreThreads) void scores_kernel(int H, int W, const T *__restrict__ pred,
                                                               const T *__restrict__ truth, double vmin,
                                                               double vrange, const double *__restrict__ window,
                                                               double *__restrict__ out) {
  extern __shared__ double sm[];
  const int n = H * W;
  double *a = sm, *b = sm + n, *win = sm + 2 * n, *red = win + kWin * kWin;
  const int64_t img = blockIdx.x;
  const int t = threadIdx.x;
  for (int k = t; k < kWin * kWin; k += kScoreThreads) win[k] = window[k];
  double sad = 0.0, ssq = 0.0;
  for (int p = t; p < n; p += kScoreThreads) {
    // norm(v) = (v - VMIN) / (VMAX - VMIN) * 2 - 1 (metrics.py:14-16)
    const double na = ((double)pred[img * n + p] - vmin) / vrange * 2.0 - 1.0;
    const double nb = ((double)truth[img * n + p] - vmin) / vrange * 2.0 - 1.0;
    const double d = na - nb;
    sad += d < 0 ? -d : d;
    ssq += d * d;
    a[p] = (na + 1.0) / 2.0;
    b[p] = (nb + 1.0) / 2.0;
  }
  __syncthreads();
  const double c1 = 0.01 * 0.01, c2 = 0.03 * 0.03;
  double ssum = 0.0;
  for (int p = t; p < n; p += kScoreThreads) {
    const int y = p / W, x = p - y * W;
    double ma = 0.0, mb = 0.0, saa = 0.0, sbb = 0.0, sab = 0.0;
    for (int dy = 0; dy < kWin; ++dy) {
      const int yy = y + dy - kWin / 2;
      if (yy < 0 || yy >= H) continue;
      for (int dx = 0; dx < kWin; ++dx) {
        const int xx = x + dx - kWin / 2;
        if (xx < 0 || xx >= W) continue;
        const double w = win[dy * kWin + dx];
        const double va = a[yy * W + xx], vb = b[yy * W + xx];
        ma += w * va;
        mb += w * vb;
        saa += w * (va * va);
        sbb += w * (vb * vb);
        sab += w * (va * vb);
      }
    }
    const double var_a = saa - ma * ma, var_b = sbb - mb * mb, cov = sab - ma * mb;
    ssum += ((2.0 * ma * mb + c1) * (2.0 * cov + c2)) / ((ma * ma + mb * mb + c1) * (var_a + var_b + c2));
  }
  const double tad = block_sum(sad, red);
  const double tsq = block_sum(ssq, red);
  const double tss = block_sum(ssum, red);
  if (t == 0) {
    out[img * 3 + 0] = tad / n;
    out[img * 3 + 1] = sqrt(tsq / n);
    out[img * 3 + 2] = tss / n;
  }
}

}  // namespace

size_t scores_lds_bytes(int H, int W) {
  return (size_t)(2 * H * W + kWin * kWin + kScoreThreads) * sizeof(double);
}

template <typename T>
int velocity_scores(int B, int H, int W, const T *pred, const T *truth, double vmin, double vmax,
                    const double *window, double *out, hipStream_t st) {
  const size_t lds = scores_lds_bytes(H, W);
  int dev = 0, maxlds = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  if (e != hipSuccess) return hip_status(e, "velocity_scores device query");
  if ((int64_t)lds > maxlds) {
    set_error("velocity_scores: map too large for one workgroup");
    return SS_EUNSUP;
  }
  e = hipFuncSetAttribute((const void *)scores_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return hip_status(e, "velocity_scores attribute");
  scores_kernel<T><<<dim3(B), dim3(kScoreThreads), lds, st>>>(H, W, pred, truth, vmin, vmax - vmin, window, out);
  return hip_status(hipGetLastError(), "velocity_scores launch");
}

template int velocity_scores<float>(int, int, int, const float *, const float *, double, double, const double *,
                                    double *, hipStream_t);
template int velocity_scores<double>(int, int, int, const double *, const double *, double, double, const double *,
                                     double *, hipStream_t);

}  // namespace ss
