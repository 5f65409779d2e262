// First-break picking for the surface-velocity anchor (background.py:30-40).
//
// For every trace (row, shot, receiver) of a (nrows, ns, nt, nx) gather set:
//   peak  = max_t |g(t)|
//   onset = first t with |g(t)| > frac * peak   (0 when no sample exceeds it)
// which is NumPy's amplitude.max(axis=0) and argmax(amplitude > frac * peak)
// on float64 data.  Comparisons run in double, so picks on FP32 gathers equal
// the reference's picks on the same data converted to float64.  One thread
// per trace; consecutive threads take consecutive receivers, so every step in
// t reads a contiguous row of nx samples.
#include "ss_kernels.h"

namespace ss {

namespace {

template <typename T>
__global__ __launch_bounds__(256) void first_break_kernel(int64_t ntraces, int nt, int nx,
                                                          const T *__restrict__ obs, double frac,
                                                          int *__restrict__ onset, T *__restrict__ peak) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ntraces) return;
  const int64_t g = k / nx;
  const int r = (int)(k - g * nx);
  const T *tr = obs + g * (int64_t)nt * nx + r;
  T amax = T(0);
  for (int t = 0; t < nt; ++t) {
    const T a = tr[(int64_t)t * nx] < T(0) ? -tr[(int64_t)t * nx] : tr[(int64_t)t * nx];
    amax = a > amax ? a : amax;
  }
  const double thr = frac * (double)amax;
  int first = 0;
  for (int t = 0; t < nt; ++t) {
    const T v = tr[(int64_t)t * nx];
    const double a = (double)(v < T(0) ? -v : v);
    if (a > thr) {
      first = t;
      break;
    }
  }
  onset[k] = first;
  peak[k] = amax;
}

}  // namespace

template <typename T>
int first_breaks(int64_t ngathers, int nt, int nx, const T *obs, double frac, int *onset, T *peak, hipStream_t st) {
  const int64_t n = ngathers * nx;
  const int64_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffff) {
    set_error("first_breaks: too many traces");
    return SS_EINVAL;
  }
  first_break_kernel<T><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(n, nt, nx, obs, frac, onset, peak);
  return hip_status(hipGetLastError(), "first_breaks launch");
}

template int first_breaks<float>(int64_t, int, int, const float *, double, int *, float *, hipStream_t);
template int first_breaks<double>(int64_t, int, int, const double *, double, int *, double *, hipStream_t);

}  // namespace ss
