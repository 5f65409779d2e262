// Hand-written kernels of the learned preconditioner: 3x3 "same" convolutions
// (forward, input gradient, weight/bias gradient), ReLU gradient mask, MSE
// loss gradient and Adam.  FP32, NCHW.  No reference counterpart (the
// reference never implemented test-time refinement, README.md:25).
//
// Convolution forward/backward-data: one block computes a 16 x 16 pixel tile
// of one image for a group of 8 output channels; the input tile (with a
// one-pixel halo) for all input channels and the group's weights are staged
// in LDS, each thread accumulates 8 outputs of one pixel.  Backward-weight:
// one block per (output channel, input channel) reduces the 9 tap products
// and (for ci = 0) the bias over the whole batch in a fixed order, so results
// are deterministic.
#include "ss_kernels.h"

namespace ss {

namespace {

constexpr int kTile = 16;
constexpr int kCoGroup = 8;
constexpr int kMaxCin = 32;

// y[n, co, :, :] = b[co] + sum_ci sum_k w[co, ci, k] * x[n, ci, shifted]  (zero padding), optional ReLU.
// For the input gradient pass, `flip` selects the transposed kernel: w is
// read as w[ci, co, 8 - k] with the channel roles of the forward swapped.
__global__ __launch_bounds__(kTile *kTile) void conv3x3_kernel(int N, int Cin, int Cout, int H, int W,
                                                               const float *__restrict__ x,
                                                               const float *__restrict__ w,
                                                               const float *__restrict__ b, float *__restrict__ y,
                                                               int relu, int flip) {
  __shared__ float xt[kMaxCin][kTile + 2][kTile + 2];
  __shared__ float wt[kCoGroup][kMaxCin][9];
  const int tiles_x = (W + kTile - 1) / kTile;
  const int ty0 = (blockIdx.x / tiles_x) * kTile, tx0 = (blockIdx.x % tiles_x) * kTile;
  const int co0 = blockIdx.y * kCoGroup;
  const int n = blockIdx.z;
  if (n >= N) return;  // whole workgroup, before any barrier
  const int tid = threadIdx.x;
  for (int k = tid; k < Cin * (kTile + 2) * (kTile + 2); k += kTile * kTile) {
    const int ci = k / ((kTile + 2) * (kTile + 2));
    const int r = k % ((kTile + 2) * (kTile + 2));
    const int yy = ty0 + r / (kTile + 2) - 1, xx = tx0 + r % (kTile + 2) - 1;
    float v = 0.0f;
    if (yy >= 0 && yy < H && xx >= 0 && xx < W) v = x[(((int64_t)n * Cin + ci) * H + yy) * W + xx];
    xt[ci][r / (kTile + 2)][r % (kTile + 2)] = v;
  }
  for (int k = tid; k < kCoGroup * Cin * 9; k += kTile * kTile) {
    const int g = k / (Cin * 9), rem = k % (Cin * 9), ci = rem / 9, tap = rem % 9;
    const int co = co0 + g;
    float v = 0.0f;
    if (co < Cout) v = flip ? w[((int64_t)ci * Cout + co) * 9 + (8 - tap)] : w[((int64_t)co * Cin + ci) * 9 + tap];
    wt[g][ci][tap] = v;
  }
  __syncthreads();
  const int py = tid / kTile, px = tid % kTile;
  const int oy = ty0 + py, ox = tx0 + px;
  float acc[kCoGroup];
#pragma unroll
  for (int g = 0; g < kCoGroup; ++g) acc[g] = (b && co0 + g < Cout) ? b[co0 + g] : 0.0f;
  for (int ci = 0; ci < Cin; ++ci) {
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const float v = xt[ci][py + tap / 3][px + tap % 3];
#pragma unroll
      for (int g = 0; g < kCoGroup; ++g) acc[g] = __builtin_fmaf(v, wt[g][ci][tap], acc[g]);
    }
  }
  if (oy < H && ox < W) {
#pragma unroll
    for (int g = 0; g < kCoGroup; ++g) {
      const int co = co0 + g;
      if (co < Cout) {
        float v = acc[g];
        if (relu) v = v > 0.0f ? v : 0.0f;
        y[(((int64_t)n * Cout + co) * H + oy) * W + ox] = v;
      }
    }
  }
}

// dw[co, ci, tap] = sum_{n, y, x} dz[n, co, y, x] * x[n, ci, y + dy, x + dx]; db[co] = sum dz[n, co].
__global__ __launch_bounds__(256) void conv3x3_wgrad_kernel(int N, int Cin, int Cout, int H, int W,
                                                            const float *__restrict__ x,
                                                            const float *__restrict__ dz, float *__restrict__ dw,
                                                            float *__restrict__ db) {
  __shared__ float red[10][256];
  const int co = blockIdx.x, ci = blockIdx.y;
  const int tid = threadIdx.x;
  float acc[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) acc[k] = 0.0f;
  const int HWp = H * W;
  for (int n = 0; n < N; ++n) {
    const float *dzp = dz + ((int64_t)n * Cout + co) * HWp;
    const float *xp = x + ((int64_t)n * Cin + ci) * HWp;
    for (int p = tid; p < HWp; p += 256) {
      const int yy = p / W, xx = p % W;
      const float d = dzp[p];
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        const int sy = yy + tap / 3 - 1, sx = xx + tap % 3 - 1;
        const float v = (sy >= 0 && sy < H && sx >= 0 && sx < W) ? xp[sy * W + sx] : 0.0f;
        acc[tap] = __builtin_fmaf(d, v, acc[tap]);
      }
      acc[9] += d;
    }
  }
#pragma unroll
  for (int k = 0; k < 10; ++k) red[k][tid] = acc[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) {
#pragma unroll
      for (int k = 0; k < 10; ++k) red[k][tid] += red[k][tid + s];
    }
    __syncthreads();
  }
  if (tid < 9) dw[((int64_t)co * Cin + ci) * 9 + tid] = red[tid][0];
  if (tid == 9 && ci == 0 && db) db[co] = red[9][0];
}

__global__ void relu_mask_kernel(int64_t n, const float *__restrict__ dy, const float *__restrict__ y,
                                 float *__restrict__ dz) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dz[i] = y[i] > 0.0f ? dy[i] : 0.0f;
}

// loss partials of 0.5*mean((p - t)^2) and its gradient (p - t) / n
__global__ void mse_kernel(int64_t n, const float *__restrict__ p, const float *__restrict__ t,
                           float *__restrict__ grad, float *__restrict__ partial) {
  __shared__ float red[256];
  float s = 0.0f;
  const float inv = 1.0f / (float)n;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const float d = p[i] - t[i];
    grad[i] = d * inv;
    s += d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void adam_kernel(int64_t n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

__device__ float block_reduce(float v, float *red, bool is_max) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      const float o = red[threadIdx.x + k];
      red[threadIdx.x] = is_max ? (o > red[threadIdx.x] ? o : red[threadIdx.x]) : red[threadIdx.x] + o;
    }
    __syncthreads();
  }
  const float r = red[0];
  __syncthreads();
  return r;
}

// One block per model: CNN input channels = (sum over shots of the misfit
// gradient, sum over shots of the illumination), each divided by its
// maximum magnitude over the model grid; gscale[b] = that gradient maximum.
__global__ __launch_bounds__(256) void features_kernel(int ns, int nzx, const float *__restrict__ grad,
                                                       const float *__restrict__ illum, float *__restrict__ x,
                                                       float *__restrict__ gscale) {
  __shared__ float red[256];
  const int b = blockIdx.x;
  const float *gb = grad + (int64_t)b * ns * nzx;
  const float *ib = illum + (int64_t)b * ns * nzx;
  float *xg = x + (int64_t)b * 2 * nzx, *xi = xg + nzx;
  float gm = 0.0f, im = 0.0f;
  for (int c = threadIdx.x; c < nzx; c += 256) {
    float gs = 0.0f, is = 0.0f;
    for (int s = 0; s < ns; ++s) {
      gs += gb[(int64_t)s * nzx + c];
      is += ib[(int64_t)s * nzx + c];
    }
    xg[c] = gs;
    xi[c] = is;
    gm = fmaxf(gm, fabsf(gs));
    im = fmaxf(im, fabsf(is));
  }
  gm = block_reduce(gm, red, true);
  im = block_reduce(im, red, true);
  const float gi = gm > 0.0f ? 1.0f / gm : 0.0f, ii = im > 0.0f ? 1.0f / im : 0.0f;
  for (int c = threadIdx.x; c < nzx; c += 256) {
    xg[c] *= gi;
    xi[c] *= ii;
  }
  if (threadIdx.x == 0) gscale[b] = gm;
}

// out[b] = sum_i (a[b, i] - c[b, i])^2, one block per row, fixed order.
__global__ __launch_bounds__(256) void sumsq_diff_kernel(int64_t n, const float *__restrict__ a,
                                                         const float *__restrict__ c, float *__restrict__ out) {
  __shared__ float red[256];
  const int64_t off = (int64_t)blockIdx.x * n;
  float s = 0.0f;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    const float d = a[off + i] - c[off + i];
    s += d * d;
  }
  s = block_reduce(s, red, false);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// out[b, i] = clamp(v[b, i] + alpha[b] * d[b, i], lo, hi)
__global__ void update_kernel(int B, int64_t n, const float *__restrict__ v, const float *__restrict__ d,
                              const float *__restrict__ alpha, float lo, float hi, float *__restrict__ out) {
  const int64_t total = (int64_t)B * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const float u = v[i] + alpha[i / n] * d[i];
    out[i] = u < lo ? lo : (u > hi ? hi : u);
  }
}

unsigned grid1(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

}  // namespace

int conv3x3(int N, int Cin, int Cout, int H, int W, const float *x, const float *w, const float *b, float *y,
            int relu, int flip, hipStream_t st) {
  if (Cin > kMaxCin) {
    set_error("conv3x3: at most 32 input channels");
    return SS_EINVAL;
  }
  const int tiles = ((H + kTile - 1) / kTile) * ((W + kTile - 1) / kTile);
  const dim3 grid(tiles, (Cout + kCoGroup - 1) / kCoGroup, N);
  conv3x3_kernel<<<grid, kTile * kTile, 0, st>>>(N, Cin, Cout, H, W, x, w, b, y, relu, flip);
  return hip_status(hipGetLastError(), "conv3x3 launch");
}

int conv3x3_wgrad(int N, int Cin, int Cout, int H, int W, const float *x, const float *dz, float *dw, float *db,
                  hipStream_t st) {
  conv3x3_wgrad_kernel<<<dim3(Cout, Cin), 256, 0, st>>>(N, Cin, Cout, H, W, x, dz, dw, db);
  return hip_status(hipGetLastError(), "conv3x3_wgrad launch");
}

int relu_mask(int64_t n, const float *dy, const float *y, float *dz, hipStream_t st) {
  relu_mask_kernel<<<grid1(n), 256, 0, st>>>(n, dy, y, dz);
  return hip_status(hipGetLastError(), "relu_mask launch");
}

int mse(int64_t n, const float *p, const float *t, float *grad, float *partial, int nblocks, hipStream_t st) {
  mse_kernel<<<nblocks, 256, 0, st>>>(n, p, t, grad, partial);
  return hip_status(hipGetLastError(), "mse launch");
}

int adam(int64_t n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
         int step, hipStream_t st) {
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  adam_kernel<<<grid1(n), 256, 0, st>>>(n, p, g, m, v, lr, b1, b2, eps, c1, c2);
  return hip_status(hipGetLastError(), "adam launch");
}

int precond_features(int B, int ns, int nzx, const float *grad, const float *illum, float *x, float *gscale,
                     hipStream_t st) {
  features_kernel<<<B, 256, 0, st>>>(ns, nzx, grad, illum, x, gscale);
  return hip_status(hipGetLastError(), "precond_features launch");
}

int sumsq_diff(int B, int64_t n, const float *a, const float *c, float *out, hipStream_t st) {
  sumsq_diff_kernel<<<B, 256, 0, st>>>(n, a, c, out);
  return hip_status(hipGetLastError(), "sumsq_diff launch");
}

int velocity_update(int B, int64_t n, const float *v, const float *d, const float *alpha, float lo, float hi,
                    float *out, hipStream_t st) {
  update_kernel<<<grid1((int64_t)B * n), 256, 0, st>>>(B, n, v, d, alpha, lo, hi, out);
  return hip_status(hipGetLastError(), "velocity_update launch");
}

}  // namespace ss
