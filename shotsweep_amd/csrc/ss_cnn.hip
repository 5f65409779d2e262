// Hand-written kernels of the learned preconditioner: 3x3 "same" convolutions
// (forward, input gradient, weight/bias gradient), ReLU gradient mask, MSE
// loss gradient and Adam.  FP32, NCHW.  No reference counterpart (the
// reference never implemented test-time refinement, README.md:25).
//
// Convolution forward/backward-data with 16-32 output channels: an implicit
// GEMM on the matrix cores (conv3x3_mfma_kernel).  With fewer output channels:
// one block computes a 16 x 16 pixel tile of one image for a group of 8
// output channels; the input tile (with a one-pixel halo) for all input
// channels and the group's weights are staged in LDS, each thread accumulates
// 8 outputs of one pixel (conv3x3_kernel).  Backward-weight
// (at most 32 channels each way): a reduction over batch x pixels of
// dz[co] x[ci] shifted by each tap, i.e. nine 32 x 32 products with a long K,
// on the matrix cores (v_mfma_f32_32x32x2f32, FP32 in and out): a workgroup
// stages a two-row band of one image (x with a one-pixel halo, dz) in LDS, each
// wave feeds pixel pairs to nine per-tap accumulators, the four waves of a
// workgroup are summed in a fixed order and the per-workgroup partials by a
// second kernel in a fixed order, so results are deterministic.  Wider layers
// run the VALU kernel (one block per channel pair).
#include <algorithm>

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

// Convolution (forward, or with `flip` the input gradient) on the matrix cores:
// an implicit GEMM D[co][pixel] = sum over (tap, ci) of w x shifted.  A
// workgroup stages a band of kCvRows image rows of one image (all input
// channels, one-pixel halo) in LDS; each wave holds the whole weight matrix in
// registers (A operands: lane l gives w[co = l % 32][ci = 2 m + l / 32][tap]
// for K step (tap, m)) and runs 32-pixel tiles of the band (B operands: lane l
// gives x[ci = 2 m + l / 32] at pixel l % 32 of the tile, shifted by the tap).
// Each wave runs two tiles at a time (two accumulator chains).  The
// accumulators start at the bias; ReLU on the way out.  Layers with fewer than
// 16 output channels run the VALU kernel.
typedef float v16f __attribute__((ext_vector_type(16)));
constexpr int kCvRows = 6;
constexpr int kCvThreads = 256;
__host__ __device__ constexpr int cv_xs(int W) { return ((kCvRows + 2) * (W + 2)) | 1; }
size_t cv_lds_bytes(int CINP, int W) { return (size_t)2 * CINP * cv_xs(W) * sizeof(float); }

template <int CINP>  // input-channel pairs (Cin <= 2 CINP)
__global__ __launch_bounds__(kCvThreads) __attribute__((amdgpu_waves_per_eu(CINP == 1 ? 2 : 1))) void conv3x3_mfma_kernel(
    int N, int Cin, int Cout, int H, int W, const float *__restrict__ x, const float *__restrict__ w,
    const float *__restrict__ b, float *__restrict__ y, int relu, int flip) {
  extern __shared__ float sm[];
  const int XS = cv_xs(W);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int j32 = lane & 31, h = lane >> 5;
  constexpr int kWaves = kCvThreads / 64;
  float A[9 * CINP];
#pragma unroll
  for (int tap = 0; tap < 9; ++tap)
#pragma unroll
    for (int m = 0; m < CINP; ++m) {
      const int co = j32, ci = 2 * m + h;
      float v = 0.0f;
      if (co < Cout && ci < Cin) v = flip ? w[((int64_t)ci * Cout + co) * 9 + (8 - tap)] : w[((int64_t)co * Cin + ci) * 9 + tap];
      A[tap * CINP + m] = v;
    }
  const int bands = (H + kCvRows - 1) / kCvRows, nchunks = N * bands;
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int n = c / bands, r0 = (c - n * bands) * kCvRows;
    const int rows = min(kCvRows, H - r0), npix = rows * W;
    __syncthreads();  // the previous band's reads are done
    for (int rowi = wave; rowi < 2 * CINP * (kCvRows + 2); rowi += kWaves) {
      const int ci = rowi / (kCvRows + 2), rr = rowi - ci * (kCvRows + 2), yy = r0 - 1 + rr;
      const bool live = ci < Cin && yy >= 0 && yy < H;
      const float *src = x + (((int64_t)n * Cin + (live ? ci : 0)) * H + (live ? yy : 0)) * W;
      for (int cc = lane; cc < W + 2; cc += 64) {
        const int xx = cc - 1;
        sm[ci * XS + rr * (W + 2) + cc] = (live && xx >= 0 && xx < W) ? src[xx] : 0.0f;
      }
    }
    __syncthreads();
    // two tiles per pass (two independent accumulator chains)
    for (int tile = 2 * wave; tile * 32 < npix; tile += 2 * kWaves) {
      int kp[2];
      const float *xb[2];
      v16f acc[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        kp[u] = (tile + u) * 32 + j32;                  // this lane's pixel in the band
        const int kk = kp[u] < npix ? kp[u] : npix - 1;  // tail lanes read a real pixel and are not stored
        const int row = kk / W, col = kk - row * W;
        xb[u] = sm + h * XS + row * (W + 2) + col;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int co = (r / 4) * 8 + h * 4 + r % 4;
          acc[u][r] = (b && co < Cout) ? b[co] : 0.0f;
        }
      }
#pragma unroll
      for (int tap = 0; tap < 9; ++tap)
#pragma unroll
        for (int m = 0; m < CINP; ++m) {
          const int o = 2 * m * XS + (tap / 3) * (W + 2) + tap % 3;
          const float a = A[tap * CINP + m];
          acc[0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, xb[0][o], acc[0], 0, 0, 0);
          acc[1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, xb[1][o], acc[1], 0, 0, 0);
        }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        if (kp[u] < npix) {
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int co = (r / 4) * 8 + h * 4 + r % 4;
            if (co < Cout) {
              float v = acc[u][r];
              if (relu) v = v > 0.0f ? v : 0.0f;
              y[(((int64_t)n * Cout + co) * H + r0) * W + kp[u]] = v;
            }
          }
        }
      }
    }
  }
}

// Backward-weight on the matrix cores.  Per 32 x 32 x 2 MFMA, lane l gives
// A[i][k] = dz[co = l % 32][pixel 2 pp + l / 32] and B[k][j] = x[ci = l % 32]
// at that pixel shifted by the tap, so D[co][ci] accumulates the tap's sum; D
// element (i, j) sits in register r = 4 (i / 8) + i % 4 of lane j + 32 ((i / 4) % 2).
constexpr int kWgRows = 2;      // image rows per staged band
constexpr int kWgThreads = 256; // four waves
constexpr int kWgAcc = 9 * 16 * 64;      // accumulator slots of a workgroup partial (tap, register, lane)
constexpr int kWgPart = kWgAcc + 64;     // + per-lane bias partials
static_assert(kWgAcc % 64 == 0 && kWgPart % 64 == 0, "the finishing blocks cover 64 slots each");

__host__ __device__ constexpr int wg_xs(int W) { return ((kWgRows + 2) * (W + 2)) | 1; }  // odd: no bank conflicts
__host__ __device__ constexpr int wg_ds(int W) { return (kWgRows * W + 1) | 1; }           // + the pad pixel
// the band tiles, also reused for the wave sums of the partial
size_t wg_lds_bytes(int W) {
  return std::max<size_t>((size_t)32 * (wg_xs(W) + wg_ds(W)), (size_t)kWgPart) * sizeof(float);
}

__global__ __launch_bounds__(kWgThreads) __attribute__((amdgpu_waves_per_eu(2))) void conv3x3_wgrad_mfma_kernel(
    int N, int Cin, int Cout, int H, int W, const float *__restrict__ x, const float *__restrict__ dz,
    float *__restrict__ part) {
  extern __shared__ float sm[];
  const int XS = wg_xs(W), DS = wg_ds(W);
  float *xs = sm;            // [ci][kWgRows + 2][W + 2], channels >= Cin zero
  float *ds = sm + 32 * XS;  // [co][kWgRows * W + pad], channels >= Cout zero
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int i32 = lane & 31, h = lane >> 5;
  constexpr int kWaves = kWgThreads / 64;
  const int bands = (H + kWgRows - 1) / kWgRows, nchunks = N * bands;
  const int RW = kWgRows * W, npairs = (RW + 1) / 2;
  v16f acc[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) acc[k] = (v16f)(0.0f);
  float dbp = 0.0f;
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int n = c / bands, r0 = (c - n * bands) * kWgRows;
    __syncthreads();  // the previous band's reads are done
    // one tile row per wave at a time (row indices are wave-uniform), lanes along the row
    for (int rowi = wave; rowi < 32 * (kWgRows + 2); rowi += kWaves) {
      const int ci = rowi / (kWgRows + 2), rr = rowi - ci * (kWgRows + 2), yy = r0 - 1 + rr;
      const bool live = ci < Cin && yy >= 0 && yy < H;
      const float *src = x + (((int64_t)n * Cin + (live ? ci : 0)) * H + (live ? yy : 0)) * W;
      for (int cc = lane; cc < W + 2; cc += 64) {
        const int xx = cc - 1;
        xs[ci * XS + rr * (W + 2) + cc] = (live && xx >= 0 && xx < W) ? src[xx] : 0.0f;
      }
    }
    for (int rowi = wave; rowi < 32 * kWgRows; rowi += kWaves) {
      const int co = rowi / kWgRows, rr = rowi - co * kWgRows;
      const bool live = co < Cout && r0 + rr < H;
      const float *src = dz + (((int64_t)n * Cout + (live ? co : 0)) * H + (live ? r0 + rr : 0)) * W;
      for (int cc = lane; cc < W; cc += 64) ds[co * DS + rr * W + cc] = live ? src[cc] : 0.0f;
    }
    if (t < 32) ds[t * DS + RW] = 0.0f;  // the pad pixel of an odd band
    __syncthreads();
    // this lane's pixel k = 2 pp + h, advanced by 2 * kWaves per pair
    int k = 2 * wave + h;
    int row = k / W, col = k - row * W;
    for (int pp = wave; pp < npairs; pp += kWaves) {
      const int kr = k < RW ? row * (W + 2) + col : (kWgRows - 1) * (W + 2) + W - 1;  // pad: any real pixel
      const float a = ds[i32 * DS + k];
      dbp += a;
      const float *xb = xs + i32 * XS + kr;
      float b[9];
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) b[tap] = xb[(tap / 3) * (W + 2) + tap % 3];
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) acc[tap] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b[tap], acc[tap], 0, 0, 0);
      k += 2 * kWaves;
      col += 2 * kWaves;
      while (col >= W) {
        col -= W;
        ++row;
      }
    }
  }
  // waves summed in a fixed order: ((w0 + w1) + w2) + w3
  float *red = sm;
  for (int w = 0; w < kWaves; ++w) {
    __syncthreads();
    if (wave == w) {
#pragma unroll
      for (int tap = 0; tap < 9; ++tap)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          float &o = red[(tap * 16 + r) * 64 + lane];
          o = w ? o + acc[tap][r] : acc[tap][r];
        }
      float &o = red[kWgAcc + lane];
      o = w ? o + dbp : dbp;
    }
  }
  __syncthreads();
  for (int i = t; i < kWgPart; i += kWgThreads) part[(int64_t)blockIdx.x * kWgPart + i] = red[i];
}

// dw, db from the workgroup partials: four threads per slot sum consecutive
// quarters of the workgroups, then the quarters are added in order.
__global__ __launch_bounds__(256) void conv3x3_wgrad_finish_kernel(int G, int Cin, int Cout,
                                                                   const float *__restrict__ part,
                                                                   float *__restrict__ dw, float *__restrict__ db) {
  __shared__ float q4[4][64];
  const int slot = blockIdx.x * 64 + (threadIdx.x & 63), quarter = threadIdx.x >> 6;
  const int g0 = (int)((int64_t)G * quarter / 4), g1 = (int)((int64_t)G * (quarter + 1) / 4);
  float s = 0.0f;
  if (slot < kWgPart)
    for (int g = g0; g < g1; ++g) s += part[(int64_t)g * kWgPart + slot];
  q4[quarter][threadIdx.x & 63] = s;
  __syncthreads();
  const int l = threadIdx.x & 63;
  float v = 0.0f;
  if (quarter == 0 && slot < kWgPart) {
    v = ((q4[0][l] + q4[1][l]) + q4[2][l]) + q4[3][l];
    if (slot < kWgAcc) {
      const int tap = slot / 1024, r = (slot / 64) % 16, ln = slot % 64;
      const int co = (r / 4) * 8 + (ln / 32) * 4 + r % 4, ci = ln % 32;
      if (co < Cout && ci < Cin) dw[((int64_t)co * Cin + ci) * 9 + tap] = v;
    }
  }
  if (blockIdx.x * 64 != kWgAcc) return;  // the last block holds the per-lane bias partials
  __syncthreads();
  if (quarter == 0) q4[0][l] = v;
  __syncthreads();
  // lanes co and co + 32 hold the same output channel
  if (quarter == 0 && l < 32 && l < Cout && db) db[l] = q4[0][l] + q4[0][l + 32];
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

static int device_attr(hipDeviceAttribute_t a, int fallback);

int conv3x3(int N, int Cin, int Cout, int H, int W, const float *x, const float *w, const float *b, float *y,
            int relu, int flip, hipStream_t st) {
  if (Cin > kMaxCin) {
    set_error("conv3x3: at most 32 input channels");
    return SS_EINVAL;
  }
  const int CINP = Cin <= 2 ? 1 : 16;
  const size_t lds = cv_lds_bytes(CINP, W);
  if (N >= 1 && Cout >= 16 && Cout <= 32 &&
      lds <= (size_t)device_attr(hipDeviceAttributeMaxSharedMemoryPerBlock, 65536)) {
    auto kern = CINP == 1 ? conv3x3_mfma_kernel<1> : conv3x3_mfma_kernel<16>;
    hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return hip_status(e, "conv3x3 attribute");
    const int64_t nchunks = (int64_t)N * ((H + kCvRows - 1) / kCvRows);
    // one resident workgroup per CU for the wide kernel (its registers allow one wave per SIMD), two otherwise;
    // each loads the weights once and walks bands
    const int per_cu = CINP == 1 ? 2 : 1;
    const int G = (int)std::min<int64_t>(nchunks, per_cu * (int64_t)device_attr(hipDeviceAttributeMultiprocessorCount, 256));
    kern<<<G, kCvThreads, lds, st>>>(N, Cin, Cout, H, W, x, w, b, y, relu, flip);
    return hip_status(hipGetLastError(), "conv3x3 launch");
  }
  const int tiles = ((H + kTile - 1) / kTile) * ((W + kTile - 1) / kTile);
  const dim3 grid(tiles, (Cout + kCoGroup - 1) / kCoGroup, N);
  conv3x3_kernel<<<grid, kTile * kTile, 0, st>>>(N, Cin, Cout, H, W, x, w, b, y, relu, flip);
  return hip_status(hipGetLastError(), "conv3x3 launch");
}

static int device_attr(hipDeviceAttribute_t a, int fallback) {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, a, dev) != hipSuccess) return fallback;
  return v;
}

int conv3x3_wgrad(int N, int Cin, int Cout, int H, int W, const float *x, const float *dz, float *dw, float *db,
                  hipStream_t st) {
  const size_t lds = wg_lds_bytes(W);
  if (N < 1 || Cin > 32 || Cout > 32 || lds > (size_t)device_attr(hipDeviceAttributeMaxSharedMemoryPerBlock, 65536)) {
    conv3x3_wgrad_kernel<<<dim3(Cout, Cin), 256, 0, st>>>(N, Cin, Cout, H, W, x, dz, dw, db);
    return hip_status(hipGetLastError(), "conv3x3_wgrad launch");
  }
  const int64_t nchunks = (int64_t)N * ((H + kWgRows - 1) / kWgRows);
  const int G = (int)std::min<int64_t>(nchunks, 2 * (int64_t)device_attr(hipDeviceAttributeMultiprocessorCount, 256));
  // per-workgroup partials in stream-ordered scratch (freed once the finishing kernel has run)
  float *part = nullptr;
  hipError_t e = hipMallocAsync((void **)&part, (size_t)G * kWgPart * sizeof(float), st);
  if (e != hipSuccess) return hip_status(e, "conv3x3_wgrad scratch");
  e = hipFuncSetAttribute((const void *)conv3x3_wgrad_mfma_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)lds);
  if (e == hipSuccess) {
    conv3x3_wgrad_mfma_kernel<<<G, kWgThreads, lds, st>>>(N, Cin, Cout, H, W, x, dz, part);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    conv3x3_wgrad_finish_kernel<<<kWgPart / 64, 256, 0, st>>>(G, Cin, Cout, part, dw, db);
    e = hipGetLastError();
  }
  const hipError_t f = hipFreeAsync(part, st);
  if (e == hipSuccess) e = f;
  return hip_status(e, "conv3x3_wgrad launch");
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
