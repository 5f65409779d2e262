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
// table; each thread keeps m1, m2, the imaging accumulator and the packed
// descriptors of its cells in registers (512 threads, 44 cells each), streams
// the forward Laplacian from HBM a few cells ahead of use, and re-reads c2
// from the L2-resident c2 plane, also a few cells ahead, when it publishes q.
//
// Cell ownership: cell i of the H*W plane belongs to thread i % threads as its
// slot j = i / threads, so every LDS and HBM access of a wavefront touches 64
// consecutive cells.  Sponge coefficients A, B (gpu.py:85-90) depend only on
// the distance (dz, dx) into the sponge, so they come from an LDS table of
// (pad+1)^2 entries.  Each cell carries a 16-bit descriptor (table index or,
// for model cells, the illumination index, plus interior and model flags),
// two per register, decoded with two bit operations per step.
//
// The 150 x 150 plane (70 x 70 model, pad 40) runs column-strip kernels
// compiled with its shape as constants (fwd_strip_kernel, adj_strip_kernel):
// each thread owns one column pair (2c, 2c+1) in one 15-row strip and keeps
// the strip's state in registers.  Their pair descriptors hold the sponge-table
// byte offset directly, and the table sits at LDS address 0 ((A, B) entries for
// the forward, (2A, A*B) for the adjoint, A = 0.5 * 2A exactly).  Other shapes
// run the cell kernels with runtime dimensions.
#include <algorithm>

#include "ss_device.h"
#include "ss_kernels.h"

namespace ss {

namespace {

#ifndef SS_ADJ_THREADS
#define SS_ADJ_THREADS 512
#endif
#ifndef SS_GROUP
#define SS_GROUP 2
#endif
#ifndef SS_PREFETCH
#define SS_PREFETCH 4
#endif
constexpr int kFwdThreads = 1024;            // forward: p(t-1), c2 + descriptors = 2.5 registers per cell
constexpr int kAdjThreads = SS_ADJ_THREADS;  // adjoint: m1, m2, imaging sum = 3 registers per cell
constexpr int kFwdMaxSlots = 24;
constexpr int kAdjMaxSlots = 48;
constexpr int kPrefetch = SS_PREFETCH;       // adjoint: Laplacian loads issued this many cells ahead
// slots of the 150 x 150 plane, rounded up to a multiple of the prefetch distance
constexpr int kAdj150 = ((150 * 150 + kAdjThreads - 1) / kAdjThreads + kPrefetch - 1) / kPrefetch * kPrefetch;

// 16-bit cell descriptor
constexpr uint32_t kIdxMask = 0x1fff;  // sponge table index, or illumination index for model cells
constexpr uint32_t kInterior = 1u << 13;
constexpr uint32_t kModel = 1u << 14;

// Cells are processed in groups of kGroup.  Before each group an empty asm
// "redefines" the base offset that its LDS and buffer addresses are formed
// from, and its descriptors, so that no work of a later group can be hoisted
// above it: otherwise the compiler interleaves all cells of a step and runs
// out of registers for the per-cell state.
constexpr int kGroup = SS_GROUP;
__device__ __forceinline__ void group_fence(int j, int &off, uint32_t *pk) {
  if (j % kGroup == 0) {
    asm volatile("" : "+v"(off));
#pragma unroll
    for (int k = j; k < j + kGroup; k += 2) asm volatile("" : "+v"(pk[k >> 1]));
  }
}

__device__ __forceinline__ uint32_t cell_desc(int i, int H, int W, int nz, int nx, int pad) {
  const int z = i / W, x = i - z * W;
  uint32_t v;
  if (z >= pad && z < pad + nz && x >= pad && x < pad + nx)
    v = kModel | (uint32_t)((z - pad) * nx + (x - pad));
  else
    v = (uint32_t)(sponge_dist(z, H, pad) * (pad + 1) + sponge_dist(x, W, pad));
  if (z >= 2 && z < H - 2 && x >= 2 && x < W - 2) v |= kInterior;
  return v;
}


// Per-cell descriptors are loop-invariant; left visible, the compiler hoists
// every value derived from them (table addresses, flag masks) out of the time
// loop and spills.  An empty asm "redefines" them once per step instead.
template <int N>
__device__ __forceinline__ void opaque(uint32_t (&pk)[N]) {
#pragma unroll
  for (int j = 0; j < N; ++j) asm volatile("" : "+v"(pk[j]));
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
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const int dz = k / (pad + 1), dx = k - dz * (pad + 1);
    const int c = (pad - dz) * W + (pad - dx);
    const float a = A[c], b = B[c];
    tab[k] = make_float2(a, mode ? a * b : b);
  }
}


// Workgroup -> propagation.  Workgroups are dealt round-robin to the 8 XCDs
// (blockIdx % 8), so XCD x receives blocks x, x+8, ...; mapping those to the
// contiguous propagation range [x*per, (x+1)*per) keeps the shots of one
// velocity row (which share their c2 plane) inside one XCD's L2.
__host__ __device__ __forceinline__ int64_t xcd_blocks(int64_t nprop) {
  return (nprop + kXcds - 1) / kXcds * kXcds;
}
__device__ __forceinline__ int64_t xcd_prop(const Geom &g) {
  const int64_t nprop = (int64_t)g.nb * g.ns;
  const int64_t per = xcd_blocks(nprop) / kXcds;
  const int64_t p = (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  return p < nprop ? p : -1;
}

// LDS layout: [guard | plane H*W | guard | (align) | sponge table | illumination]
__host__ __device__ __forceinline__ int table_offset(int H, int W) {
  const int n = H * W + 2 * (2 * W + 2);
  return (n + 1) & ~1;
}

size_t fwd_lds_bytes(const Geom &g) {
  return (size_t)table_offset(g.H, g.W) * sizeof(float) + (size_t)(g.pad + 1) * (g.pad + 1) * sizeof(float2);
}

size_t adj_lds_bytes(const Geom &g) {
  return fwd_lds_bytes(g) + (size_t)g.nz * g.nx * sizeof(float);
}

// Strip kernels: [pair table | illumination (adjoint) | guard | plane | guard].
size_t fwd_strip_lds_bytes(const Geom &g) {
  return (size_t)(pair_table_floats(g.pad) + g.H * g.W + 2 * pair_guard(g.W)) * sizeof(float);
}
size_t adj_strip_lds_bytes(const Geom &g) {
  return (size_t)(pair_table_floats(g.pad) + g.nz * g.nx + g.H * g.W + 2 * pair_guard(g.W)) * sizeof(float);
}

// NZ, NX, PAD > 0 and ROW >= 0 fix the geometry at compile time (0 / -1: runtime).
template <int K, int NZ, int NX, int PAD, int ROW>
__global__ __launch_bounds__(kFwdThreads) void fwd_resident_kernel(Geom g, int t0, int t1, const float *__restrict__ c2,
                                                               const float *__restrict__ A,
                                                               const float *__restrict__ B,
                                                               const float *__restrict__ w,
                                                               const float *__restrict__ amp, float *__restrict__ p0g,
                                                               float *__restrict__ p1g, float *__restrict__ gathers,
                                                               float *__restrict__ hist, float *__restrict__ ckpt,
                                                               int ckpt_every) {
  extern __shared__ float lds[];
  const int nz = NZ ? NZ : g.nz, nx = NX ? NX : g.nx, pad = PAD ? PAD : g.pad, row = ROW >= 0 ? ROW : g.row;
  const int H = nz + 2 * pad, W = nx + 2 * pad;
  const int HW = H * W;
  const int guard = 2 * W + 2;
  float *P = lds + guard;
  float2 *tab = reinterpret_cast<float2 *>(lds + table_offset(H, W));
  const int64_t prop = xcd_prop(g);
  if (prop < 0) return;  // padding workgroup (whole workgroup exits before any barrier)
  const int b = (int)(prop / g.ns), s = (int)(prop - (int64_t)b * g.ns);
  const int t = threadIdx.x;

  for (int k = t; k < guard; k += kFwdThreads) {
    lds[k] = 0.0f;
    P[HW + k] = 0.0f;
  }
  fill_table(tab, A, B, W, pad, 0);

  float pn[K], cc[K];
  uint32_t pk[(K + 1) / 2];
  const float *p0p = p0g + prop * HW;
  const float *p1p = p1g + prop * HW;
  const float *c2b = c2 + (int64_t)b * HW;
#pragma unroll
  for (int j = 0; j < (K + 1) / 2; ++j) pk[j] = 0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int i = t + j * kFwdThreads;
    pn[j] = 0.0f;
    cc[j] = 0.0f;
    if (i < HW) {
      pn[j] = p0p[i];
      P[i] = p1p[i];
      cc[j] = c2b[i];
      pk[j >> 1] |= cell_desc(i, H, W, nz, nx, pad) << (16 * (j & 1));
    }
  }
  const int rec0 = (pad + row) * W + pad;
  const int srci = rec0 + g.src[s];
  const bool src_owner = (srci % kFwdThreads) == t;
  const float ampv = amp[prop];
  const int nslots = t1 - t0;
  const int nck = ckpt ? (nslots + ckpt_every - 1) / ckpt_every : 0;
  const bool rec_thread = gathers && t < nx;
  // Per-propagation buffer resources: every global access inside the time
  // loop is a buffer instruction with a per-thread byte offset and a uniform
  // (scalar) offset, so no per-cell 64-bit address stays live across steps.
  const __amdgpu_buffer_rsrc_t hrs = make_rsrc(hist ? hist + prop * nslots * HW : p0g, (int64_t)nslots * HW);
  const __amdgpu_buffer_rsrc_t grs =
      make_rsrc(gathers ? gathers + prop * g.nt * nx : p0g, (int64_t)g.nt * nx);
  const __amdgpu_buffer_rsrc_t crs = make_rsrc(ckpt ? ckpt + prop * nck * 2 * HW : p0g, nck * 2 * HW);
  __syncthreads();

  for (int it = t0; it < t1; ++it) {
    const int slot = it - t0;
    opaque(pk);
    if (rec_thread && it > t0)  // p(it) after injection is the sample of step it-1
      for (int x = t; x < nx; x += kFwdThreads) buf_store(P[rec0 + x], grs, x * 4, (it - 1) * nx * 4);
    if (ckpt && slot % ckpt_every == 0) {
      const int base = (slot / ckpt_every) * 2 * HW;
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const int i = t + j * kFwdThreads;
        if (i < HW) {
          buf_store(pn[j], crs, t * 4, (base + j * kFwdThreads) * 4);
          buf_store(P[i], crs, t * 4, (base + HW + j * kFwdThreads) * 4);
        }
      }
    }
    int off = t;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      group_fence(j, off, pk);
      if (t + j * kFwdThreads < HW) {
        const int i = off + j * kFwdThreads;
        const uint32_t d = desc_of(pk, j);
        const float c = P[i];
        const float sum = lap_sum(P[i - 2 * W], P[i - W], c, P[i + W], P[i + 2 * W], P[i - 2], P[i - 1], P[i + 1],
                                  P[i + 2]);
        const float Lq = div12(sum);
        const float L = (d & kInterior) ? Lq : 0.0f;
        const float2 ab = sponge(tab, d);
        pn[j] = leapfrog(c, pn[j], ab.y, cc[j], L, ab.x);
        if (hist) buf_store(L, hrs, off * 4, (slot * HW + j * kFwdThreads) * 4);
        asm volatile("" : "+v"(pn[j]));  // finish this cell before the next group starts
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int i = t + j * kFwdThreads;
      if (i < HW) {
        const float c = P[i];
        P[i] = pn[j];
        pn[j] = c;
      }
    }
    if (src_owner) P[srci] = P[srci] + w[it] * ampv;  // gpu.py:114, added after damping
    __syncthreads();
  }
  if (rec_thread)
    for (int x = t; x < nx; x += kFwdThreads) buf_store(P[rec0 + x], grs, x * 4, (t1 - 1) * nx * 4);
  float *p0o = p0g + prop * HW;
  float *p1o = p1g + prop * HW;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int i = t + j * kFwdThreads;
    if (i < HW) {
      p0o[i] = pn[j];
      p1o[i] = P[i];
    }
  }
}


// Column-strip forward for the fixed 150 x 150 plane.  Thread t < 750 owns
// the column pair (2c, 2c+1), c = t % 75, in the 15 rows r0 .. r0+14 of strip
// t / 75 (r0 = 15 (t / 75)), and keeps p(t-1), p(t) and c2 of those rows in
// registers: two register sets whose roles swap every step (the time loop runs
// two steps per iteration), p(t+1) overwriting p(t-1).  A row's nine taps come
// from registers except the two rows above and below the strip (four 8-byte
// LDS reads per strip and step) and the left and right neighbour pairs (two
// 8-byte reads per row, issued one row ahead together with the row's 16-byte
// sponge entry), where reading all nine taps takes seven reads per pair.  After
// every wave has read p(t) the strip publishes p(t+1) to LDS; the source owner
// first adds the source to its register copy (gpu.py:114, after damping) and
// the receiver row (gpu.py:134) is stored from registers.  The barriers order
// LDS accesses only: the history, receiver and checkpoint stores are read by
// later launches.  Lanes t >= 750 mirror thread 749's reads and store nothing;
// buffer resources stay uniform (a per-lane resource turns every access into a
// loop over lanes), idle lanes storing past every resource's size, which the
// hardware drops.
constexpr int kStripRows = 15, kStripThreads = 768;

// One time step of a strip: PC holds p(t) of the strip's rows, PO p(t-1) on entry and
// p(t+1) on exit.  Only the two rows above and below the strip and the left and right
// neighbour pairs of every row come from LDS.
template <int R, int W, int RK, bool RIM>
__device__ __forceinline__ void strip_step(float2 (&PO)[R], float2 (&PC)[R], const float2 (&cc)[R],
                                           const uint32_t (&tb)[R], const bool (&inner)[R], uint32_t a,
                                           __amdgpu_buffer_rsrc_t hrs, int hv, int hoff,
                                           bool src_owner, bool src_hi, float sv, bool rec_thread,
                                           __amdgpu_buffer_rsrc_t grs, int gv, int goff, bool act) {
  constexpr uint32_t kRow = 4u * W;
  asm volatile("" : "+v"(a));
  const float2 up2 = lds_ld2(a), up1 = lds_ld2(a + kRow);
  const float2 dn1 = lds_ld2(a + (R + 2) * kRow), dn2 = lds_ld2(a + (R + 3) * kRow);
  float2 lf = lds_ld2(a + 2 * kRow - 8), rt = lds_ld2(a + 2 * kRow + 8);
  float4 ab = lds_ld4(tb[0]);
#pragma unroll
  for (int k = 0; k < R; ++k) {
    asm volatile("" : "+v"(a));  // one row at a time
    const uint32_t ar = a + (uint32_t)(k + 2) * kRow;
    float2 lfn = make_float2(0.0f, 0.0f), rtn = lfn;
    float4 abn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (k + 1 < R) {  // next row's neighbours and coefficients, one row ahead
      lfn = lds_ld2(ar + kRow - 8);
      rtn = lds_ld2(ar + kRow + 8);
      abn = lds_ld4(tb[k + 1]);
    }
    const float2 zm2 = k >= 2 ? PC[k >= 2 ? k - 2 : 0] : (k == 1 ? up1 : up2);
    const float2 zm1 = k >= 1 ? PC[k >= 1 ? k - 1 : 0] : up1;
    const float2 zc = PC[k];
    const float2 zp1 = k + 1 < R ? PC[k + 1 < R ? k + 1 : 0] : dn1;
    const float2 zp2 = k + 2 < R ? PC[k + 2 < R ? k + 2 : 0] : (k + 2 == R ? dn1 : dn2);
    const float s0 = lap_sum(zm2.x, zm1.x, zc.x, zp1.x, zp2.x, lf.x, lf.y, zc.y, rt.x);
    const float s1 = lap_sum(zm2.y, zm1.y, zc.y, zp1.y, zp2.y, lf.y, zc.x, rt.x, rt.y);
    const float2 Lq = div12x2(s0, s1);
    const float L0 = !RIM || inner[k] ? Lq.x : 0.0f, L1 = !RIM || inner[k] ? Lq.y : 0.0f;
    PO[k].x = leapfrog(zc.x, PO[k].x, ab.y, cc[k].x, L0, ab.x);
    PO[k].y = leapfrog(zc.y, PO[k].y, ab.w, cc[k].y, L1, ab.z);
    buf_store2(make_float2(L0, L1), hrs, hv, hoff + k * (int)kRow);
    asm volatile("" : "+v"(PO[k].x), "+v"(PO[k].y));  // finish this row before the next starts
    lf = lfn;
    rt = rtn;
    ab = abn;
  }
  lds_barrier();
  if (src_owner) {  // gpu.py:114, added after damping
    if (src_hi)
      PO[RK].y = PO[RK].y + sv;
    else
      PO[RK].x = PO[RK].x + sv;
  }
  if (rec_thread) buf_store2(PO[RK], grs, gv, goff);  // p(it+1) after injection
  if (act) {
#pragma unroll
    for (int k = 0; k < R; ++k) lds_st2(a + (uint32_t)(k + 2) * kRow, PO[k]);
  }
  lds_barrier();
}

//
// RIM = true: the Laplacian history holds zero on the two-cell rim, as gpu.py:35
// writes it (the forward entry point's contract).  RIM = false (the native
// driver, whose c2 * A adjoint zeroes the rim's sponge entries and so never uses
// the rim's history): no per-pair select; c2 is zero on the rim pairs instead, so
// ((2 p1 - B p0) + 0 * L) A keeps the rim's wavefield at +0 as the reference's
// zero Laplacian does, whatever finite value the rim's L has.
template <int NZ, int NX, int PAD, int ROW, bool RIM>
__global__ __launch_bounds__(kStripThreads) void fwd_strip_kernel(Geom g, int t0, int t1, const float *__restrict__ c2,
                                                                  const float *__restrict__ A,
                                                                  const float *__restrict__ B,
                                                                  const float *__restrict__ w,
                                                                  const float *__restrict__ amp,
                                                                  float *__restrict__ p0g, float *__restrict__ p1g,
                                                                  float *__restrict__ gathers, float *__restrict__ hist,
                                                                  float *__restrict__ ckpt, int ckpt_every) {
  extern __shared__ float lds[];
  constexpr int nz = NZ, nx = NX, pad = PAD, row = ROW;
  constexpr int H = nz + 2 * pad, W = nx + 2 * pad, HW = H * W;
  constexpr int R = kStripRows, NC = W / 2, NS = H / R, NACT = NC * NS;
  static_assert(W % 2 == 0 && nx % 2 == 0 && pad % 2 == 0, "pairs need even widths");
  static_assert(H % R == 0 && NACT <= kStripThreads && R >= 2, "strips must tile the plane");
  static_assert((pad + 1) * (pad + 1) * 16 <= (int)kPTabMask + 16, "pair-table offsets must fit the descriptor");
  constexpr int guard = pair_guard(W), tabf = pair_table_floats(pad);
  constexpr int RS = (pad + row) / R, RK = (pad + row) % R;  // strip and slot of the receiver row
  float4 *tab = reinterpret_cast<float4 *>(lds);
  float *P = lds + tabf + guard;
  constexpr uint32_t kPAddr = (uint32_t)(tabf + guard) * 4;
  const int64_t prop = xcd_prop(g);
  if (prop < 0) return;
  const int b = (int)(prop / g.ns), s = (int)(prop - (int64_t)b * g.ns);
  const int t = threadIdx.x;
  const bool act = t < NACT;
  const int tt = act ? t : NACT - 1;
  const int strip = tt / NC, c = tt - strip * NC, r0 = strip * R;

  for (int k = t; k < guard; k += kStripThreads) {
    P[k - guard] = 0.0f;
    P[HW + k] = 0.0f;
  }
  fill_pair_table(tab, A, B, W, pad, 0);

  float2 Pa[R], Pb[R], cc[R];
  uint32_t tb[R];
  bool inner[R];
  const float2 *p0p = reinterpret_cast<const float2 *>(p0g + prop * HW);
  const float2 *p1p = reinterpret_cast<const float2 *>(p1g + prop * HW);
  const float2 *c2b = reinterpret_cast<const float2 *>(c2 + (int64_t)b * HW);
  float2 *P2 = reinterpret_cast<float2 *>(P);
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int q = (r0 + k) * NC + c;
    Pa[k] = p0p[q];
    Pb[k] = p1p[q];
    const uint32_t d = pair_desc(q, H, W, nz, nx, pad);
    tb[k] = d & kPTabMask;
    inner[k] = (d & kPInterior) != 0;
    cc[k] = RIM || inner[k] ? c2b[q] : make_float2(0.0f, 0.0f);
    if (act) P2[q] = Pb[k];
  }
  const int srcx = pad + g.src[s];
  const bool src_owner = act && strip == RS && c == (srcx >> 1);
  const bool src_hi = (srcx & 1) != 0;
  const float ampv = amp[prop];
  const bool rec_thread = act && gathers && strip == RS && c >= pad / 2 && c < (pad + nx) / 2;
  const int nslots = t1 - t0;
  const int nck = ckpt ? (nslots + ckpt_every - 1) / ckpt_every : 0;
  const __amdgpu_buffer_rsrc_t hrs =
      make_rsrc(hist ? hist + prop * nslots * HW : p0g, hist ? (int64_t)nslots * HW : 0);
  const __amdgpu_buffer_rsrc_t grs = make_rsrc(gathers ? gathers + prop * g.nt * nx : p0g, (int64_t)g.nt * nx);
  const __amdgpu_buffer_rsrc_t crs = make_rsrc(ckpt ? ckpt + prop * nck * 2 * HW : p0g, nck * 2 * HW);
  const int hv = act ? (r0 * W + 2 * c) * 4 : 0x40000000;
  const int gv = (2 * c - pad) * 4;
  const uint32_t a = kPAddr + (uint32_t)(((r0 - 2) * W + 2 * c) * 4);
  __syncthreads();

  auto ckpt_store = [&](int slot, const float2(&older)[R], const float2(&newer)[R]) {
    if (ckpt && slot % ckpt_every == 0) {
      const int base = (slot / ckpt_every) * 2 * HW * 4;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        buf_store2(older[k], crs, hv, base + k * W * 4);
        buf_store2(newer[k], crs, hv, base + (HW + k * W) * 4);
      }
    }
  };
  int it = t0;
  for (; it + 1 < t1; it += 2) {
    ckpt_store(it - t0, Pa, Pb);
    strip_step<R, W, RK, RIM>(Pa, Pb, cc, tb, inner, a, hrs, hv, (it - t0) * HW * 4, src_owner, src_hi,
                              w[it] * ampv, rec_thread, grs, gv, it * nx * 4, act);
    ckpt_store(it + 1 - t0, Pb, Pa);
    strip_step<R, W, RK, RIM>(Pb, Pa, cc, tb, inner, a, hrs, hv, (it + 1 - t0) * HW * 4, src_owner, src_hi,
                              w[it + 1] * ampv, rec_thread, grs, gv, (it + 1) * nx * 4, act);
  }
  const bool odd = it < t1;
  if (odd) {
    ckpt_store(it - t0, Pa, Pb);
    strip_step<R, W, RK, RIM>(Pa, Pb, cc, tb, inner, a, hrs, hv, (it - t0) * HW * 4, src_owner, src_hi,
                              w[it] * ampv, rec_thread, grs, gv, it * nx * 4, act);
  }
  if (act) {
    float2 *p0o = reinterpret_cast<float2 *>(p0g + prop * HW);
    float2 *p1o = reinterpret_cast<float2 *>(p1g + prop * HW);
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int q = (r0 + k) * NC + c;
      p0o[q] = odd ? Pb[k] : Pa[k];
      p1o[q] = odd ? Pa[k] : Pb[k];
    }
  }
}

// NZ, NX, PAD > 0 and ROW >= 0 fix the geometry at compile time (0 / -1: runtime).
template <int K, int NZ, int NX, int PAD, int ROW>
__global__ __launch_bounds__(kAdjThreads) void adj_resident_kernel(Geom g, int t0, int t1, const float *__restrict__ c2,
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
  const int nz = NZ ? NZ : g.nz, nx = NX ? NX : g.nx, pad = PAD ? PAD : g.pad, row = ROW >= 0 ? ROW : g.row;
  const int H = nz + 2 * pad, W = nx + 2 * pad;
  const int HW = H * W;
  const int guard = 2 * W + 2;
  float *Q = lds + guard;
  float2 *tab = reinterpret_cast<float2 *>(lds + table_offset(H, W));
  float *il = reinterpret_cast<float *>(tab + (pad + 1) * (pad + 1));
  const int64_t prop = xcd_prop(g);
  if (prop < 0) return;
  const int b = (int)(prop / g.ns), s = (int)(prop - (int64_t)b * g.ns);
  const int t = threadIdx.x;
  const int nzx = nz * nx;
  for (int k = t; k < guard; k += kAdjThreads) {
    lds[k] = 0.0f;
    Q[HW + k] = 0.0f;
  }
  fill_table(tab, A, B, W, pad, 1);
  for (int k = t; k < nzx; k += kAdjThreads) il[k] = illumg[prop * nzx + k];
  __syncthreads();
  uint32_t pk[(K + 1) / 2];
#pragma unroll
  for (int j = 0; j < (K + 1) / 2; ++j) pk[j] = 0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int i = t + j * kAdjThreads;
    if (i < HW) pk[j >> 1] |= cell_desc(i, H, W, nz, nx, pad) << ((j & 1) * 16);
  }
  const __amdgpu_buffer_rsrc_t srs1 = make_rsrc(m1g + prop * HW, HW);
  const __amdgpu_buffer_rsrc_t srs2 = make_rsrc(m2g + prop * HW, HW);
  const __amdgpu_buffer_rsrc_t srsg = make_rsrc(gaccg + prop * HW, HW);
  const float *c2b = c2 + (int64_t)b * HW;
  const __amdgpu_buffer_rsrc_t crs = make_rsrc(c2b, (int64_t)HW);
  float m1[K], m2[K], ga[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    m1[j] = m2[j] = ga[j] = 0.0f;
    if (t + j * kAdjThreads < HW) {
      m1[j] = buf_load(srs1, t * 4, j * kAdjThreads * 4);
      m2[j] = buf_load(srs2, t * 4, j * kAdjThreads * 4);
      ga[j] = buf_load(srsg, t * 4, j * kAdjThreads * 4);
    }
  }
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int i = t + j * kAdjThreads;
    if (i < HW) {
      const uint32_t d = desc_of(pk, j);
      Q[i] = (d & kInterior) ? div12((buf_load(crs, t * 4, j * kAdjThreads * 4) * sponge(tab, d).x) * m1[j]) : 0.0f;
    }
  }
  const int rec0 = (pad + row) * W + pad, rec1 = rec0 + nx;
  const int srci = rec0 + g.src[s];
  const float sdv = sd[prop];
  const bool own_src = (srci % kAdjThreads) == t;
  float sg = own_src ? src_grad[prop] : 0.0f;
  const int nslots = t1 - t0;
  const __amdgpu_buffer_rsrc_t rrs = make_rsrc(res + prop * g.nt * nx, (int64_t)g.nt * nx);
  const __amdgpu_buffer_rsrc_t hrs = make_rsrc(hist + prop * nslots * HW, (int64_t)nslots * HW);
  __syncthreads();
  constexpr int D = K < kPrefetch ? K : kPrefetch;
  static_assert(K % D == 0, "slot count must be a multiple of the prefetch distance");
  float Lb[D], Cb[D];  // history and c2 of the cells D ahead
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const bool live = t + d * kAdjThreads < HW;
    Lb[d] = live ? buf_load(hrs, t * 4, ((t1 - 1 - t0) * HW + d * kAdjThreads) * 4) : 0.0f;
    Cb[d] = live ? buf_load(crs, t * 4, d * kAdjThreads * 4) : 0.0f;
  }
  for (int it = t1 - 1; it >= t0; --it) {
    const int slot = it - t0;
    const float wt = w[it];
    int off = t;
    opaque(pk);
#pragma unroll
    for (int j = 0; j < K; ++j) {
      group_fence(j, off, pk);
      const float L = Lb[j % D];
      {
        const int jn = j + D;
        if (jn < K) {
          if (t + jn * kAdjThreads < HW) Lb[j % D] = buf_load(hrs, off * 4, (slot * HW + jn * kAdjThreads) * 4);
        } else if (it > t0 && t + (jn - K) * kAdjThreads < HW) {
          Lb[j % D] = buf_load(hrs, off * 4, ((slot - 1) * HW + (jn - K) * kAdjThreads) * 4);
        }
      }
      if (t + j * kAdjThreads < HW) {
        const int i = off + j * kAdjThreads;
        const uint32_t d = desc_of(pk, j);
        const float lt = lapt_sum(Q[i], Q[i + 2 * W], Q[i + W], Q[i - W], Q[i - 2 * W], Q[i + 2], Q[i + 1],
                                  Q[i - 1], Q[i - 2]);
        const float2 aab = sponge(tab, d);
        float mu = (2.0f * aab.x) * m1[j];
        mu = mu + lt;
        mu = mu - aab.y * m2[j];
        if (j * kAdjThreads < rec1 && (j + 1) * kAdjThreads > rec0) {
          if (i >= rec0 && i < rec1) mu = mu + buf_load(rrs, (i - rec0) * 4, it * nx * 4);
        }
        ga[j] = ga[j] + (aab.x * mu) * L;
        if (j == srci / kAdjThreads) {
          if (i == srci) sg = sg + (mu * wt) * sdv;
        }
        if (d & kModel) {
          float &v = il[d & kIdxMask];
          v = v + L * L;
        }
        m2[j] = m1[j];
        m1[j] = mu;
        asm volatile("" : "+v"(m1[j]), "+v"(m2[j]), "+v"(ga[j]));
      }
    }
    __syncthreads();
    off = t;
    asm volatile("" : "+v"(off));  // keep per-cell descriptors from being hoisted out of the time loop
    opaque(pk);
#pragma unroll
    for (int j = 0; j < K; ++j) {
      group_fence(j, off, pk);
      const float cv = Cb[j % D];
      {
        // c2 is the same every step: after the last cell, load the first cells of the next step
        const int jn = j + D < K ? j + D : j + D - K;
        if (t + jn * kAdjThreads < HW) Cb[j % D] = buf_load(crs, off * 4, jn * kAdjThreads * 4);
      }
      if (t + j * kAdjThreads < HW) {
        const uint32_t d = desc_of(pk, j);
        const float q = div12((cv * sponge(tab, d).x) * m1[j]);
        Q[off + j * kAdjThreads] = (d & kInterior) ? q : 0.0f;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (t + j * kAdjThreads < HW) {
      buf_store(m1[j], srs1, t * 4, j * kAdjThreads * 4);
      buf_store(m2[j], srs2, t * 4, j * kAdjThreads * 4);
      buf_store(ga[j], srsg, t * 4, j * kAdjThreads * 4);
    }
  }
  if (own_src) src_grad[prop] = sg;
  for (int k = t; k < nzx; k += kAdjThreads) illumg[prop * nzx + k] = il[k];
}

// Column-strip adjoint for the fixed 150 x 150 plane with the c2 * A plane
// (c2a_kernel: zero on the two-cell rim, so q is zero there, as lap_adjoint
// reads only p[2:-2, 2:-2], gpu.py:41).  The strip layout of fwd_strip_kernel:
// thread t < 750 owns column pair c = t % 75 in rows r0 .. r0+14 of strip t / 75
// and keeps m1, m2 (two register sets swapping roles every step, as the state
// arrays of the reference's recursion shift) and the imaging sum of those rows in
// registers.  LDS holds q = (c2 A m1) / 12, the operand of lap_adjoint, the
// model-cell illumination and the sponge table of (2A0, A0 B0, 2A1, A1 B1)
// entries.  Phase 1 walks the strip top to bottom with the centre q pairs of rows
// r-2 .. r+2 in registers (one new 8-byte read per row) plus the row's left and
// right neighbour pairs, and reads the row's sponge entry and forward Laplacian
// ahead of use (history three rows ahead into registers): mu = 2A m1 + lapT(q) -
// A B m2 (gpu.py:179), the residual on the receiver row (gpu.py:180), the imaging
// sum (A mu) L and, on model cells, the illumination L^2 (gpu.py:181, 183), the
// source term once per step (gpu.py:182).  Phase 2, after every wave has read q,
// publishes the next step's q from the new state, c2 A loaded through an
// eight-row ring from the L2-resident plane.  Lanes t >= 750 mirror thread 749
// (same reads, same state); they write nothing but phase 2's q, whose values
// equal thread 749's.
//
// C2A = false (callers without the c2 * A plane): `c2a` is c2 itself; phase 2
// forms c2 * A with A = 0.5 * 2A from the row's sponge entry and zeroes q on
// the rim by the interior flag.
template <int NZ, int NX, int PAD, int ROW, bool C2A>
__global__ __launch_bounds__(kStripThreads) void adj_strip_kernel(Geom g, int t0, int t1,
                                                                  const float *__restrict__ c2a,
                                                                  const float *__restrict__ A,
                                                                  const float *__restrict__ B,
                                                                  const float *__restrict__ w,
                                                                  const float *__restrict__ sd,
                                                                  const float *__restrict__ res,
                                                                  const float *__restrict__ hist,
                                                                  float *__restrict__ m1g, float *__restrict__ m2g,
                                                                  float *__restrict__ gaccg,
                                                                  float *__restrict__ illumg,
                                                                  float *__restrict__ src_grad) {
  extern __shared__ float lds[];
  constexpr int nz = NZ, nx = NX, pad = PAD, row = ROW;
  constexpr int H = nz + 2 * pad, W = nx + 2 * pad, HW = H * W;
  constexpr int R = kStripRows, NC = W / 2, NS = H / R, NACT = NC * NS;
  static_assert(W % 2 == 0 && nx % 2 == 0 && pad % 2 == 0, "pairs need even widths");
  static_assert(H % R == 0 && NACT <= kStripThreads && R >= 4, "strips must tile the plane");
  static_assert((pad + 1) * (pad + 1) * 16 <= (int)kPTabMask + 16, "pair-table offsets must fit the descriptor");
  constexpr int nzx = nz * nx;
  constexpr int guard = pair_guard(W), tabf = pair_table_floats(pad);
  constexpr int RS = (pad + row) / R, RK = (pad + row) % R;  // strip and slot of the receiver row
  constexpr uint32_t kRow = 4u * W;
  constexpr int DH = 3, DC = 8;  // history rows loaded ahead; depth of the c2 A ring
  float4 *tab = reinterpret_cast<float4 *>(lds);
  float *il = lds + tabf;
  constexpr uint32_t kIlAddr = (uint32_t)tabf * 4;
  float *Q = il + nzx + guard;
  float2 *Q2 = reinterpret_cast<float2 *>(Q);
  constexpr uint32_t kQAddr = (uint32_t)(tabf + nzx + guard) * 4;
  const int64_t prop = xcd_prop(g);
  if (prop < 0) return;
  const int b = (int)(prop / g.ns), s = (int)(prop - (int64_t)b * g.ns);
  const int t = threadIdx.x;
  const bool act = t < NACT;
  const int tt = act ? t : NACT - 1;
  const int strip = tt / NC, c = tt - strip * NC, r0 = strip * R;

  for (int k = t; k < guard; k += kStripThreads) {
    Q[k - guard] = 0.0f;
    Q[HW + k] = 0.0f;
  }
  // with the c2 * A plane (the native driver, whose forward leaves the rim's history
  // unspecified) the rim's sponge entries are zero: the imaging sum (A mu) L stays +0
  // there as with the reference's zero Laplacian, and mu on the rim, which only that
  // sum, q (zero there through c2 A) and the rim's own state read, is left unspecified
  fill_pair_table(tab, A, B, W, pad, C2A ? 2 : 1);
  for (int k = t; k < nzx; k += kStripThreads) il[k] = illumg[prop * nzx + k];

  float2 m1[R], m2[R], ga[R];
  uint32_t tb[R];
  bool mdl[R], inner[R];
  const bool colmodel = act && c >= pad / 2 && c < (pad + nx) / 2;
  const float2 *m1p = reinterpret_cast<const float2 *>(m1g + prop * HW);
  const float2 *m2p = reinterpret_cast<const float2 *>(m2g + prop * HW);
  const float2 *gap = reinterpret_cast<const float2 *>(gaccg + prop * HW);
  const float2 *cap = reinterpret_cast<const float2 *>(c2a + (int64_t)b * HW);
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int q = (r0 + k) * NC + c;
    m1[k] = m1p[q];
    m2[k] = m2p[q];
    ga[k] = gap[q];
    const uint32_t d = pair_desc(q, H, W, nz, nx, pad);
    tb[k] = d & kPTabMask;
    inner[k] = (d & kPInterior) != 0;
    mdl[k] = colmodel && r0 + k >= pad && r0 + k < pad + nz;
  }
  // q = ((c2 A) m) / 12 of row k from c2 A (C2A) or c2 and the row's (2A) entries
  auto qrow = [&](int k, float2 cv, float2 m) __attribute__((always_inline)) {
    if (C2A) return div12x2(cv.x * m.x, cv.y * m.y);
    const float4 e = lds_ld4(tb[k]);
    const float2 qq = div12x2((cv.x * (0.5f * e.x)) * m.x, (cv.y * (0.5f * e.z)) * m.y);
    return inner[k] ? qq : make_float2(0.0f, 0.0f);
  };
  __syncthreads();  // the sponge table, for qrow
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int q = (r0 + k) * NC + c;
    if (act) Q2[q] = qrow(k, cap[q], m1[k]);
  }
  const int srcx = pad + g.src[s];
  const bool own_src = act && strip == RS && c == (srcx >> 1);
  const bool src_hi = (srcx & 1) != 0;
  const float sdv = sd[prop];
  float sg = own_src ? src_grad[prop] : 0.0f;
  const bool rec_thread = act && strip == RS && c >= pad / 2 && c < (pad + nx) / 2;
  const int nslots = t1 - t0;
  // uniform resources; lanes with nothing to load read past the resource's end (zeros)
  const __amdgpu_buffer_rsrc_t rrs = make_rsrc(res + prop * g.nt * nx, (int64_t)g.nt * nx);
  const __amdgpu_buffer_rsrc_t hrs = make_rsrc(hist + prop * nslots * HW, (int64_t)nslots * HW);
  const __amdgpu_buffer_rsrc_t crs = make_rsrc(c2a + (int64_t)b * HW, (int64_t)HW);
  const int hv = (r0 * W + 2 * c) * 4;              // the thread's first pair in a plane (bytes)
  const uint32_t a = kQAddr + (uint32_t)(((r0 - 2) * W + 2 * c) * 4);
  const uint32_t ia0 = kIlAddr + (uint32_t)(((r0 - pad) * nx + 2 * c - pad) * 4);  // model rows only
  __syncthreads();

  // history of a step's first DH rows and the receiver lanes' residual pair, loaded at
  // the end of the previous step
  float2 Lh[DH], rv;
  auto load_step = [&](int it) __attribute__((always_inline)) {
#pragma unroll
    for (int d = 0; d < DH; ++d) Lh[d] = buf_load2(hrs, hv, (it - t0) * HW * 4 + d * (int)kRow);
    // receiver lanes (strip RS) read their pair of the residual row; other lanes read
    // some other in-range word or past the resource (zero), which they never use
    rv = buf_load2(rrs, hv - (RS * R * W + pad) * 4, it * nx * 4);
  };
  load_step(t1 - 1);
  auto step = [&](float2(&M1)[R], float2(&M2)[R], const int it) __attribute__((always_inline)) {
    const int slot = it - t0;
    const float wt = w[it];
    uint32_t aa = a;
    asm volatile("" : "+v"(aa));
    float2 z0 = lds_ld2(aa), z1 = lds_ld2(aa + kRow), z2 = lds_ld2(aa + 2 * kRow), z3 = lds_ld2(aa + 3 * kRow);
    float2 z4 = lds_ld2(aa + 4 * kRow);
    float2 lf = lds_ld2(aa + 2 * kRow - 8), rt = lds_ld2(aa + 2 * kRow + 8);
    float4 e = lds_ld4(tb[0]);
    float2 msrc = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int k = 0; k < R; ++k) {
      asm volatile("" : "+v"(aa));  // one row at a time
      const uint32_t ar = aa + (uint32_t)(k + 2) * kRow;
      float2 z5 = make_float2(0.0f, 0.0f), lfn = z5, rtn = z5;
      float4 en = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (k + 1 < R) {  // next row's centre (row r+3), neighbours and coefficients, one row ahead
        z5 = lds_ld2(ar + 3 * kRow);
        lfn = lds_ld2(ar + kRow - 8);
        rtn = lds_ld2(ar + kRow + 8);
        en = lds_ld4(tb[k + 1]);
      }
      const float2 L = Lh[k % DH];
      if (k + DH < R) Lh[k % DH] = buf_load2(hrs, hv, slot * HW * 4 + (k + DH) * (int)kRow);
      const float lt0 = lapt_sum(z2.x, z4.x, z3.x, z1.x, z0.x, rt.x, z2.y, lf.y, lf.x);
      const float lt1 = lapt_sum(z2.y, z4.y, z3.y, z1.y, z0.y, rt.y, rt.x, z2.x, lf.y);
      float mu0 = e.x * M1[k].x;
      float mu1 = e.z * M1[k].y;
      mu0 = mu0 + lt0;
      mu1 = mu1 + lt1;
      mu0 = mu0 - e.y * M2[k].x;
      mu1 = mu1 - e.w * M2[k].y;
      if (k == RK && rec_thread) {  // gpu.py:180
        mu0 = mu0 + rv.x;
        mu1 = mu1 + rv.y;
      }
      ga[k].x = ga[k].x + ((0.5f * e.x) * mu0) * L.x;  // A = 0.5 (2A), exact
      ga[k].y = ga[k].y + ((0.5f * e.z) * mu1) * L.y;
      if (k == RK) msrc = make_float2(mu0, mu1);
      if (mdl[k]) {
        const uint32_t ia = ia0 + (uint32_t)(k * nx * 4);
        float2 v = lds_ld2(ia);
        v.x = v.x + L.x * L.x;
        v.y = v.y + L.y * L.y;
        lds_st2(ia, v);
      }
      M2[k] = make_float2(mu0, mu1);  // the new state; M1 (the old state) becomes the next step's m2
      asm volatile("" : "+v"(ga[k].x), "+v"(ga[k].y), "+v"(M2[k].x), "+v"(M2[k].y));
      z0 = z1;
      z1 = z2;
      z2 = z3;
      z3 = z4;
      z4 = z5;
      lf = lfn;
      rt = rtn;
      e = en;
    }
    if (own_src) sg = sg + ((src_hi ? msrc.y : msrc.x) * wt) * sdv;  // gpu.py:182, in step order
    // Loads complete in issue order (one counter for all of them), so a wait for a
    // c2 A row also waits for every history load issued before it: phase 2 reads c2 A
    // from a ring DC rows deep whose first DC loads are issued before the barrier, and
    // the next step's first history rows are issued only after phase 2.
    float2 cb[DC];
#pragma unroll
    for (int k = 0; k < DC; ++k) cb[k] = buf_load2(crs, hv, k * (int)kRow);
    lds_barrier();
    // Idle lanes store too: they mirror thread 749 and write its values to its cells.
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const float2 cv = cb[k % DC];
      if (k + DC < R) cb[k % DC] = buf_load2(crs, hv, (k + DC) * (int)kRow);
      lds_st2(aa + (uint32_t)(k + 2) * kRow, qrow(k, cv, M2[k]));
    }
    if (it - 1 >= t0) load_step(it - 1);
    lds_barrier();
  };
  int it = t1 - 1;
  for (; it - 1 >= t0; it -= 2) {
    step(m1, m2, it);
    step(m2, m1, it - 1);
  }
  const bool odd = it >= t0;  // one step left: the current state is then in m2
  if (odd) step(m1, m2, it);
  if (act) {
    float2 *m1o = reinterpret_cast<float2 *>(m1g + prop * HW);
    float2 *m2o = reinterpret_cast<float2 *>(m2g + prop * HW);
    float2 *gao = reinterpret_cast<float2 *>(gaccg + prop * HW);
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int q = (r0 + k) * NC + c;
      m1o[q] = odd ? m2[k] : m1[k];
      m2o[q] = odd ? m1[k] : m2[k];
      gao[q] = ga[k];
    }
  }
  if (own_src) src_grad[prop] = sg;
  for (int k = t; k < nzx; k += kStripThreads) illumg[prop * nzx + k] = il[k];
}

// c2a = c2 * A on the interior cells of every plane of c2 (nb, H, W), zero on the
// two-cell rim: the first product of the adjoint's q = ((c2 A) m1) / 12 (gpu.py:179),
// rounded once as there.  Zero on the rim makes q zero there, as lap_adjoint only
// reads p[2:-2, 2:-2] (gpu.py:41).
__global__ void c2a_kernel(int64_t n, int H, int W, const float *__restrict__ c2, const float *__restrict__ A,
                           float *__restrict__ out) {
  const int64_t HW = (int64_t)H * W;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n * HW; i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % HW), z = c / W, x = c - z * W;
    out[i] = (z >= 2 && z < H - 2 && x >= 2 && x < W - 2) ? c2[i] * A[c] : 0.0f;
  }
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

int slots_for(const Geom &g, int threads) { return (g.H * g.W + threads - 1) / threads; }

// The OpenFWI geometry (70 x 70 model, pad 40, sources/receivers on row 1)
// is compiled with its dimensions as constants; everything else runs a
// runtime-shape instantiation with the smallest slot count >= needed.
bool is_openfwi(const Geom &g) { return g.nz == 70 && g.nx == 70 && g.pad == 40 && g.row == 1; }
template <typename Fn>
int dispatch_fwd(const Geom &g, Fn &&fn) {
  using std::integral_constant;
  if (is_openfwi(g)) return fn(integral_constant<int, 22>(), std::true_type());
  // slot counts whose register allocation is spill-free in the time loop
  const int k = slots_for(g, kFwdThreads);
  if (k <= 4) return fn(integral_constant<int, 4>(), std::false_type());
  if (k <= 8) return fn(integral_constant<int, 8>(), std::false_type());
  return fn(integral_constant<int, kFwdMaxSlots>(), std::false_type());
}

template <typename Fn>
int dispatch_adj(const Geom &g, Fn &&fn) {
  using std::integral_constant;
  if (is_openfwi(g)) return fn(integral_constant<int, kAdj150>(), std::true_type());
  const int k = slots_for(g, kAdjThreads);
  if (k <= 8) return fn(integral_constant<int, 8>(), std::false_type());
  return fn(integral_constant<int, kAdjMaxSlots>(), std::false_type());
}

// The paired kernels address their sponge table at LDS address 0, which holds
// only when the kernel has no static LDS ahead of its dynamic LDS.
bool no_static_lds(const void *kern) {
  hipFuncAttributes attr;
  return hipFuncGetAttributes(&attr, kern) == hipSuccess && attr.sharedSizeBytes == 0;
}

}  // namespace

int64_t resident_max_cells() {
  const int64_t a = (int64_t)kFwdMaxSlots * kFwdThreads, b = (int64_t)kAdjMaxSlots * kAdjThreads;
  return a < b ? a : b;
}

bool resident_fits(const Geom &g) {
  if (slots_for(g, kFwdThreads) > kFwdMaxSlots || slots_for(g, kAdjThreads) > kAdjMaxSlots) return false;
  if ((g.pad + 1) * (g.pad + 1) > (int)kIdxMask + 1 || g.nz * g.nx > (int)kIdxMask + 1) return false;
  if ((int64_t)g.nt * g.H * g.W > 0x1fffffff) return false;  // history slots addressed with 32-bit byte offsets
  const size_t need = is_openfwi(g) ? std::max(adj_strip_lds_bytes(g), fwd_strip_lds_bytes(g)) : adj_lds_bytes(g);
  return need <= (size_t)device_lds_limit();
}

int forward_resident(const Geom &g, int t0, int t1, const float *c2, const float *A, const float *B, const float *w,
                     const float *amp, float *p0, float *p1, float *gathers, float *hist, float *ckpt,
                     int ckpt_every, hipStream_t st, bool exact_rim) {
  const size_t lds = is_openfwi(g) ? fwd_strip_lds_bytes(g) : fwd_lds_bytes(g);
  return dispatch_fwd(g, [&](auto kc, auto fixed) -> int {
    constexpr bool F = decltype(fixed)::value;
    // the fixed 150 x 150 plane runs the column-strip kernel
    auto kern = F ? (exact_rim ? fwd_strip_kernel<70, 70, 40, 1, true> : fwd_strip_kernel<70, 70, 40, 1, false>)
                  : fwd_resident_kernel<decltype(kc)::value, F ? 70 : 0, F ? 70 : 0, F ? 40 : 0, F ? 1 : -1>;
    hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return hip_status(e, "forward_resident attribute");
    if (F && !no_static_lds((const void *)kern)) return hip_status(hipErrorInvalidValue, "forward_resident static LDS");
    kern<<<dim3((unsigned)xcd_blocks((int64_t)g.nb * g.ns)), dim3(F ? kStripThreads : kFwdThreads), lds, st>>>(
        g, t0, t1, c2, A, B, w, amp, p0,
                                                                                 p1, gathers, hist, ckpt, ckpt_every);
    return hip_status(hipGetLastError(), "forward_resident launch");
  });
}

int adjoint_resident(const Geom &g, int t0, int t1, const float *c2, const float *A, const float *B, const float *w,
                     const float *sd, const float *res, const float *hist, float *m1, float *m2, float *gacc,
                     float *illum, float *src_grad, hipStream_t st, const float *c2a) {
  const bool pre = c2a != nullptr;
  const size_t lds = is_openfwi(g) ? adj_strip_lds_bytes(g) : adj_lds_bytes(g);
  return dispatch_adj(g, [&](auto kc, auto fixed) -> int {
    constexpr bool F = decltype(fixed)::value;
    // the fixed 150 x 150 plane runs the column-strip kernel, with the c2 * A plane
    // when the caller has one
    auto kern = F ? (pre ? adj_strip_kernel<70, 70, 40, 1, true> : adj_strip_kernel<70, 70, 40, 1, false>)
                  : adj_resident_kernel<decltype(kc)::value, F ? 70 : 0, F ? 70 : 0, F ? 40 : 0, F ? 1 : -1>;
    hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return hip_status(e, "adjoint_resident attribute");
    if (F && !no_static_lds((const void *)kern)) return hip_status(hipErrorInvalidValue, "adjoint_resident static LDS");
    kern<<<dim3((unsigned)xcd_blocks((int64_t)g.nb * g.ns)), dim3(F ? kStripThreads : kAdjThreads), lds, st>>>(
        g, t0, t1, F && pre ? c2a : c2, A, B, w, sd, res, hist, m1, m2, gacc, illum, src_grad);
    return hip_status(hipGetLastError(), "adjoint_resident launch");
  });
}

bool resident_uses_c2a(const Geom &g) { return is_openfwi(g); }

int c2a_plane(int64_t nb, int H, int W, const float *c2, const float *A, float *out, hipStream_t st) {
  const int64_t n = nb * H * W;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
  c2a_kernel<<<dim3(blocks), dim3(256), 0, st>>>(nb, H, W, c2, A, out);
  return hip_status(hipGetLastError(), "c2a launch");
}

}  // namespace ss
