// Mixed-size sweep: tiled, temporally blocked FP32 forward and adjoint over
// several grid-size groups in one launch per step range.
//
// A group is a batch of propagations on one padded H x W grid with its own
// step count, shot count and source columns.  A work item is one tile of one
// propagation.  One launch advances every active item of every group by its
// group's temporal depth T: the workgroup loads the item's extended region
// (the tile plus a halo of 2T cells, clipped to the grid) into LDS and
// registers and steps it T times.  After step s only cells at least 2s cells
// inside the region's cut edges are exact; after T steps that is the tile.
// Halo cells are recomputed redundantly, source and receiver injections
// included, so every copy of a cell follows the same arithmetic; only the
// tile's own cells write the Laplacian history, receiver samples, imaging
// sum and illumination.  The state is double-buffered in HBM between
// launches (launch k reads buffer k % 2 and writes the other one), and
// launch 0 starts from zero state and zero accumulators without reading them.
//
// Groups with more launches come first in the item list, so the items active
// in launch k are a prefix of it; groups drop out as their time range ends.
// The launch sequence is captured once into a hipGraph per batch layout and
// replayed.  Tile shape and depth per group come from the caller, from the
// autotuner (trial launches timed with HIP events), or from a cost model.
//
// When all plane widths, pads and model widths are even, paired variants run
// (two adjacent cells per thread slot, 8-byte LDS and buffer accesses); the
// cell kernels handle every other layout.  Arithmetic per cell is that of the
// stepwise kernels (ss_common.h), so the results are bit-identical to them:
// forward gpu.py:111-115 and 126-135, adjoint gpu.py:177-184.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "ss_device.h"
#include "ss_kernels.h"

namespace ss {

namespace {

constexpr int kThreads = 1024;
constexpr int kMaxGroups = SS_SWEEP_MAX_GROUPS;
constexpr int kFwdMaxK = 16;  // cells per thread: forward keeps p(t-1), c2 and a descriptor
constexpr int kAdjMaxK = 12;  // adjoint keeps m1, m2, the imaging sum, c2*A and a descriptor
constexpr int kMaxLdsFloats = 160 * 1024 / 4;

// 32-bit cell descriptor: bits 0-16 offset in the padded plane, bits 17-27
// sponge table index, then flags.
constexpr uint32_t kOffMask = 0x1ffff;
constexpr int kTabShift = 17;
constexpr uint32_t kTabMask = 0x7ff;
constexpr uint32_t kInt = 1u << 28;  // Laplacian interior [2, H-3] x [2, W-3]
constexpr uint32_t kOwn = 1u << 29;  // inside the item's own tile
constexpr uint32_t kRow = 1u << 30;  // adjoint: receiver row, model column (every copy)
constexpr uint32_t kSrc = 1u << 31;  // source cell (forward: every copy; adjoint: own cell)
constexpr int kMaxPlane = 1 << 17;
constexpr int kMaxTable = 1 << 11;
// A raw buffer access whose offset is at least the resource's size is
// discarded (stores) or reads zero (loads).  Every resource here covers fewer
// than 2^31 bytes, so this offset masks an access without a branch.
constexpr int kNoAccess = (int)0x80000000u;

struct SGroup {
  const float *c2, *A, *B, *w, *amp, *res;  // amp: forward source amplitude / adjoint source derivative
  float *st;                                // (2 buffers, 2 planes, nprop, H, W)
  float *gathers, *hist, *gacc, *illum, *sgrad;
  int ns, nz, nx, pad, nt, row, H, W;
  int th, tw, T, ntx, ntiles, nlaunch, item0, nitems;
  int tlo, thi;        // step range of this call; hist holds steps [tlo, thi)
  int resume, phase;   // resume: launch 0 reads state and sums; launch k reads state buffer (k + phase) % 2
  int src[SS_MAX_SHOTS];
};

// LDS layout (floats): [sponge table | guard | plane 0 | guard | plane 1 | guard | illumination].
// The table starts at LDS address 0 (tab_off 0): 8-byte (A, B), (A, A*B) or,
// for the paired adjoint, (2A, A*B) entries.
struct SArgs {
  SGroup g[kMaxGroups];
  int ngroups, k, nactive;
  int plane0, plane1, tab_off, il_off;
};

struct ItemView {
  int prop, b, s;
  int y0, x0, y1, x1;  // own cells
  int ey0, ex0, EH, EW;
};

__device__ __forceinline__ int find_group(const SArgs &a, int item) {
  int gi = 0;
#pragma unroll
  for (int i = 1; i < kMaxGroups; ++i)
    if (i < a.ngroups && item >= a.g[i].item0) gi = i;
  return gi;
}

__device__ __forceinline__ ItemView item_view(const SGroup &G, int item) {
  ItemView v;
  const int local = item - G.item0;
  v.prop = local / G.ntiles;
  const int tile = local - v.prop * G.ntiles;
  v.b = v.prop / G.ns;
  v.s = v.prop - v.b * G.ns;
  const int ty = tile / G.ntx, tx = tile - ty * G.ntx;
  v.y0 = ty * G.th;
  v.x0 = tx * G.tw;
  v.y1 = min(v.y0 + G.th, G.H);
  v.x1 = min(v.x0 + G.tw, G.W);
  v.ey0 = max(0, v.y0 - 2 * G.T);
  v.ex0 = max(0, v.x0 - 2 * G.T);
  v.EH = min(G.H, v.y1 + 2 * G.T) - v.ey0;
  v.EW = min(G.W, v.x1 + 2 * G.T) - v.ex0;
  return v;
}

// Workgroup -> item: blocks are dealt round-robin to the 8 XCDs; XCD x takes
// the contiguous item range [x*per, (x+1)*per), so the tiles of one
// propagation (which read each other's halos) and the shots of one velocity
// row (which share c2) stay in one XCD's L2.
__device__ __forceinline__ int xcd_item() {
  const int per = (int)gridDim.x / kXcds;
  return (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
}

__device__ __forceinline__ void sweep_table(float2 *tab, const SGroup &G, int mode) {
  const int n = (G.pad + 1) * (G.pad + 1);
  for (int k = threadIdx.x; k < n; k += kThreads) {
    const int dz = k / (G.pad + 1), dx = k - dz * (G.pad + 1);
    const int c = (G.pad - dz) * G.W + (G.pad - dx);
    const float a = G.A[c], b = G.B[c];
    tab[k] = make_float2(a, mode ? a * b : b);
  }
}

// Zero the stencil overhang (2 rows + 2 cells) before and after both planes.
__device__ __forceinline__ void zero_guards(float *P0, float *P1, int EHW, int EW) {
  const int gw = 2 * EW + 2;
  for (int q = threadIdx.x; q < gw; q += kThreads) {
    P0[-1 - q] = 0.0f;
    P0[EHW + q] = 0.0f;
    P1[-1 - q] = 0.0f;
    P1[EHW + q] = 0.0f;
  }
}

// Forward: plane Pc holds p(t); each thread keeps p(t-1), c2 and descriptors
// of its cells in registers and writes p(t+1) to the other plane, so one
// barrier per step suffices.  Receiver samples of the own tile are copied
// from LDS by the first threads at the start of the next step.
template <int K>
__global__ __launch_bounds__(kThreads) void sweep_fwd_kernel(SArgs a) {
  extern __shared__ float lds[];
  const int item = xcd_item();
  if (item >= a.nactive) return;  // padding workgroup: exits before any barrier
  const SGroup &G = a.g[find_group(a, item)];
  const ItemView v = item_view(G, item);
  const int H = G.H, W = G.W, pad = G.pad, HW = H * W;
  const int EW = v.EW, EHW = v.EH * v.EW;
  const int t0 = G.tlo + a.k * G.T, n = min(G.T, G.thi - t0);
  const int64_t PS = (int64_t)(G.nitems / G.ntiles) * HW;
  const int ib = (a.k + G.phase) & 1;
  const float *p0i = G.st + 2 * ib * PS + (int64_t)v.prop * HW, *p1i = p0i + PS;
  float *p0o = G.st + 2 * (ib ^ 1) * PS + (int64_t)v.prop * HW, *p1o = p0o + PS;
  float *Pc = lds + a.plane0, *Pn = lds + a.plane1;
  float2 *tab = reinterpret_cast<float2 *>(lds + a.tab_off);
  const int t = threadIdx.x;
  zero_guards(Pc, Pn, EHW, EW);
  sweep_table(tab, G, 0);

  const uint32_t magic = 0xffffffffu / (uint32_t)EW + 1u;
  const int r = pad + G.row, sxc = pad + G.src[v.s];
  const float *c2b = G.c2 + (int64_t)v.b * HW;
  float pn[K], cc[K];
  uint32_t dsc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int e = t + j * kThreads;
    pn[j] = 0.0f;
    cc[j] = 0.0f;
    dsc[j] = 0;
    if (e < EHW) {
      const int q = (int)magic_div((uint32_t)e, magic);
      const int z = v.ey0 + q, x = v.ex0 + (e - q * EW);
      const int gi = z * W + x;
      uint32_t d = (uint32_t)gi | ((uint32_t)(sponge_dist(z, H, pad) * (pad + 1) + sponge_dist(x, W, pad)) << kTabShift);
      if (z >= 2 && z < H - 2 && x >= 2 && x < W - 2) d |= kInt;
      if (z >= v.y0 && z < v.y1 && x >= v.x0 && x < v.x1) d |= kOwn;
      if (z == r && x == sxc) d |= kSrc;
      dsc[j] = d;
      cc[j] = c2b[gi];
      if (a.k > 0 || G.resume) {
        pn[j] = p0i[gi];
        Pc[e] = p1i[gi];
      } else {
        Pc[e] = 0.0f;
      }
    }
  }
  const int64_t hspan = G.thi - G.tlo;  // history slots of this call
  // without a history buffer (a forward that only advances the state) the
  // resource has size 0 and the hardware drops the Laplacian stores
  const __amdgpu_buffer_rsrc_t hrs =
      make_rsrc(G.hist ? G.hist + (int64_t)v.prop * hspan * HW : G.st, G.hist ? hspan * HW : 0);
  // own receivers: row r, model columns [xa, xb) of this tile
  const int xa = max(v.x0, pad), xb = min(v.x1, pad + G.nx);
  const bool rec = G.gathers != nullptr && r >= v.y0 && r < v.y1 && t < xb - xa;
  const int rec_e = (r - v.ey0) * EW + (xa + t - v.ex0);
  float *grec = G.gathers ? G.gathers + (int64_t)v.prop * G.nt * G.nx + (xa - pad) + t : nullptr;
  const float ampv = G.amp[v.prop];
  __syncthreads();

  for (int st = 0; st < n; ++st) {
    const int it = t0 + st;
    if (rec && st > 0) grec[(int64_t)(it - 1) * G.nx] = Pc[rec_e];  // p(it) is the sample of step it-1
    const float wsrc = G.w[it] * ampv;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int e = t + j * kThreads;
      if (e < EHW) {
        const uint32_t d = dsc[j];
        const float c = Pc[e];
        const float s = lap_sum(Pc[e - 2 * EW], Pc[e - EW], c, Pc[e + EW], Pc[e + 2 * EW], Pc[e - 2], Pc[e - 1],
                                Pc[e + 1], Pc[e + 2]);
        const float L = (d & kInt) ? div12(s) : 0.0f;
        const float2 ab = tab[(d >> kTabShift) & kTabMask];
        float p2 = leapfrog(c, pn[j], ab.y, cc[j], L, ab.x);
        const float ps = p2 + wsrc;
        p2 = (d & kSrc) ? ps : p2;
        buf_store(L, hrs, (d & kOwn) ? (int)(d & kOffMask) * 4 : kNoAccess, (it - G.tlo) * HW * 4);
        Pn[e] = p2;
        pn[j] = c;
      }
    }
    lds_barrier();  // the history and receiver accesses stay in flight
    float *tmp = Pc;
    Pc = Pn;
    Pn = tmp;
  }
  if (rec) grec[(int64_t)(t0 + n - 1) * G.nx] = Pc[rec_e];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int e = t + j * kThreads;
    if (e < EHW && (dsc[j] & kOwn)) {
      const int gi = (int)(dsc[j] & kOffMask);
      p0o[gi] = pn[j];
      p1o[gi] = Pc[e];
    }
  }
}

// Adjoint: plane Qc holds q = (c2 A m1)/12 on the interior (zero on the
// two-cell border); each thread keeps m1, m2, c2*A, the imaging sum and
// descriptors in registers and writes the next q to the other plane.
template <int K>
__global__ __launch_bounds__(kThreads) void sweep_adj_kernel(SArgs a) {
  extern __shared__ float lds[];
  const int item = xcd_item();
  if (item >= a.nactive) return;
  const SGroup &G = a.g[find_group(a, item)];
  const ItemView v = item_view(G, item);
  const int H = G.H, W = G.W, pad = G.pad, HW = H * W;
  const int EW = v.EW, EHW = v.EH * v.EW;
  const int hi = G.thi - a.k * G.T, n = min(G.T, hi - G.tlo);
  const int64_t PS = (int64_t)(G.nitems / G.ntiles) * HW;
  const int ib = (a.k + G.phase) & 1;
  const float *m1i = G.st + 2 * ib * PS + (int64_t)v.prop * HW, *m2i = m1i + PS;
  float *m1o = G.st + 2 * (ib ^ 1) * PS + (int64_t)v.prop * HW, *m2o = m1o + PS;
  float *Qc = lds + a.plane0, *Qn = lds + a.plane1;
  float2 *tab = reinterpret_cast<float2 *>(lds + a.tab_off);
  float *il = lds + a.il_off;
  const int t = threadIdx.x;
  zero_guards(Qc, Qn, EHW, EW);
  sweep_table(tab, G, 1);
  __syncthreads();

  const uint32_t magic = 0xffffffffu / (uint32_t)EW + 1u;
  const int r = pad + G.row, sxc = pad + G.src[v.s];
  const float *c2b = G.c2 + (int64_t)v.b * HW;
  float *gac = G.gacc + (int64_t)v.prop * HW;
  float *ilg = G.illum + (int64_t)v.prop * G.nz * G.nx;
  float m1[K], m2[K], ga[K], ca[K];
  uint32_t dsc[K];
  float sg = 0.0f;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int e = t + j * kThreads;
    m1[j] = m2[j] = ga[j] = ca[j] = 0.0f;
    dsc[j] = 0;
    if (e < EHW) {
      const int q = (int)magic_div((uint32_t)e, magic);
      const int z = v.ey0 + q, x = v.ex0 + (e - q * EW);
      const int gi = z * W + x;
      const uint32_t ti = (uint32_t)(sponge_dist(z, H, pad) * (pad + 1) + sponge_dist(x, W, pad));
      uint32_t d = (uint32_t)gi | (ti << kTabShift);
      const bool interior = z >= 2 && z < H - 2 && x >= 2 && x < W - 2;
      if (interior) d |= kInt;
      const bool own = z >= v.y0 && z < v.y1 && x >= v.x0 && x < v.x1;
      const bool model = z >= pad && z < pad + G.nz && x >= pad && x < pad + G.nx;
      if (own) d |= kOwn;
      if (z == r && x >= pad && x < pad + G.nx) d |= kRow;
      if (own && z == r && x == sxc) {
        d |= kSrc;
        sg = (a.k > 0 || G.resume) ? G.sgrad[v.prop] : 0.0f;
      }
      dsc[j] = d;
      ca[j] = c2b[gi] * tab[ti].x;  // c2 * A (gpu.py:179)
      float ilv = 0.0f;
      if (a.k > 0 || G.resume) {
        m1[j] = m1i[gi];
        m2[j] = m2i[gi];
        if (own) ga[j] = gac[gi];
        if (own && model) ilv = ilg[(z - pad) * G.nx + (x - pad)];
      }
      il[e] = ilv;
      Qc[e] = interior ? div12(ca[j] * m1[j]) : 0.0f;
    }
  }
  const int64_t hspan = G.thi - G.tlo;  // history slots of this call
  const __amdgpu_buffer_rsrc_t hrs = make_rsrc(G.hist + (int64_t)v.prop * hspan * HW, hspan * HW);
  const __amdgpu_buffer_rsrc_t rrs = make_rsrc(G.res + (int64_t)v.prop * G.nt * G.nx, (int64_t)G.nt * G.nx);
  const int rec0 = r * W + pad;
  const float sdv = G.amp[v.prop];
  __syncthreads();

  for (int st = 0; st < n; ++st) {
    const int it = hi - 1 - st;
    const float wt = G.w[it];
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int e = t + j * kThreads;
      if (e < EHW) {
        const uint32_t d = dsc[j];
        const int gi = (int)(d & kOffMask);
        const float L = buf_load(hrs, (d & kOwn) ? gi * 4 : kNoAccess, (it - G.tlo) * HW * 4);  // 0 off the own tile
        const float lt = lapt_sum(Qc[e], Qc[e + 2 * EW], Qc[e + EW], Qc[e - EW], Qc[e - 2 * EW], Qc[e + 2],
                                  Qc[e + 1], Qc[e - 1], Qc[e - 2]);
        const float2 aab = tab[(d >> kTabShift) & kTabMask];  // (A, A*B)
        float mu = (2.0f * aab.x) * m1[j];
        mu = mu + lt;
        mu = mu - aab.y * m2[j];
        if (d & kRow) mu = mu + buf_load(rrs, (gi - rec0) * 4, it * G.nx * 4);
        const float gn = ga[j] + (aab.x * mu) * L;
        ga[j] = (d & kOwn) ? gn : ga[j];
        il[e] = il[e] + L * L;  // unchanged off the own tile (L = 0)
        if (d & kSrc) sg = sg + (mu * wt) * sdv;
        m2[j] = m1[j];
        m1[j] = mu;
        Qn[e] = (d & kInt) ? div12(ca[j] * mu) : 0.0f;
      }
    }
    lds_barrier();  // the history and receiver accesses stay in flight
    float *tmp = Qc;
    Qc = Qn;
    Qn = tmp;
  }
  const uint32_t magic_w = 0xffffffffu / (uint32_t)W + 1u;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int e = t + j * kThreads;
    const uint32_t d = dsc[j];
    if (e < EHW && (d & kOwn)) {
      const int gi = (int)(d & kOffMask);
      m1o[gi] = m1[j];
      m2o[gi] = m2[j];
      gac[gi] = ga[j];
      const int z = (int)__umulhi((uint32_t)gi, magic_w), x = gi - z * W;
      if (z >= pad && z < pad + G.nz && x >= pad && x < pad + G.nx) ilg[(z - pad) * G.nx + (x - pad)] = il[e];
      if (d & kSrc) G.sgrad[v.prop] = sg;
    }
  }
}

// Paired variants, used when every group has an even plane width, pad and
// model width and every tile width is even: thread t owns the cell pairs
// (2p, 2p+1) of the extended region, p = t + 1024 j.  The nine taps of both
// cells come from seven aligned 8-byte LDS reads, the history moves two cells
// per buffer instruction, and interior, own-tile and receiver-row flags are the
// same for both cells (their ranges start on even columns).  Two descriptor
// words per pair: A = plane offset of the first cell | flags (one source flag
// per cell), B = the two cells' sponge-table byte offsets.  Per cell the
// arithmetic is that of the cell kernels above.  Step s (1-based) computes only
// the rows at least 2s rows inside a cut edge of the extended region (an edge
// inside the grid): they are the rows still exact after step s, they read only
// rows computed at step s-1, and they contain the own tile up to s = T.  The
// rows left out are never read again and never stored, so whole waves skip
// them while the band shrinks by two rows per step at each cut edge.
constexpr uint32_t kQInt = 1u << 17, kQOwn = 1u << 18, kQRow = 1u << 19, kQOwnModel = 1u << 20;
constexpr uint32_t kTabOff = 0x3ff8u;  // table byte offsets of 8-byte entries

// Adjoint table of the paired kernel: (2A, A*B); A = 0.5 * 2A exactly (A is a
// normal number in (0, 1]), so one 8-byte read gives all three coefficients.
__device__ __forceinline__ void sweep_table2a(float2 *tab, const SGroup &G) {
  const int n = (G.pad + 1) * (G.pad + 1);
  for (int k = threadIdx.x; k < n; k += kThreads) {
    const int dz = k / (G.pad + 1), dx = k - dz * (G.pad + 1);
    const int c = (G.pad - dz) * G.W + (G.pad - dx);
    const float a = G.A[c], b = G.B[c];
    tab[k] = make_float2(2.0f * a, a * b);
  }
}

template <int KP>
__global__ __launch_bounds__(kThreads) void sweep_fwd_pair_kernel(SArgs a) {
  extern __shared__ float lds[];
  const int item = xcd_item();
  if (item >= a.nactive) return;  // padding workgroup: exits before any barrier
  const SGroup &G = a.g[find_group(a, item)];
  const ItemView v = item_view(G, item);
  const int H = G.H, W = G.W, pad = G.pad, HW = H * W;
  const int EW = v.EW, EW2 = EW / 2, NPE = v.EH * v.EW / 2;
  const int t0 = G.tlo + a.k * G.T, n = min(G.T, G.thi - t0);
  const int64_t PS = (int64_t)(G.nitems / G.ntiles) * HW;
  const int ib = (a.k + G.phase) & 1;
  const float *p0i = G.st + 2 * ib * PS + (int64_t)v.prop * HW, *p1i = p0i + PS;
  float *p0o = G.st + 2 * (ib ^ 1) * PS + (int64_t)v.prop * HW, *p1o = p0o + PS;
  float *Pc = lds + a.plane0, *Pn = lds + a.plane1;
  float2 *tab = reinterpret_cast<float2 *>(lds + a.tab_off);
  const int t = threadIdx.x;
  zero_guards(Pc, Pn, 2 * NPE, EW);
  sweep_table(tab, G, 0);

  const uint32_t magic = 0xffffffffu / (uint32_t)EW + 1u;
  const int r = pad + G.row, sxc = pad + G.src[v.s];
  // the source cell's pair in the extended region (every item that computes the cell injects), uniform
  const int se = (r - v.ey0) * EW + (sxc - v.ex0);
  const bool src_in = r >= v.ey0 && r < v.ey0 + v.EH && sxc >= v.ex0 && sxc < v.ex0 + EW;
  const int src_j = src_in ? (se >> 1) / kThreads : -1, src_t = (se >> 1) % kThreads, src_hi = se & 1;
  const float2 *c2b = reinterpret_cast<const float2 *>(G.c2 + (int64_t)v.b * HW);
  float2 pn[KP], cc[KP];
  uint32_t da[KP], db[KP];
#pragma unroll
  for (int j = 0; j < KP; ++j) {
    const int p = t + j * kThreads;
    pn[j] = cc[j] = make_float2(0.0f, 0.0f);
    da[j] = db[j] = 0;
    if (p < NPE) {
      const int e = 2 * p;
      const int q = (int)magic_div((uint32_t)e, magic);
      const int z = v.ey0 + q, x = v.ex0 + (e - q * EW);
      const int gi = z * W + x;
      uint32_t d = (uint32_t)gi;
      if (z >= 2 && z < H - 2 && x >= 2 && x < W - 2) d |= kQInt;
      if (z >= v.y0 && z < v.y1 && x >= v.x0 && x < v.x1) d |= kQOwn;
      da[j] = d;
      const int sz = sponge_dist(z, H, pad) * (pad + 1);
      db[j] = ((uint32_t)(sz + sponge_dist(x, W, pad)) << 3) | ((uint32_t)(sz + sponge_dist(x + 1, W, pad)) << 19);
      cc[j] = c2b[gi >> 1];
      float2 *P2 = reinterpret_cast<float2 *>(Pc);
      if (a.k > 0 || G.resume) {
        pn[j] = reinterpret_cast<const float2 *>(p0i)[gi >> 1];
        P2[p] = reinterpret_cast<const float2 *>(p1i)[gi >> 1];
      } else {
        P2[p] = make_float2(0.0f, 0.0f);
      }
    }
  }
  const int64_t hspan = G.thi - G.tlo;  // history slots of this call
  // without a history buffer (a forward that only advances the state) the
  // resource has size 0 and the hardware drops the Laplacian stores
  const __amdgpu_buffer_rsrc_t hrs =
      make_rsrc(G.hist ? G.hist + (int64_t)v.prop * hspan * HW : G.st, G.hist ? hspan * HW : 0);
  const int xa = max(v.x0, pad), xb = min(v.x1, pad + G.nx);
  const bool rec = G.gathers != nullptr && r >= v.y0 && r < v.y1 && t < xb - xa;
  const int rec_e = (r - v.ey0) * EW + (xa + t - v.ex0);
  float *grec = G.gathers ? G.gathers + (int64_t)v.prop * G.nt * G.nx + (xa - pad) + t : nullptr;
  const float ampv = G.amp[v.prop];
  const int cut_lo = v.ey0 > 0 ? 2 * EW2 : 0, cut_hi = v.ey0 + v.EH < H ? 2 * EW2 : 0;
  __syncthreads();

  for (int st = 0; st < n; ++st) {
    const int it = t0 + st;
    if (rec && st > 0) grec[(int64_t)(it - 1) * G.nx] = Pc[rec_e];  // p(it) is the sample of step it-1
    const float wsrc = G.w[it] * ampv;
    const float2 *C2 = reinterpret_cast<const float2 *>(Pc);
    float2 *N2 = reinterpret_cast<float2 *>(Pn);
    const int plo = cut_lo * (st + 1), phi = NPE - cut_hi * (st + 1);  // shrinking row band
#pragma unroll
    for (int j = 0; j < KP; ++j) {
      const int p = t + j * kThreads;
      if (p >= plo && p < phi) {
        const uint32_t d = da[j];
        const float2 r0 = C2[p - 1], r1 = C2[p], r2 = C2[p + 1];
        const float2 um = C2[p - EW2], u2m = C2[p - EW], up = C2[p + EW2], u2p = C2[p + EW];
        const float s0 = lap_sum(u2m.x, um.x, r1.x, up.x, u2p.x, r0.x, r0.y, r1.y, r2.x);
        const float s1 = lap_sum(u2m.y, um.y, r1.y, up.y, u2p.y, r0.y, r1.x, r2.x, r2.y);
        const float2 Lq = div12x2(s0, s1);
        const float L0 = keep_if<17>(Lq.x, d), L1 = keep_if<17>(Lq.y, d);  // interior (kQInt)
        const float2 ab0 = lds_table2(db[j] & kTabOff), ab1 = lds_table2((db[j] >> 16) & kTabOff);
        float q0 = leapfrog(r1.x, pn[j].x, ab0.y, cc[j].x, L0, ab0.x);
        float q1 = leapfrog(r1.y, pn[j].y, ab1.y, cc[j].y, L1, ab1.x);
        buf_store2(make_float2(L0, L1), hrs, (d & kQOwn) ? (int)(d & kOffMask) * 4 : kNoAccess, (it - G.tlo) * HW * 4);
        N2[p] = make_float2(q0, q1);
        pn[j] = r1;
      }
    }
    // the source, added after damping (gpu.py:114) by the thread that just wrote the source's pair
    // (once per step, not a select in every pair slot)
    if (src_j >= 0 && t == src_t) {
      float *ps = reinterpret_cast<float *>(N2 + t + src_j * kThreads) + src_hi;
      *ps = *ps + wsrc;
    }
    lds_barrier();  // the history and receiver accesses stay in flight
    float *tmp = Pc;
    Pc = Pn;
    Pn = tmp;
  }
  if (rec) grec[(int64_t)(t0 + n - 1) * G.nx] = Pc[rec_e];
#pragma unroll
  for (int j = 0; j < KP; ++j) {
    const int p = t + j * kThreads;
    if (p < NPE && (da[j] & kQOwn)) {
      const int gi = (int)(da[j] & kOffMask);
      reinterpret_cast<float2 *>(p0o)[gi >> 1] = pn[j];
      reinterpret_cast<float2 *>(p1o)[gi >> 1] = reinterpret_cast<const float2 *>(Pc)[p];
    }
  }
}

template <int KP>
__global__ __launch_bounds__(kThreads) void sweep_adj_pair_kernel(SArgs a) {
  extern __shared__ float lds[];
  const int item = xcd_item();
  if (item >= a.nactive) return;
  const SGroup &G = a.g[find_group(a, item)];
  const ItemView v = item_view(G, item);
  const int H = G.H, W = G.W, pad = G.pad, HW = H * W;
  const int EW = v.EW, EW2 = EW / 2, NPE = v.EH * v.EW / 2;
  const int hi = G.thi - a.k * G.T, n = min(G.T, hi - G.tlo);
  const int64_t PS = (int64_t)(G.nitems / G.ntiles) * HW;
  const int ib = (a.k + G.phase) & 1;
  const float2 *m1i = reinterpret_cast<const float2 *>(G.st + 2 * ib * PS + (int64_t)v.prop * HW);
  const float2 *m2i = reinterpret_cast<const float2 *>(G.st + (2 * ib + 1) * PS + (int64_t)v.prop * HW);
  float2 *m1o = reinterpret_cast<float2 *>(G.st + 2 * (ib ^ 1) * PS + (int64_t)v.prop * HW);
  float2 *m2o = reinterpret_cast<float2 *>(G.st + (2 * (ib ^ 1) + 1) * PS + (int64_t)v.prop * HW);
  float *Qc = lds + a.plane0, *Qn = lds + a.plane1;
  float2 *tab = reinterpret_cast<float2 *>(lds + a.tab_off);
  float2 *il = reinterpret_cast<float2 *>(lds + a.il_off);
  const int t = threadIdx.x;
  zero_guards(Qc, Qn, 2 * NPE, EW);
  sweep_table2a(tab, G);
  __syncthreads();

  const uint32_t magic = 0xffffffffu / (uint32_t)EW + 1u;
  const int r = pad + G.row, sxc = pad + G.src[v.s];
  // the source cell's pair when the source lies in the own tile (its owner sums the source gradient), uniform
  const int se = (r - v.ey0) * EW + (sxc - v.ex0);
  const bool src_own = r >= v.y0 && r < v.y1 && sxc >= v.x0 && sxc < v.x1;
  const int src_j = src_own ? (se >> 1) / kThreads : -1, src_t = (se >> 1) % kThreads, src_hi = se & 1;
  const float2 *c2b = reinterpret_cast<const float2 *>(G.c2 + (int64_t)v.b * HW);
  float2 *gac = reinterpret_cast<float2 *>(G.gacc + (int64_t)v.prop * HW);
  float2 *ilg = reinterpret_cast<float2 *>(G.illum + (int64_t)v.prop * G.nz * G.nx);
  float2 m1[KP], m2[KP], ga[KP], ca[KP];
  uint32_t da[KP], db[KP];
  float sg = 0.0f;
#pragma unroll
  for (int j = 0; j < KP; ++j) {
    const int p = t + j * kThreads;
    m1[j] = m2[j] = ga[j] = ca[j] = make_float2(0.0f, 0.0f);
    da[j] = db[j] = 0;
    if (p < NPE) {
      const int e = 2 * p;
      const int q = (int)magic_div((uint32_t)e, magic);
      const int z = v.ey0 + q, x = v.ex0 + (e - q * EW);
      const int gi = z * W + x;
      uint32_t d = (uint32_t)gi;
      const bool interior = z >= 2 && z < H - 2 && x >= 2 && x < W - 2;
      const bool own = z >= v.y0 && z < v.y1 && x >= v.x0 && x < v.x1;
      const bool model = z >= pad && z < pad + G.nz && x >= pad && x < pad + G.nx;
      if (interior) d |= kQInt;
      if (own) d |= kQOwn;
      if (own && model) d |= kQOwnModel;
      if (z == r && x >= pad && x < pad + G.nx) d |= kQRow;
      da[j] = d;
      const int sz = sponge_dist(z, H, pad) * (pad + 1);
      const uint32_t o0 = (uint32_t)(sz + sponge_dist(x, W, pad)) << 3;
      const uint32_t o1 = (uint32_t)(sz + sponge_dist(x + 1, W, pad)) << 3;
      db[j] = o0 | (o1 << 16);
      const float2 cv = c2b[gi >> 1];
      // c2 * A (gpu.py:179), A = 0.5 * 2A
      ca[j] = make_float2(cv.x * (0.5f * lds_table_a(o0)), cv.y * (0.5f * lds_table_a(o1)));
      float2 ilv = make_float2(0.0f, 0.0f);
      if (a.k > 0 || G.resume) {
        m1[j] = m1i[gi >> 1];
        m2[j] = m2i[gi >> 1];
        if (own) ga[j] = gac[gi >> 1];
        if (own && model) ilv = ilg[((z - pad) * G.nx + (x - pad)) >> 1];
      }
      il[p] = ilv;
      reinterpret_cast<float2 *>(Qc)[p] = make_float2(interior ? div12(ca[j].x * m1[j].x) : 0.0f,
                                                       interior ? div12(ca[j].y * m1[j].y) : 0.0f);
    }
  }
  if (src_j >= 0 && t == src_t) sg = (a.k > 0 || G.resume) ? G.sgrad[v.prop] : 0.0f;
  const int64_t hspan = G.thi - G.tlo;  // history slots of this call
  const __amdgpu_buffer_rsrc_t hrs = make_rsrc(G.hist + (int64_t)v.prop * hspan * HW, hspan * HW);
  const __amdgpu_buffer_rsrc_t rrs = make_rsrc(G.res + (int64_t)v.prop * G.nt * G.nx, (int64_t)G.nt * G.nx);
  const int rec0 = r * W + pad;
  const float sdv = G.amp[v.prop];
  // Laplacian history of the pairs D ahead.  The loads are unconditional: off
  // the own tile and past the last pair (descriptor 0) they are masked and read 0.
  constexpr int D = KP % 2 == 0 ? 2 : 1;
  float2 Lb[D];
#pragma unroll
  for (int j = 0; j < D; ++j)
    Lb[j] = buf_load2(hrs, (da[j] & kQOwn) ? (int)(da[j] & kOffMask) * 4 : kNoAccess, (hi - 1 - G.tlo) * HW * 4);
  const int cut_lo = v.ey0 > 0 ? 2 * EW2 : 0, cut_hi = v.ey0 + v.EH < H ? 2 * EW2 : 0;
  __syncthreads();

  for (int st = 0; st < n; ++st) {
    const int it = hi - 1 - st;
    const float wt = G.w[it];
    float2 msrc = make_float2(0.0f, 0.0f);  // the new state of the source's pair (its owner's lane)
    const float2 *C2 = reinterpret_cast<const float2 *>(Qc);
    float2 *N2 = reinterpret_cast<float2 *>(Qn);
    const int plo = cut_lo * (st + 1), phi = NPE - cut_hi * (st + 1);  // shrinking row band
#pragma unroll
    for (int j = 0; j < KP; ++j) {
      const int p = t + j * kThreads;
      const float2 L = Lb[j % D];
      {
        // wraps to the first pairs of the next step (the last step reloads its own: unused)
        const int jn = j + D < KP ? j + D : j + D - KP;
        const int itn = (j + D < KP || st + 1 == n) ? it : it - 1;
        Lb[j % D] = buf_load2(hrs, (da[jn] & kQOwn) ? (int)(da[jn] & kOffMask) * 4 : kNoAccess, (itn - G.tlo) * HW * 4);
      }
      if (p >= plo && p < phi) {
        const uint32_t d = da[j];
        const int gi = (int)(d & kOffMask);
        const float2 r0 = C2[p - 1], r1 = C2[p], r2 = C2[p + 1];
        const float2 um = C2[p - EW2], u2m = C2[p - EW], up = C2[p + EW2], u2p = C2[p + EW];
        const float lt0 = lapt_sum(r1.x, u2p.x, up.x, um.x, u2m.x, r2.x, r1.y, r0.y, r0.x);
        const float lt1 = lapt_sum(r1.y, u2p.y, up.y, um.y, u2m.y, r2.y, r2.x, r1.x, r0.y);
        const float2 e0 = lds_table2(db[j] & kTabOff), e1 = lds_table2((db[j] >> 16) & kTabOff);  // (2A, A B)
        float mu0 = e0.x * m1[j].x;
        float mu1 = e1.x * m1[j].y;
        mu0 = mu0 + lt0;
        mu1 = mu1 + lt1;
        mu0 = mu0 - e0.y * m2[j].x;
        mu1 = mu1 - e1.y * m2[j].y;
        if (d & kQRow) {
          const float2 rv = buf_load2(rrs, (gi - rec0) * 4, it * G.nx * 4);
          mu0 = mu0 + rv.x;
          mu1 = mu1 + rv.y;
        }
        // off the own tile L = 0: the imaging sum there is never stored, the illumination unchanged
        ga[j].x = ga[j].x + ((0.5f * e0.x) * mu0) * L.x;  // A = 0.5 (2A), exact
        ga[j].y = ga[j].y + ((0.5f * e1.x) * mu1) * L.y;
        if (d & kQOwnModel) {  // illumination of the own model cells only (the rest is never stored)
          const float2 iv = il[p];
          il[p] = make_float2(iv.x + L.x * L.x, iv.y + L.y * L.y);
        }
        if (j == src_j) msrc = make_float2(mu0, mu1);  // the source's pair slot (uniform): two selects
        m2[j] = m1[j];
        m1[j] = make_float2(mu0, mu1);
        const float2 qq = div12x2(ca[j].x * mu0, ca[j].y * mu1);
        N2[p] = make_float2(keep_if<17>(qq.x, d), keep_if<17>(qq.y, d));  // interior (kQInt)
      }
    }
    if (src_j >= 0 && t == src_t) sg = sg + ((src_hi ? msrc.y : msrc.x) * wt) * sdv;  // gpu.py:182, step order
    lds_barrier();  // the history and receiver accesses stay in flight
    float *tmp = Qc;
    Qc = Qn;
    Qn = tmp;
  }
  const uint32_t magic_w = 0xffffffffu / (uint32_t)W + 1u;
#pragma unroll
  for (int j = 0; j < KP; ++j) {
    const int p = t + j * kThreads;
    const uint32_t d = da[j];
    if (p < NPE && (d & kQOwn)) {
      const int gi = (int)(d & kOffMask);
      m1o[gi >> 1] = m1[j];
      m2o[gi >> 1] = m2[j];
      gac[gi >> 1] = ga[j];
      const int z = (int)__umulhi((uint32_t)gi, magic_w), x = gi - z * W;
      if (z >= pad && z < pad + G.nz && x >= pad && x < pad + G.nx) ilg[((z - pad) * G.nx + (x - pad)) >> 1] = il[p];
      if (j == src_j && t == src_t) G.sgrad[v.prop] = sg;
    }
  }
}

// ------------------------------------------------------------------ planning

struct Plan {
  SArgs args;
  int dir = 0;  // 0 forward, 1 adjoint
  int K = 0;           // cells per thread (cell kernels) or pairs per thread (paired kernels)
  bool paired = false;
  size_t lds = 0;
  int nlaunch = 0;
  std::vector<int> nactive;
};

struct TileChoice {
  int th, tw, T;
};

int ext_width(int W, const TileChoice &c) { return c.tw >= W ? W : std::min(W, c.tw + 4 * c.T); }

int ext_cells(int H, int W, const TileChoice &c) {
  const int eh = c.th >= H ? H : std::min(H, c.th + 4 * c.T);
  return eh * ext_width(W, c);
}

// LDS floats for the largest extended region ehw (widest ew) and sponge width pad.
// The table comes first (LDS address 0); an adjoint reserves 16 bytes per entry
// (it uses 8), which keeps the tile limits the autotuner ranks against.
int lds_floats(int ehw, int ew, int pad, int dir, int *plane0, int *plane1, int *il_off) {
  const int guard = 2 * ew + 2;
  const int tab = (dir ? 4 : 2) * (pad + 1) * (pad + 1);
  const int p0 = tab + guard;
  const int p1 = p0 + ehw + guard;
  const int il = p1 + ehw + guard;
  if (plane0) *plane0 = p0;
  if (plane1) *plane1 = p1;
  if (il_off) *il_off = il;
  return il + (dir ? ehw : 0);
}

// Cells loaded per launch over all tiles, divided by H*W (halo redundancy).
double redundancy(int H, int W, const TileChoice &c) {
  int64_t tot = 0;
  for (int y0 = 0; y0 < H; y0 += c.th) {
    const int y1 = std::min(H, y0 + c.th);
    const int eh = std::min(H, y1 + 2 * c.T) - std::max(0, y0 - 2 * c.T);
    for (int x0 = 0; x0 < W; x0 += c.tw) {
      const int x1 = std::min(W, x0 + c.tw);
      tot += (int64_t)eh * (std::min(W, x1 + 2 * c.T) - std::max(0, x0 - 2 * c.T));
    }
  }
  return (double)tot / ((double)H * W);
}

// Candidate tilings of an H x W grid that fit a workgroup, best modelled first.
// Model: per step, every loaded cell is updated once, and each launch pays
// about two extra updates' worth of time for state loads and stores.
std::vector<TileChoice> candidates(int H, int W, int pad, int dir) {
  const int cap = (dir ? kAdjMaxK : kFwdMaxK) * kThreads;
  std::vector<std::pair<double, TileChoice>> c;
  for (int T : {1, 2, 3, 4, 6, 8, 12}) {
    for (int nty = 1; nty <= 12; ++nty) {
      const int th = (H + nty - 1) / nty;
      if (nty > 1 && th < 8) break;
      for (int ntx = 1; ntx <= 12; ++ntx) {
        int tw = (W + ntx - 1) / ntx;
        if (W % 2 == 0 && tw < W) tw += tw & 1;  // even widths keep cell pairs inside tiles
        if (ntx > 1 && tw < 16) break;
        const TileChoice tc{th, tw, T};
        // one thread per receiver column of a tile: at most kThreads columns
        if (tw > kThreads ||
            ext_cells(H, W, tc) > cap ||
            lds_floats(ext_cells(H, W, tc), ext_width(W, tc), pad, dir, nullptr, nullptr, nullptr) > kMaxLdsFloats)
          continue;
        c.push_back({redundancy(H, W, tc) * (1.0 + 2.0 / T), tc});
      }
    }
  }
  std::stable_sort(c.begin(), c.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
  std::vector<TileChoice> out;
  for (auto &p : c) out.push_back(p.second);
  return out;
}

std::mutex g_mutex;
std::map<std::string, TileChoice> g_tuned;  // "dir:H:W:pad:propagations" -> tuned tiling

// Tuned tilings are kept per direction, grid and propagation count: a tiling
// that fills the GPU with a few hundred propagations is not the one that
// spreads 80 of them over all CUs.
std::string tune_key(int dir, int H, int W, int pad, int64_t nprop) {
  char buf[96];
  snprintf(buf, sizeof buf, "%d:%d:%d:%d:%lld", dir, H, W, pad, (long long)nprop);
  return buf;
}

// Polls of a band exchange before it counts as failed: about a second, against a
// few microseconds per exchange when every workgroup of a team is resident.  The
// environment variable SHOTSWEEP_BAND_SPIN overrides it (a test forces the failure
// path with 0: the launch still ends, the missed halos become NaN).
constexpr int kBandSpin = 1 << 20;
int band_spin() {
  const char *e = std::getenv("SHOTSWEEP_BAND_SPIN");
  if (!e || !*e) return kBandSpin;
  const long v = std::strtol(e, nullptr, 10);
  return v < 0 ? 0 : (v > kBandSpin ? kBandSpin : (int)v);
}

// A whole-record call from zero state with no tiling asked for (the full-history
// sweeps, which run alone on their stream).
bool whole_record(const ss_sweep_group &in, const Geom &g) {
  return in.tile_h == 0 && in.tile_w == 0 && in.depth == 0 && !in.resume && in.t0 == 0 &&
         (in.t1 == 0 || in.t1 == g.nt);
}

// Such a call on the 150 x 150 plane (70 x 70 models, pad 40, row 1) runs on the
// column-strip kernels, one workgroup per propagation for the whole record; the
// state stays in state buffer `phase` (no launch of the tiled sequence).
bool use_resident(const ss_sweep_group &in, Geom *g) {
  return check_geom(&in.geom, g) == SS_OK && resident_uses_c2a(*g) && whole_record(in, *g);
}

__global__ void zero_kernel(float *p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

// zero fill as a kernel (not a memset node, should a caller capture the call)
int zero_fill(float *p, int64_t n, hipStream_t st) {
  if (n <= 0) return SS_OK;
  const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  zero_kernel<<<dim3(blocks), dim3(256), 0, st>>>(p, n);
  return hip_status(hipGetLastError(), "sweep zero fill");
}

int resident_run(const ss_sweep_group &in, const Geom &g, int dir, hipStream_t st) {
  const int64_t nprop = (int64_t)g.nb * g.ns, PS = nprop * g.H * g.W;
  float *s0 = static_cast<float *>(in.state) + 2 * (in.phase & 1) * PS, *s1 = s0 + PS;
  const float *c2 = static_cast<const float *>(in.c2), *A = static_cast<const float *>(in.A),
              *B = static_cast<const float *>(in.B), *w = static_cast<const float *>(in.w),
              *amp = static_cast<const float *>(in.amp);
  int rc = zero_fill(s0, 2 * PS, st);
  if (rc != SS_OK) return rc;
  if (dir == 0)
    return forward_resident(g, 0, g.nt, c2, A, B, w, amp, s0, s1, static_cast<float *>(in.gathers),
                            static_cast<float *>(in.hist), nullptr, 1, st, true);
  float *gacc = static_cast<float *>(in.gacc), *illum = static_cast<float *>(in.illum),
        *sg = static_cast<float *>(in.src_grad);
  if (rc == SS_OK) rc = zero_fill(gacc, PS, st);
  if (rc == SS_OK) rc = zero_fill(illum, nprop * g.nz * g.nx, st);
  if (rc == SS_OK) rc = zero_fill(sg, nprop, st);
  if (rc != SS_OK) return rc;
  return adjoint_resident(g, 0, g.nt, c2, A, B, w, amp, static_cast<const float *>(in.res),
                          static_cast<const float *>(in.hist), s0, s1, gacc, illum, sg, st, nullptr);
}

// A stream being captured into a graph cannot take a cooperative launch with its
// co-residency guarantee: captured calls keep to the tiled kernels.
bool capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone;
}

// A group runs on the band kernels when a band layout exists for its grid, the
// caller asks for no particular tiling, and the call covers the whole record from
// zero state (the full-history sweeps, which run alone on their stream).
bool use_band(const ss_sweep_group &in, int dir, Geom *g) {
  if (check_geom(&in.geom, g) != SS_OK || !band_fits(*g, dir)) return false;
  // tags of the exchange words count at most 2^20 steps per team (at least 8 teams)
  const int64_t rounds = ((int64_t)g->nb * g->ns + 7) / 8;
  return in.tile_h == 0 && in.tile_w == 0 && in.depth == 0 && !in.resume && in.t0 == 0 &&
         (in.t1 == 0 || in.t1 == g->nt) && rounds * (g->nt + 2) < (1 << 20);
}

int fill_group(const ss_sweep_group &in, int dir, SGroup *o, char *err, size_t errn) {
  Geom g;
  int rc = check_geom(&in.geom, &g);
  if (rc != SS_OK) return rc;
  const int64_t HW = (int64_t)g.H * g.W;
  if (HW > kMaxPlane || (g.pad + 1) * (g.pad + 1) > kMaxTable || g.nt * HW > (int64_t)0x1fffffff ||
      (int64_t)g.nt * g.nx > (int64_t)0x1fffffff) {
    snprintf(err, errn, "sweep: grid %dx%d, pad %d, nt %d exceeds the tiled kernels' limits "
             "(H*W <= %d, pad <= 44, nt*H*W < 2^29)", g.H, g.W, g.pad, g.nt, kMaxPlane);
    return SS_EUNSUP;
  }
  if (in.t0 < 0 || in.t1 < 0 || (in.t1 > 0 && (in.t1 > g.nt || in.t0 >= in.t1)) || (in.t1 == 0 && in.t0 >= g.nt)) {
    snprintf(err, errn, "sweep: step range [%d, %d) outside [0, %d)", in.t0, in.t1, g.nt);
    return SS_EINVAL;
  }
  const void *need[] = {in.c2, in.A, in.B, in.w, in.amp, in.state, dir ? in.hist : in.state};
  for (const void *p : need)
    if (!p) {
      snprintf(err, errn, "sweep: c2, A, B, w, amp/sd, state (and for the adjoint hist) are required");
      return SS_EINVAL;
    }
  if (dir && (!in.res || !in.gacc || !in.illum || !in.src_grad)) {
    snprintf(err, errn, "sweep adjoint: res, gacc, illum and src_grad are required");
    return SS_EINVAL;
  }
  TileChoice tc{in.tile_h, in.tile_w, in.depth};
  if (tc.th <= 0 || tc.tw <= 0 || tc.T <= 0) {
    std::lock_guard<std::mutex> lock(g_mutex);
    auto it = g_tuned.find(tune_key(dir, g.H, g.W, g.pad, (int64_t)g.nb * g.ns));
    if (it != g_tuned.end()) {
      tc = it->second;
    } else {
      const auto c = candidates(g.H, g.W, g.pad, dir);
      if (c.empty()) {
        snprintf(err, errn, "sweep: no tiling of %dx%d fits a workgroup", g.H, g.W);
        return SS_EUNSUP;
      }
      tc = c.front();
    }
  }
  tc.th = std::min(tc.th, g.H);
  tc.tw = std::min(tc.tw, g.W);
  if (g.W % 2 == 0 && tc.tw < g.W) tc.tw += tc.tw & 1;  // even widths keep cell pairs inside tiles
  const int cap = (dir ? kAdjMaxK : kFwdMaxK) * kThreads;
  if (std::min(tc.tw, g.nx) > kThreads) {  // one thread per receiver column of a tile
    snprintf(err, errn, "sweep: tile width %d records more than %d receiver columns", tc.tw, kThreads);
    return SS_EUNSUP;
  }
  if (ext_cells(g.H, g.W, tc) > cap ||
      lds_floats(ext_cells(g.H, g.W, tc), ext_width(g.W, tc), g.pad, dir, nullptr, nullptr, nullptr) >
          kMaxLdsFloats) {
    snprintf(err, errn, "sweep: tile %dx%d at depth %d needs %d cells per workgroup (max %d)", tc.th, tc.tw, tc.T,
             ext_cells(g.H, g.W, tc), cap);
    return SS_EUNSUP;
  }
  const int ntx = (g.W + tc.tw - 1) / tc.tw, nty = (g.H + tc.th - 1) / tc.th;
  const int64_t nprop = (int64_t)g.nb * g.ns;
  if (nprop * ntx * nty > (1 << 30)) {
    snprintf(err, errn, "sweep: too many work items");
    return SS_EINVAL;
  }
  o->c2 = static_cast<const float *>(in.c2);
  o->A = static_cast<const float *>(in.A);
  o->B = static_cast<const float *>(in.B);
  o->w = static_cast<const float *>(in.w);
  o->amp = static_cast<const float *>(in.amp);
  o->res = static_cast<const float *>(in.res);
  o->st = static_cast<float *>(in.state);
  o->gathers = dir ? nullptr : static_cast<float *>(in.gathers);
  o->hist = static_cast<float *>(in.hist);
  o->gacc = static_cast<float *>(in.gacc);
  o->illum = static_cast<float *>(in.illum);
  o->sgrad = static_cast<float *>(in.src_grad);
  o->ns = g.ns;
  o->nz = g.nz;
  o->nx = g.nx;
  o->pad = g.pad;
  o->nt = g.nt;
  o->row = g.row;
  o->H = g.H;
  o->W = g.W;
  o->th = tc.th;
  o->tw = tc.tw;
  o->T = tc.T;
  o->ntx = ntx;
  o->ntiles = ntx * nty;
  o->tlo = in.t0;
  o->thi = in.t1 > 0 ? in.t1 : g.nt;
  o->resume = in.resume ? 1 : 0;
  o->phase = in.phase & 1;
  o->nlaunch = (o->thi - o->tlo + tc.T - 1) / tc.T;
  o->nitems = (int)(nprop * o->ntiles);
  for (int s = 0; s < SS_MAX_SHOTS; ++s) o->src[s] = g.src[s];
  return SS_OK;
}

int make_plan(int ngroups, const ss_sweep_group *groups, int dir, Plan *p) {
  char err[256];
  if (ngroups < 1 || ngroups > kMaxGroups || !groups) {
    snprintf(err, sizeof err, "sweep: need 1..%d groups", kMaxGroups);
    set_error(err);
    return SS_EINVAL;
  }
  std::vector<SGroup> gs(ngroups);
  for (int i = 0; i < ngroups; ++i) {
    err[0] = 0;
    const int rc = fill_group(groups[i], dir, &gs[i], err, sizeof err);
    if (rc != SS_OK) {
      if (err[0]) set_error(std::string(err) + " (group " + std::to_string(i) + ")");
      return rc;
    }
  }
  std::stable_sort(gs.begin(), gs.end(), [](const SGroup &x, const SGroup &y) { return x.nlaunch > y.nlaunch; });
  std::memset(&p->args, 0, sizeof p->args);
  int64_t item0 = 0;
  int ehw = 0, ew = 0, pad = 0;
  for (int i = 0; i < ngroups; ++i) {
    gs[i].item0 = (int)item0;
    item0 += gs[i].nitems;
    p->args.g[i] = gs[i];
    ehw = std::max(ehw, ext_cells(gs[i].H, gs[i].W, {gs[i].th, gs[i].tw, gs[i].T}));
    ew = std::max(ew, ext_width(gs[i].W, {gs[i].th, gs[i].tw, gs[i].T}));
    pad = std::max(pad, gs[i].pad);
  }
  if (item0 > (1 << 30)) {
    set_error("sweep: too many work items");
    return SS_EINVAL;
  }
  p->args.ngroups = ngroups;
  p->dir = dir;
  p->nlaunch = gs[0].nlaunch;
  p->nactive.assign(p->nlaunch, 0);
  for (int k = 0; k < p->nlaunch; ++k)
    for (int i = 0; i < ngroups; ++i)
      if (gs[i].nlaunch > k) p->nactive[k] = gs[i].item0 + gs[i].nitems;
  int plane0, plane1, il_off;
  const int nlds = lds_floats(ehw, ew, pad, dir, &plane0, &plane1, &il_off);
  if (nlds > kMaxLdsFloats) {
    set_error("sweep: the groups' tilings together need more than 160 KiB of LDS");
    return SS_EUNSUP;
  }
  p->args.plane0 = plane0;
  p->args.plane1 = plane1;
  p->args.tab_off = 0;
  p->args.il_off = il_off;
  p->lds = (size_t)nlds * sizeof(float);
  // cell pairs when every plane width, pad and model width is even and every tile width is even or whole
  p->paired = true;
  for (int i = 0; i < ngroups; ++i) {
    const SGroup &G = gs[i];
    if (G.W % 2 || G.pad % 2 || G.nx % 2 || (G.tw % 2 && G.tw < G.W)) p->paired = false;
  }
  if (p->paired) {
    const int kp = (ehw / 2 + kThreads - 1) / kThreads;
    p->K = kp <= 2 ? 2 : kp <= 4 ? 4 : kp <= 6 ? 6 : 8;
  } else {
    const int kneed = (ehw + kThreads - 1) / kThreads;
    p->K = kneed <= 4 ? 4 : kneed <= 8 ? 8 : kneed <= 12 ? 12 : 16;
  }
  return SS_OK;
}

using KernelFn = void (*)(SArgs);

KernelFn kernel_of(const Plan &p) {
  if (p.paired) {
    if (p.dir == 0) {
      switch (p.K) {
        case 2: return sweep_fwd_pair_kernel<2>;
        case 4: return sweep_fwd_pair_kernel<4>;
        case 6: return sweep_fwd_pair_kernel<6>;
        default: return sweep_fwd_pair_kernel<8>;
      }
    }
    return p.K <= 2 ? sweep_adj_pair_kernel<2> : p.K <= 4 ? sweep_adj_pair_kernel<4> : sweep_adj_pair_kernel<6>;
  }
  if (p.dir == 0) {
    switch (p.K) {
      case 4: return sweep_fwd_kernel<4>;
      case 8: return sweep_fwd_kernel<8>;
      case 12: return sweep_fwd_kernel<12>;
      default: return sweep_fwd_kernel<16>;
    }
  }
  return p.K <= 4 ? sweep_adj_kernel<4> : p.K <= 8 ? sweep_adj_kernel<8> : sweep_adj_kernel<12>;
}

int prepare_kernel(const Plan &p, KernelFn *fn) {
  *fn = kernel_of(p);
  int dev = 0, maxlds = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  if (e != hipSuccess) return hip_status(e, "sweep device query");
  if ((int64_t)p.lds > maxlds) {
    set_error("sweep: tiles need more LDS than the device has");
    return SS_EUNSUP;
  }
  // the sponge table is addressed at LDS address 0: no static LDS may precede the dynamic LDS
  hipFuncAttributes attr;
  e = hipFuncGetAttributes(&attr, (const void *)*fn);
  if (e != hipSuccess) return hip_status(e, "sweep attributes");
  if (attr.sharedSizeBytes != 0) {
    set_error("sweep: kernel has static LDS ahead of its sponge table");
    return SS_EUNSUP;
  }
  return hip_status(hipFuncSetAttribute((const void *)*fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds),
                    "sweep attribute");
}

void launch_range(const Plan &p, KernelFn fn, int k0, int k1, hipStream_t st) {
  SArgs a = p.args;
  for (int k = k0; k < k1; ++k) {
    a.k = k;
    a.nactive = p.nactive[k];
    const unsigned grid = (unsigned)((a.nactive + kXcds - 1) / kXcds * kXcds);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(kThreads), p.lds, st, a);
  }
}

// ------------------------------------------------------------------- graphs

// Executable graphs by batch layout, least recently used evicted beyond
// kMaxGraphs.  A checkpointed gradient replays two graphs per segment (forward
// recompute and adjoint, each with its own step range), so the cache holds the
// 40 layouts of 1000 steps in segments of 50 several times over.
//
// Every launch of a cached graph records the graph's own event on the
// launching stream.  An evicted graph may still be queued there, so it is not
// destroyed at once: it moves to a retired list with its event and is
// destroyed by a later call once that event has completed (hipEventQuery), so
// eviction never waits for the device or for other streams.  Only when more
// than kMaxRetired graphs are pending does a call wait, and then only for the
// oldest retired graph's own event.
constexpr size_t kMaxGraphs = 256;
constexpr size_t kMaxRetired = 64;
struct CachedGraph {
  hipGraphExec_t exec;
  hipEvent_t done;     // recorded after each launch of exec
  uint64_t last_use;
};
struct GraphCache {
  std::map<std::string, CachedGraph> exec;
  std::map<int, hipStream_t> capture_stream;
  std::deque<CachedGraph> retired;
  uint64_t clock = 0;
};
GraphCache g_graphs;

void destroy_graph(const CachedGraph &c) {
  (void)hipGraphExecDestroy(c.exec);
  (void)hipEventDestroy(c.done);
}

// Destroy the retired graphs whose last launch has completed (caller holds g_mutex).
int reap_retired(bool wait_oldest) {
  if (wait_oldest && !g_graphs.retired.empty()) {
    hipError_t e = hipEventSynchronize(g_graphs.retired.front().done);
    if (e != hipSuccess) return hip_status(e, "sweep graph retire");
  }
  for (auto it = g_graphs.retired.begin(); it != g_graphs.retired.end();) {
    hipError_t q = hipEventQuery(it->done);
    if (q == hipSuccess) {
      destroy_graph(*it);
      it = g_graphs.retired.erase(it);
    } else if (q == hipErrorNotReady) {
      ++it;
    } else {
      return hip_status(q, "sweep graph retire query");
    }
  }
  return SS_OK;
}

std::string graph_key(const Plan &p, int dev) {
  std::string key(reinterpret_cast<const char *>(&p.args), sizeof p.args);
  key.append(reinterpret_cast<const char *>(&p.dir), sizeof p.dir);
  key.append(reinterpret_cast<const char *>(&dev), sizeof dev);
  return key;
}

int run_plan(const Plan &p, int use_graph, hipStream_t st) {
  KernelFn fn;
  int rc = prepare_kernel(p, &fn);
  if (rc != SS_OK) return rc;
  if (!use_graph) {
    launch_range(p, fn, 0, p.nlaunch, st);
    return hip_status(hipGetLastError(), "sweep launch");
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_status(e, "sweep device");
  std::lock_guard<std::mutex> lock(g_mutex);
  const std::string key = graph_key(p, dev);
  auto it = g_graphs.exec.find(key);
  rc = SS_OK;
  hipGraphExec_t exec = nullptr;
  if (it != g_graphs.exec.end()) {
    exec = it->second.exec;
    it->second.last_use = ++g_graphs.clock;
  } else {
    hipStream_t cs;
    auto sit = g_graphs.capture_stream.find(dev);
    if (sit == g_graphs.capture_stream.end()) {
      e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
      if (e != hipSuccess) return hip_status(e, "sweep capture stream");
      g_graphs.capture_stream[dev] = cs;
    } else {
      cs = sit->second;
    }
    if (g_graphs.exec.size() >= kMaxGraphs) {
      auto old = g_graphs.exec.begin();
      for (auto i = g_graphs.exec.begin(); i != g_graphs.exec.end(); ++i)
        if (i->second.last_use < old->second.last_use) old = i;
      g_graphs.retired.push_back(old->second);  // destroyed once its last launch has completed
      g_graphs.exec.erase(old);
    }
    rc = reap_retired(g_graphs.retired.size() > kMaxRetired);
    if (rc != SS_OK) return rc;
    hipGraph_t graph = nullptr;
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_status(e, "sweep begin capture");
    launch_range(p, fn, 0, p.nlaunch, cs);
    hipError_t le = hipGetLastError();
    e = hipStreamEndCapture(cs, &graph);
    if (le != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return hip_status(le, "sweep capture launch");
    }
    if (e != hipSuccess) return hip_status(e, "sweep end capture");
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_status(e, "sweep graph instantiate");
    hipEvent_t done;
    e = hipEventCreateWithFlags(&done, hipEventDisableTiming);
    if (e != hipSuccess) {
      (void)hipGraphExecDestroy(exec);
      return hip_status(e, "sweep graph event");
    }
    it = g_graphs.exec.emplace(key, CachedGraph{exec, done, ++g_graphs.clock}).first;
  }
  e = hipGraphLaunch(exec, st);
  if (e != hipSuccess) return hip_status(e, "sweep graph launch");
  return hip_status(hipEventRecord(it->second.done, st), "sweep graph event record");
}

// Kernel variant and LDS floors a group's items get in a joint launch: the
// joint plan runs every item with the largest slot count and LDS size of all
// its groups.
struct Floors {
  bool set = false, paired = false;
  int K = 0;
  size_t lds = 0;
  int ehw = 0, ew = 0, pad = 0;  // largest extended region, widest one and widest sponge so far
};

// The joint launch sizes its LDS from the largest extended region, the widest one
// and the widest sponge of all its groups, which may come from different groups.
bool joint_lds_fits(int H, int W, int pad, int dir, const TileChoice &c, const Floors &fl) {
  return lds_floats(std::max(ext_cells(H, W, c), fl.ehw), std::max(ext_width(W, c), fl.ew), std::max(pad, fl.pad),
                    dir, nullptr, nullptr, nullptr) <= kMaxLdsFloats;
}

// Time the first (up to) six launches of one group alone under a tiling, the
// faster of two runs after one untimed launch, with the kernel variant and LDS
// raised to the floors (as in the joint launch); ms per cell-step, negative on
// error.
double time_choice(const ss_sweep_group &grp, int dir, const TileChoice &tc, hipStream_t st, hipEvent_t e0,
                   hipEvent_t e1, const Floors &fl) {
  ss_sweep_group g1 = grp;
  g1.tile_h = tc.th;
  g1.tile_w = tc.tw;
  g1.depth = tc.T;
  Plan p;
  if (make_plan(1, &g1, dir, &p) != SS_OK) return -1.0;
  if (fl.set && fl.paired == p.paired) {
    p.K = std::max(p.K, fl.K);
    p.lds = std::max(p.lds, fl.lds);
  }
  KernelFn fn;
  if (prepare_kernel(p, &fn) != SS_OK) return -1.0;
  const SGroup &G = p.args.g[0];
  const int m = std::min(p.nlaunch, 6);
  launch_range(p, fn, 0, 1, st);
  float best = -1.0f;
  for (int rep = 0; rep < 2; ++rep) {  // the faster of two timings of the first m launches
    if (hipEventRecord(e0, st) != hipSuccess) return -1.0;
    launch_range(p, fn, 0, m, st);
    if (hipEventRecord(e1, st) != hipSuccess) return -1.0;
    if (hipEventSynchronize(e1) != hipSuccess) return -1.0;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.0;
    if (best < 0.0f || ms < best) best = ms;
  }
  const double steps = (double)std::min(G.thi - G.tlo, m * G.T);
  return best / (steps * (double)G.nitems / G.ntiles * G.H * G.W);
}

}  // namespace

int sweep_run(int ngroups, const ss_sweep_group *groups, int dir, int use_graph, hipStream_t st) {
  if (ngroups < 1 || ngroups > kMaxGroups || !groups) {
    set_error("sweep: need 1..4 groups");
    return SS_EINVAL;
  }
  // band groups run whole in one cooperative launch each, the others in the tiled launches
  std::vector<ss_sweep_group> tiled;
  const bool cap = capturing(st);
  for (int i = 0; i < ngroups; ++i) {
    Geom g;
    if (!cap && use_band(groups[i], dir, &g)) {
      const int rc = band_run(groups[i], g, dir, band_spin(), st);
      if (rc != SS_OK) return rc;
    } else if (use_resident(groups[i], &g)) {
      const int rc = resident_run(groups[i], g, dir, st);
      if (rc != SS_OK) return rc;
    } else {
      tiled.push_back(groups[i]);
    }
  }
  if (tiled.empty()) return SS_OK;
  Plan p;
  int rc = make_plan((int)tiled.size(), tiled.data(), dir, &p);
  if (rc != SS_OK) return rc;
  return run_plan(p, use_graph, st);
}

bool sweep_tuned(int dir, int H, int W, int pad, int64_t nprop) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return g_tuned.count(tune_key(dir, H, W, pad, nprop)) != 0;
}

int sweep_launches(const ss_sweep_group &group, int dir, hipStream_t st) {
  Geom g;
  if (!capturing(st) && use_band(group, dir, &g)) return 1;
  if (use_resident(group, &g)) return 0;  // state read and written in place: buffer `phase`
  SGroup o;
  char err[256] = {0};
  const int rc = fill_group(group, dir, &o, err, sizeof err);
  if (rc != SS_OK) {
    if (err[0]) set_error(err);
    return rc;
  }
  return o.nlaunch;
}

bool sweep_fits(const Geom &g) {
  const int64_t HW = (int64_t)g.H * g.W;
  return HW <= kMaxPlane && (g.pad + 1) * (g.pad + 1) <= kMaxTable && g.nt * HW <= (int64_t)0x1fffffff &&
         (int64_t)g.nt * g.nx <= (int64_t)0x1fffffff;
}

int sweep_autotune(int ngroups, ss_sweep_group *groups, int dir, int max_trials, hipStream_t st) {
  if (ngroups < 1 || ngroups > kMaxGroups || !groups) {
    set_error("sweep autotune: need 1..4 groups");
    return SS_EINVAL;
  }
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return hip_status(e, "autotune event");
  e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(e0);
    return hip_status(e, "autotune event");
  }
  int rc = SS_OK;
  // Largest groups first (cell-steps): in the joint launch their tilings set the kernel
  // variant and LDS size that every smaller group's items run with, so those are timed
  // under the same floors.
  std::vector<int> order(ngroups);
  std::vector<double> work(ngroups, 0.0);
  for (int i = 0; i < ngroups; ++i) {
    order[i] = i;
    const ss_geom &q = groups[i].geom;
    work[i] = (double)q.nb * q.ns * q.nt * (q.nz + 2.0 * q.pad) * (q.nx + 2.0 * q.pad);
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
  Floors fl;
  for (int oi = 0; oi < ngroups && rc == SS_OK; ++oi) {
    const int i = order[oi];
    Geom g;
    if (use_band(groups[i], dir, &g) || use_resident(groups[i], &g)) continue;  // no tiling: whole planes held
    rc = check_geom(&groups[i].geom, &g);
    if (rc != SS_OK) break;
    const auto cand = candidates(g.H, g.W, g.pad, dir);
    double best = -1.0;
    TileChoice pick{0, 0, 0};
    const int ntry = max_trials > 0 ? max_trials : 12;
    for (int c = 0, tried = 0; c < (int)cand.size() && tried < ntry; ++c) {
      if (!joint_lds_fits(g.H, g.W, g.pad, dir, cand[c], fl)) continue;  // would not fit beside the larger groups
      ++tried;
      const double tpc = time_choice(groups[i], dir, cand[c], st, e0, e1, fl);
      if (tpc > 0.0 && (best < 0.0 || tpc < best)) {
        best = tpc;
        pick = cand[c];
      }
    }
    if (best < 0.0) {
      set_error("sweep autotune: no candidate tiling ran");
      rc = SS_EUNSUP;
      break;
    }
    groups[i].tile_h = pick.th;
    groups[i].tile_w = pick.tw;
    groups[i].depth = pick.T;
    Plan pp;
    if (make_plan(1, &groups[i], dir, &pp) == SS_OK) {
      if (!fl.set) {
        fl.set = true;
        fl.paired = pp.paired;
      }
      if (fl.paired == pp.paired) {
        fl.K = std::max(fl.K, pp.K);
        fl.lds = std::max(fl.lds, pp.lds);
      }
    }
    fl.ehw = std::max(fl.ehw, ext_cells(g.H, g.W, pick));
    fl.ew = std::max(fl.ew, ext_width(g.W, pick));
    fl.pad = std::max(fl.pad, g.pad);
    std::lock_guard<std::mutex> lock(g_mutex);
    g_tuned[tune_key(dir, g.H, g.W, g.pad, (int64_t)g.nb * g.ns)] = pick;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc == SS_OK) rc = hip_status(hipGetLastError(), "autotune");
  return rc;
}

}  // namespace ss
