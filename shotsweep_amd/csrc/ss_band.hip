// Band kernels: a plane too large for one workgroup's registers and LDS, held
// resident for a whole step range by a team of workgroups, one horizontal band
// of rows each, that trade their two boundary rows every step.
//
// The tiled sweep (ss_sweep.hip) reloads every tile's state from HBM every T
// steps and recomputes a halo of 2T cells around it; here nothing is reloaded
// and nothing recomputed.  A band is laid out like the column-strip kernels
// (ss_resident.hip): thread t < NACT owns column pair c = t mod NC in the R rows
// of strip t / NC and keeps p(t-1), p(t) and c2 of those rows in registers (two
// register sets that swap roles every step), LDS holds p(t) of the band plus two
// halo rows above and below it.  One extra wave per workgroup, the exchange
// wave, copies the band's first and last two rows of p(t) to the team's
// exchange buffer and the neighbours' rows into the halo rows while the compute
// waves update the rows whose taps lie inside the band (rows 2 .. R-3 of every
// strip); after a barrier they update the strip's first and last two rows.  The
// exchange uses agent-scope atomic loads and stores (coherent across the XCDs'
// L2 caches) and a per-band step counter: the writer waits for its stores to
// complete before it publishes the counter, the reader polls the counter before
// it loads.  A team's workgroups are dealt to one XCD, and the launch is
// cooperative, so every workgroup of a team is resident while the others wait
// for it.  Each wait is bounded: past the bound the band's halo becomes NaN and
// an error counter is raised, so a failure shows in the outputs instead of
// hanging the device.
//
// Per cell the arithmetic is that of the stepwise and strip kernels
// (ss_common.h): forward gpu.py:111-115 and 126-135.
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "ss_device.h"
#include "ss_kernels.h"

namespace ss {

namespace {

constexpr int kBandThreads = 768;

template <int NZ, int NX, int PAD, int ROW, int NB, int RR>
struct BandShape {
  static constexpr int nz = NZ, nx = NX, pad = PAD, row = ROW, R = RR;
  static constexpr int H = NZ + 2 * PAD, W = NX + 2 * PAD, HW = H * W;
  static constexpr int BH = H / NB;                 // rows per band
  static constexpr int NC = W / 2, NS = BH / R;     // pairs per row, strips per band
  static constexpr int NACT = NC * NS;              // compute lanes
  static constexpr int NCW = (NACT + 63) / 64;      // compute waves; wave NCW exchanges
  static constexpr int RZ = PAD + ROW;              // receiver and source row
  static constexpr int RB = RZ / BH, RS = (RZ % BH) / R, RK = (RZ % BH) % R;
  static constexpr int XSIDE = 2 * W;                 // exchange words of one side: two rows
  static constexpr int XTEAM = NB * 2 * 2 * XSIDE;    // per team: bands x sides x parities
  static_assert(H % NB == 0 && BH % R == 0, "bands and strips must tile the plane");
  static_assert(W % 2 == 0 && NX % 2 == 0 && PAD % 2 == 0, "pairs need even widths");
  static_assert((NCW + 1) * 64 <= kBandThreads && R >= 6, "strips of at least six rows, one exchange wave");
  static_assert((PAD + 1) * (PAD + 1) * 16 <= (int)kPTabMask + 16, "pair-table offsets must fit the descriptor");
  // the source is added after the band's edge rows are published: keep it off them
  static_assert(RZ % BH >= 2 && RZ % BH < BH - 2, "source row on a band's edge rows");
};

// 200 x 200 models, pad 40, receivers and sources on row 1: 280 x 280 planes in
// four bands of 70 rows, strips of 14 rows (700 compute lanes, 11 waves).
using Band280 = BandShape<200, 200, 40, 1, 4, 14>;
// 100 x 140 models, pad 40, row 1: 180 x 220 planes in two bands of 90 rows, strips
// of 15 rows (660 compute lanes, 11 waves).
using Band180 = BandShape<100, 140, 40, 1, 2, 15>;

template <class S>
constexpr bool band_is(const Geom &g) {
  return g.nz == S::nz && g.nx == S::nx && g.pad == S::pad && g.row == S::row;
}

struct BandArgs {
  const float *c2, *A, *B, *w, *amp;
  const float *sin;  // state in: p0 planes at sin, p1 planes at sin + PS
  float *sout;       // state out, same layout
  float *hist, *gathers;
  int64_t PS;        // floats of one state plane set (nprop * H * W)
  int nprop, ns, nt, tlo, thi, resume;
  int src[SS_MAX_SHOTS];
  uint64_t *xbuf;    // per team: exchange words
  int *err;          // waits past the bound
  int epoch;         // launch count of this buffer set (mod 2^11), the tags' high bits
  int spin_limit;
};

// Exchange words: one float and the step count it belongs to in one 64-bit word,
// written and read by agent-scope atomic accesses (single-copy atomic and coherent
// across the XCDs' L2 caches).  A reader polls the words themselves until every
// tag matches the step it needs, so no separate counter, and no wait for the
// stores to complete on the writer's side, sits on the path between two bands.
__device__ __forceinline__ void x_put(uint64_t *p, float v, int tag) {
  const uint64_t u = ((uint64_t)(uint32_t)tag << 32) | __builtin_bit_cast(uint32_t, v);
  __hip_atomic_store(p, u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t x_get(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exchange slots of a team: side 0 of a band holds its first two rows (read by the
// band above), side 1 its last two (read by the band below), two parities (the
// step count's lowest bit), 2 W words each.
template <class S>
__device__ __forceinline__ uint64_t *x_side(uint64_t *xb, int band, int side, int tag) {
  return xb + ((band * 2 + side) * 2 + (tag & 1)) * S::XSIDE;
}

// The exchange wave: the neighbours' rows of p(t) (tag) into the halo rows, LDS
// rows 0, 1 (from the band above) and BH + 2, BH + 3 (from the band below).  Every
// lane loads its words, keeps those whose tag matches and reloads the rest; past
// spin_limit rounds the missing words become NaN and the error counter rises.
template <class S>
__device__ void x_receive(uint64_t *xb, int band, int tag, uint32_t band_addr, int lane, int spin_limit, int *err,
                          bool &failed) {
  constexpr int NB = S::H / S::BH, NW = S::XSIDE, NWL = (NW + 63) / 64;
  constexpr uint32_t kRow = 4u * S::W;
#pragma unroll
  for (int sd = 0; sd < 2; ++sd) {
    if (sd == 0 ? band == 0 : band == NB - 1) continue;
    const uint64_t *src = x_side<S>(xb, sd == 0 ? band - 1 : band + 1, sd == 0 ? 1 : 0, tag);
    const uint32_t dst = band_addr + (sd == 0 ? 0u : (uint32_t)(S::BH + 2) * kRow);
    uint32_t pending = 0;
#pragma unroll
    for (int k = 0; k < NWL; ++k)
      if (lane + 64 * k < NW) pending |= 1u << k;
    for (int spins = 0;; ++spins) {
      uint64_t v[NWL];
#pragma unroll
      for (int k = 0; k < NWL; ++k) v[k] = (pending >> k) & 1u ? x_get(src + lane + 64 * k) : 0;
#pragma unroll
      for (int k = 0; k < NWL; ++k) {
        if (((pending >> k) & 1u) && (int)(uint32_t)(v[k] >> 32) == tag) {
          lds_st1(dst + 4u * (uint32_t)(lane + 64 * k), __builtin_bit_cast(float, (uint32_t)v[k]));
          pending &= ~(1u << k);
        }
      }
      if (!__builtin_amdgcn_ballot_w64(pending != 0)) break;
      if (failed || spins >= spin_limit) {
        if (!failed && lane == 0) atomicAdd(err, 1);
        failed = true;
#pragma unroll
        for (int k = 0; k < NWL; ++k)
          if ((pending >> k) & 1u) lds_st1(dst + 4u * (uint32_t)(lane + 64 * k), __builtin_nanf(""));
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
}

// One row k of a strip: p(t+1) into PO[k] from p(t) (PC, the vertical taps zm2 ..
// zp2 and the left / right neighbour pairs lf, rt) and p(t-1) (PO[k]); its
// Laplacian to the history (zero on the two-cell rim, gpu.py:35).
template <int R, int W>
__device__ __forceinline__ void fwd_row(float2 (&PO)[R], const float2 (&PC)[R], const float2 (&cc)[R],
                                        const bool (&inner)[R], int k, float2 zm2, float2 zm1, float2 zp1,
                                        float2 zp2, float2 lf, float2 rt, float4 ab, __amdgpu_buffer_rsrc_t hrs,
                                        int hv, int hoff) {
  const float2 zc = PC[k];
  const float s0 = lap_sum(zm2.x, zm1.x, zc.x, zp1.x, zp2.x, lf.x, lf.y, zc.y, rt.x);
  const float s1 = lap_sum(zm2.y, zm1.y, zc.y, zp1.y, zp2.y, lf.y, zc.x, rt.x, rt.y);
  const float2 Lq = div12x2(s0, s1);
  const float L0 = inner[k] ? Lq.x : 0.0f, L1 = inner[k] ? Lq.y : 0.0f;
  PO[k].x = leapfrog(zc.x, PO[k].x, ab.y, cc[k].x, L0, ab.x);
  PO[k].y = leapfrog(zc.y, PO[k].y, ab.w, cc[k].y, L1, ab.z);
  buf_store2(make_float2(L0, L1), hrs, hv, hoff + k * 4 * W);
  asm volatile("" : "+v"(PO[k].x), "+v"(PO[k].y));  // finish this row before the next starts
}

// LDS layout: [pair table | guard | band region: (BH + 4) rows x W | guard]; band
// region rows 0, 1 and BH + 2, BH + 3 are the halo rows (zero at the plane's top
// and bottom, where they lie outside the plane).
template <class S>
constexpr size_t band_lds_bytes() {
  return (size_t)(pair_table_floats(S::pad) + 2 * pair_guard(S::W) + (S::BH + 4) * S::W) * sizeof(float);
}

template <class S>
__global__ __launch_bounds__(kBandThreads) void fwd_band_kernel(BandArgs a) {
  extern __shared__ float lds[];
  constexpr int H = S::H, W = S::W, HW = S::HW, BH = S::BH, R = S::R, NC = S::NC, NACT = S::NACT;
  constexpr int NB = H / BH, nz = S::nz, nx = S::nx, pad = S::pad;
  constexpr int tabf = pair_table_floats(pad), guard = pair_guard(W);
  constexpr uint32_t kRow = 4u * W;
  constexpr uint32_t kBand = (uint32_t)(tabf + guard) * 4;  // LDS byte address of band region row 0
  float4 *tab = reinterpret_cast<float4 *>(lds);
  float *P = lds + tabf + guard;
  float2 *P2 = reinterpret_cast<float2 *>(P);
  const int t = threadIdx.x;
  const bool xw = (t >> 6) == S::NCW;  // the exchange wave
  const int per_xcd = (int)gridDim.x / kXcds;
  const int j = (int)blockIdx.x / kXcds;
  const int band = j % NB, team = (int)(blockIdx.x % kXcds) * (per_xcd / NB) + j / NB;
  const int nteams = (int)gridDim.x / NB;
  const bool act = t < NACT;
  const int tt = act ? t : NACT - 1;
  const int strip = tt / NC, c = tt - strip * NC, r0 = strip * R, z0 = band * BH + r0;

  for (int k = t; k < guard; k += kBandThreads) {
    P[k - guard] = 0.0f;
    P[(BH + 4) * W + k] = 0.0f;
  }
  for (int k = t; k < 2 * W; k += kBandThreads) {
    P[k] = 0.0f;
    P[(BH + 2) * W + k] = 0.0f;
  }
  fill_pair_table(tab, a.A, a.B, W, pad, 0);
  uint32_t tb[R];
  bool inner[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const uint32_t d = pair_desc((z0 + k) * NC + c, H, W, nz, nx, pad);
    tb[k] = d & kPTabMask;
    inner[k] = (d & kPInterior) != 0;
  }
  const int hv = act ? (z0 * W + 2 * c) * 4 : 0x40000000;  // idle lanes store past every resource
  const int gv = act ? (2 * c - pad) * 4 : 0x40000000;
  const uint32_t aa0 = kBand + (uint32_t)((r0 * W + 2 * c) * 4);  // band region row r0: the strip's row -2
  const int nslots = a.thi - a.tlo;
  uint64_t *xb = a.xbuf + (int64_t)team * S::XTEAM;
  bool xfailed = false;  // the exchange wave: a wait ran past the bound
  // edge strips publish the band's first two rows (strip 0) and last two (strip NS - 1)
  const bool pub_up = act && strip == 0 && band > 0, pub_dn = act && strip == S::NS - 1 && band < NB - 1;
  auto publish = [&](const float2(&Pv)[R], int tag) __attribute__((always_inline)) {
    if (pub_up) {
      uint64_t *d = x_side<S>(xb, band, 0, tag) + 2 * c;
      x_put(d, Pv[0].x, tag);
      x_put(d + 1, Pv[0].y, tag);
      x_put(d + W, Pv[1].x, tag);
      x_put(d + W + 1, Pv[1].y, tag);
    }
    if (pub_dn) {
      uint64_t *d = x_side<S>(xb, band, 1, tag) + 2 * c;
      x_put(d, Pv[R - 2].x, tag);
      x_put(d + 1, Pv[R - 2].y, tag);
      x_put(d + W, Pv[R - 1].x, tag);
      x_put(d + W + 1, Pv[R - 1].y, tag);
    }
  };

  for (int prop = team; prop < a.nprop; prop += nteams) {  // the same planes for every band of a team
    // tag of slot 0: launch count, then the step count.  A plane's steps carry tags
    // gbase .. gbase + nslots; the next plane starts two higher, so its first words
    // (same parity) replace the last plane's final rows, which no step reads, and
    // never rows a neighbour may still be reading.
    const int gbase = (a.epoch << 20) + (prop / nteams) * (nslots + 2);
    const int b = prop / a.ns, s = prop - b * a.ns;
    float2 Pa[R], Pb[R], cc[R];
    {
      const float2 *p0p = reinterpret_cast<const float2 *>(a.sin + (int64_t)prop * HW);
      const float2 *p1p = reinterpret_cast<const float2 *>(a.sin + a.PS + (int64_t)prop * HW);
      const float2 *c2b = reinterpret_cast<const float2 *>(a.c2 + (int64_t)b * HW);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int q = (z0 + k) * NC + c;
        Pa[k] = a.resume ? p0p[q] : make_float2(0.0f, 0.0f);
        Pb[k] = a.resume ? p1p[q] : make_float2(0.0f, 0.0f);
        cc[k] = c2b[q];
        if (act) P2[(r0 + 2 + k) * NC + c] = Pb[k];
      }
    }
    publish(Pb, gbase);  // p(t0) of the edge rows
    const int srcx = pad + a.src[s];
    const bool src_owner = act && band == S::RB && strip == S::RS && c == (srcx >> 1);
    const bool src_hi = (srcx & 1) != 0;
    const float ampv = a.amp[prop];
    const bool rec = act && a.gathers && band == S::RB && strip == S::RS && c >= pad / 2 && c < (pad + nx) / 2;
    const __amdgpu_buffer_rsrc_t hrs = make_rsrc(a.hist ? a.hist + (int64_t)prop * nslots * HW : a.c2,
                                                 a.hist ? (int64_t)nslots * HW : 0);
    const __amdgpu_buffer_rsrc_t grs = make_rsrc(a.gathers ? a.gathers + (int64_t)prop * a.nt * nx : a.c2,
                                                 a.gathers ? (int64_t)a.nt * nx : 0);
    __syncthreads();

    // one step: PO holds p(t-1) on entry and p(t+1) on exit, PC p(t)
    auto step = [&](float2(&PO)[R], const float2(&PC)[R], int it) __attribute__((always_inline)) {
      const int g = gbase + (it - a.tlo), hoff = (it - a.tlo) * HW * 4;
      if (xw) {
        x_receive<S>(xb, band, g, kBand, t & 63, a.spin_limit, a.err, xfailed);
      } else {  // rows 2 .. R-3: taps inside the strip, neighbours from LDS one row ahead
        uint32_t aa = aa0;
        asm volatile("" : "+v"(aa));
        float2 lf = lds_ld2(aa + 4 * kRow - 8), rt = lds_ld2(aa + 4 * kRow + 8);
        float4 ab = lds_ld4(tb[2]);
#pragma unroll
        for (int k = 2; k <= R - 3; ++k) {
          asm volatile("" : "+v"(aa));
          float2 lfn = make_float2(0.0f, 0.0f), rtn = lfn;
          float4 abn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          if (k + 1 <= R - 3) {
            lfn = lds_ld2(aa + (uint32_t)(k + 3) * kRow - 8);
            rtn = lds_ld2(aa + (uint32_t)(k + 3) * kRow + 8);
            abn = lds_ld4(tb[k + 1]);
          }
          fwd_row<R, W>(PO, PC, cc, inner, k, PC[k - 2], PC[k - 1], PC[k + 1], PC[k + 2], lf, rt, ab, hrs, hv,
                        hoff);
          lf = lfn;
          rt = rtn;
          ab = abn;
        }
      }
      lds_barrier();  // halo rows of p(t) in place
      if (!xw) {      // rows 0, 1, R-2, R-1: taps from the neighbouring strips or the halo rows
        uint32_t aa = aa0;
        asm volatile("" : "+v"(aa));
        const float2 up2 = lds_ld2(aa), up1 = lds_ld2(aa + kRow);
        const float2 dn1 = lds_ld2(aa + (uint32_t)(R + 2) * kRow), dn2 = lds_ld2(aa + (uint32_t)(R + 3) * kRow);
        const float2 l0 = lds_ld2(aa + 2 * kRow - 8), r0v = lds_ld2(aa + 2 * kRow + 8);
        const float2 l1 = lds_ld2(aa + 3 * kRow - 8), r1v = lds_ld2(aa + 3 * kRow + 8);
        const float2 l2 = lds_ld2(aa + (uint32_t)R * kRow - 8), r2v = lds_ld2(aa + (uint32_t)R * kRow + 8);
        const float2 l3 = lds_ld2(aa + (uint32_t)(R + 1) * kRow - 8), r3v = lds_ld2(aa + (uint32_t)(R + 1) * kRow + 8);
        const float4 e0 = lds_ld4(tb[0]), e1 = lds_ld4(tb[1]), e2 = lds_ld4(tb[R - 2]), e3 = lds_ld4(tb[R - 1]);
        fwd_row<R, W>(PO, PC, cc, inner, 0, up2, up1, PC[1], PC[2], l0, r0v, e0, hrs, hv, hoff);
        fwd_row<R, W>(PO, PC, cc, inner, 1, up1, PC[0], PC[2], PC[3], l1, r1v, e1, hrs, hv, hoff);
        fwd_row<R, W>(PO, PC, cc, inner, R - 2, PC[R - 4], PC[R - 3], PC[R - 1], dn1, l2, r2v, e2, hrs, hv, hoff);
        fwd_row<R, W>(PO, PC, cc, inner, R - 1, PC[R - 3], PC[R - 2], dn1, dn2, l3, r3v, e3, hrs, hv, hoff);
        publish(PO, g + 1);  // p(t+1) of the edge rows, final (the source row is not among them)
      }
      lds_barrier();  // every wave has read p(t)
      if (!xw) {
        if (src_owner) {  // gpu.py:114, added after damping
          const float sv = a.w[it] * ampv;
          if (src_hi)
            PO[S::RK].y = PO[S::RK].y + sv;
          else
            PO[S::RK].x = PO[S::RK].x + sv;
        }
        if (rec) buf_store2(PO[S::RK], grs, gv, it * nx * 4);  // p(it+1) after injection (gpu.py:134)
        if (act) {
#pragma unroll
          for (int k = 0; k < R; ++k) lds_st2(aa0 + (uint32_t)(k + 2) * kRow, PO[k]);
        }
      }
      lds_barrier();
    };
    int it = a.tlo;
    for (; it + 1 < a.thi; it += 2) {
      step(Pa, Pb, it);
      step(Pb, Pa, it + 1);
    }
    const bool odd = it < a.thi;
    if (odd) step(Pa, Pb, it);
    if (act && !xw) {
      float2 *p0o = reinterpret_cast<float2 *>(a.sout + (int64_t)prop * HW);
      float2 *p1o = reinterpret_cast<float2 *>(a.sout + a.PS + (int64_t)prop * HW);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int q = (z0 + k) * NC + c;
        p0o[q] = odd ? Pb[k] : Pa[k];
        p1o[q] = odd ? Pa[k] : Pb[k];
      }
    }
    __syncthreads();  // LDS is reloaded for the next plane
  }
}

// Exchange words and the error counter, one set per device and
// stream: launches on one stream run in order, launches on different streams
// never share a set.
struct BandBuffers {
  int *err = nullptr;
  uint64_t *xbuf = nullptr;
  int64_t words = 0;
  int epoch = 0;
};
std::mutex g_band_mutex;
std::map<std::pair<int, hipStream_t>, BandBuffers> g_band_buffers;

int band_buffers(int dev, hipStream_t st, int64_t words, BandBuffers *out) {
  std::lock_guard<std::mutex> lock(g_band_mutex);
  BandBuffers &b = g_band_buffers[{dev, st}];
  if (b.words < words) {
    // a stream may still use the old buffers: keep them (a few MB), allocate new ones
    BandBuffers n;
    hipError_t e = hipMalloc(&n.err, sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&n.xbuf, (size_t)words * sizeof(uint64_t));
    if (e != hipSuccess) return hip_status(e, "band exchange buffers");
    e = hipMemset(n.err, 0, sizeof(int));
    if (e != hipSuccess) return hip_status(e, "band error counter");
    e = hipMemset(n.xbuf, 0, (size_t)words * sizeof(uint64_t));
    if (e != hipSuccess) return hip_status(e, "band exchange buffers");
    n.words = words;
    b = n;
  }
  b.epoch = (b.epoch + 1) & 0x7ff;  // words left by an earlier launch never carry this launch's tags
  *out = b;
  return SS_OK;
}

template <class S>
int band_forward(const ss_sweep_group &grp, const Geom &g, int spin_limit, hipStream_t st) {
  constexpr int NB = S::H / S::BH;
  auto kern = fwd_band_kernel<S>;
  constexpr size_t lds = band_lds_bytes<S>();
  int dev = 0, ncu = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBandThreads, lds);
  if (e != hipSuccess) return hip_status(e, "band launch setup");
  // every team on one XCD: a multiple of 8 XCDs x NB bands, all resident at once
  const int grid = (ncu * std::min(per_cu, 1)) / (kXcds * NB) * (kXcds * NB);
  if (grid < kXcds * NB) {
    set_error("band kernels: the device cannot hold a team of workgroups");
    return SS_EUNSUP;
  }
  const int nteams = grid / NB;
  const int64_t nprop = (int64_t)g.nb * g.ns;
  const int nslots = (grp.t1 > 0 ? grp.t1 : g.nt) - grp.t0;
  if ((nprop + nteams - 1) / nteams * (nslots + 2) >= (1 << 20)) {
    set_error("band kernels: more than 2^20 steps per team in one call");
    return SS_EUNSUP;
  }
  BandBuffers bb;
  int rc = band_buffers(dev, st, (int64_t)nteams * S::XTEAM, &bb);
  if (rc != SS_OK) return rc;
  BandArgs a{};
  a.c2 = static_cast<const float *>(grp.c2);
  a.A = static_cast<const float *>(grp.A);
  a.B = static_cast<const float *>(grp.B);
  a.w = static_cast<const float *>(grp.w);
  a.amp = static_cast<const float *>(grp.amp);
  a.PS = nprop * S::HW;
  const int phase = grp.phase & 1;
  float *state = static_cast<float *>(grp.state);
  a.sin = state + 2 * phase * a.PS;
  a.sout = state + 2 * (phase ^ 1) * a.PS;
  a.hist = static_cast<float *>(grp.hist);
  a.gathers = static_cast<float *>(grp.gathers);
  a.nprop = (int)nprop;
  a.ns = g.ns;
  a.nt = g.nt;
  a.tlo = grp.t0;
  a.thi = grp.t1 > 0 ? grp.t1 : g.nt;
  a.resume = grp.resume ? 1 : 0;
  for (int s = 0; s < SS_MAX_SHOTS; ++s) a.src[s] = g.src[s];
  a.xbuf = bb.xbuf;
  a.epoch = bb.epoch;
  a.err = bb.err;
  a.spin_limit = spin_limit;
  void *args[] = {&a};
  e = hipLaunchCooperativeKernel((const void *)kern, dim3(grid), dim3(kBandThreads), args, (unsigned)lds, st);
  return hip_status(e, "band launch");
}

}  // namespace

bool band_fits(const Geom &g, int dir) { return dir == 0 && (band_is<Band280>(g) || band_is<Band180>(g)); }

int band_run(const ss_sweep_group &grp, const Geom &g, int dir, int spin_limit, hipStream_t st) {
  if (dir == 0 && band_is<Band280>(g)) return band_forward<Band280>(grp, g, spin_limit, st);
  if (dir == 0 && band_is<Band180>(g)) return band_forward<Band180>(grp, g, spin_limit, st);
  set_error("band kernels: no band layout for this grid");
  return SS_EUNSUP;
}

int band_errors(int *count) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_status(e, "band errors");
  std::vector<int *> ps;
  {
    std::lock_guard<std::mutex> lock(g_band_mutex);
    for (auto &kv : g_band_buffers)
      if (kv.first.first == dev) ps.push_back(kv.second.err);
  }
  *count = 0;
  for (int *p : ps) {
    int n = 0;
    e = hipMemcpy(&n, p, sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_status(e, "band errors");
    *count += n;
  }
  return SS_OK;
}

}  // namespace ss
