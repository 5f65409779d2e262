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
  const float *c2, *A, *B, *w, *amp;  // amp: forward source amplitude / adjoint source derivative
  const float *res;                   // adjoint: muted residual (nprop, nt, nx)
  const float *c2a;                   // adjoint: c2 * A per velocity row, zero on the two-cell rim
  float *gacc, *illum, *sgrad;        // adjoint outputs
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
typedef uint64_t *XR;  // a team's exchange region
// Pause of the exchange waves between polling rounds (units of 64 clocks, at most
// 127): polls share the memory system with the history stream.  Measured on the
// uneven batch: the adjoint's sweep 3.5 % shorter at 96 than at 1 (127: slower
// again); the forward's no better than at 1 between 16 and 32 and slower at 64.
constexpr int kFwdSleep = 32, kAdjSleep = 96;
__device__ __forceinline__ void x_put(uint64_t *p, float v, int tag) {
  const uint64_t u = ((uint64_t)(uint32_t)tag << 32) | __builtin_bit_cast(uint32_t, v);
  __hip_atomic_store(p, u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t x_get(XR r, int voff, int soff) {
  return __hip_atomic_load(r + (soff + voff) / 8, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Byte offset of an exchange side in a team's region: side 0 of a band holds its
// first two rows (read by the band above), side 1 its last two (read by the band
// below), two parities (the tag's lowest bit), 2 W words each.
template <class S>
__device__ __forceinline__ int x_side(int band, int side, int tag) {
  return ((band * 2 + side) * 2 + (tag & 1)) * S::XSIDE * 8;
}
template <class S>
__device__ __forceinline__ XR x_rsrc(uint64_t *xbuf, int team) { return xbuf + (int64_t)team * S::XTEAM; }

// The exchange wave: the neighbours' rows of p(t) (tag) into the halo rows, LDS
// rows 0, 1 (from the band above) and BH + 2, BH + 3 (from the band below).  Every
// lane loads its words of both sides at once, keeps those whose tag matches and
// reloads the rest, sleeping SLEEP x 64 clocks between rounds (fewer polls in
// the memory system under the history stream); past spin_limit rounds the
// missing words become NaN and the error counter rises.  `conv` turns a received
// word into the halo's value (side, word index, value).
template <class S, int SLEEP, class Conv>
__device__ void x_receive_sides(XR xr, int band, int tag, uint32_t halo_addr, int lane, int spin_limit, int *err,
                                bool &failed, Conv conv) {
  constexpr int NB = S::H / S::BH, NW = S::XSIDE, NWL = (NW + 63) / 64;
  constexpr uint32_t kRow = 4u * S::W;
  const bool has[2] = {band > 0, band < NB - 1};
  const int src[2] = {x_side<S>(band - 1, 1, tag), x_side<S>(band + 1, 0, tag)};
  const uint32_t dst[2] = {halo_addr, halo_addr + (uint32_t)(S::BH + 2) * kRow};
  uint64_t pending = 0;  // bit sd * NWL + k: word lane + 64 k of side sd still missing
#pragma unroll
  for (int sd = 0; sd < 2; ++sd)
#pragma unroll
    for (int k = 0; k < NWL; ++k)
      if (has[sd] && lane + 64 * k < NW) pending |= 1ull << (sd * NWL + k);
  for (int spins = 0; pending != 0 || __builtin_amdgcn_ballot_w64(pending != 0); ++spins) {
    uint64_t v[2 * NWL];
#pragma unroll
    for (int sd = 0; sd < 2; ++sd)
#pragma unroll
      for (int k = 0; k < NWL; ++k)
        v[sd * NWL + k] = (pending >> (sd * NWL + k)) & 1u ? x_get(xr, (lane + 64 * k) * 8, src[sd]) : 0;
#pragma unroll
    for (int sd = 0; sd < 2; ++sd)
#pragma unroll
      for (int k = 0; k < NWL; ++k) {
        const int b = sd * NWL + k;
        if (((pending >> b) & 1u) && (int)(uint32_t)(v[b] >> 32) == tag) {
          const int w = lane + 64 * k;
          lds_st1(dst[sd] + 4u * (uint32_t)w, conv(sd, w, __builtin_bit_cast(float, (uint32_t)v[b])));
          pending &= ~(1ull << b);
        }
      }
    if (!__builtin_amdgcn_ballot_w64(pending != 0)) break;
    if (failed || spins >= spin_limit) {
      if (!failed && lane == 0) atomicAdd(err, 1);
      failed = true;
#pragma unroll
      for (int sd = 0; sd < 2; ++sd)
#pragma unroll
        for (int k = 0; k < NWL; ++k)
          if ((pending >> (sd * NWL + k)) & 1u) lds_st1(dst[sd] + 4u * (uint32_t)(lane + 64 * k), __builtin_nanf(""));
      break;
    }
    __builtin_amdgcn_s_sleep(SLEEP);
  }
}

template <class S>
__device__ void x_receive(XR xr, int band, int tag, uint32_t band_addr, int lane, int spin_limit, int *err,
                          bool &failed) {
  x_receive_sides<S, kFwdSleep>(xr, band, tag, band_addr, lane, spin_limit, err, failed,
                     [](int, int, float v) { return v; });
}

// One row k of a strip: p(t+1) into PO[k] from p(t) (PC, the vertical taps zm2 ..
// zp2 and the left / right neighbour pairs lf, rt) and p(t-1) (PO[k]); its
// Laplacian to the history (zero on the two-cell rim, gpu.py:35).
template <int R, int W>
__device__ __forceinline__ void fwd_row(float2 (&PO)[R], const float2 (&PC)[R], const float2 (&cc)[R], bool inner,
                                        int k, float2 zm2, float2 zm1, float2 zp1, float2 zp2, float2 lf, float2 rt,
                                        float4 ab, __amdgpu_buffer_rsrc_t hrs, int hv, int hoff) {
  const float2 zc = PC[k];
  const float s0 = lap_sum(zm2.x, zm1.x, zc.x, zp1.x, zp2.x, lf.x, lf.y, zc.y, rt.x);
  const float s1 = lap_sum(zm2.y, zm1.y, zc.y, zp1.y, zp2.y, lf.y, zc.x, rt.x, rt.y);
  const float2 Lq = div12x2(s0, s1);
  const float L0 = inner ? Lq.x : 0.0f, L1 = inner ? Lq.y : 0.0f;
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
#pragma unroll
  for (int k = 0; k < R; ++k) tb[k] = pair_desc((z0 + k) * NC + c, H, W, nz, nx, pad) & kPTabMask;
  // Laplacian interior: the two-cell rim is the first and last column pair, the
  // first two rows of the top band's first strip and the last two rows of the
  // bottom band's last strip (all four among the rows updated after the halo)
  const bool in_mid = c > 0 && c < NC - 1;
  const bool in_top = in_mid && !(band == 0 && strip == 0), in_bot = in_mid && !(band == NB - 1 && strip == S::NS - 1);
  const int hv = act ? (z0 * W + 2 * c) * 4 : 0x40000000;  // idle lanes store past every resource
  const int gv = act ? (2 * c - pad) * 4 : 0x40000000;
  const uint32_t aa0 = kBand + (uint32_t)((r0 * W + 2 * c) * 4);  // band region row r0: the strip's row -2
  const int nslots = a.thi - a.tlo;
  const XR xr = x_rsrc<S>(a.xbuf, team);
  // tag of a plane's slot 0: launch count, then the step count.  A plane's steps
  // carry tags gbase .. gbase + nslots; the next plane starts two higher, so its
  // first words (same parity) replace the last plane's final rows, which no step
  // reads, and never rows a neighbour may still be reading.
  auto tag0 = [&](int prop) { return (a.epoch << 20) + (prop / nteams) * (nslots + 2); };
  if (xw) {  // the exchange wave: one receive per step, the compute waves' barriers
    __builtin_amdgcn_s_setprio(2);  // its halo work ahead of the compute waves' issue (uneven step -1 %)
    bool failed = false;
    for (int prop = team; prop < a.nprop; prop += nteams) {
      const int gbase = tag0(prop);
      __syncthreads();
      for (int it = a.tlo; it < a.thi; ++it) {
        x_receive<S>(xr, band, gbase + (it - a.tlo), kBand, t & 63, a.spin_limit, a.err, failed);
        lds_barrier();
        lds_barrier();
        lds_barrier();
      }
      __syncthreads();
    }
    return;
  }
  // edge strips publish the band's first two rows (strip 0) and last two (strip NS - 1)
  const bool pub_up = act && strip == 0 && band > 0, pub_dn = act && strip == S::NS - 1 && band < NB - 1;
  uint64_t *const xl = xr + 2 * c;  // the lane's first word in a side; the side's offset is uniform
  auto publish = [&](const float2(&Pv)[R], int tag) __attribute__((always_inline)) {
    if (pub_up) {
      uint64_t *d = xl + x_side<S>(band, 0, tag) / 8;
      x_put(d, Pv[0].x, tag);
      x_put(d + 1, Pv[0].y, tag);
      x_put(d + W, Pv[1].x, tag);
      x_put(d + W + 1, Pv[1].y, tag);
    }
    if (pub_dn) {
      uint64_t *d = xl + x_side<S>(band, 1, tag) / 8;
      x_put(d, Pv[R - 2].x, tag);
      x_put(d + 1, Pv[R - 2].y, tag);
      x_put(d + W, Pv[R - 1].x, tag);
      x_put(d + W + 1, Pv[R - 1].y, tag);
    }
  };

  for (int prop = team; prop < a.nprop; prop += nteams) {  // the same planes for every band of a team
    const int gbase = tag0(prop);
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
      {  // rows 2 .. R-3: taps inside the strip, neighbours from LDS one row ahead
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
          fwd_row<R, W>(PO, PC, cc, in_mid, k, PC[k - 2], PC[k - 1], PC[k + 1], PC[k + 2], lf, rt, ab, hrs, hv,
                        hoff);
          lf = lfn;
          rt = rtn;
          ab = abn;
        }
      }
      lds_barrier();  // halo rows of p(t) in place
      {  // rows 0, 1, R-2, R-1: taps from the neighbouring strips or the halo rows
        uint32_t aa = aa0;
        asm volatile("" : "+v"(aa));
        const float2 up2 = lds_ld2(aa), up1 = lds_ld2(aa + kRow);
        const float2 dn1 = lds_ld2(aa + (uint32_t)(R + 2) * kRow), dn2 = lds_ld2(aa + (uint32_t)(R + 3) * kRow);
        const float2 l0 = lds_ld2(aa + 2 * kRow - 8), r0v = lds_ld2(aa + 2 * kRow + 8);
        const float2 l1 = lds_ld2(aa + 3 * kRow - 8), r1v = lds_ld2(aa + 3 * kRow + 8);
        const float2 l2 = lds_ld2(aa + (uint32_t)R * kRow - 8), r2v = lds_ld2(aa + (uint32_t)R * kRow + 8);
        const float2 l3 = lds_ld2(aa + (uint32_t)(R + 1) * kRow - 8), r3v = lds_ld2(aa + (uint32_t)(R + 1) * kRow + 8);
        const float4 e0 = lds_ld4(tb[0]), e1 = lds_ld4(tb[1]), e2 = lds_ld4(tb[R - 2]), e3 = lds_ld4(tb[R - 1]);
        fwd_row<R, W>(PO, PC, cc, in_top, 0, up2, up1, PC[1], PC[2], l0, r0v, e0, hrs, hv, hoff);
        fwd_row<R, W>(PO, PC, cc, in_top, 1, up1, PC[0], PC[2], PC[3], l1, r1v, e1, hrs, hv, hoff);
        fwd_row<R, W>(PO, PC, cc, in_bot, R - 2, PC[R - 4], PC[R - 3], PC[R - 1], dn1, l2, r2v, e2, hrs, hv, hoff);
        fwd_row<R, W>(PO, PC, cc, in_bot, R - 1, PC[R - 3], PC[R - 2], dn1, dn2, l3, r3v, e3, hrs, hv, hoff);
        publish(PO, g + 1);  // p(t+1) of the edge rows, final (the source row is not among them)
      }
      lds_barrier();  // every wave has read p(t)
      {
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
    if (act) {
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

// ------------------------------------------------------------------- adjoint
//
// The band version of the column-strip adjoint (ss_resident.hip): registers hold
// m1, m2 (two sets swapping roles) and the imaging sum of the strip's rows, LDS
// holds q = (c2 A m1) / 12 of the band plus two halo rows above and below, the
// illumination of the band's model rows and the band's sponge entries
// (2A0, A0 B0, 2A1, A1 B1).  Phase 1 updates rows 2 .. R-3 of every strip while
// the exchange wave turns the neighbours' new edge-row states into the halo's q,
// then rows 0, 1, R-2, R-1; the edge strips publish the new state of the band's
// edge rows; phase 2 publishes q from the new state.  q = ((c2 A) m) / 12 comes
// from the c2 * A plane of the velocity row (c2a_kernel: c2 A rounded once, as
// the reference's first product, and zero on the two-cell rim, so q is zero
// there as lap_adjoint reads only p[2:-2, 2:-2], gpu.py:41).
//
// LDS: [sponge table | illumination | halo coefficients c2 A (4 rows) | guard |
// q: BH + 4 rows | guard].  A band inside the plane's vertical extent away from
// the top and bottom sponges needs only the table's dz = 0 entries, and the
// bands that touch those sponges have fewer model rows: both fit below the
// coefficient rows, which start at a fixed address.
template <class S>
struct AdjBandLayout {
  static constexpr bool zsponge(int b) { return b * S::BH < S::pad || (b + 1) * S::BH > S::H - S::pad; }
  static constexpr int tab_floats(int b) { return 4 * (zsponge(b) ? (S::pad + 1) * (S::pad + 1) : S::pad + 1); }
  static constexpr int model_lo(int b) { return b * S::BH > S::pad ? b * S::BH : S::pad; }
  static constexpr int model_rows(int b) {
    const int hi = (b + 1) * S::BH < S::pad + S::nz ? (b + 1) * S::BH : S::pad + S::nz;
    return hi > model_lo(b) ? hi - model_lo(b) : 0;
  }
  static constexpr int base() {
    int m = 0;
    for (int b = 0; b < S::H / S::BH; ++b) {
      const int v = tab_floats(b) + model_rows(b) * S::nx;
      m = v > m ? v : m;
    }
    return (m + 3) / 4 * 4;
  }
  static constexpr int coef = base();                              // 4 rows x W
  static constexpr int q0 = coef + 4 * S::W + pair_guard(S::W);    // q region row 0 (upper halo)
  static constexpr int total = q0 + (S::BH + 4) * S::W + pair_guard(S::W);
};

template <class S>
constexpr size_t adj_band_lds_bytes() {
  return (size_t)AdjBandLayout<S>::total * sizeof(float);
}

// The adjoint's exchange wave: the neighbours' new edge-row states m (tag) into the
// halo's q = ((c2 A) m) / 12, c2 A from the coefficient rows (zero on the rim,
// where q is +0).
template <class S>
__device__ void x_receive_q(XR xr, int band, int tag, uint32_t q_addr, uint32_t coef_addr, int lane, int spin_limit,
                            int *err, bool &failed) {
  constexpr uint32_t kRow = 4u * S::W;
  x_receive_sides<S, kAdjSleep>(xr, band, tag, q_addr, lane, spin_limit, err, failed, [&](int sd, int w, float m) {
    const float ca = lds_table_a(coef_addr + (uint32_t)sd * 2u * kRow + 4u * (uint32_t)w);
    const float q = div12(ca * m);
    return ca != 0.0f ? q : 0.0f;
  });
}

template <class S>
__global__ __launch_bounds__(kBandThreads) void adj_band_kernel(BandArgs a) {
  extern __shared__ float lds[];
  using L = AdjBandLayout<S>;
  constexpr int H = S::H, W = S::W, HW = S::HW, BH = S::BH, R = S::R, NC = S::NC, NACT = S::NACT;
  constexpr int NB = H / BH, nz = S::nz, nx = S::nx, pad = S::pad;
  constexpr uint32_t kRow = 4u * W;
  constexpr uint32_t kQ = (uint32_t)L::q0 * 4, kCoef = (uint32_t)L::coef * 4;
  constexpr int DH = 3, DC = 8;  // history rows loaded ahead; depth of the c2 A ring
  static_assert(R - 4 >= DH && R >= DC, "strips too short for the prefetch depths");
  float4 *tab = reinterpret_cast<float4 *>(lds);
  float *Q = lds + L::q0;
  float2 *Q2 = reinterpret_cast<float2 *>(Q);
  const int t = threadIdx.x;
  const bool xw = (t >> 6) == S::NCW;
  const int per_xcd = (int)gridDim.x / kXcds;
  const int j = (int)blockIdx.x / kXcds;
  const int band = j % NB, team = (int)(blockIdx.x % kXcds) * (per_xcd / NB) + j / NB;
  const int nteams = (int)gridDim.x / NB;
  const bool act = t < NACT;
  const int tt = act ? t : NACT - 1;
  const int strip = tt / NC, c = tt - strip * NC, r0 = strip * R, z0 = band * BH + r0;

  {  // the band's sponge entries: every (dz, dx0) where the band reaches a vertical sponge, else dz = 0
    const int n = L::zsponge(band) ? (pad + 1) * (pad + 1) : pad + 1;
    for (int k = t; k < n; k += kBandThreads) {
      const int dz = k / (pad + 1), dx0 = k - dz * (pad + 1);
      const int x0 = dx0 == 0 ? pad : ((dx0 & 1) ? W - 1 - pad + dx0 : pad - dx0);
      const int cc = (pad - dz) * W + x0;
      const float a0 = a.A[cc], b0 = a.B[cc], a1 = a.A[cc + 1], b1 = a.B[cc + 1];
      tab[k] = make_float4(2.0f * a0, a0 * b0, 2.0f * a1, a1 * b1);
    }
  }
  const int il_f = L::tab_floats(band), mz0 = L::model_lo(band), nil = L::model_rows(band) * nx;
  const uint32_t il_addr = (uint32_t)il_f * 4;
  uint32_t mdl = 0;  // bit k: row k of the strip is a model row in a model column
  const bool colmodel = act && c >= pad / 2 && c < (pad + nx) / 2;
#pragma unroll
  for (int k = 0; k < R; ++k)
    if (colmodel && z0 + k >= pad && z0 + k < pad + nz) mdl |= 1u << k;
  // sponge-table byte offset of row k's pair (the pair descriptor's entry index,
  // (dz (pad + 1) + dx0) 16), formed per use: dz = max(pad - z, z - (H - 1 - pad), 0)
  const uint32_t tdx = pair_desc(z0 * NC + c, H, W, nz, nx, pad) & kPTabMask;
  const uint32_t tx0 = tdx - (uint32_t)(sponge_dist(z0, H, pad) * (pad + 1) * 16);  // the dx0 part
  // (loop-invariant: the step redefines its inputs through an empty asm, or the
  // compiler keeps all R offsets live across the time loop and spills them)
  int tz = z0;
  uint32_t tx = tx0;
  auto tbk = [&](int k) __attribute__((always_inline)) {
    const int z = tz + k;
    const int dz = max(max(pad - z, z - (H - 1 - pad)), 0);
    return tx + (uint32_t)(dz * (pad + 1) * 16);
  };
  const int hv = act ? (z0 * W + 2 * c) * 4 : 0x40000000;  // idle lanes read zeros and write nothing
  const uint32_t aa0 = kQ + (uint32_t)((r0 * W + 2 * c) * 4);  // q region row r0: the strip's row -2
  const uint32_t ia0 = il_addr + (uint32_t)(((z0 - mz0) * nx + 2 * c - pad) * 4);  // model rows only
  const int nslots = a.thi - a.tlo;
  const XR xr = x_rsrc<S>(a.xbuf, team);
  auto tag0 = [&](int prop) { return (a.epoch << 20) + (prop / nteams) * (nslots + 2); };  // as the forward's
  // every wave clears the plane's LDS state and fills the halo's c2 A rows
  auto plane_setup = [&](int prop) __attribute__((always_inline)) {
    const int b = prop / a.ns;
    for (int k = t; k < (BH + 4) * W + 2 * pair_guard(W); k += kBandThreads) Q[k - pair_guard(W)] = 0.0f;
    for (int k = t; k < nil; k += kBandThreads) lds[il_f + k] = 0.0f;
    const float *cab = a.c2a + (int64_t)b * HW;  // c2 A of the halo rows: rows -2, -1 and BH, BH + 1
    for (int k = t; k < 4 * W; k += kBandThreads) {
      const int i = k / W, x = k - i * W;
      const int z = band * BH + (i < 2 ? i - 2 : BH + i - 2);
      lds[L::coef + k] = z >= 0 && z < H ? cab[z * W + x] : 0.0f;
    }
  };
  if (xw) {  // the exchange wave: one receive per step, the compute waves' barriers
    __builtin_amdgcn_s_setprio(2);  // its halo work ahead of the compute waves' issue (uneven step -1 %)
    bool failed = false;
    for (int prop = team; prop < a.nprop; prop += nteams) {
      const int gbase = tag0(prop);
      plane_setup(prop);
      __syncthreads();
      for (int it = a.thi - 1; it >= a.tlo; --it) {
        x_receive_q<S>(xr, band, gbase + (a.thi - 1 - it), kQ, kCoef, t & 63, a.spin_limit, a.err, failed);
        lds_barrier();
        lds_barrier();
        lds_barrier();
      }
      __syncthreads();
    }
    return;
  }
  const bool pub_up = act && strip == 0 && band > 0, pub_dn = act && strip == S::NS - 1 && band < NB - 1;
  uint64_t *const xl = xr + 2 * c;  // the lane's first word in a side; the side's offset is uniform
  auto publish = [&](const float2(&Mv)[R], int tag) __attribute__((always_inline)) {
    if (pub_up) {
      uint64_t *d = xl + x_side<S>(band, 0, tag) / 8;
      x_put(d, Mv[0].x, tag);
      x_put(d + 1, Mv[0].y, tag);
      x_put(d + W, Mv[1].x, tag);
      x_put(d + W + 1, Mv[1].y, tag);
    }
    if (pub_dn) {
      uint64_t *d = xl + x_side<S>(band, 1, tag) / 8;
      x_put(d, Mv[R - 2].x, tag);
      x_put(d + 1, Mv[R - 2].y, tag);
      x_put(d + W, Mv[R - 1].x, tag);
      x_put(d + W + 1, Mv[R - 1].y, tag);
    }
  };

  for (int prop = team; prop < a.nprop; prop += nteams) {
    const int gbase = tag0(prop);
    const int b = prop / a.ns, s = prop - b * a.ns;
    plane_setup(prop);  // zero state: q, the halo rows and the illumination start at zero
    float2 m1[R], m2[R], ga[R];
#pragma unroll
    for (int k = 0; k < R; ++k) m1[k] = m2[k] = ga[k] = make_float2(0.0f, 0.0f);
    publish(m1, gbase);  // the edge rows' zero state
    const int srcx = pad + a.src[s];
    const bool own_src = act && band == S::RB && strip == S::RS && c == (srcx >> 1);
    const bool src_hi = (srcx & 1) != 0;
    const float sdv = a.amp[prop];
    float sg = 0.0f;
    const bool rec = act && band == S::RB && strip == S::RS && c >= pad / 2 && c < (pad + nx) / 2;
    const int rvv = rec ? (2 * c - pad) * 4 : 0x40000000;
    const __amdgpu_buffer_rsrc_t rrs = make_rsrc(a.res + (int64_t)prop * a.nt * nx, (int64_t)a.nt * nx);
    const __amdgpu_buffer_rsrc_t hrs = make_rsrc(a.hist + (int64_t)prop * nslots * HW, (int64_t)nslots * HW);
    const __amdgpu_buffer_rsrc_t crs = make_rsrc(a.c2a + (int64_t)b * HW, (int64_t)HW);
    __syncthreads();

    // history of the first DH rows of a step (rows 2, 3, 4: phase 1's order starts at row 2)
    // and the receiver lanes' residual pair, loaded at the end of the previous step
    float2 Lh[DH], rv;
    auto load_step = [&](int it) __attribute__((always_inline)) {
#pragma unroll
      for (int d = 0; d < DH; ++d) Lh[d] = buf_load2(hrs, hv, (it - a.tlo) * HW * 4 + (2 + d) * (int)kRow);
      rv = buf_load2(rrs, rvv, it * nx * 4);
    };
    load_step(a.thi - 1);
    // row k of phase 1: mu (gpu.py:179-180), imaging sum and illumination (gpu.py:181, 183)
    auto row = [&](float2(&M1)[R], float2(&M2)[R], int k, float2 zm2, float2 zm1, float2 zc, float2 zp1,
                   float2 zp2, float2 lf, float2 rt, float4 e, float2 Lv, float2 &msrc)
                   __attribute__((always_inline)) {
      const float lt0 = lapt_sum(zc.x, zp2.x, zp1.x, zm1.x, zm2.x, rt.x, zc.y, lf.y, lf.x);
      const float lt1 = lapt_sum(zc.y, zp2.y, zp1.y, zm1.y, zm2.y, rt.y, rt.x, zc.x, lf.y);
      float mu0 = e.x * M1[k].x;
      float mu1 = e.z * M1[k].y;
      mu0 = mu0 + lt0;
      mu1 = mu1 + lt1;
      mu0 = mu0 - e.y * M2[k].x;
      mu1 = mu1 - e.w * M2[k].y;
      if (k == S::RK && rec) {  // gpu.py:180
        mu0 = mu0 + rv.x;
        mu1 = mu1 + rv.y;
      }
      ga[k].x = ga[k].x + ((0.5f * e.x) * mu0) * Lv.x;  // A = 0.5 (2A), exact
      ga[k].y = ga[k].y + ((0.5f * e.z) * mu1) * Lv.y;
      if (k == S::RK) msrc = make_float2(mu0, mu1);
      if ((mdl >> k) & 1u) {
        const uint32_t ia = ia0 + (uint32_t)(k * nx * 4);
        float2 iv = lds_ld2(ia);
        iv.x = iv.x + Lv.x * Lv.x;
        iv.y = iv.y + Lv.y * Lv.y;
        lds_st2(ia, iv);
      }
      M2[k] = make_float2(mu0, mu1);  // the new state; M1 (the old one) becomes the next step's m2
      asm volatile("" : "+v"(ga[k].x), "+v"(ga[k].y), "+v"(M2[k].x), "+v"(M2[k].y));
    };
    // processing order: rows 2 .. R-3, then 0, 1, R-2, R-1; history DH rows ahead in that order
    auto ord = [](int i) constexpr { return i < R - 4 ? i + 2 : (i == R - 4 ? 0 : i == R - 3 ? 1 : i); };
    auto step = [&](float2(&M1)[R], float2(&M2)[R], const int it) __attribute__((always_inline)) {
      const int g = gbase + (a.thi - 1 - it), slot = (it - a.tlo) * HW * 4;
      const float wt = a.w[it];
      float2 msrc = make_float2(0.0f, 0.0f);
      uint32_t aa = aa0;
      asm volatile("" : "+v"(aa));
      tz = z0;
      tx = tx0;
      asm volatile("" : "+v"(tz), "+v"(tx));
      {  // phase 1a: rows 2 .. R-3, a window of five q rows sliding down the strip
        float2 z0v = lds_ld2(aa + 2 * kRow), z1v = lds_ld2(aa + 3 * kRow), z2v = lds_ld2(aa + 4 * kRow),
               z3v = lds_ld2(aa + 5 * kRow), z4v = lds_ld2(aa + 6 * kRow);
        float2 lf = lds_ld2(aa + 4 * kRow - 8), rt = lds_ld2(aa + 4 * kRow + 8);
        float4 e = lds_ld4(tbk(2));
#pragma unroll
        for (int i = 0; i < R - 4; ++i) {
          const int k = i + 2;
          asm volatile("" : "+v"(aa));
          float2 z5v = make_float2(0.0f, 0.0f), lfn = z5v, rtn = z5v;
          float4 en = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          if (k + 1 <= R - 3) {
            z5v = lds_ld2(aa + (uint32_t)(k + 5) * kRow);
            lfn = lds_ld2(aa + (uint32_t)(k + 3) * kRow - 8);
            rtn = lds_ld2(aa + (uint32_t)(k + 3) * kRow + 8);
            en = lds_ld4(tbk(k + 1));
          }
          const float2 Lv = Lh[i % DH];
          if (i + DH < R) Lh[i % DH] = buf_load2(hrs, hv, slot + ord(i + DH) * (int)kRow);
          row(M1, M2, k, z0v, z1v, z2v, z3v, z4v, lf, rt, e, Lv, msrc);
          z0v = z1v;
          z1v = z2v;
          z2v = z3v;
          z3v = z4v;
          z4v = z5v;
          lf = lfn;
          rt = rtn;
          e = en;
        }
      }
      lds_barrier();  // the halo's q in place
      {  // phase 1b: rows 0, 1, R-2, R-1 (taps from the neighbouring strips or the halo)
        constexpr int i0 = R - 4;  // ring position of row 0; the ring's last load is row R-1
        // rows 0 and 1: q rows -2 .. 3
        {
          const float2 u0 = lds_ld2(aa), u1 = lds_ld2(aa + kRow), u2 = lds_ld2(aa + 2 * kRow),
                       u3 = lds_ld2(aa + 3 * kRow), u4 = lds_ld2(aa + 4 * kRow);
          const float2 l0 = lds_ld2(aa + 2 * kRow - 8), r0v = lds_ld2(aa + 2 * kRow + 8);
          const float4 e0 = lds_ld4(tbk(0));
          const float2 u5 = lds_ld2(aa + 5 * kRow), l1 = lds_ld2(aa + 3 * kRow - 8), r1v = lds_ld2(aa + 3 * kRow + 8);
          const float4 e1 = lds_ld4(tbk(1));
          const float2 L0 = Lh[i0 % DH];
          Lh[i0 % DH] = buf_load2(hrs, hv, slot + ord(i0 + DH) * (int)kRow);
          row(M1, M2, 0, u0, u1, u2, u3, u4, l0, r0v, e0, L0, msrc);
          row(M1, M2, 1, u1, u2, u3, u4, u5, l1, r1v, e1, Lh[(i0 + 1) % DH], msrc);
        }
        // rows R-2 and R-1: q rows R-4 .. R+1
        {
          const float2 d0 = lds_ld2(aa + (uint32_t)(R - 2) * kRow), d1 = lds_ld2(aa + (uint32_t)(R - 1) * kRow),
                       d2 = lds_ld2(aa + (uint32_t)R * kRow), d3 = lds_ld2(aa + (uint32_t)(R + 1) * kRow),
                       d4 = lds_ld2(aa + (uint32_t)(R + 2) * kRow);
          const float2 l2 = lds_ld2(aa + (uint32_t)R * kRow - 8), r2v = lds_ld2(aa + (uint32_t)R * kRow + 8);
          const float4 e2 = lds_ld4(tbk(R - 2));
          const float2 d5 = lds_ld2(aa + (uint32_t)(R + 3) * kRow), l3 = lds_ld2(aa + (uint32_t)(R + 1) * kRow - 8),
                       r3v = lds_ld2(aa + (uint32_t)(R + 1) * kRow + 8);
          const float4 e3 = lds_ld4(tbk(R - 1));
          row(M1, M2, R - 2, d0, d1, d2, d3, d4, l2, r2v, e2, Lh[(i0 + 2) % DH], msrc);
          row(M1, M2, R - 1, d1, d2, d3, d4, d5, l3, r3v, e3, Lh[(i0 + 3) % DH], msrc);
        }
        if (own_src) sg = sg + ((src_hi ? msrc.y : msrc.x) * wt) * sdv;  // gpu.py:182, in step order
        publish(M2, g + 1);  // the new state of the band's edge rows
      }
      // Phase 2 reads c2 A through a ring DC rows deep whose first loads precede the
      // barrier; loads complete in issue order, so the next step's history rows are
      // issued after phase 2.
      float2 cb[DC];
#pragma unroll
      for (int k = 0; k < DC; ++k) cb[k] = buf_load2(crs, hv, k * (int)kRow);
      lds_barrier();  // every wave has read q
      {
#pragma unroll
        for (int k = 0; k < R; ++k) {
          const float2 cv = cb[k % DC];
          if (k + DC < R) cb[k % DC] = buf_load2(crs, hv, (k + DC) * (int)kRow);
          if (act) lds_st2(aa + (uint32_t)(k + 2) * kRow, div12x2(cv.x * M2[k].x, cv.y * M2[k].y));
        }
        if (it - 1 >= a.tlo) load_step(it - 1);
      }
      lds_barrier();
    };
    int it = a.thi - 1;
    for (; it - 1 >= a.tlo; it -= 2) {
      step(m1, m2, it);
      step(m2, m1, it - 1);
    }
    const bool odd = it >= a.tlo;  // one step left: the current state is then in m2
    if (odd) step(m1, m2, it);
    if (act) {
      float2 *m1o = reinterpret_cast<float2 *>(a.sout + (int64_t)prop * HW);
      float2 *m2o = reinterpret_cast<float2 *>(a.sout + a.PS + (int64_t)prop * HW);
      float2 *gao = reinterpret_cast<float2 *>(a.gacc + (int64_t)prop * HW);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int q = (z0 + k) * NC + c;
        m1o[q] = odd ? m2[k] : m1[k];
        m2o[q] = odd ? m1[k] : m2[k];
        gao[q] = ga[k];
      }
    }
    if (own_src) a.sgrad[prop] = sg;
    float *ilg = a.illum + (int64_t)prop * nz * nx + (int64_t)(mz0 - pad) * nx;
    for (int k = t; k < nil; k += S::NCW * 64) ilg[k] = lds[il_f + k];  // the compute waves' lanes
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
  float *c2a = nullptr;  // adjoint: c2 * A of the call's velocity rows
  int64_t c2a_floats = 0;
  int epoch = 0;
};
std::mutex g_band_mutex;
std::map<std::pair<int, hipStream_t>, BandBuffers> g_band_buffers;

int band_buffers(int dev, hipStream_t st, int64_t words, int64_t c2a_floats, BandBuffers *out) {
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
  if (b.c2a_floats < c2a_floats) {
    // earlier launches on this stream may still read the old plane
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess && b.c2a) e = hipFree(b.c2a);
    b.c2a = nullptr;
    b.c2a_floats = 0;
    if (e == hipSuccess) e = hipMalloc(&b.c2a, (size_t)c2a_floats * sizeof(float));
    if (e != hipSuccess) return hip_status(e, "band c2 * A plane");
    b.c2a_floats = c2a_floats;
  }
  b.epoch = (b.epoch + 1) & 0x7ff;  // words left by an earlier launch never carry this launch's tags
  *out = b;
  return SS_OK;
}

// Workgroups of a band launch: one per CU, a multiple of 8 XCDs x NB bands (every
// team on one XCD), all resident at once; the kernel's LDS attribute is set.
template <class S>
int band_grid(int dir, int *dev, int *grid) {
  constexpr int NB = S::H / S::BH;
  auto kern = dir ? adj_band_kernel<S> : fwd_band_kernel<S>;
  const size_t lds = dir ? adj_band_lds_bytes<S>() : band_lds_bytes<S>();
  int ncu = 0, per_cu = 0;
  hipError_t e = hipGetDevice(dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, *dev);
  if (e == hipSuccess) e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBandThreads, lds);
  if (e != hipSuccess) return hip_status(e, "band launch setup");
  *grid = (ncu * std::min(per_cu, 1)) / (kXcds * NB) * (kXcds * NB);
  if (*grid < kXcds * NB) {
    set_error("band kernels: the device cannot hold a team of workgroups");
    return SS_EUNSUP;
  }
  return SS_OK;
}

template <class S>
int band_launch(const ss_sweep_group &grp, const Geom &g, int dir, int spin_limit, hipStream_t st) {
  constexpr int NB = S::H / S::BH;
  auto kern = dir ? adj_band_kernel<S> : fwd_band_kernel<S>;
  const size_t lds = dir ? adj_band_lds_bytes<S>() : band_lds_bytes<S>();
  int dev = 0, grid = 0;
  int rc = band_grid<S>(dir, &dev, &grid);
  if (rc != SS_OK) return rc;
  hipError_t e = hipSuccess;
  const int nteams = grid / NB;
  const int64_t nprop = (int64_t)g.nb * g.ns;
  const int nslots = (grp.t1 > 0 ? grp.t1 : g.nt) - grp.t0;
  if ((nprop + nteams - 1) / nteams * (nslots + 2) >= (1 << 20)) {
    set_error("band kernels: more than 2^20 steps per team in one call");
    return SS_EUNSUP;
  }
  BandBuffers bb;
  rc = band_buffers(dev, st, (int64_t)nteams * S::XTEAM, dir ? (int64_t)g.nb * S::HW : 0, &bb);
  if (rc != SS_OK) return rc;
  if (dir) {
    rc = c2a_plane(g.nb, S::H, S::W, static_cast<const float *>(grp.c2), static_cast<const float *>(grp.A), bb.c2a,
                   st);
    if (rc != SS_OK) return rc;
  }
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
  a.gathers = dir ? nullptr : static_cast<float *>(grp.gathers);
  a.res = static_cast<const float *>(grp.res);
  a.c2a = bb.c2a;
  a.gacc = static_cast<float *>(grp.gacc);
  a.illum = static_cast<float *>(grp.illum);
  a.sgrad = static_cast<float *>(grp.src_grad);
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

bool band_fits(const Geom &g, int dir) { return band_is<Band280>(g) || band_is<Band180>(g); }

int band_run(const ss_sweep_group &grp, const Geom &g, int dir, int spin_limit, hipStream_t st) {
  if (band_is<Band280>(g)) return band_launch<Band280>(grp, g, dir, spin_limit, st);
  if (band_is<Band180>(g)) return band_launch<Band180>(grp, g, dir, spin_limit, st);
  set_error("band kernels: no band layout for this grid");
  return SS_EUNSUP;
}

template <class S>
int band_teams(const Geom &) {
  int dev = 0, g0 = 0, g1 = 0;
  int rc = band_grid<S>(0, &dev, &g0);
  if (rc == SS_OK) rc = band_grid<S>(1, &dev, &g1);
  return rc != SS_OK ? rc : std::min(g0, g1) / (S::H / S::BH);
}

int band_round(const Geom &g) {
  if (band_is<Band280>(g)) return band_teams<Band280>(g);
  if (band_is<Band180>(g)) return band_teams<Band180>(g);
  if (resident_uses_c2a(g)) {  // the column-strip kernels: one workgroup per CU
    int dev = 0, ncu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    return e == hipSuccess ? ncu : hip_status(e, "sweep round");
  }
  return 0;
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
