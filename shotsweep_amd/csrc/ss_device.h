// Device-only helpers shared by the LDS-resident and tiled FP32 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ss {

constexpr int kXcds = 8;                    // MI355X: 8 XCDs, workgroups dealt round-robin
constexpr uint32_t kBufWord3 = 0x00020000;  // raw buffer descriptor word 3 for the gfx9 family

// Distance of index i into the absorbing layer of an axis of n cells
// (gpu.py:86-87: max(pad - i, i - (n - 1 - pad)) clipped at 0).
__host__ __device__ __forceinline__ int sponge_dist(int i, int n, int pad) {
  int d = pad - i;
  int e = i - (n - 1 - pad);
  d = d > e ? d : e;
  return d > 0 ? d : 0;
}

// Buffer resource over nfloats floats at base (range capped at 2^31 - 1 bytes).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const float *base, int64_t nfloats) {
  const int64_t bytes = nfloats * 4;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base), (short)0,
                                           (int)(bytes > 0x7fffffff ? 0x7fffffff : bytes), (int)kBufWord3);
}

// Float views of the 32-bit raw buffer intrinsics (which move untyped bits).
__device__ __forceinline__ void buf_store(float v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
}
__device__ __forceinline__ float buf_load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// Two consecutive floats (8-byte aligned offsets) as one 64-bit buffer access.
// The 64-bit builtins move a vector of two 32-bit words (word 0 = first float).
// The words are packed and unpacked through one uint64 bit cast: reading the
// vector's elements individually after the load is mis-optimised by this
// compiler (the load is narrowed to one dword that feeds both elements).
typedef unsigned int ss_words2 __attribute__((__vector_size__(2 * sizeof(unsigned int))));
__device__ __forceinline__ void buf_store2(float2 v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  const uint64_t u = (uint64_t)__builtin_bit_cast(uint32_t, v.x) | ((uint64_t)__builtin_bit_cast(uint32_t, v.y) << 32);
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(ss_words2, u), r, voff, soff, 0);
}
__device__ __forceinline__ float2 buf_load2(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  const uint64_t u = __builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  return make_float2(__builtin_bit_cast(float, (uint32_t)u), __builtin_bit_cast(float, (uint32_t)(u >> 32)));
}

// Sponge tables at LDS address 0.  Kernels without static LDS (checked on the
// host before launch) start their dynamic LDS at address 0, so a table placed
// first is read at the LDS address equal to the entry's byte offset, with no
// base add.  Tables hold 8-byte entries: (A, B) for the forward, (A, A*B) or
// (2A, A*B) for the adjoint (2A is exact, and A = 0.5 * 2A).
typedef float ss_v2f __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) const float ss_lds_cfloat;
typedef __attribute__((address_space(3))) const ss_v2f ss_lds_cv2f;
__device__ __forceinline__ float2 lds_table2(uint32_t byte_off) {
  const ss_v2f v = *reinterpret_cast<ss_lds_cv2f *>((uintptr_t)byte_off);
  return make_float2(v.x, v.y);
}
// 8-byte read at an LDS byte address (same addressing as the tables).
__device__ __forceinline__ float2 lds_ld2(uint32_t byte_addr) { return lds_table2(byte_addr); }
// 8-byte write and 16-byte read at LDS byte addresses.
typedef float ss_v4f __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) ss_v2f ss_lds_v2f;
typedef __attribute__((address_space(3))) const ss_v4f ss_lds_cv4f;
__device__ __forceinline__ void lds_st2(uint32_t byte_addr, float2 v) {
  ss_v2f w;
  w.x = v.x;
  w.y = v.y;
  *reinterpret_cast<ss_lds_v2f *>((uintptr_t)byte_addr) = w;
}
typedef __attribute__((address_space(3))) float ss_lds_float;
__device__ __forceinline__ void lds_st1(uint32_t byte_addr, float v) {
  *reinterpret_cast<ss_lds_float *>((uintptr_t)byte_addr) = v;
}
__device__ __forceinline__ float4 lds_ld4(uint32_t byte_addr) {
  const ss_v4f v = *reinterpret_cast<ss_lds_cv4f *>((uintptr_t)byte_addr);
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float lds_table_a(uint32_t byte_off) {
  return *reinterpret_cast<ss_lds_cfloat *>((uintptr_t)byte_off);
}

// v if bit BIT of d is set, else +0: the sign-extended bit as an AND mask.  The
// empty asm keeps the compiler from turning the AND back into compare + select.
template <int BIT>
__device__ __forceinline__ float keep_if(float v, uint32_t d) {
  uint32_t m = (uint32_t)((int32_t)(d << (31 - BIT)) >> 31);
  asm("" : "+v"(m));
  return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, v) & m);
}

// Barrier for loops whose cross-wave traffic is LDS only: this wave's LDS
// accesses complete, then s_barrier.  Unlike __syncthreads() it does not wait
// for outstanding buffer loads and stores (history streams, receiver samples),
// which later launches or later loop iterations consume.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// floor(n / d) with magic = 0xffffffff / d + 1, exact whenever n * d < 2^32:
// n * magic / 2^32 exceeds n / d by less than n / 2^32 < 1/d, which is no more
// than the distance from n / d up to the next integer.
__device__ __forceinline__ uint32_t magic_div(uint32_t n, uint32_t magic) { return __umulhi(n, magic); }

// Pair descriptor of the paired kernels (32 bits, one per cell pair (2q, 2q+1)):
//   bit 0       Laplacian interior (both cells share it: the two-cell rim has even width)
//   bits 4-14   byte offset of the pair's 16-byte sponge entry in the pair table,
//               which starts at LDS address 0 (a lookup is one mask and one read)
//   bits 16-30  model pairs: byte offset of the pair's illumination in the model plane
//   bit 31      model pair (both cells are model cells: the model starts on an even
//               column and has even width), so one signed compare tests it
// With pad and the model width even, a pair starts on an even column and the
// sponge depth of its first cell, dx0, fixes the second's: 0 in the model
// columns, dx0 - 1 in the left sponge (dx0 even), dx0 + 1 in the right one
// (dx0 odd).  So the pair table has one entry per (dz, dx0), dz, dx0 in
// [0, pad], holding both cells' coefficients: (A0, B0, A1, B1) for the forward,
// (2A0, A0 B0, 2A1, A1 B1) for the adjoint (2A is exact and A = 0.5 * 2A).
// Model pairs sit at sponge depth 0, entry 0 (A = B = 1).
constexpr uint32_t kPInterior = 1u, kPTabMask = 0x7ff0u, kPIlMask = 0x7fffu;
__device__ __forceinline__ uint32_t pair_desc(int q, int H, int W, int nz, int nx, int pad) {
  const int c = 2 * q, z = c / W, x = c - z * W;
  uint32_t v = (uint32_t)(sponge_dist(z, H, pad) * (pad + 1) + sponge_dist(x, W, pad)) << 4;
  if (z >= 2 && z < H - 2 && x >= 2 && x < W - 2) v |= kPInterior;
  if (z >= pad && z < pad + nz && x >= pad && x < pad + nx)
    v |= 0x80000000u | ((uint32_t)(((z - pad) * nx + (x - pad)) * 4) << 16);
  return v;
}

// Pair table entry (dz, dx0) from the full A, B planes at a representative pair
// of that depth in the top sponge (row pad - dz): columns pad (dx0 = 0),
// pad - dx0 (dx0 even, left) or W - 1 - pad + dx0 (dx0 odd, right).
// mode 0: (A0, B0, A1, B1); mode 1: (2A0, A0 B0, 2A1, A1 B1); mode 2: as mode 1
// but zero for the pairs on the two-cell rim (dz or dx0 at least pad - 1).
__device__ inline void fill_pair_table(float4 *tab, const float *A, const float *B, int W, int pad, int mode) {
  const int n = (pad + 1) * (pad + 1);
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const int dz = k / (pad + 1), dx0 = k - dz * (pad + 1);
    const int x0 = dx0 == 0 ? pad : ((dx0 & 1) ? W - 1 - pad + dx0 : pad - dx0);
    const int c = (pad - dz) * W + x0;
    const float a0 = A[c], b0 = B[c], a1 = A[c + 1], b1 = B[c + 1];
    tab[k] = mode ? make_float4(2.0f * a0, a0 * b0, 2.0f * a1, a1 * b1) : make_float4(a0, b0, a1, b1);
    if (mode == 2 && (dz >= pad - 1 || dx0 >= pad - 1)) tab[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

// LDS floats of the stencil overhang (two rows and two cells) on each side of a
// plane of width W, and of a pair table for sponge width pad.  The table sits at
// LDS address 0, so a pair's entry is read at the byte offset its descriptor
// holds, with no base add.
constexpr int pair_guard(int W) { return 2 * W + 2; }
constexpr int pair_table_floats(int pad) { return (pad + 1) * (pad + 1) * 4; }

}  // namespace ss
