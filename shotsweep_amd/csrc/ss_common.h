// Shared device helpers for the shotsweep HIP kernels (gfx950).
//
// Arithmetic contract: every helper reproduces the operation order of
// openfwi/physics/gpu.py in the working type.  The library is compiled with
// -ffp-contract=off, so a product and a sum that the reference rounds
// separately stay separate; an explicit fma is used only where the product is
// exact (multiplication by 2 or 16), which makes the fused form bit-identical.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "shotsweep.h"

namespace ss {

// Device-side copy of the batch geometry, passed by value to kernels.
struct Geom {
  int nb, ns, nz, nx, pad, nt, row;
  int H, W;  // padded grid
  int src[SS_MAX_SHOTS];
};

void set_error(const std::string &msg);
int check_geom(const ss_geom *g, Geom *out);
int hip_status(hipError_t e, const char *what);

template <typename T>
__host__ __device__ __forceinline__ T fma_exact(T a, T b, T c);
template <>
__host__ __device__ __forceinline__ float fma_exact<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
template <>
__host__ __device__ __forceinline__ double fma_exact<double>(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}

// lap (gpu.py:33-37) at one interior cell from its nine taps.
// zm2..zp2 are the column taps (rows -2..+2), xm2..xp2 the row taps.
// Order: -zm2 + 16 zm1 - 30 c + 16 zp1 - zp2 - xm2 + 16 xm1 - 30 c + 16 xp1 - xp2, then / 12.
// 16*x is exact, so "s + 16*x" equals fma(16, x, s) bit for bit.
template <typename T>
__host__ __device__ __forceinline__ T lap_sum(T zm2, T zm1, T c, T zp1, T zp2, T xm2, T xm1, T xp1, T xp2) {
  const T c30 = T(30) * c;
  T s = fma_exact(T(16), zm1, -zm2);
  s = s - c30;
  s = fma_exact(T(16), zp1, s);
  s = s - zp2;
  s = s - xm2;
  s = fma_exact(T(16), xm1, s);
  s = s - c30;
  s = fma_exact(T(16), xp1, s);
  s = s - xp2;
  return s;
}

// lap_adjoint (gpu.py:40-53) gathered at one cell: the reference scatters
// -60 q(c), then -q, +16q, +16q, -q along z (destination shifts -2,-1,+1,+2),
// then the same along x.  Gathered, destination shift d reads q at c - d.
// q is zero outside the interior, so absent taps contribute an exact +-0.
// The reference's first term is 0 - 60 q (it subtracts from a zero buffer);
// -60 q is the same number for every q except q = +0, where it is -0 instead
// of +0.  A zero's sign never changes a nonzero sum or product, so every
// nonzero value downstream is the same bit for bit, and zeros compare equal.
template <typename T>
__host__ __device__ __forceinline__ T lapt_sum(T qc, T qzp2, T qzp1, T qzm1, T qzm2, T qxp2, T qxp1, T qxm1,
                                               T qxm2) {
  T s = T(-60) * qc;
  s = s - qzp2;
  s = fma_exact(T(16), qzp1, s);
  s = fma_exact(T(16), qzm1, s);
  s = s - qzm2;
  s = s - qxp2;
  s = fma_exact(T(16), qxp1, s);
  s = fma_exact(T(16), qxm1, s);
  s = s - qxm2;
  return s;
}

// Correctly rounded s/12 from the reciprocal product and one FMA correction.
// Whenever that result is zero or normal it equals IEEE division.  When it is
// subnormal (|s| below about 2^-122) it can be wrong only at an exact tie of
// the subnormal grid, s/12 = (n + 1/2) 2^-149, which the FMA rounds to the
// neighbour the correction's sign points at rather than to the even one; for
// infinite s it is NaN.  One class test on the result sends those cases to
// div12_fix (checked on the host for all 2^32 inputs by
// tests/native/div12_check.cpp).  Exact zeros, common ahead of the wavefront,
// stay on the fast path.
constexpr int kZeroOrNormal = 0x168;  // __builtin_isfpclass: -normal, -zero, +zero, +normal

// The remainder is formed as -(12 q0 - s) rather than s - 12 q0: the same value
// for every s except zeros, where the negated form carries the sign that makes
// the final FMA return s's signed zero (no separate copysign).
__host__ __device__ __forceinline__ float div12_fast(float s) {
  const float R = 0x1.555556p-4f;  // RN(1/12)
  const float q0 = s * R;
  const float r = -__builtin_fmaf(q0, 12.0f, -s);
  return __builtin_fmaf(r, R, q0);
}

// q = div12_fast(s) when it is neither zero nor normal.  Subnormal q: d = 12 q - s
// is exact (both are multiples of 2^-149 and d is small); d = +-6 * 2^-149 marks
// an exact tie, and an odd q then moves one grid step towards s/12, to its even
// neighbour: q - d * RN(1/6) is q -+ 2^-149 plus a perturbation far below half a
// grid step, so the FMA rounds it to q -+ 2^-149 exactly.  Non-finite q (infinite
// or NaN s): s * RN(1/12) is +-inf or NaN, as IEEE division gives.
__host__ __device__ __forceinline__ float div12_fix(float s, float q) {
  const float d = __builtin_fmaf(q, 12.0f, -s);
  const bool tie = __builtin_fabsf(d) == 0x1.8p-147f && (__builtin_bit_cast(uint32_t, q) & 1u) != 0u;
  q = tie ? __builtin_fmaf(-d, 0x1.555556p-3f, q) : q;
  if (q != q) q = s * 0x1.555556p-4f;
  return q;
}

__host__ __device__ __forceinline__ float div12(float s) {
  float q = div12_fast(s);
  if (!__builtin_isfpclass(q, kZeroOrNormal)) q = div12_fix(s, q);
  return q;
}

// div12 of two values with one branch: both keep the fast result unless either
// needs the fix, and then each is fixed only if it needs it.  Results equal
// div12(s0), div12(s1).
__host__ __device__ __forceinline__ float2 div12x2(float s0, float s1) {
  float2 q = make_float2(div12_fast(s0), div12_fast(s1));
  const bool ok0 = __builtin_isfpclass(q.x, kZeroOrNormal), ok1 = __builtin_isfpclass(q.y, kZeroOrNormal);
  if (!(ok0 && ok1)) {
    if (!ok0) q.x = div12_fix(s0, q.x);
    if (!ok1) q.y = div12_fix(s1, q.y);
  }
  return q;
}

// Damped leapfrog update (gpu.py:113): ((2 p1 - B p0) + c2 L) A.
// 2*p1 is exact, so 2 p1 - (B p0) equals fma(2, p1, -(B p0)).
template <typename T>
__host__ __device__ __forceinline__ T leapfrog(T p1, T p0, T B, T c2, T L, T A) {
  T t = fma_exact(T(2), p1, -(B * p0));
  t = t + c2 * L;
  return t * A;
}

// Adjoint update (gpu.py:179): (2A) m1 + lapT - (A B) m2, with 2A exact.
template <typename T>
__host__ __device__ __forceinline__ T adjoint_mu(T A, T B, T m1, T lapt, T m2) {
  T t = (T(2) * A) * m1;
  t = t + lapt;
  return t - (A * B) * m2;
}

__host__ __device__ __forceinline__ int64_t plane_index(int b, int s, int ns) { return (int64_t)b * ns + s; }

}  // namespace ss
