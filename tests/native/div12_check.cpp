// Host check of ss::div12 (shotsweep_amd/csrc/ss_common.h), the division by 12
// used by the resident kernels: compared bitwise with IEEE division over every
// mantissa at three binades, the whole range below 2^-116 (subnormal results
// and the fallback boundary), the largest binades and the special values; and
// ss::div12x2 (two values, one branch) against div12 on mixed pairs.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "ss_common.h"

static uint64_t check_range(uint32_t lo, uint32_t hi) {
  volatile float twelve = 12.0f;
  uint64_t bad = 0;
  for (uint64_t u = lo; u < hi; ++u) {
    for (uint32_t sign = 0; sign < 2; ++sign) {
      uint32_t b = (uint32_t)u | (sign << 31);
      float x;
      std::memcpy(&x, &b, 4);
      float e = x / twelve, c = ss::div12(x);
      uint32_t be, bc;
      std::memcpy(&be, &e, 4);
      std::memcpy(&bc, &c, 4);
      if (be != bc && !(std::isnan(e) && std::isnan(c))) {
        if (bad < 5) std::printf("mismatch x=%a ieee=%a trick=%a\n", x, e, c);
        ++bad;
      }
    }
  }
  return bad;
}

// div12x2 against div12 on pairs mixing fast-path and fallback values.
static uint64_t check_pairs() {
  const float vals[] = {0.0f, -0.0f, 1.0f, -3.5f, 0x1p-119f, 0x1.fffffep-120f, 0x1p-140f, -0x1p-149f, 0x1p-126f,
                        1e30f, 0x1.fffffep+127f, INFINITY, -INFINITY, NAN, 12.0f, 7.25e-20f};
  uint64_t bad = 0;
  for (float a : vals)
    for (float b : vals) {
      const float2 q = ss::div12x2(a, b);
      const float ea = ss::div12(a), eb = ss::div12(b);
      uint32_t x0, x1, y0, y1;
      std::memcpy(&x0, &q.x, 4);
      std::memcpy(&x1, &q.y, 4);
      std::memcpy(&y0, &ea, 4);
      std::memcpy(&y1, &eb, 4);
      const bool nan_ok = std::isnan(q.x) && std::isnan(ea) && std::isnan(q.y) == std::isnan(eb);
      if ((x0 != y0 && !(std::isnan(q.x) && std::isnan(ea))) || (x1 != y1 && !(std::isnan(q.y) && std::isnan(eb)))) {
        if (bad < 5 && !nan_ok) std::printf("pair mismatch a=%a b=%a\n", a, b);
        ++bad;
      }
    }
  return bad;
}

int main() {
  uint64_t bad = check_pairs();
  const uint32_t mant = 1u << 23;
  bad += check_range(0, 11u * mant);             // exponent fields 0..10: below 2^-116
  bad += check_range(127u * mant, 130u * mant);  // [1, 8): every mantissa, three binades
  bad += check_range(250u * mant, 256u * mant);  // largest binades, inf and NaN
  std::printf("mismatches=%llu\n", (unsigned long long)bad);
  return bad != 0;
}
