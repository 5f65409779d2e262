// Host check of ss::div12 (shotsweep_amd/csrc/ss_common.h), the division by 12
// used by the resident kernels: compared bitwise with IEEE division over every
// mantissa at three binades, the whole range below 2^-116 (subnormal results
// and the fallback boundary), the largest binades and the special values.
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

int main() {
  uint64_t bad = 0;
  const uint32_t mant = 1u << 23;
  bad += check_range(0, 11u * mant);             // exponent fields 0..10: below 2^-116
  bad += check_range(127u * mant, 130u * mant);  // [1, 8): every mantissa, three binades
  bad += check_range(250u * mant, 256u * mant);  // largest binades, inf and NaN
  std::printf("mismatches=%llu\n", (unsigned long long)bad);
  return bad != 0;
}
