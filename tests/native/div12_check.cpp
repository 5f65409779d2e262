// Host check of ss::div12 (shotsweep_amd/csrc/ss_common.h), the division by 12
// of the FD kernels: compared bitwise with IEEE division for every one of the
// 2^32 float inputs (NaN results compare as equal whatever their payload), and
// ss::div12x2 (two values, one branch) against div12 on mixed pairs.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "ss_common.h"

static uint64_t check_range(uint64_t lo, uint64_t hi) {
  volatile float twelve = 12.0f;
  uint64_t bad = 0;
  for (uint64_t u = lo; u < hi; ++u) {
    const uint32_t b = (uint32_t)u;
    float x;
    std::memcpy(&x, &b, 4);
    const float e = x / twelve, c = ss::div12(x);
    uint32_t be, bc;
    std::memcpy(&be, &e, 4);
    std::memcpy(&bc, &c, 4);
    if (be != bc && !(std::isnan(e) && std::isnan(c))) {
      if (bad < 5) std::printf("mismatch x=%a ieee=%a fast=%a\n", x, e, c);
      ++bad;
    }
  }
  return bad;
}

// div12x2 against div12 on pairs mixing fast-path and fallback values.
static uint64_t check_pairs() {
  const float vals[] = {0.0f,        -0.0f,      1.0f,       -3.5f,     0x1p-119f, 0x1.fffffep-120f, 0x1p-140f,
                        -0x1p-149f,  0x1p-126f,  0x1p-122f,  0x1p-123f, 1e30f,     0x1.fffffep+127f, INFINITY,
                        -INFINITY,   NAN,        12.0f,      7.25e-20f,
                        0x1.8p-147f, -0x1.2p-145f, 0x1.ep-145f, -0x1.5p-144f};  // exact subnormal ties
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
      if ((x0 != y0 && !(std::isnan(q.x) && std::isnan(ea))) || (x1 != y1 && !(std::isnan(q.y) && std::isnan(eb)))) {
        if (bad < 5) std::printf("pair mismatch a=%a b=%a\n", a, b);
        ++bad;
      }
    }
  return bad;
}

int main(int argc, char **argv) {
  uint64_t bad = check_pairs();
  const int nthreads = argc > 1 ? std::atoi(argv[1]) : 8;
  const uint64_t total = 1ull << 32, chunk = total / nthreads;
  std::vector<uint64_t> part(nthreads, 0);
  std::vector<std::thread> pool;
  for (int i = 0; i < nthreads; ++i)
    pool.emplace_back([&, i] { part[i] = check_range(i * chunk, i + 1 == nthreads ? total : (i + 1) * chunk); });
  for (auto &t : pool) t.join();
  for (uint64_t p : part) bad += p;
  std::printf("mismatches=%llu\n", (unsigned long long)bad);
  return bad != 0;
}
