// Checks lrtd::FastDiv (litert_amd/csrc/kernels/fastdiv.hpp) against the
// integer division of the host compiler: divisors at every power of two and
// its neighbours plus random ones, numerators at 0, near every multiple of the
// divisor close to the top of the range, at 2^32 - 1 and random.
#define __host__
#define __device__
#include <cstdio>
#include <random>
#include <vector>
#include "../../litert_amd/csrc/kernels/fastdiv.hpp"

static bool check(uint32_t d, uint32_t n) {
  const lrtd::FastDiv f = lrtd::make_fastdiv(d);
  if (f.div(n) != n / d) {
    std::printf("FAIL d=%u n=%u got %u want %u\n", d, n, f.div(n), n / d);
    return false;
  }
  return true;
}

int main() {
  std::mt19937_64 g(12345);
  std::vector<uint32_t> ds;
  for (int k = 0; k < 32; k++) {
    const uint64_t p = 1ull << k;
    for (int64_t o = -2; o <= 2; o++)
      if ((int64_t)p + o >= 1 && (int64_t)p + o <= 0xFFFFFFFFll) ds.push_back((uint32_t)(p + o));
  }
  ds.push_back(0xFFFFFFFFu);
  for (int i = 0; i < 3000; i++) ds.push_back((uint32_t)(g() >> (32 + g() % 32)) | 1u);
  for (uint32_t d : ds) {
    if (d == 0) continue;
    const uint64_t top = 0xFFFFFFFFull;
    std::vector<uint64_t> ns = {0, 1, d - 1ull, d, d + 1ull, 2ull * d - 1, 2ull * d, top, top - 1};
    const uint64_t kmax = top / d;
    for (uint64_t k = kmax > 3 ? kmax - 3 : 0; k <= kmax; k++)
      for (int64_t o = -1; o <= 1; o++) ns.push_back(k * d + o);
    for (int i = 0; i < 2000; i++) ns.push_back(g() & top);
    for (uint64_t n : ns)
      if (n <= top && !check(d, (uint32_t)n)) return 1;
  }
  for (uint32_t d = 1; d <= 300; d++)  // every numerator below 2^20 for small divisors
    for (uint32_t n = 0; n < (1u << 20); n += 7)
      if (!check(d, n)) return 1;
  std::printf("ok %zu divisors\n", ds.size());
  return 0;
}
