// HostPool (ngap_amd/csrc/host_pool.h) on the CPU: every part of a range runs
// exactly once, parts tile the range, copies are exact, and runs back to back
// (the pool's helpers picking up a new job while stragglers of the last one
// find nothing left) stay exact.  Prints PASS or the first failure.
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../ngap_amd/csrc/host_pool.h"

int main() {
  using ngapamd::HostPool;
  HostPool pool;
  std::mt19937_64 rng(7);
  const size_t sizes[] = {0, 1, 4095, 4096, HostPool::kPart - 1, HostPool::kPart,
                          2 * HostPool::kPart + 17, 3u << 20, (8u << 20) + 5, 33u << 20};
  for (int rep = 0; rep < 8; rep++)
    for (size_t n : sizes) {
      std::vector<uint8_t> src(n), dst(n, 0);
      for (auto &x : src) x = (uint8_t)rng();
      std::vector<std::atomic<int>> hits(n ? (n + 4095) / 4096 : 1);
      for (auto &h : hits) h = 0;
      std::atomic<size_t> covered{0};
      pool.run_parts(n, [&](size_t, size_t a, size_t b) {
        if (a % 4096 || b > n || a >= b) {
          std::printf("FAIL bad part [%zu, %zu) of %zu\n", a, b, n);
          std::exit(1);
        }
        for (size_t p = a; p < b; p += 4096) hits[p / 4096]++;
        covered += b - a;
      });
      if (covered != n) {
        std::printf("FAIL covered %zu of %zu\n", covered.load(), n);
        return 1;
      }
      for (size_t i = 0; n && i < hits.size(); i++)
        if (hits[i] != 1) {
          std::printf("FAIL block %zu ran %d times (n %zu)\n", i, hits[i].load(), n);
          return 1;
        }
      pool.copy(dst.data(), src.data(), n);
      if (n && std::memcmp(dst.data(), src.data(), n)) {
        std::printf("FAIL copy of %zu bytes\n", n);
        return 1;
      }
    }
  std::printf("PASS width %zu\n", pool.width());
  return 0;
}
