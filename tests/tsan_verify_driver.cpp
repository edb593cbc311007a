// Host-only driver of the ThreadSanitizer test: six threads call
// fs_verify_factors at once, 200 times each, on a list of real factors (read
// from stdin as "p nc" lines) and on the same list with one sign flipped; every
// call must return what a lone caller gets (-1, or the flipped factor's index).
// Beside them, two threads build setup programs and power tables for changing
// nmin (the shared prime-list cache), checked against a lone caller's counts.
#include <stdint.h>
#include <stdio.h>
#include <atomic>
#include <thread>
#include <vector>
#include "factsieve.h"

int main() {
  std::vector<fs_factor> good;
  unsigned long long p;
  int nc;
  while (scanf("%llu %d", &p, &nc) == 2) good.push_back(fs_factor{(uint64_t)p, (int32_t)nc, 0});
  if (good.size() < 8) {
    fprintf(stderr, "need at least 8 factors on stdin\n");
    return 2;
  }
  std::vector<fs_factor> bad = good;
  bad[7].nc = -bad[7].nc;
  std::atomic<int> wrong{0};
  std::vector<std::thread> threads;
  for (int t = 0; t < 6; ++t)
    threads.emplace_back([&, t] {
      for (int k = 0; k < 200; ++k) {
        const bool b = ((t + k) & 1) != 0;
        const std::vector<fs_factor>& l = b ? bad : good;
        if (fs_verify_factors(l.data(), l.size(), 1 + (t + k) % 8) != (b ? 7 : -1)) wrong++;
      }
    });
  const uint32_t nmins[] = {1000, 5000, 20000, 1000, 77777, 5000};
  std::vector<int64_t> want;
  for (uint32_t n : nmins) {
    want.push_back(fs_setup_program(n, FS_PATH_FP64, nullptr));
    want.push_back(fs_setup_program(n, FS_PATH_HYBRID, nullptr));
    want.push_back(fs_power_table(n, nullptr, nullptr, nullptr));
  }
  for (int t = 0; t < 2; ++t)
    threads.emplace_back([&, t] {
      for (int k = 0; k < 30; ++k) {
        const size_t i = (size_t)(t + k) % 6;
        const uint32_t n = nmins[i];
        if (fs_setup_program(n, FS_PATH_FP64, nullptr) != want[3 * i] ||
            fs_setup_program(n, FS_PATH_HYBRID, nullptr) != want[3 * i + 1] ||
            fs_power_table(n, nullptr, nullptr, nullptr) != want[3 * i + 2])
          wrong++;
      }
    });
  for (auto& th : threads) th.join();
  printf("wrong answers: %d\n", wrong.load());
  return wrong.load() ? 1 : 0;
}
