// Host-only driver of the ThreadSanitizer test: six threads call
// fs_verify_factors at once, 200 times each, on a list of real factors (read
// from stdin as "p nc" lines) and on the same list with one sign flipped; every
// call must return what a lone caller gets (-1, or the flipped factor's index).
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
  for (auto& th : threads) th.join();
  printf("wrong answers: %d\n", wrong.load());
  return wrong.load() ? 1 : 0;
}
