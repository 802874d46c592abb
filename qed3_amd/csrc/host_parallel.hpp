// Minimal host thread fan-out for independent index ranges.
#pragma once
#include <algorithm>
#include <thread>
#include <vector>

namespace qed3 {

template <class F>
void parallel_ranges(size_t n, F&& f) {  // f(begin, end)
  const unsigned nt = std::max(1u, std::min(std::thread::hardware_concurrency(), 64u));
  if (n < 8192 || nt == 1) {
    f(size_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  const size_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; t++) {
    const size_t a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([a, b, &f] { f(a, b); });
  }
  for (auto& t : th) t.join();
}

}  // namespace qed3
