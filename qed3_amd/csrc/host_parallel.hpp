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

// f(task) for ntasks heavy, independent tasks (e.g. one time slice each): one thread per
// task up to the core count, with no small-size cutoff.
template <class F>
void parallel_tasks(size_t ntasks, F&& f) {
  const size_t nt = std::max<size_t>(1, std::min<size_t>(std::thread::hardware_concurrency(), 64));
  if (ntasks <= 1 || nt == 1) {
    for (size_t i = 0; i < ntasks; i++) f(i);
    return;
  }
  std::vector<std::thread> th;
  const size_t nthreads = std::min(nt, ntasks);
  for (size_t t = 0; t < nthreads; t++)
    th.emplace_back([t, nthreads, ntasks, &f] {
      for (size_t i = t; i < ntasks; i += nthreads) f(i);
    });
  for (auto& t : th) t.join();
}

}  // namespace qed3
