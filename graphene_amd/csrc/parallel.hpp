// Host worker threads for the setup pipelines (matrix layout, stream
// packing). parallelFor(n, grain, f) calls f(i) for every i in [0, n), taking
// chunks of `grain` indices at a time on up to hostThreads() threads; the
// calling thread works too. The first exception thrown by f is rethrown on
// the calling thread once every worker has stopped.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace gph {

// Worker count: the hardware threads, or OMP_NUM_THREADS when set, at most 16.
inline int hostThreads() {
  static const int n = [] {
    int t = (int)std::thread::hardware_concurrency();
    if (const char* e = std::getenv("OMP_NUM_THREADS")) {
      const int v = std::atoi(e);
      if (v > 0) t = v;
    }
    return std::max(1, std::min(t, 16));
  }();
  return n;
}

template <typename F>
void parallelFor(int64_t n, int64_t grain, F&& f) {
  if (n <= 0) return;
  if (grain < 1) grain = 1;
  const int nt = (int)std::min<int64_t>(hostThreads(), (n + grain - 1) / grain);
  if (nt <= 1) {
    for (int64_t i = 0; i < n; i++) f(i);
    return;
  }
  std::atomic<int64_t> next{0};
  std::exception_ptr err;
  std::mutex em;
  auto work = [&] {
    try {
      for (;;) {
        const int64_t b = next.fetch_add(grain);
        if (b >= n) break;
        const int64_t e = std::min(n, b + grain);
        for (int64_t i = b; i < e; i++) f(i);
      }
    } catch (...) {
      std::lock_guard<std::mutex> lk(em);
      if (!err) err = std::current_exception();
      next.store(n);
    }
  };
  std::vector<std::thread> pool;
  pool.reserve((size_t)nt - 1);
  for (int t = 1; t < nt; t++) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  if (err) std::rethrow_exception(err);
}

}  // namespace gph
