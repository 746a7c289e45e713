// Data-parallel loops over index ranges on host threads, for the Mesh and
// Domain construction code.
#pragma once

#include <algorithm>
#include <cstddef>
#include <exception>
#include <memory>
#include <new>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

namespace fem2d {

// threads used when a caller passes 0: the machine's cores, at most 16
inline int default_threads() {
  return int(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
}

// f(begin, end) over [0, n), split into contiguous ranges of at least `grain`
// indices on up to `threads` threads. If ranges throw, the exception of the
// lowest failing range is rethrown after every range has ended, so a loop
// whose iterations fail independently reports the failure a sequential loop
// would have met first.
template <class F>
void parallel_ranges(size_t n, int threads, size_t grain, F&& f) {
  const size_t nt = std::min<size_t>(size_t(std::max(1, threads)), (n + grain - 1) / std::max<size_t>(1, grain));
  if (nt <= 1) {
    if (n) f(size_t(0), n);
    return;
  }
  std::vector<std::exception_ptr> err(nt);
  std::vector<std::thread> pool;
  pool.reserve(nt - 1);
  auto run = [&](size_t k) {
    try {
      f(n * k / nt, n * (k + 1) / nt);
    } catch (...) {
      err[k] = std::current_exception();
    }
  };
  for (size_t k = 1; k < nt; ++k) pool.emplace_back(run, k);
  run(0);
  for (auto& t : pool) t.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

// An allocator that default-initialises: resize() of a vector of a trivially
// default-constructible type leaves the new elements unwritten, so a parallel
// loop, not the resizing thread, is the first to touch (and page in) them.
template <class T>
struct default_init_allocator : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = default_init_allocator<U>;
  };
  default_init_allocator() = default;
  template <class U>
  default_init_allocator(const default_init_allocator<U>&) noexcept {}
  template <class U>
  void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) {
    ::new (static_cast<void*>(p)) U;
  }
  template <class U, class... Args>
  void construct(U* p, Args&&... args) {
    ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
  }
};

template <class T>
using uninit_vector = std::vector<T, default_init_allocator<T>>;

// dst = src element by element, in parallel ranges
template <class T>
void parallel_copy(std::vector<T>& dst, const std::vector<T>& src, int threads) {
  dst.clear();
  dst.resize(src.size());
  parallel_ranges(src.size(), threads, 1024, [&](size_t b, size_t e) {
    for (size_t k = b; k < e; ++k) dst[k] = src[k];
  });
}

}  // namespace fem2d
