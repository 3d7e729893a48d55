// pool.hip — the device memory pool of DeviceBuffer (host code only).
#include <algorithm>

#include "internal.h"

namespace fpmhip {

size_t pool_size_class(size_t n) {
  size_t p = 1;
  while (p <= n / 2) p *= 2;  // largest power of two <= n
  const size_t g = std::max<size_t>(256, p / 8);
  return (n + g - 1) / g * g;
}

DevicePool*& thread_pool() {
  thread_local DevicePool* pool = nullptr;
  return pool;
}

DevicePool::~DevicePool() {
  std::lock_guard<std::mutex> lock(mu_);
  for (auto& kv : idle_)
    for (void* p : kv.second) (void)hipFree(p);
  for (auto& b : pending_) (void)hipFree(b.first);
  idle_.clear();
  pending_.clear();
  idle_bytes_ = 0;
}

void DevicePool::free_idle_locked(size_t keep) {
  // largest classes first: the fewest hipFree calls for the bytes returned
  for (auto it = idle_.rbegin(); it != idle_.rend() && idle_bytes_ > keep; ++it) {
    std::vector<void*>& blocks = it->second;
    while (!blocks.empty() && idle_bytes_ > keep) {
      (void)hipFree(blocks.back());
      blocks.pop_back();
      idle_bytes_ -= it->first;
    }
  }
}

void* DevicePool::acquire(size_t bytes, size_t* cls) {
  const size_t c = pool_size_class(bytes);
  *cls = c;
  {
    std::lock_guard<std::mutex> lock(mu_);
    auto it = idle_.find(c);
    if (it != idle_.end() && !it->second.empty()) {
      void* p = it->second.back();
      it->second.pop_back();
      idle_bytes_ -= c;
      reuses_++;
      return p;
    }
  }
  void* p = nullptr;
  if (hipMalloc(&p, c) != hipSuccess) {
    (void)hipGetLastError();
    {
      std::lock_guard<std::mutex> lock(mu_);
      free_idle_locked(0);
    }
    HIP_CHECK(hipMalloc(&p, c));
  }
  allocations_++;
  return p;
}

void DevicePool::give_back(void* p, size_t cls) {
  std::lock_guard<std::mutex> lock(mu_);
  pending_.emplace_back(p, cls);
}

void DevicePool::recycle() {
  std::lock_guard<std::mutex> lock(mu_);
  for (auto& b : pending_) {
    idle_[b.second].push_back(b.first);
    idle_bytes_ += b.second;
  }
  pending_.clear();
  free_idle_locked(max_cached_);
}

}  // namespace fpmhip
