#include "scratch.hpp"

#include <mutex>
#include <thread>
#include <vector>

namespace momms {

hipError_t stream_scratch(hipStream_t st, size_t bytes, void** out, int slot) {
  struct Entry {
    int dev;
    hipStream_t st;
    std::thread::id thread;
    int slot;
    void* p;
    size_t bytes;
  };
  static std::mutex mu;
  static std::vector<Entry> pool;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const std::thread::id self = std::this_thread::get_id();
  std::lock_guard<std::mutex> lock(mu);
  for (Entry& x : pool) {
    if (x.dev != dev || x.st != st || x.thread != self || x.slot != slot) continue;
    if (x.bytes < bytes) {
      if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
      if ((e = hipFree(x.p)) != hipSuccess) return e;
      x.p = nullptr;
      x.bytes = 0;
      if ((e = hipMalloc(&x.p, bytes)) != hipSuccess) return e;
      x.bytes = bytes;
    }
    *out = x.p;
    return hipSuccess;
  }
  void* p = nullptr;
  if ((e = hipMalloc(&p, bytes)) != hipSuccess) return e;
  pool.push_back({dev, st, self, slot, p, bytes});
  *out = p;
  return hipSuccess;
}

}  // namespace momms
