// Grow-only device workspaces keyed by (device, stream, slot).
//
// A call that needs scratch memory takes a lease: the entry's mutex is held
// from the lease until the caller has enqueued every kernel that uses the
// memory, so
//   * two calls on different streams never share memory (different keys), and
//   * a call on the same stream that needs more memory cannot free the buffer
//     while another thread is still enqueueing work on it; growth waits for
//     the device to drain before freeing the old buffer (a destroyed stream's
//     handle can be reused while its last kernels still run).
// Calls on one stream are ordered by the stream itself.
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "workspace.h"

namespace psmma {

namespace {
struct WsEntry {
  std::mutex mu;
  void* p = nullptr;
  size_t cap = 0;
};
std::mutex g_map_mu;
std::map<std::tuple<int, hipStream_t, int>, std::unique_ptr<WsEntry>> g_map;
}  // namespace

void* ws_lease(int slot, size_t bytes, hipStream_t s, std::unique_lock<std::mutex>& lk) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  WsEntry* e;
  {
    std::lock_guard<std::mutex> g(g_map_mu);
    auto& slotp = g_map[std::make_tuple(dev, s, slot)];
    if (!slotp) slotp.reset(new WsEntry);
    e = slotp.get();
  }
  lk = std::unique_lock<std::mutex>(e->mu);
  if (bytes == 0) bytes = 256;
  if (e->cap < bytes) {
    if (e->p) {
      hipDeviceSynchronize();
      hipFree(e->p);
    }
    e->p = nullptr;
    e->cap = 0;
    if (hipMalloc(&e->p, bytes) != hipSuccess) {
      lk.unlock();
      return nullptr;
    }
    e->cap = bytes;
  }
  return e->p;
}

}  // namespace psmma
