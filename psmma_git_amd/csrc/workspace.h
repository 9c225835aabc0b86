// Device workspaces of the library's multi-kernel products (see workspace.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>

namespace psmma {

// Workspace slots (one grow-only buffer per device, stream and slot).
enum WsSlot : int {
  WS_SPLIT = 0,     // split-precision products (single and batched eager calls)
  WS_BLR_FIXED = 1, // compressed products: block lists and compressor outputs
  WS_BLR_GEMM = 2,  // compressed products: rank-dependent GEMM operands
  WS_RING = 3,      // ring steps and redistributions: index lists, offset tables
  WS_TUNE = 4,      // split-depth measurement: sampled FP64 and split results
  WS_HSS_TABLES = 5,  // HSS stages: problem tables of the persistent launches
  WS_HSS_WORK = 6,    // HSS product: the up- and down-pass matrices G and F
  WS_BLR_TABLES = 7,  // compressed products: problem tables of the batch launches
};

// Returns a buffer of at least `bytes` for (current device, s, slot), or nullptr
// when the allocation fails.  On success `lk` holds the entry until the caller
// releases it: keep it until every kernel using the buffer has been enqueued.
void* ws_lease(int slot, size_t bytes, hipStream_t s, std::unique_lock<std::mutex>& lk);

}  // namespace psmma
