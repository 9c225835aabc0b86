// Several launches of the persistent FP64 batch kernel (sgemm_f64_batch_kernel)
// whose problem tables reach the device in one copy: the multi-stage products
// (nested HSS, the compressed one-level products) enqueue their stages between
// kernels of their own.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "psmma_common.h"

namespace psmma {

// Kind sets of the batch kernel: the public kinds, the HSS kinds (Cauchy kinds,
// their transposes, explicit matrices), and the leaf set (generated Cauchy-kind
// blocks with an explicit tail A2 * H2 on 256 x 128 tiles).
enum BatchKset : int { BKSET_PUBLIC = 0, BKSET_HSS = 1, BKSET_LEAF = 2 };

class BatchStages {
 public:
  BatchStages() = default;
  BatchStages(const BatchStages&) = delete;
  BatchStages& operator=(const BatchStages&) = delete;
  // A caller that returns early (an error after upload) still must not free the
  // tables' host copy or the lease while the stream may read them.
  ~BatchStages() {
    if (lk_.owns_lock()) (void)hipStreamSynchronize(s_);
  }
  // Tiles every stage (problems with min(m, n) >= big_min take the large tile
  // class of their set) and copies all tables to a buffer leased from `slot` of
  // the stream's workspaces.  Returns 0, 1002 (no memory) or 1000 + hipError.
  int upload(const std::vector<std::vector<BatchProblem>>& stages, const std::vector<int>& ksets,
             int big_min, int slot, hipStream_t s);
  // Enqueues stage i (nothing for an empty stage).  Returns 0 or 1000 + hipError.
  int launch(size_t i);
  // Synchronises the stream: the tables stay leased until the launches are done.
  int finish();
  // Stage i's device problem table (after upload) and its problem count, for a
  // caller that runs that stage on a kernel of its own.
  const BatchProblem* table(size_t i) const { return (const BatchProblem*)(d_ + offp_[i]); }
  int problems(size_t i) const { return nprob_[i]; }
  hipStream_t stream() const { return s_; }

 private:
  std::vector<int> ntiles_, nprob_, kset_;
  std::vector<size_t> offp_, offt_, offc_;
  std::vector<char> host_;  // the tables' host copy, kept until finish()
  char* d_ = nullptr;
  hipStream_t s_ = nullptr;
  int cus_ = 256;
  std::unique_lock<std::mutex> lk_;
};

}  // namespace psmma
