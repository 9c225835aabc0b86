// BatchStages: problem tables of several persistent batch launches in one upload.
#include "batch_stages.h"

#include <cstring>

#include "workspace.h"

namespace psmma {

int batch_tiles(const std::vector<BatchProblem>& in, int big_min, std::vector<BatchProblem>& out,
                std::vector<int>& tbeg, int kset);
hipError_t launch_batch_f64(const BatchProblem* dprobs, const int* dtbeg, int nprob, int ntiles,
                            unsigned* ctr, int nworkgroups, hipStream_t s, int kset);

int BatchStages::upload(const std::vector<std::vector<BatchProblem>>& stages,
                        const std::vector<int>& ksets, int big_min, int slot, hipStream_t s) {
  const size_t ns = stages.size();
  std::vector<std::vector<BatchProblem>> probs(ns);
  std::vector<std::vector<int>> tbeg(ns);
  ntiles_.assign(ns, 0);
  nprob_.assign(ns, 0);
  kset_.assign(ns, BKSET_HSS);
  offp_.assign(ns, 0);
  offt_.assign(ns, 0);
  offc_.assign(ns, 0);
  s_ = s;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t bytes = 0;
  for (size_t i = 0; i < ns; ++i) {
    if (i < ksets.size()) kset_[i] = ksets[i];
    ntiles_[i] = batch_tiles(stages[i], big_min, probs[i], tbeg[i], kset_[i]);
    nprob_[i] = (int)probs[i].size();
    bytes += al(sizeof(BatchProblem) * probs[i].size()) + al(sizeof(int) * tbeg[i].size()) +
             al(2 * sizeof(unsigned));
  }
  host_.assign(bytes, 0);
  char* host = host_.data();
  size_t o = 0;
  for (size_t i = 0; i < ns; ++i) {
    offp_[i] = o;
    if (!probs[i].empty())
      memcpy(host + o, probs[i].data(), sizeof(BatchProblem) * probs[i].size());
    o += al(sizeof(BatchProblem) * probs[i].size());
    offt_[i] = o;
    if (!tbeg[i].empty()) memcpy(host + o, tbeg[i].data(), sizeof(int) * tbeg[i].size());
    o += al(sizeof(int) * tbeg[i].size());
    offc_[i] = o;  // the two work counters of the launch: zero
    o += al(2 * sizeof(unsigned));
  }
  d_ = (char*)ws_lease(slot, bytes, s, lk_);
  if (!d_) return 1002;
  hipError_t e = hipMemcpyAsync(d_, host, bytes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return 1000 + (int)e;
  int dev = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev);
  return 0;
}

int BatchStages::launch(size_t i) {
  if (i >= ntiles_.size() || ntiles_[i] == 0) return 0;
  // the leaf set runs one 512-thread workgroup per CU, the others two per CU
  const int nwg = kset_[i] == BKSET_LEAF ? cus_ : 2 * cus_;
  hipError_t e = launch_batch_f64((const BatchProblem*)(d_ + offp_[i]), (const int*)(d_ + offt_[i]),
                                  nprob_[i], ntiles_[i], (unsigned*)(d_ + offc_[i]), nwg, s_,
                                  kset_[i]);
  return e == hipSuccess ? 0 : 1000 + (int)e;
}

int BatchStages::finish() {
  hipError_t e = hipStreamSynchronize(s_);
  if (lk_.owns_lock()) lk_.unlock();
  return e == hipSuccess ? 0 : 1000 + (int)e;
}

}  // namespace psmma
