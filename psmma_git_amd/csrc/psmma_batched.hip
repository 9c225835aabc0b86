// Batched uneven structured x dense products: one-shot calls and hipGraph
// plans.  Declarations and contracts: include/psmma.h (section 3).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <numeric>
#include <vector>

#include "../../include/psmma.h"
#include "psmma_common.h"

namespace psmma {
hipError_t launch_grouped_f64(int cfg, const BatchProblem* dprobs, int nprob, int ntiles,
                              int nworkgroups, hipStream_t s);
void grouped_tile_shape(int cfg, int* bm, int* bn);
int launch_right_split(int kind, int prec, int slices, int M, int N, int K, double alpha,
                       const double* A, int lda, const GenPtrs& g, int r0, int c0, double beta,
                       double* C, int ldc, hipStream_t s);
size_t split_workspace_bytes(int prec, int slices, int M, int N, int K);
void* split_workspace(size_t bytes);
int split_moduli_for_bits(int K, int bits);
}  // namespace psmma

using namespace psmma;

namespace {

constexpr int kSplitBits = 43;  // bits per operand for a relative error <= 1e-12

// Validates one problem (same rules as the single-problem device calls) and
// converts it to the kernel's descriptor.  Returns 0 or -(field index + 1).
int to_batch(const psmma_problem& q, BatchProblem& b) {
  if (q.kind < PSMMA_KIND_CAUCHY || q.kind > PSMMA_KIND_CAUCHYEIG) return -1;
  if (q.m < 0 || q.n < 0 || q.k < 0) return -2;
  if (q.lda < std::max(1, q.m) || q.ldc < std::max(1, q.m)) return -6;
  if (q.r0 < 0 || q.c0 < 0) return -17;
  if (q.m > 0 && q.n > 0 && (!q.c || (q.k > 0 && (!q.a || !q.g0)))) return -7;
  GenPtrs g{q.g0, q.g1, q.g2, q.g3, q.g4, q.korder, q.r0 + q.k, q.c0 + q.n};
  if (q.m > 0 && q.n > 0 && q.k > 0) {
    if (q.kind == PSMMA_KIND_TOEPLITZ) {
      long long lo = (long long)q.korder - 1 + q.c0 - (q.r0 + q.k - 1);
      long long hi = (long long)q.korder - 1 + (q.c0 + q.n - 1) - q.r0;
      if (lo < 0 || hi >= q.lg) return -15;
    } else if (q.kind == PSMMA_KIND_HANKEL) {
      if ((long long)q.r0 + q.k - 1 + q.c0 + q.n - 1 >= q.lg) return -15;
    } else if (q.kind == PSMMA_KIND_CAUCHY) {
      if (!q.g1 || !q.g2 || !q.g3) return -7;
    } else {
      if (!q.g1 || !q.g2 || !q.g3 || !q.g4) return -7;
      if (q.r0 + q.k > q.lg || q.c0 + q.n > q.lg) return -15;
      g.korder = q.lg;
    }
  }
  b = BatchProblem{q.kind, q.m, q.n, q.k, q.alpha, q.beta, q.a, q.lda, g, q.r0, q.c0, q.c, q.ldc,
                   0, 0, 0};
  return 0;
}

int device_workgroups(int per_cu) {
  int dev = 0, cus = 256;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  return cus * per_cu;
}

// Tile list for a grouped launch: non-empty problems sorted by decreasing K.
int build_table(const std::vector<BatchProblem>& in, int cfg, std::vector<BatchProblem>& out) {
  int bm, bn;
  grouped_tile_shape(cfg, &bm, &bn);
  out.clear();
  for (const auto& b : in)
    if (b.m > 0 && b.n > 0) out.push_back(b);
  std::stable_sort(out.begin(), out.end(),
                   [](const BatchProblem& x, const BatchProblem& y) { return x.k > y.k; });
  int t = 0;
  for (auto& b : out) {
    b.ntm = (b.m + bm - 1) / bm;
    b.ntn = (b.n + bn - 1) / bn;
    b.tile_begin = t;
    t += b.ntm * b.ntn;
  }
  return t;
}

struct Plan {
  int prec = 0, slices = 0;
  int cfg = 0;
  int nodes = 0;
  BatchProblem* dprobs = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

// Enqueues the whole batch on stream s (FP64: grouped launch from the device
// table dtab; split: the per-problem pipelines).  Returns 0 or an error.
int enqueue(int prec, int slices, const std::vector<BatchProblem>& probs, const BatchProblem* dtab,
            int ntab, int ntiles, int cfg, hipStream_t s, int* nodes) {
  if (prec == PSMMA_PREC_FP64) {
    if (nodes) *nodes = ntiles > 0 ? 1 : 0;
    hipError_t e = launch_grouped_f64(cfg, dtab, ntab, ntiles, device_workgroups(cfg == 1 ? 1 : 2), s);
    return e == hipSuccess ? 0 : 1000 + (int)e;
  }
  int n = 0;
  for (const auto& b : probs) {
    if (b.m == 0 || b.n == 0) continue;
    int t = slices;
    if (prec == PSMMA_PREC_SPLIT_I8 && t == 0) t = split_moduli_for_bits(b.k, kSplitBits);
    if (prec == PSMMA_PREC_SPLIT_BF16 && t == 0) t = 6;
    int st = launch_right_split(b.kind, prec, t, b.m, b.n, b.k, b.alpha, b.A, b.lda, b.g, b.r0,
                                b.c0, b.beta, b.C, b.ldc, s);
    if (st) return st;
    ++n;
  }
  if (nodes) *nodes = n;
  return 0;
}

size_t max_split_bytes(int prec, int slices, const std::vector<BatchProblem>& probs) {
  size_t mx = 0;
  for (const auto& b : probs) {
    if (b.m == 0 || b.n == 0 || b.k == 0) continue;
    int t = slices;
    if (prec == PSMMA_PREC_SPLIT_I8 && t == 0) t = split_moduli_for_bits(b.k, kSplitBits);
    if (prec == PSMMA_PREC_SPLIT_BF16 && t == 0) t = 6;
    mx = std::max(mx, split_workspace_bytes(prec, t, b.m, b.n, b.k));
  }
  return mx;
}

int collect(int nprob, const psmma_problem* probs, std::vector<BatchProblem>& out) {
  if (nprob < 0 || (nprob > 0 && !probs)) return -1;
  out.resize(nprob);
  for (int i = 0; i < nprob; ++i) {
    int st = to_batch(probs[i], out[i]);
    if (st) {
      fprintf(stderr, "psmma_batched: problem %d: invalid field (%d)\n", i, st);
      return -1000 - i;
    }
  }
  return 0;
}

int check_prec(int prec, int slices) {
  if (prec != PSMMA_PREC_FP64 && prec != PSMMA_PREC_SPLIT_I8 && prec != PSMMA_PREC_SPLIT_BF16)
    return -3;
  if (slices < 0 || slices > 16 || (prec == PSMMA_PREC_SPLIT_BF16 && slices > 7)) return -4;
  return 0;
}

}  // namespace

extern "C" int psmma_batched_dev(int nprob, const psmma_problem* probs, int prec, int slices,
                                 void* stream) {
  int st = check_prec(prec, slices);
  if (st) return st;
  std::vector<BatchProblem> in;
  if ((st = collect(nprob, probs, in))) return st;
  hipStream_t s = (hipStream_t)stream;
  if (prec != PSMMA_PREC_FP64) return enqueue(prec, slices, in, nullptr, 0, 0, 0, s, nullptr);
  std::vector<BatchProblem> tab;
  int cfg = 0;
  int ntiles = build_table(in, cfg, tab);
  if (ntiles == 0) return 0;
  BatchProblem* d = nullptr;
  if (hipMalloc(&d, sizeof(BatchProblem) * tab.size()) != hipSuccess) return 1002;
  hipError_t e = hipMemcpy(d, tab.data(), sizeof(BatchProblem) * tab.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) st = enqueue(prec, slices, in, d, (int)tab.size(), ntiles, cfg, s, nullptr);
  else st = 1000 + (int)e;
  // the table must outlive the kernel: wait, then release it
  hipStreamSynchronize(s);
  hipFree(d);
  return st;
}

extern "C" int psmma_batched_plan_create(int nprob, const psmma_problem* probs, int prec,
                                         int slices, void** plan_out) {
  if (!plan_out) return -5;
  *plan_out = nullptr;
  int st = check_prec(prec, slices);
  if (st) return st;
  std::vector<BatchProblem> in;
  if ((st = collect(nprob, probs, in))) return st;
  Plan* P = new Plan;
  P->prec = prec;
  P->slices = slices;
  hipStream_t cs = nullptr;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
    delete P;
    return 1001;
  }
  std::vector<BatchProblem> tab;
  int ntiles = 0;
  if (prec == PSMMA_PREC_FP64) {
    // Tile-shape tuning: time both grouped shapes on this batch when rerunning
    // is harmless (every beta == 0), else take the shape with more tiles per CU.
    bool idempotent = std::all_of(in.begin(), in.end(), [](const BatchProblem& b) { return b.beta == 0.0; });
    std::vector<BatchProblem> t0, t1;
    int n0 = build_table(in, 0, t0), n1 = build_table(in, 1, t1);
    int cfg = (n1 >= 4 * device_workgroups(1)) ? 1 : 0;
    if (idempotent && n0 > 0) {
      BatchProblem *d0 = nullptr, *d1 = nullptr;
      hipMalloc(&d0, sizeof(BatchProblem) * t0.size());
      hipMalloc(&d1, sizeof(BatchProblem) * t1.size());
      hipMemcpy(d0, t0.data(), sizeof(BatchProblem) * t0.size(), hipMemcpyHostToDevice);
      hipMemcpy(d1, t1.data(), sizeof(BatchProblem) * t1.size(), hipMemcpyHostToDevice);
      double best = 1e30;
      for (int c = 0; c < 2; ++c) {
        const BatchProblem* d = c ? d1 : d0;
        int nt = c ? n1 : n0, nb = (int)(c ? t1.size() : t0.size());
        enqueue(prec, 0, in, d, nb, nt, c, cs, nullptr);  // warm-up
        hipEvent_t a, b;
        hipEventCreate(&a);
        hipEventCreate(&b);
        hipEventRecord(a, cs);
        for (int r = 0; r < 3; ++r) enqueue(prec, 0, in, d, nb, nt, c, cs, nullptr);
        hipEventRecord(b, cs);
        hipEventSynchronize(b);
        float ms = 0.f;
        hipEventElapsedTime(&ms, a, b);
        hipEventDestroy(a);
        hipEventDestroy(b);
        if (ms < best) {
          best = ms;
          cfg = c;
        }
      }
      hipFree(d0);
      hipFree(d1);
    }
    P->cfg = cfg;
    ntiles = build_table(in, cfg, tab);
    if (!tab.empty()) {
      if (hipMalloc(&P->dprobs, sizeof(BatchProblem) * tab.size()) != hipSuccess ||
          hipMemcpy(P->dprobs, tab.data(), sizeof(BatchProblem) * tab.size(),
                    hipMemcpyHostToDevice) != hipSuccess) {
        hipStreamDestroy(cs);
        psmma_batched_plan_destroy(P);
        return 1002;
      }
    }
  } else {
    // every pipeline shares one workspace: size it now, capture allocates nothing
    size_t need = max_split_bytes(prec, slices, in);
    if (need && !split_workspace(need)) {
      hipStreamDestroy(cs);
      delete P;
      return 1002;
    }
  }
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess)
    st = enqueue(prec, slices, in, P->dprobs, (int)tab.size(), ntiles, P->cfg, cs, &P->nodes);
  hipError_t e2 = hipStreamEndCapture(cs, &P->graph);
  if (e == hipSuccess && e2 == hipSuccess && st == 0)
    e = hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0);
  hipStreamDestroy(cs);
  if (e != hipSuccess || e2 != hipSuccess || st != 0) {
    fprintf(stderr, "psmma_batched_plan_create: capture failed (%s / %s / %d)\n",
            hipGetErrorString(e), hipGetErrorString(e2), st);
    psmma_batched_plan_destroy(P);
    return st ? st : 1003;
  }
  *plan_out = P;
  return 0;
}

extern "C" int psmma_batched_plan_launch(void* plan, void* stream) {
  Plan* P = (Plan*)plan;
  if (!P || !P->exec) return -1;
  hipError_t e = hipGraphLaunch(P->exec, (hipStream_t)stream);
  return e == hipSuccess ? 0 : 1000 + (int)e;
}

extern "C" int psmma_batched_plan_info(void* plan, int* tile_cfg, int* nodes) {
  Plan* P = (Plan*)plan;
  if (!P) return -1;
  if (tile_cfg) *tile_cfg = P->cfg;
  if (nodes) *nodes = P->nodes;
  return 0;
}

extern "C" int psmma_batched_plan_destroy(void* plan) {
  Plan* P = (Plan*)plan;
  if (!P) return 0;
  if (P->exec) hipGraphExecDestroy(P->exec);
  if (P->graph) hipGraphDestroy(P->graph);
  if (P->dprobs) {
    hipDeviceSynchronize();
    hipFree(P->dprobs);
  }
  delete P;
  return 0;
}
