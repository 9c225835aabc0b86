// Batched uneven structured x dense products: one-shot calls and hipGraph
// plans.  Declarations and contracts: include/psmma.h (section 3).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <vector>

#include "../../include/psmma.h"
#include "psmma_common.h"
#include "workspace.h"

namespace psmma {
hipError_t launch_batch_f64(const BatchProblem* dprobs, const int* dtbeg, int nprob, int ntiles,
                            unsigned* ctr, int nworkgroups, hipStream_t s, int kset);
void batch_tile_shape(int kset, int kind, int cls, int* bm, int* bn);
void* gsplit_create(const BatchProblem* probs, int nprob, int t, size_t* ws_bytes);
int gsplit_bind(void* h, void* ws);
int gsplit_enqueue(void* h, hipStream_t s, int* nodes);
void gsplit_destroy(void* h);
int gsplit_moduli(const BatchProblem* probs, int nprob, int bits);
}  // namespace psmma

using namespace psmma;

namespace {

// Bits per operand the batched split form keeps: 43 bits measure <= 3e-13 on the
// N = 8192 bench shape; two more bits cover wider-range problems of a batch.
constexpr int kSplitBits = 45;

}  // namespace

namespace psmma {
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
}  // namespace psmma

namespace {

int device_workgroups(int per_cu) {
  int dev = 0, cus = 256;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  return cus * per_cu;
}

// FP64 batch layout: one tile list over every problem (all kinds, both tile
// classes), problems ordered by the cost of one tile (K x tile area, longest
// first), each problem's tiles contiguous; one persistent launch runs it.
// A problem uses the large tiles when min(m, n) >= big_min.
struct F64Table {
  std::vector<BatchProblem> probs;
  std::vector<int> tbeg;
  int ntiles = 0;
};
struct F64Dev {
  BatchProblem* probs = nullptr;
  int* tbeg = nullptr;
  unsigned* ctr = nullptr;  // work counter and exit counter of the persistent launch
};

}  // namespace

namespace psmma {
// Tile list of a persistent batch launch: problems with m, n > 0 get their tile
// class (large when min(m, n) >= big_min), tile counts and first tile, ordered
// by the cost of one tile, longest first.  Returns the number of tiles.
int batch_tiles(const std::vector<BatchProblem>& in, int big_min, std::vector<BatchProblem>& out,
                std::vector<int>& tbeg, int kset) {
  std::vector<BatchProblem> all;
  for (auto b : in) {
    if (b.m == 0 || b.n == 0) continue;
    b.cls = std::min(b.m, b.n) >= big_min ? 1 : 0;
    int bm, bn;
    batch_tile_shape(kset, b.kind, b.cls, &bm, &bn);
    b.ntm = (b.m + bm - 1) / bm;
    b.ntn = (b.n + bn - 1) / bn;
    all.push_back(b);
  }
  auto cost = [kset](const BatchProblem& b) {
    int bm, bn;
    batch_tile_shape(kset, b.kind, b.cls, &bm, &bn);
    return (long long)std::max(b.k + b.k2, 1) * bm * bn;
  };
  std::stable_sort(all.begin(), all.end(), [&](const BatchProblem& x, const BatchProblem& y) {
    return cost(x) > cost(y);
  });
  out.clear();
  tbeg.clear();
  int ntiles = 0;
  for (auto& b : all) {
    b.tile_begin = ntiles;
    out.push_back(b);
    tbeg.push_back(ntiles);
    ntiles += b.ntm * b.ntn;
  }
  return ntiles;
}
}  // namespace psmma

namespace {

void build_f64(const std::vector<BatchProblem>& in, int big_min, F64Table& T) {
  T = F64Table{};
  T.ntiles = batch_tiles(in, big_min, T.probs, T.tbeg, 0);
}

void free_f64(F64Dev& D) {
  if (D.probs) hipFree(D.probs);
  if (D.tbeg) hipFree(D.tbeg);
  if (D.ctr) hipFree(D.ctr);
  D = F64Dev{};
}

int upload_f64(const F64Table& T, F64Dev& D) {
  if (T.probs.empty()) return 0;
  if (hipMalloc(&D.probs, sizeof(BatchProblem) * T.probs.size()) != hipSuccess ||
      hipMalloc(&D.tbeg, sizeof(int) * T.tbeg.size()) != hipSuccess ||
      hipMalloc(&D.ctr, sizeof(unsigned) * 2) != hipSuccess)
    return 1002;
  if (hipMemcpy(D.probs, T.probs.data(), sizeof(BatchProblem) * T.probs.size(),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(D.tbeg, T.tbeg.data(), sizeof(int) * T.tbeg.size(), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemset(D.ctr, 0, sizeof(unsigned) * 2) != hipSuccess)
    return 1003;
  return 0;
}

// Enqueues the batch's one persistent launch on stream s; *nodes = launches made.
int enqueue_f64(const F64Table& T, const F64Dev& D, hipStream_t s, int* nodes) {
  int n = 0;
  if (T.ntiles > 0) {
    hipError_t e = launch_batch_f64(D.probs, D.tbeg, (int)T.probs.size(), T.ntiles, D.ctr,
                                    device_workgroups(2), s, 0);
    if (e != hipSuccess) return 1000 + (int)e;
    n = 1;
  }
  if (nodes) *nodes = n;
  return 0;
}

// Times one FP64 layout (3 launches after a warm-up) on stream s, in ms.
float time_f64(const F64Table& T, const F64Dev& D, hipStream_t s) {
  if (enqueue_f64(T, D, s, nullptr)) return 1e30f;
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  hipEventRecord(a, s);
  for (int r = 0; r < 3; ++r) enqueue_f64(T, D, s, nullptr);
  hipEventRecord(b, s);
  hipEventSynchronize(b);
  float ms = 1e30f;
  hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  return ms;
}

struct Plan {
  int prec = 0, slices = 0;
  int cfg = 0;  // FP64: min(m, n) from which a problem uses the large tiles
  int nodes = 0;
  F64Table f64;
  F64Dev f64d;
  void* gsplit = nullptr;  // grouped split batch (PSMMA_PREC_SPLIT_I8)
  void* ws = nullptr;      // the plan's own split workspace
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

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

// Batched forms: plain FP64 and the int8 modular split (slices = moduli, 0 = by model).
int check_prec(int prec, int slices) {
  if (prec != PSMMA_PREC_FP64 && prec != PSMMA_PREC_SPLIT_I8) return -3;
  if (slices < 0 || slices == 1 || slices > 16) return -4;
  return 0;
}

// Split depth of a plan created with slices = 0: the largest depth the
// autotuner measures at 4e-13 relative error (psmma_split_slices_dev) over the
// longest-K problem of each kind.  The longest contraction keeps the fewest
// bits per operand, so that problem roughly bounds the batch's error; the rest
// of the 1e-12 target covers the columns not sampled, shorter problems whose
// operand rounding happens to be larger (up to 1.25x the longest one's on the
// 512-problem batch) and the fused form's fraction rounding (at most about
// 1.5e-13 there).  0 when a measurement fails.
int measured_moduli(int nprob, const psmma_problem* probs, hipStream_t s) {
  int pick[4] = {-1, -1, -1, -1};
  for (int i = 0; i < nprob; ++i) {
    const psmma_problem& q = probs[i];
    if (q.m == 0 || q.n == 0 || q.k == 0) continue;
    int& j = pick[q.kind];
    if (j < 0 || q.k > probs[j].k || (q.k == probs[j].k && q.n > probs[j].n)) j = i;
  }
  int t = 0;
  for (int kind = 0; kind < 4; ++kind) {
    if (pick[kind] < 0) continue;
    int tk = 0;
    double err = 0.0;
    if (psmma_split_slices_dev(&probs[pick[kind]], PSMMA_PREC_SPLIT_I8, 4e-13, &tk, &err, s) != 0)
      return 0;
    t = std::max(t, tk);
  }
  return t;
}

// The int8 form's int32 accumulators are exact for K <= PSMMA_SPLIT_I8_KMAX only.
int check_split_k(int prec, const std::vector<BatchProblem>& in) {
  if (prec != PSMMA_PREC_SPLIT_I8) return 0;
  for (size_t i = 0; i < in.size(); ++i)
    if (in[i].k > PSMMA_SPLIT_I8_KMAX) {
      fprintf(stderr, "psmma_batched: problem %zu: k = %d exceeds the int8 split limit %d\n", i,
              in[i].k, PSMMA_SPLIT_I8_KMAX);
      return -1000 - (int)i;
    }
  return 0;
}

}  // namespace

extern "C" int psmma_batched_dev(int nprob, const psmma_problem* probs, int prec, int slices,
                                 void* stream) {
  int st = check_prec(prec, slices);
  if (st) return st;
  std::vector<BatchProblem> in;
  if ((st = collect(nprob, probs, in))) return st;
  if ((st = check_split_k(prec, in))) return st;
  hipStream_t s = (hipStream_t)stream;
  if (prec != PSMMA_PREC_FP64) {
    int t = slices ? slices : gsplit_moduli(in.data(), nprob, kSplitBits);
    size_t bytes = 0;
    void* gs = gsplit_create(in.data(), nprob, t, &bytes);
    std::unique_lock<std::mutex> lk;
    void* ws = bytes ? ws_lease(WS_SPLIT, bytes, s, lk) : nullptr;
    st = (bytes && !ws) ? 1002 : gsplit_bind(gs, ws);
    if (st == 0) st = gsplit_enqueue(gs, s, nullptr);
    hipStreamSynchronize(s);  // the table is released below
    gsplit_destroy(gs);
    return st;
  }
  F64Table T;
  build_f64(in, 256, T);
  if (T.ntiles == 0) return 0;
  F64Dev D;
  st = upload_f64(T, D);
  if (st == 0) st = enqueue_f64(T, D, s, nullptr);
  hipStreamSynchronize(s);  // the table must outlive the kernels
  free_f64(D);
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
  if ((st = check_split_k(prec, in))) return st;
  Plan* P = new Plan;
  P->prec = prec;
  P->slices = slices;
  hipStream_t cs = nullptr;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
    delete P;
    return 1001;
  }
  auto fail = [&](int code) {
    hipStreamDestroy(cs);
    psmma_batched_plan_destroy(P);
    return code;
  };
  if (prec == PSMMA_PREC_FP64) {
    // Tile-class tuning: time the candidate large-tile thresholds on this batch
    // when rerunning is harmless (every beta == 0); else use 256.
    bool idempotent =
        std::all_of(in.begin(), in.end(), [](const BatchProblem& b) { return b.beta == 0.0; });
    int best_min = 256;
    if (idempotent) {
      float best = 1e30f;
      for (int cand : {256, 512, 1 << 30}) {
        F64Table T;
        F64Dev D;
        build_f64(in, cand, T);
        if (T.ntiles == 0 || upload_f64(T, D)) {
          free_f64(D);
          continue;
        }
        float ms = time_f64(T, D, cs);
        free_f64(D);
        if (ms < best) {
          best = ms;
          best_min = cand;
        }
      }
    }
    P->cfg = best_min;
    build_f64(in, best_min, P->f64);
    if ((st = upload_f64(P->f64, P->f64d))) return fail(st);
  } else {
    // the plan owns its workspace and table: capture allocates nothing, and no
    // later call can move memory the graph points to
    // measured on the default stream: its workspaces are reused by every plan
    // (the capture stream is new per plan)
    int t = slices ? slices : measured_moduli(nprob, probs, nullptr);
    if (t <= 0) t = gsplit_moduli(in.data(), nprob, kSplitBits);
    size_t bytes = 0;
    P->gsplit = gsplit_create(in.data(), nprob, t, &bytes);
    P->slices = t;
    if (bytes && hipMalloc(&P->ws, bytes) != hipSuccess) return fail(1002);
    if ((st = gsplit_bind(P->gsplit, P->ws))) return fail(st);
  }
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess)
    st = (prec == PSMMA_PREC_FP64) ? enqueue_f64(P->f64, P->f64d, cs, &P->nodes)
                                   : gsplit_enqueue(P->gsplit, cs, &P->nodes);
  hipError_t e2 = hipStreamEndCapture(cs, &P->graph);
  if (e == hipSuccess && e2 == hipSuccess && st == 0)
    e = hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0);
  if (e != hipSuccess || e2 != hipSuccess || st != 0) {
    fprintf(stderr, "psmma_batched_plan_create: capture failed (%s / %s / %d)\n",
            hipGetErrorString(e), hipGetErrorString(e2), st);
    return fail(st ? st : 1003);
  }
  hipStreamDestroy(cs);
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
  if (tile_cfg) *tile_cfg = (P->prec == PSMMA_PREC_FP64) ? P->cfg : P->slices;
  if (nodes) *nodes = P->nodes;
  return 0;
}

extern "C" int psmma_batched_plan_destroy(void* plan) {
  Plan* P = (Plan*)plan;
  if (!P) return 0;
  if (P->exec) hipGraphExecDestroy(P->exec);
  if (P->graph) hipGraphDestroy(P->graph);
  if (P->f64d.probs || P->ws) hipDeviceSynchronize();
  free_f64(P->f64d);
  if (P->gsplit) gsplit_destroy(P->gsplit);
  if (P->ws) hipFree(P->ws);
  delete P;
  return 0;
}
