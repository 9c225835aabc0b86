// The project's autotuner.
//
//  * FP64 tile configurations of the single-problem structured kernel: every
//    configuration (0..6) is timed on the caller's shape, bucketed by
//    (kind, ceil log2 of M, N, K), on synthetic generators and operands in
//    scratch memory; the fastest is cached for the process and consulted by
//    every later launch of that bucket (psmma_autotune_f64; automatically at a
//    bucket's first launch when PSMMA_AUTOTUNE=1).
//  * Split depth: the fewest int8 moduli (or bf16 slices) whose relative
//    Frobenius error, measured on a block of the caller's own columns against
//    the FP64 kernel, stays within the target (psmma_split_slices_dev); the
//    batched split plans use it for slices = 0.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

#include "../../include/psmma.h"
#include "psmma_common.h"
#include "workspace.h"

namespace psmma {
hipError_t launch_right_f64(int kind, int cfg, int M, int N, int K, double alpha, const double* A,
                            int lda, const GenPtrs& g, int r0, int c0, double beta, double* C,
                            int ldc, hipStream_t s);
int launch_right_split(int kind, int prec, int slices, int M, int N, int K, double alpha,
                       const double* A, int lda, const GenPtrs& g, int r0, int c0, double beta,
                       double* C, int ldc, hipStream_t s);
int split_moduli_for_bits(int K, int bits);
int to_batch(const psmma_problem& q, BatchProblem& b);
constexpr int kF64Configs = 7;  // configurations 0..6 of launch_right_f64
}  // namespace psmma

using namespace psmma;

namespace {

int clog2(int x) {
  int l = 0;
  while ((1ll << l) < (long long)x) ++l;
  return l;
}

using TuneKey = std::tuple<int, int, int, int, int>;  // device, kind, log2 M, N, K
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tuned;

TuneKey key_of(int kind, int M, int N, int K) {
  int dev = 0;
  hipGetDevice(&dev);
  return TuneKey{dev, kind, clog2(M), clog2(N), clog2(K)};
}

// Synthetic operands for timing: A = 1; generators with nonzero pole
// differences (Cauchy: d_i = 2i, f_j = 2j + 1; eigen-Cauchy: F_i = i,
// DIFL = -1/2, DIFR = 1/2), u = v = 1, Toeplitz / Hankel values 1/2^10.
__global__ void tune_fill_kernel(long long n, double* a, double v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void tune_gens_kernel(int n, double* p0, double* p1, double* p2, double* p3, double* p4,
                                 int kind) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (kind == KIND_CAUCHY) {
    p0[i] = 2.0 * i;
    p1[i] = 2.0 * i + 1.0;
    p2[i] = 1.0;
    p3[i] = 1.0;
  } else if (kind == KIND_CAUCHYEIG) {
    p0[i] = (double)i;
    p1[i] = 1.0;
    p2[i] = 1.0;
    p3[i] = -0.5;
    p4[i] = 0.5;
  } else {
    p0[i] = 1.0 / 1024.0;
  }
}

// Relative Frobenius error of x against ref (m x n, leading dimension ld), one
// workgroup: out[0] = sum (x - ref)^2, out[1] = sum ref^2.
__global__ __launch_bounds__(1024) void rel_err_kernel(int m, int n, const double* __restrict__ x,
                                                       const double* __restrict__ ref, int ld,
                                                       double* __restrict__ out) {
  double d2 = 0.0, r2 = 0.0;
  const long long tot = (long long)m * n;
  for (long long e = threadIdx.x; e < tot; e += blockDim.x) {
    const long long i = e % m, j = e / m;
    const double r = ref[j * ld + i], d = x[j * ld + i] - r;
    d2 += d * d;
    r2 += r * r;
  }
  __shared__ double s1[1024], s2[1024];
  s1[threadIdx.x] = d2;
  s2[threadIdx.x] = r2;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s1[0];
    out[1] = s2[0];
  }
}

}  // namespace

namespace psmma {

// Cached configuration of a bucket, or -1 (tunes first when PSMMA_AUTOTUNE=1).
int tuned_right_cfg(int kind, int M, int N, int K) {
  static const bool automatic = [] {
    const char* e = getenv("PSMMA_AUTOTUNE");
    return e && atoi(e) != 0;
  }();
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(key_of(kind, M, N, K));
    if (it != g_tuned.end()) return it->second;
  }
  if (!automatic) return -1;
  int cfg = -1;
  return psmma_autotune_f64(kind, M, N, K, &cfg, nullptr) == 0 ? cfg : -1;
}

}  // namespace psmma

extern "C" int psmma_autotune_f64(int kind, int m, int n, int k, int* cfg_out, float* ms_out) {
  if (kind < PSMMA_KIND_CAUCHY || kind > PSMMA_KIND_CAUCHYEIG) return -1;
  if (m < 1) return -2;
  if (n < 1) return -3;
  if (k < 1) return -4;
  const int nglob = std::max(k, n) + 1;
  const size_t abytes = sizeof(double) * (size_t)m * k, cbytes = sizeof(double) * (size_t)m * n;
  const size_t gbytes = sizeof(double) * (size_t)(2 * nglob + 1);
  char* mem = nullptr;
  if (hipMalloc(&mem, abytes + cbytes + 5 * gbytes) != hipSuccess) return 1002;
  double* A = (double*)mem;
  double* C = (double*)(mem + abytes);
  double* gp[5];
  for (int i = 0; i < 5; ++i) gp[i] = (double*)(mem + abytes + cbytes + i * gbytes);
  hipStream_t s = nullptr;
  const long long na = (long long)m * k;
  hipLaunchKernelGGL(tune_fill_kernel, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, s, na, A,
                     1.0 / 1024.0);
  const int ng = 2 * nglob + 1;
  hipLaunchKernelGGL(tune_gens_kernel, dim3((ng + 255) / 256), dim3(256), 0, s, ng, gp[0], gp[1],
                     gp[2], gp[3], gp[4], kind);
  // generator layout per kind, as the device API passes it (block at r0 = c0 = 0)
  GenPtrs g{gp[0], gp[1], gp[2], gp[3], gp[4], 0, k, n};
  if (kind == KIND_TOEPLITZ) g.korder = k;        // g has k + n - 1 entries from index 0
  if (kind == KIND_CAUCHYEIG) g.korder = nglob;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  int best = -1;
  float best_ms = 1e30f;
  for (int cfg = 0; cfg < kF64Configs; ++cfg) {
    if (launch_right_f64(kind, cfg, m, n, k, 1.0, A, m, g, 0, 0, 0.0, C, m, s) != hipSuccess)
      continue;  // warm-up
    hipEventRecord(e0, s);
    for (int r = 0; r < 3; ++r) launch_right_f64(kind, cfg, m, n, k, 1.0, A, m, g, 0, 0, 0.0, C, m, s);
    hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess) continue;
    float ms = 1e30f;
    hipEventElapsedTime(&ms, e0, e1);
    ms /= 3.0f;
    if (ms < best_ms) {
      best_ms = ms;
      best = cfg;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipStreamSynchronize(s);
  hipFree(mem);
  if (best < 0) return 1003;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[key_of(kind, m, n, k)] = best;
  }
  if (cfg_out) *cfg_out = best;
  if (ms_out) *ms_out = best_ms;
  return 0;
}

extern "C" int psmma_split_slices_dev(const psmma_problem* p, int prec, double target,
                                      int* slices_out, double* err_out, void* stream) {
  if (!p) return -1;
  if (prec != PSMMA_PREC_SPLIT_I8 && prec != PSMMA_PREC_SPLIT_BF16) return -2;
  if (!(target > 0.0)) return -3;
  if (!slices_out) return -4;
  const psmma_problem& q = *p;
  BatchProblem checked;
  if (to_batch(q, checked) != 0 || q.m < 1 || q.n < 1 || q.k < 1) return -1;
  if (prec == PSMMA_PREC_SPLIT_I8 && q.k > PSMMA_SPLIT_I8_KMAX) return -1;
  // a block of up to 64 columns from the middle of the problem's column range
  const int ncols = std::min(q.n, 64);
  const int c0 = q.c0 + (q.n - ncols) / 2;
  GenPtrs g{q.g0, q.g1, q.g2, q.g3, q.g4, q.korder, q.r0 + q.k, c0 + ncols};
  if (q.kind == PSMMA_KIND_CAUCHYEIG) g.korder = q.lg;
  hipStream_t s = (hipStream_t)stream;
  const size_t blk = sizeof(double) * (size_t)q.m * ncols;
  std::unique_lock<std::mutex> lk;
  char* ws = (char*)ws_lease(WS_TUNE, 2 * blk + 256, s, lk);
  if (!ws) return 1002;
  double* ref = (double*)ws;
  double* out = (double*)(ws + blk);
  double* sums = (double*)(ws + 2 * blk);
  hipError_t e = launch_right_f64(q.kind, -1, q.m, ncols, q.k, 1.0, q.a, q.lda, g, q.r0, c0, 0.0,
                                  ref, q.m, s);
  if (e != hipSuccess) return 1000 + (int)e;
  const int lo = prec == PSMMA_PREC_SPLIT_I8 ? 2 : 1, hi = prec == PSMMA_PREC_SPLIT_I8 ? 16 : 7;
  // start a little below the bit model's choice (int8) to save trials
  int t0 = lo;
  if (prec == PSMMA_PREC_SPLIT_I8) t0 = std::max(lo, split_moduli_for_bits(q.k, 30));
  double err = INFINITY;
  int chosen = -1;
  for (int t = t0; t <= hi; ++t) {
    // the split call leases the stream's split workspace (another slot) itself
    int st = launch_right_split(q.kind, prec, t, q.m, ncols, q.k, 1.0, q.a, q.lda, g, q.r0, c0,
                                0.0, out, q.m, s);
    if (st < 0) continue;  // depth not available for this K
    if (st > 0) return st;
    hipLaunchKernelGGL(rel_err_kernel, dim3(1), dim3(1024), 0, s, q.m, ncols, out, ref, q.m, sums);
    double h[2] = {0.0, 0.0};
    if (hipMemcpyAsync(h, sums, sizeof h, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return 1003;
    err = h[1] > 0.0 ? std::sqrt(h[0] / h[1]) : (h[0] > 0.0 ? INFINITY : 0.0);
    if (err <= target) {
      chosen = t;
      break;
    }
  }
  if (err_out) *err_out = err;
  if (chosen < 0) {
    *slices_out = hi;
    return 1;  // no depth reaches the target on the sample
  }
  *slices_out = chosen;
  return 0;
}
