// The ring step of PS*_COMPUTE on one GPU, and the block-cyclic redistribution
// of REDIST, for callers that keep the reference's P x Q process grid.
//
//  * psmma_constindex / psmma_is_intersect: the global index lists of a process
//    column (ConstIndex) and the step classification (is_intersect).
//  * psmma_ring_step_dev: one step's product C += alpha * A * H(RIndex, CIndex)
//    (the dense branch of SRC/pscauchy_compute.f90:391-407 and siblings) on the
//    FP64 kernel, the structured tile generated through the index lists.
//  * psmma_redistribute_dev: PDGEMR2D between two block-cyclic layouts of one
//    grid (NB-cyclic <-> blocks of ceil(N/P), SRC/psmma_cauchy.f90:113-162) as a
//    device permutation of the ranks' local arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <vector>

#include "../../include/psmma.h"
#include "psmma_common.h"
#include "workspace.h"

namespace psmma {
hipError_t launch_indexed_f64(int kind, int M, int N, int K, double alpha, const double* A,
                              int lda, const GenPtrs& g, double beta, double* C, int ldc,
                              hipStream_t s);
}  // namespace psmma

using namespace psmma;

namespace {

// Local extent of a block-cyclic dimension (ScaLAPACK NUMROC with source 0).
int local_count(int n, int nb, int iproc, int nprocs) {
  if (n <= 0) return 0;
  const int nblocks = n / nb;
  int cnt = (nblocks / nprocs) * nb;
  const int extra = nblocks % nprocs;
  if (iproc < extra) cnt += nb;
  else if (iproc == extra) cnt += n % nb;
  return cnt;
}

}  // namespace

// ConstIndex (SRC/Auxil_2p5D.f90:2-66): the 1-based global indices owned by
// process column `mycol` of `npcol` in a block-cyclic layout of n with block nb,
// in increasing order.  Computed in closed form: local index l maps to
// (l / nb) * npcol * nb + mycol * nb + l % nb + 1 for l < NUMROC(n, nb, mycol).
extern "C" int psmma_constindex(int n, int nb, int mycol, int npcol, int* index) {
  if (n < 0 || nb < 1 || npcol < 1 || mycol < 0 || mycol >= npcol) return -1;
  const int len = local_count(n, nb, mycol, npcol);
  if (index)
    for (int l = 0; l < len; ++l) index[l] = (l / nb) * npcol * nb + mycol * nb + l % nb + 1;
  return len;
}

// is_intersect (SRC/Auxil_2p5D.f90:752-792) with the reference's sense: returns 1
// when the two index sets share NO element (PS*_COMPUTE then takes its low-rank
// branch, SRC/pscauchy_compute.f90:368-371), 0 otherwise.  The reference probes
// the shorter list's first elements against the longer list; for the lists
// ConstIndex produces (two process columns' sets: equal or disjoint) that is the
// full test, which is what this computes (both lists increasing: one merge pass).
extern "C" int psmma_is_intersect(int m, const int* seta, int n, const int* setb) {
  int i = 0, j = 0;
  while (i < m && j < n) {
    if (seta[i] == setb[j]) return 0;
    if (seta[i] < setb[j]) ++i;
    else ++j;
  }
  return 1;
}

extern "C" int psmma_ring_step_dev(const psmma_problem* p, const int* rindex, const int* cindex,
                                   void* stream) {
  if (!p) return -1;
  const psmma_problem& q = *p;
  if (q.kind < PSMMA_KIND_CAUCHY || q.kind > PSMMA_KIND_CAUCHYEIG) return -1;
  if (q.m < 0 || q.n < 0 || q.k < 0) return -1;
  if (q.lda < std::max(1, q.m) || q.ldc < std::max(1, q.m)) return -1;
  if (q.m == 0 || q.n == 0) return 0;
  if (!q.c || (q.k > 0 && (!q.a || !q.g0 || !rindex)) || !cindex) return -1;
  // generator extents: the largest global row / column the lists may name
  int rows_max = 0, cols_max = 0;
  switch (q.kind) {
    case PSMMA_KIND_TOEPLITZ: rows_max = cols_max = q.korder; break;   // B(K + j - i)
    case PSMMA_KIND_HANKEL: rows_max = cols_max = (q.lg + 1) / 2; break;
    case PSMMA_KIND_CAUCHY:
      if (!q.g1 || !q.g2 || !q.g3) return -1;
      rows_max = cols_max = q.lg;  // length of each of D, F, U, V
      break;
    default:
      if (!q.g1 || !q.g2 || !q.g3 || !q.g4) return -1;
      rows_max = cols_max = q.lg;
  }
  for (int i = 0; i < q.k; ++i)
    if (rindex[i] < 1 || rindex[i] > rows_max) return -2;
  for (int j = 0; j < q.n; ++j)
    if (cindex[j] < 1 || cindex[j] > cols_max) return -3;
  if (q.kind == PSMMA_KIND_TOEPLITZ && q.k > 0) {
    // G(K + C_j - R_i) must lie in 1 .. lg
    const int rmin = *std::min_element(rindex, rindex + q.k);
    const int rmax = *std::max_element(rindex, rindex + q.k);
    const int cmin = *std::min_element(cindex, cindex + q.n);
    const int cmax = *std::max_element(cindex, cindex + q.n);
    if ((long long)q.korder + cmin - rmax < 1 || (long long)q.korder + cmax - rmin > q.lg)
      return -2;
  }
  if (q.kind == PSMMA_KIND_HANKEL && q.k > 0) {
    const int rmax = *std::max_element(rindex, rindex + q.k);
    const int cmax = *std::max_element(cindex, cindex + q.n);
    if ((long long)rmax + cmax - 1 > q.lg) return -2;
  }
  hipStream_t s = (hipStream_t)stream;
  std::unique_lock<std::mutex> lk;
  int* dix = (int*)ws_lease(WS_RING, sizeof(int) * ((size_t)q.k + q.n), s, lk);
  if (!dix) return 1002;
  hipError_t e = hipSuccess;
  if (q.k > 0) e = hipMemcpyAsync(dix, rindex, sizeof(int) * q.k, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dix + q.k, cindex, sizeof(int) * q.n, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return 1000 + (int)e;
  GenPtrs g{q.g0, q.g1, q.g2, q.g3, q.g4, q.korder, rows_max, cols_max};
  if (q.kind == PSMMA_KIND_CAUCHYEIG) g.korder = q.lg;
  g.ridx = dix;
  g.cidx = dix + q.k;
  e = launch_indexed_f64(q.kind, q.m, q.n, q.k, q.alpha, q.a, q.lda, g, q.beta, q.c, q.ldc, s);
  return e == hipSuccess ? 0 : 1000 + (int)e;
}

// ---------------------------------------------------------------------------
// Block-cyclic redistribution.

namespace {

struct Layout {
  int mb, nb;
};

// One destination rank per blockIdx.y; threads walk its local array (column
// major, ld = its local rows) and gather each element from the source rank
// that owns the same global entry.
__global__ __launch_bounds__(256) void redistribute_kernel(int m, int n, int P, int Q, Layout src,
                                                           Layout dst,
                                                           const double* __restrict__ a,
                                                           const long long* __restrict__ aoff,
                                                           double* __restrict__ b,
                                                           const long long* __restrict__ boff) {
  const int r = blockIdx.y;
  const int pr = r / Q, pc = r % Q;
  const long long base = boff[r];
  const long long cnt = boff[r + 1] - base;
  if (cnt == 0) return;
  // local rows of this destination rank (NUMROC)
  const int nblk = m / dst.mb;
  int mloc = (nblk / P) * dst.mb;
  if (pr < nblk % P) mloc += dst.mb;
  else if (pr == nblk % P) mloc += m % dst.mb;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < cnt;
       e += (long long)gridDim.x * blockDim.x) {
    const int li = (int)(e % mloc), lj = (int)(e / mloc);
    const int i = (li / dst.mb) * P * dst.mb + pr * dst.mb + li % dst.mb;  // global, 0-based
    const int j = (lj / dst.nb) * Q * dst.nb + pc * dst.nb + lj % dst.nb;
    const int sr = (i / src.mb) % P, sc = (j / src.nb) % Q;
    const int si = (i / (P * src.mb)) * src.mb + i % src.mb;
    const int sj = (j / (Q * src.nb)) * src.nb + j % src.nb;
    // the source rank's local rows
    const int snblk = m / src.mb;
    int smloc = (snblk / P) * src.mb;
    if (sr < snblk % P) smloc += src.mb;
    else if (sr == snblk % P) smloc += m % src.mb;
    b[base + e] = a[aoff[sr * Q + sc] + (long long)sj * smloc + si];
  }
}

// Offsets of the P*Q local arrays (rank-major, r = pr * Q + pc) of an m x n
// matrix in layout (mb, nb); offs has P*Q + 1 entries.
std::vector<long long> local_offsets(int m, int n, int P, int Q, int mb, int nb) {
  std::vector<long long> offs(P * Q + 1, 0);
  for (int pr = 0; pr < P; ++pr)
    for (int pc = 0; pc < Q; ++pc) {
      const int r = pr * Q + pc;
      offs[r + 1] = offs[r] + (long long)local_count(m, mb, pr, P) * local_count(n, nb, pc, Q);
    }
  return offs;
}

}  // namespace

extern "C" long long psmma_local_offsets(int m, int n, int nprow, int npcol, int mb, int nb,
                                         long long* offsets) {
  if (m < 0 || n < 0 || nprow < 1 || npcol < 1 || mb < 1 || nb < 1) return -1;
  std::vector<long long> o = local_offsets(m, n, nprow, npcol, mb, nb);
  if (offsets) std::copy(o.begin(), o.end(), offsets);
  return o.back();
}

extern "C" int psmma_redistribute_dev(int m, int n, int nprow, int npcol, int src_mb, int src_nb,
                                      const double* a, int dst_mb, int dst_nb, double* b,
                                      void* stream) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (nprow < 1) return -3;
  if (npcol < 1) return -4;
  if (src_mb < 1) return -5;
  if (src_nb < 1) return -6;
  if (dst_mb < 1) return -8;
  if (dst_nb < 1) return -9;
  if (m == 0 || n == 0) return 0;
  if (!a) return -7;
  if (!b) return -10;
  if ((long long)m * n >= (1ll << 31) || (long long)nprow * npcol > 65535) return -1;
  std::vector<long long> ao = local_offsets(m, n, nprow, npcol, src_mb, src_nb);
  std::vector<long long> bo = local_offsets(m, n, nprow, npcol, dst_mb, dst_nb);
  const int nr = nprow * npcol;
  hipStream_t s = (hipStream_t)stream;
  std::unique_lock<std::mutex> lk;
  long long* doff = (long long*)ws_lease(WS_RING, sizeof(long long) * 2 * (nr + 1), s, lk);
  if (!doff) return 1002;
  std::vector<long long> both(ao);
  both.insert(both.end(), bo.begin(), bo.end());
  hipError_t e = hipMemcpyAsync(doff, both.data(), sizeof(long long) * both.size(),
                                hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return 1000 + (int)e;
  long long maxcnt = 0;
  for (int r = 0; r < nr; ++r) maxcnt = std::max(maxcnt, bo[r + 1] - bo[r]);
  const unsigned gx = (unsigned)std::min<long long>((maxcnt + 255) / 256, 4096);
  hipLaunchKernelGGL(redistribute_kernel, dim3(gx, nr), dim3(256), 0, s, m, n, nprow, npcol,
                     Layout{src_mb, src_nb}, Layout{dst_mb, dst_nb}, a, doff, b, doff + nr + 1);
  e = hipGetLastError();
  return e == hipSuccess ? 0 : 1000 + (int)e;
}
