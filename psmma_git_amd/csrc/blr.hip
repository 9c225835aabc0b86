// Compressed form of the eigenvector-merge product X = alpha * A * C + beta * X,
// C the eigen-Cauchy matrix of psmma_cauchyeig (and of the plain Cauchy product,
// C(i,j) = U_i V_j / (D_i - F_j), the same way): C is split into leaf x leaf
// blocks; every off-diagonal block is compressed by the batched truncated RRQR
// (Compress_RRQR semantics, tolerance tol) straight from the generators, and
// the product runs as
//   low-rank blocks C(I,J) ~ Q_IJ R_IJ:  T_I = A(:,I) [Q_I1 .. Q_In]
//   diagonal blocks with the low-rank columns:
//       X(:,J) = alpha (A(:,J) C(J,J) + [T_1J .. T_nJ] [R_1J; ..; R_nJ]) + beta X(:,J)
//   incompressible blocks: X(:,J) += alpha A(:,I) C(I,J)   (structured FP64 MFMA)
// Both GEMM stages are single launches of the persistent FP64 batch kernel on its
// 256 x 128 leaf tiles: the first over the block rows (A(:, I) U_I as the explicit
// segment of an otherwise empty contraction), the second over the block columns,
// whose diagonal block is generated in LDS and whose low-rank columns follow as an
// explicit tail of the same contraction (T'_J R_J, each block column at its own
// width).
// The first stage stores each column of T_I straight at its place in the block
// column order (J, I) through a column map, so each block column's tail is one
// contiguous panel of T' and no regrouping pass runs between the stages.
// This is the reference's off-diagonal ring step (PSTOEPLITZ_COMPUTE's
// RRQR branch, SRC/pstoeplitz_compute.f90:263-288, which compresses every
// off-diagonal tile with Compress_RRQR and multiplies A by [Q | R P^T] with two
// DGEMMs) applied on one GPU to a grid of leaf-sized tiles, with the GEMMs
// grouped per block row and column.
// The compressor is the batched RRQR (Compress_RRQR) or, as an alternative,
// the generator-only RRLU (Compress_cauchyEig2, SRC/CauchyHssEig.f90:2640-2829,
// the reference's compressor of off-diagonal eigen-Cauchy ring tiles), whose
// row ID C(I,J) ~ Q Y fills the same Q and R slots.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/psmma.h"
#include "batch_stages.h"
#include "psmma_common.h"
#include "workspace.h"

namespace psmma {
hipError_t launch_right_f64(int kind, int cfg, int M, int N, int K, double alpha, const double* A,
                            int lda, const GenPtrs& g, int r0, int c0, double beta, double* C,
                            int ldc, hipStream_t s);
hipError_t launch_rrqr_structured(int kind, const GenPtrs& g, const int* r0, const int* c0,
                                  int nblk, int m, int n, double tol, int kmax, int* rank,
                                  int* jpvt, long long jstride, double* q, int ldq,
                                  long long qstride, double* rpt, int ldr, long long rstride,
                                  double* cache, hipStream_t s);
}  // namespace psmma

using namespace psmma;

namespace {

constexpr int kKmax = 32;

// One low-rank block: where its Q and R P^T slots are, and where its columns
// go in the two GEMM stages.
struct PackEntry {
  int slot;   // block index of the RRQR outputs
  int r;      // rank
  int offa;   // column offset in U (block-row order (I, J))
  int offb;   // column offset in T' and row offset in R (block-column order (J, I))
};

// U(:, offa + c) = Q_slot(:, c) and R(offb + c, :) = (R P^T)_slot(c, :), both row-major
// (U: leaf x S, R: rows of leaf entries), the layout of the batch kernel's explicit
// tail operand; one workgroup per block.
__global__ __launch_bounds__(256) void blr_pack_kernel(const PackEntry* __restrict__ tab, int leaf,
                                                       const double* __restrict__ qs,
                                                       const double* __restrict__ rs,
                                                       double* __restrict__ U, double* __restrict__ R,
                                                       long long S) {
  const PackEntry e = tab[blockIdx.x];
  const double* q = qs + (size_t)e.slot * leaf * kKmax;
  const double* rp = rs + (size_t)e.slot * kKmax * leaf;
  for (int x = threadIdx.x; x < leaf * e.r; x += 256) {
    const int i = x % leaf, c = x / leaf;
    U[(size_t)i * S + e.offa + c] = q[(size_t)c * leaf + i];
    // R P^T is stored column-major with leading dimension kKmax
    R[(size_t)(e.offb + c) * leaf + i] = rp[(size_t)i * kKmax + c];
  }
}

// Global 1-based rows / columns of block b (leaf consecutive indices from r0[b] / c0[b]).
__global__ __launch_bounds__(256) void blr_index_kernel(const int* __restrict__ r0,
                                                        const int* __restrict__ c0, int leaf,
                                                        int* __restrict__ pl, int* __restrict__ pu) {
  const int b = blockIdx.x, t = threadIdx.x;
  if (t < leaf) {
    pl[(size_t)b * leaf + t] = r0[b] + t + 1;
    pu[(size_t)b * leaf + t] = c0[b] + t + 1;
  }
}

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

namespace {

// The compressed product for one structured kind (KIND_CAUCHYEIG or KIND_CAUCHY);
// arguments already checked, m > 0, nglob > 0.
int blr_product(int kind, const GenPtrs& g, int m, int nglob, double alpha, const double* a,
                int lda, int leaf, double tol, int compressor, double beta, double* c, int ldc,
                long long* stats, hipStream_t s) {
  const int nb = nglob / leaf;
  const int nblk = nb * (nb - 1);
  const bool eig = kind == KIND_CAUCHYEIG;
  const bool rrlu = compressor == PSMMA_BLR_RRLU;
  // workspace: block lists, compressor outputs (Q and R P^T, or Q and Y), and for the
  // RRQR of eigen-Cauchy blocks the evaluated blocks (each generated once, then read by
  // every pass)
  const size_t qbytes = (size_t)nblk * leaf * kKmax * sizeof(double);
  const size_t cbytes = (rrlu || !eig) ? 0 : (size_t)nblk * leaf * leaf * sizeof(double);
  size_t off_r0 = 0, off_c0 = align_up(off_r0 + sizeof(int) * nblk);
  size_t off_rank = align_up(off_c0 + sizeof(int) * nblk);
  size_t off_jp = align_up(off_rank + sizeof(int) * nblk);  // RRQR pivots or RRLU rows
  size_t off_pu = align_up(off_jp + sizeof(int) * (size_t)nblk * leaf);  // RRLU columns
  size_t off_q = align_up(off_pu + (rrlu ? sizeof(int) * (size_t)nblk * leaf : 0));
  size_t off_r = align_up(off_q + qbytes);
  size_t off_cache = align_up(off_r + qbytes);
  size_t fixed = align_up(off_cache + cbytes);
  std::unique_lock<std::mutex> lk_fixed, lk_gemm;
  char* ws = (char*)ws_lease(WS_BLR_FIXED, fixed, s, lk_fixed);
  if (!ws) return 1002;
  std::vector<int> hr0(nblk), hc0(nblk);
  for (int I = 0, b = 0; I < nb; ++I)
    for (int J = 0; J < nb; ++J)
      if (I != J) {
        hr0[b] = I * leaf;
        hc0[b] = J * leaf;
        ++b;
      }
  int* dr0 = (int*)(ws + off_r0);
  int* dc0 = (int*)(ws + off_c0);
  int* drank = (int*)(ws + off_rank);
  double* dq = (double*)(ws + off_q);
  double* drp = (double*)(ws + off_r);
  hipError_t e = hipSuccess;
  if (nblk > 0) {
    e = hipMemcpyAsync(dr0, hr0.data(), sizeof(int) * nblk, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dc0, hc0.data(), sizeof(int) * nblk, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !rrlu)
      e = launch_rrqr_structured(kind, g, dr0, dc0, nblk, leaf, leaf, tol, kKmax, drank,
                                 (int*)(ws + off_jp), leaf, dq, leaf, (long long)leaf * kKmax, drp,
                                 kKmax, (long long)kKmax * leaf,
                                 eig ? (double*)(ws + off_cache) : nullptr, s);
    if (e == hipSuccess && rrlu) {
      int* dpl = (int*)(ws + off_jp);
      int* dpu = (int*)(ws + off_pu);
      blr_index_kernel<<<nblk, 256, 0, s>>>(dr0, dc0, leaf, dpl, dpu);
      e = hipGetLastError();
      if (e == hipSuccess) {
        // generators: eigen-Cauchy p0..p4 = F, U, V, DIFL, DIFR; Cauchy p0..p3 = D, F, U, V
        const int st =
            eig ? psmma_rrlu_cauchyeig_batched_dev(nblk, leaf, leaf, dpl, dpu, g.p0, g.p1, g.p2,
                                                   g.p3, g.p4, nglob, tol, kKmax, drank, nullptr,
                                                   dq, leaf, (long long)leaf * kKmax, drp, kKmax,
                                                   (long long)kKmax * leaf, s)
                : psmma_rrlu_cauchy_batched_dev(nblk, leaf, leaf, dpl, dpu, g.p0, g.p1, g.p2,
                                                g.p3, nglob, nglob, tol, kKmax, drank, dq, leaf,
                                                (long long)leaf * kKmax, drp, kKmax,
                                                (long long)kKmax * leaf, s);
        if (st) return st > 0 ? st : 1006;
      }
    }
  }
  std::vector<int> hrank(nblk);
  if (e == hipSuccess && nblk > 0) {
    e = hipMemcpyAsync(hrank.data(), drank, sizeof(int) * nblk, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  if (e != hipSuccess) {
    fprintf(stderr, "compressed structured product: %s\n", hipGetErrorString(e));
    return 1000 + (int)e;
  }

  // classify the blocks; offsets of the low-rank columns in both stage orders
  std::vector<std::vector<int>> rowlist(nb), collist(nb);  // low-rank block ids per I / per J
  std::vector<int> dense;
  long long ranksum = 0, nzero = 0;
  for (int b = 0; b < nblk; ++b) {
    const int r = hrank[b];
    if (r == 0) {
      ++nzero;
    } else if (r < 0 || r >= leaf || r > kKmax) {  // RRLU ranks above kKmax: no outputs
      dense.push_back(b);
    } else {
      rowlist[hr0[b] / leaf].push_back(b);
      collist[hc0[b] / leaf].push_back(b);
      ranksum += r;
    }
  }
  // Column slots: block row I owns wA columns of U and T (its blocks' ranks first),
  // block column J owns wB columns of T' and rows of R, with wA / wB the widest block
  // row / column; each stage problem takes only its own row's or column's width.
  std::vector<PackEntry> tab(nblk);
  std::vector<int> wrow(nb, 0), wcol(nb, 0);
  for (int I = 0; I < nb; ++I)
    for (int b : rowlist[I]) wrow[I] += hrank[b];
  for (int J = 0; J < nb; ++J)
    for (int b : collist[J]) wcol[J] += hrank[b];
  const int wA = *std::max_element(wrow.begin(), wrow.end());
  const int wB = *std::max_element(wcol.begin(), wcol.end());
  for (int I = 0; I < nb; ++I) {
    int o = I * wA;
    for (int b : rowlist[I]) {
      tab[b].slot = b;
      tab[b].r = hrank[b];
      tab[b].offa = o;
      o += hrank[b];
    }
  }
  for (int J = 0; J < nb; ++J) {
    int o = J * wB;
    for (int b : collist[J]) {
      tab[b].offb = o;
      o += hrank[b];
    }
  }
  const long long S = (long long)nb * wA;   // columns of U and T
  const long long SR = (long long)nb * wB;  // columns of T', rows of R
  std::vector<PackEntry> packed;
  for (int I = 0; I < nb; ++I)
    for (int b : rowlist[I]) packed.push_back(tab[b]);
  // column map of the first stage: column offa + c of the block-row order goes to
  // column offb + c of T'
  std::vector<int> cmap((size_t)std::max<long long>(S, 1), 0);
  for (const PackEntry& pe : packed)
    for (int c = 0; c < pe.r; ++c) cmap[(size_t)pe.offa + c] = pe.offb + c;
  if (stats) {
    stats[0] = nblk;
    stats[1] = (long long)packed.size();
    stats[2] = (long long)dense.size();
    stats[3] = nzero;
    stats[4] = ranksum;
  }

  // second workspace part: pack table, column map (S entries), U (leaf x S, row-major),
  // R (SR x leaf, row-major), T' (m x SR)
  const size_t off_tab = 0;
  const size_t off_map = align_up(off_tab + sizeof(PackEntry) * std::max<size_t>(1, packed.size()));
  const size_t off_u = align_up(off_map + sizeof(int) * cmap.size());
  const size_t off_rr = align_up(off_u + sizeof(double) * (size_t)leaf * S);
  const size_t off_tp = align_up(off_rr + sizeof(double) * (size_t)SR * leaf);
  const size_t total = align_up(off_tp + sizeof(double) * (size_t)m * SR);
  char* ws2 = (char*)ws_lease(WS_BLR_GEMM, total, s, lk_gemm);
  if (!ws2) return 1002;
  PackEntry* dtab = (PackEntry*)(ws2 + off_tab);
  int* dmap = (int*)(ws2 + off_map);
  double* U = (double*)(ws2 + off_u);
  double* R = (double*)(ws2 + off_rr);
  double* Tp = (double*)(ws2 + off_tp);
  // the tables are copied before any launch of this call (the stream is idle here)
  if (!packed.empty()) {
    e = hipMemcpy(dtab, packed.data(), sizeof(PackEntry) * packed.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(dmap, cmap.data(), sizeof(int) * cmap.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return 1000 + (int)e;
  }

  // stage problems: T_I = A(:, I) U_I per block row; X(:, J) = alpha (A(:, J) C(J, J) +
  // T'_J R_J) + beta X(:, J) per block column (generated diagonal block, explicit tail)
  const bool lowrank = !packed.empty();
  std::vector<std::vector<BatchProblem>> stg(2);
  auto problem = [&](int kd, int n, int k, double al, const double* A, int ld, double be,
                     double* C, int ldC) {
    BatchProblem b{};
    b.kind = kd;
    b.m = m;
    b.n = n;
    b.k = k;
    b.alpha = al;
    b.beta = be;
    b.A = A;
    b.lda = ld;
    b.C = C;
    b.ldc = ldC;
    return b;
  };
  // (each block row and column takes its own width: the padding is never multiplied).
  // T_I is a plain product on the leaf tiles: an empty generated segment (whose
  // generators are then never read; the kind only picks the tile code) and
  // A(:, I) U_I as the tail.
  const bool window = kind == KIND_TOEPLITZ || kind == KIND_HANKEL;
  const int tail_kind = window ? KIND_CAUCHY : kind;
  if (lowrank)
    for (int I = 0; I < nb; ++I) {
      BatchProblem b = problem(tail_kind, wrow[I], 0, 1.0, a, lda, 0.0, Tp, m);
      b.cmap = dmap + (size_t)I * wA;  // stored in block-column order
      if (!window) b.g = g;
      b.A2 = a + (size_t)I * leaf * lda;
      b.lda2 = lda;
      b.H2 = U + (size_t)I * wA;
      b.ldh2 = (int)S;
      b.k2 = leaf;
      stg[0].push_back(b);
    }
  // The diagonal blocks: the Cauchy kinds take the low-rank columns as the tail of
  // the same contraction; the window kinds (Toeplitz / Hankel tiles take no tail)
  // run them on their own window tiles, then the tails as a third stage.
  stg.resize(window ? 3 : 2);
  for (int J = 0; J < nb; ++J) {
    BatchProblem b = problem(kind, leaf, leaf, alpha, a + (size_t)J * leaf * lda, lda, beta,
                             c + (size_t)J * leaf * ldc, ldc);
    b.g = g;
    b.g.nrows = (J + 1) * leaf;
    b.g.ncols = (J + 1) * leaf;
    b.r0 = b.c0 = J * leaf;
    BatchProblem tl = problem(tail_kind, leaf, 0, alpha, a, lda, 1.0, b.C, ldc);
    BatchProblem& tb = window ? tl : b;
    if (lowrank) {
      tb.A2 = Tp + (size_t)J * m * wB;
      tb.lda2 = m;
      tb.H2 = R + (size_t)J * wB * leaf;
      tb.ldh2 = leaf;
      tb.k2 = wcol[J];
    }
    stg[1].push_back(b);
    if (window && lowrank && wcol[J] > 0) stg[2].push_back(tl);
  }
  BatchStages bs;
  int st = window
               ? bs.upload(stg, {BKSET_LEAF, BKSET_PUBLIC, BKSET_LEAF}, 256, WS_BLR_TABLES, s)
               : bs.upload(stg, {BKSET_LEAF, BKSET_LEAF}, 256, WS_BLR_TABLES, s);
  if (st) return st;
  if (lowrank) {
    {
      blr_pack_kernel<<<(unsigned)packed.size(), 256, 0, s>>>(dtab, leaf, dq, drp, U, R, S);
      e = hipGetLastError();
    }
    if (e != hipSuccess) {
      fprintf(stderr, "compressed structured product: %s\n", hipGetErrorString(e));
      return 1000 + (int)e;
    }
    if ((st = bs.launch(0))) return st;  // T_I = A(:, I) U_I, stored as T'
  }
  if ((st = bs.launch(1))) return st;  // the diagonal blocks carry beta (and the tails)
  if (window && (st = bs.launch(2))) return st;  // the window kinds' tails, after beta
  // incompressible blocks one by one (they accumulate)
  for (size_t x = 0; x < dense.size() && e == hipSuccess; ++x) {
    const int b = dense[x], I = hr0[b] / leaf, J = hc0[b] / leaf;
    e = launch_right_f64(kind, -1, m, leaf, leaf, alpha, a + (size_t)I * leaf * lda,
                         lda, g, I * leaf, J * leaf, 1.0, c + (size_t)J * leaf * ldc, ldc, s);
  }
  if (e != hipSuccess) return 1000 + (int)e;
  return bs.finish();
}

}  // namespace

extern "C" int psmma_cauchyeig_blr_ex_dev(int m, int nglob, double alpha, const double* a,
                                          int lda, const double* fpole, const double* u,
                                          const double* v, const double* difl,
                                          const double* difr, int leaf, double tol,
                                          int compressor, double beta, double* c, int ldc,
                                          long long* stats, void* stream) {
  if (m < 0) return -1;
  if (nglob < 0) return -2;
  if (lda < std::max(1, m)) return -5;
  if (leaf < 1 || leaf > 256 || (nglob % leaf) != 0) return -11;
  if (!(tol >= 0.0)) return -12;
  if (compressor != PSMMA_BLR_RRQR && compressor != PSMMA_BLR_RRLU) return -13;
  if (ldc < std::max(1, m)) return -16;
  if (m == 0 || nglob == 0) return 0;
  if (!a || !c || !fpole || !u || !v || !difl || !difr) return -4;
  GenPtrs g{fpole, u, v, difl, difr, nglob, nglob, nglob};
  return blr_product(KIND_CAUCHYEIG, g, m, nglob, alpha, a, lda, leaf, tol, compressor, beta, c,
                     ldc, stats, (hipStream_t)stream);
}

extern "C" int psmma_cauchy_blr_dev(int m, int nglob, double alpha, const double* a, int lda,
                                    const double* d, const double* f, const double* u,
                                    const double* v, int leaf, double tol, int compressor,
                                    double beta, double* c, int ldc, long long* stats,
                                    void* stream) {
  if (m < 0) return -1;
  if (nglob < 0) return -2;
  if (lda < std::max(1, m)) return -5;
  if (leaf < 1 || leaf > 256 || (nglob % leaf) != 0) return -10;
  if (!(tol >= 0.0)) return -11;
  if (compressor != PSMMA_BLR_RRQR && compressor != PSMMA_BLR_RRLU) return -12;
  if (ldc < std::max(1, m)) return -15;
  if (m == 0 || nglob == 0) return 0;
  if (!a || !c || !d || !f || !u || !v) return -4;
  GenPtrs g{d, f, u, v, nullptr, nglob, nglob, nglob};
  return blr_product(KIND_CAUCHY, g, m, nglob, alpha, a, lda, leaf, tol, compressor, beta, c, ldc,
                     stats, (hipStream_t)stream);
}

extern "C" int psmma_cauchyeig_blr_dev(int m, int nglob, double alpha, const double* a, int lda,
                                       const double* fpole, const double* u, const double* v,
                                       const double* difl, const double* difr, int leaf,
                                       double tol, double beta, double* c, int ldc,
                                       long long* stats, void* stream) {
  const int st = psmma_cauchyeig_blr_ex_dev(m, nglob, alpha, a, lda, fpole, u, v, difl, difr, leaf,
                                            tol, PSMMA_BLR_RRQR, beta, c, ldc, stats, stream);
  return st == -16 ? -15 : st;  // argument numbers of this entry point
}

// Compressed Toeplitz / Hankel products: the reference's off-diagonal Toeplitz ring
// step (ConstToeplitzlowrank -> Compress_RRQR -> two DGEMMs, or the regenerated tile
// at full rank, SRC/pstoeplitz_compute.f90:263-295, SRC/Auxil_2p5D.f90:968-1189) on
// a grid of leaf tiles of the n x n block H(0:n, 0:n).
extern "C" int psmma_toeplitz_blr_dev(int m, int n, double alpha, const double* a, int lda,
                                      const double* g, int lg, int korder, int leaf, double tol,
                                      double beta, double* c, int ldc, long long* stats,
                                      void* stream) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max(1, m)) return -5;
  if (leaf < 1 || leaf > 256 || (n % leaf) != 0) return -9;
  if (!(tol >= 0.0)) return -10;
  if (ldc < std::max(1, m)) return -13;
  if (m == 0 || n == 0) return 0;
  if (!a || !c) return -4;
  // H(r, c) = g[korder - 1 + c - r] over 0 <= r, c < n
  if (!g || korder - n < 0 || (long long)korder - 1 + n - 1 >= lg) return -6;
  GenPtrs gp{g, nullptr, nullptr, nullptr, nullptr, korder, n, n};
  return blr_product(KIND_TOEPLITZ, gp, m, n, alpha, a, lda, leaf, tol, PSMMA_BLR_RRQR, beta, c,
                     ldc, stats, (hipStream_t)stream);
}

extern "C" int psmma_hankel_blr_dev(int m, int n, double alpha, const double* a, int lda,
                                    const double* g, int lg, int leaf, double tol, double beta,
                                    double* c, int ldc, long long* stats, void* stream) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max(1, m)) return -5;
  if (leaf < 1 || leaf > 256 || (n % leaf) != 0) return -8;
  if (!(tol >= 0.0)) return -9;
  if (ldc < std::max(1, m)) return -12;
  if (m == 0 || n == 0) return 0;
  if (!a || !c) return -4;
  // H(r, c) = g[r + c] over 0 <= r, c < n
  if (!g || 2LL * n - 2 >= lg) return -6;
  GenPtrs gp{g, nullptr, nullptr, nullptr, nullptr, 0, n, n};
  return blr_product(KIND_HANKEL, gp, m, n, alpha, a, lda, leaf, tol, PSMMA_BLR_RRQR, beta, c,
                     ldc, stats, (hipStream_t)stream);
}
