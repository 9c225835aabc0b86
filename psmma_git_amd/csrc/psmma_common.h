// Shared device-side definitions for the PSMMA structured x dense kernels (gfx950).
//
// Conventions (all kernels):
//   * Matrices are column-major with an explicit leading dimension, as in the
//     Fortran entry points they serve.
//   * A structured matrix H is never materialised in HBM: each kernel evaluates
//     the entries of the K-tile it is about to consume from the O(n) generators
//     and writes them straight into LDS, where the MFMA fragments read them.
//   * Row/column offsets (r0, c0) select a sub-block of the global structured
//     matrix, so a caller can process one block of a larger product (a ring step,
//     one problem of a batch, one HSS block).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));

namespace psmma {

// Kinds of structured operand.  Values are part of the C ABI (include/psmma.h).
enum Kind : int {
  KIND_CAUCHY = 0,     // H(i,j) = U_i V_j / (D_i - F_j)                 psmma_cauchy
  KIND_TOEPLITZ = 1,   // H(i,j) = G(K + j - i)  (1-based, G length 2K-1) psmma_toeplitz
  KIND_HANKEL = 2,     // H(i,j) = G(i + j - 1)                           psmma_hankel
  KIND_CAUCHYEIG = 3,  // H(i,j) = U_i V_j / (F_i - lambda_j), stable form psmma_cauchyeig
  // Internal kinds (not part of the C ABI), used by the HSS construction and
  // product: the transposes of the Cauchy kinds and explicit matrices.
  KIND_CAUCHYEIG_T = 4,  // H(i,j) = Ceig(j,i)
  KIND_CAUCHY_T = 5,     // H(i,j) = Ccauchy(j,i)
  KIND_DENSE = 6,        // H(i,j) = p0[j * korder + i]   (column-major, ld = korder)
  KIND_DENSE_T = 7,      // H(i,j) = p0[i * korder + j]   (the transpose of a column-major matrix)
};

// Generators of one structured operand.  Pointers are device pointers.
//   Cauchy:     p0 = D, p1 = F, p2 = U, p3 = V                  (columns of B(LDB,4))
//   Toeplitz:   p0 = G (length 2K-1), korder = K
//   Hankel:     p0 = G (length >= rows + cols - 1)
//   CauchyEig:  p0 = F (old eigenvalues, poles), p1 = U, p2 = V, p3 = DIFL, p4 = DIFR
struct GenPtrs {
  const double* p0;
  const double* p1;
  const double* p2;
  const double* p3;
  const double* p4;
  int korder;  // order K of the Toeplitz matrix (index shift); unused otherwise
  int nrows;   // global row count of H (bounds for generator reads)
  int ncols;   // global column count of H
  // Index-list blocks (ring steps): 1-based global rows / columns of the block,
  // as ConstIndex lists them; null for contiguous blocks at (r0, c0).
  const int* ridx = nullptr;
  const int* cidx = nullptr;
};

// Entry H(r, c) of the structured matrix, 0-based global indices.
template <int KIND>
__device__ __forceinline__ double gen_entry(const GenPtrs& g, int r, int c);

// SRC/Cauchylowrank.f90:509-521 (Cauchylike), via SRC/Auxil_2p5D.f90:138-144 (ConstCauchy).
template <>
__device__ __forceinline__ double gen_entry<KIND_CAUCHY>(const GenPtrs& g, int r, int c) {
  return (g.p2[r] * g.p3[c]) / (g.p0[r] - g.p1[c]);
}

// SRC/Auxil_2p5D.f90:224-231 (ConstToeplitz): T(i,j) = G(K+J1-I1), 1-based.
template <>
__device__ __forceinline__ double gen_entry<KIND_TOEPLITZ>(const GenPtrs& g, int r, int c) {
  return g.p0[g.korder - 1 + c - r];
}

// SRC/Auxil_2p5D.f90:309-316 (ConstHankel): T(i,j) = G(I1+J1-1), 1-based.
template <>
__device__ __forceinline__ double gen_entry<KIND_HANKEL>(const GenPtrs& g, int r, int c) {
  return g.p0[r + c];
}

// SRC/CauchyHssEig.f90:1758-1770 (CauchylikeEig) with PL = r+1, PU = c+1:
// the pole difference d_r - lambda_c is formed as (F_r - F_c) - DIFL_c when
// r <= c and as (F_r - F_{c+1}) - DIFR_c otherwise, never from lambda itself.
template <>
__device__ __forceinline__ double gen_entry<KIND_CAUCHYEIG>(const GenPtrs& g, int r, int c) {
  const double* F = g.p0;
  double den = (r <= c) ? ((F[r] - F[c]) - g.p3[c]) : ((F[r] - F[c + 1]) - g.p4[c]);
  return (g.p1[r] * g.p2[c]) / den;
}

template <>
__device__ __forceinline__ double gen_entry<KIND_CAUCHYEIG_T>(const GenPtrs& g, int r, int c) {
  return gen_entry<KIND_CAUCHYEIG>(g, c, r);
}
template <>
__device__ __forceinline__ double gen_entry<KIND_CAUCHY_T>(const GenPtrs& g, int r, int c) {
  return gen_entry<KIND_CAUCHY>(g, c, r);
}
template <>
__device__ __forceinline__ double gen_entry<KIND_DENSE>(const GenPtrs& g, int r, int c) {
  return g.p0[(size_t)c * g.korder + r];
}
template <>
__device__ __forceinline__ double gen_entry<KIND_DENSE_T>(const GenPtrs& g, int r, int c) {
  return g.p0[(size_t)r * g.korder + c];
}

__device__ __forceinline__ double gen_entry_rt(int kind, const GenPtrs& g, int r, int c) {
  switch (kind) {
    case KIND_CAUCHY: return gen_entry<KIND_CAUCHY>(g, r, c);
    case KIND_TOEPLITZ: return gen_entry<KIND_TOEPLITZ>(g, r, c);
    case KIND_HANKEL: return gen_entry<KIND_HANKEL>(g, r, c);
    default: return gen_entry<KIND_CAUCHYEIG>(g, r, c);
  }
}

// Remap a linear workgroup id so that consecutive output tiles land on the same
// XCD.  Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; tiles
// in one row of tiles share the same rows of the dense operand, so keeping them
// on one XCD keeps those rows in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  constexpr int NXCD = 8;
  if (nblocks % NXCD != 0) return bid;
  int per = nblocks / NXCD;
  return (bid % NXCD) * per + bid / NXCD;
}

// Patch rasterisation: workgroup b runs on XCD b % 8 as that XCD's (b / 8)-th
// workgroup.  Consecutive local indices of one XCD are laid out over a PR x PC
// patch of output tiles (patches dealt round-robin to the XCDs), so the
// workgroups an XCD runs at the same time share PR row blocks of the first
// operand and PC column blocks of the second in its L2.  Returns (bm, bn).
template <int PR, int PC>
__device__ __forceinline__ void patch_remap(int bid, int nbm, int nbn, int& bm, int& bn) {
  constexpr int NXCD = 8, P = PR * PC;
  const int npr = nbm / PR, npc = nbn / PC;
  const int npatch = npr * npc;
  if (nbm % PR == 0 && nbn % PC == 0 && npatch % NXCD == 0) {
    int x = bid % NXCD, li = bid / NXCD;
    int q = (li / P) * NXCD + x;  // global patch id
    int p = li % P;
    bm = (q / npc) * PR + p / PC;
    bn = (q % npc) * PC + p % PC;
  } else {
    int t = xcd_remap(bid, nbm * nbn);
    bm = t / nbn;
    bn = t % nbn;
  }
}

}  // namespace psmma

namespace psmma {

// One problem of a batched launch (device-side table entry).
struct BatchProblem {
  int kind, m, n, k;
  double alpha, beta;
  const double* A;
  int lda;
  GenPtrs g;
  int r0, c0;
  double* C;
  int ldc;
  int tile_begin;  // first index of this problem's tiles in the batch tile list
  int ntm, ntn;    // tiles along m and n
  int cls;         // tile class: 0 small (64 x 64), 1 large (shape by kind)
  // Optional explicit tail of the contraction (generated kinds of the HSS kind set):
  // C = alpha * (A * H + A2 * H2) + beta * C with A2 m x k2 (ld lda2) and
  // H2(k, c) = H2[k * ldh2 + c]; k must then be a multiple of the tile's slab depth.
  const double* A2;
  int lda2;
  const double* H2;
  int ldh2;
  int k2;
  // Optional column map of C (tail form only): output column c of the product is
  // stored in column cmap[c] of C.
  const int* cmap;
};

// The explicit tail segment as the tile body takes it.
struct TailSeg {
  const double* A = nullptr;
  int lda = 1;
  const double* H = nullptr;
  int ldh = 1;
  int K = 0;
  const int* cmap = nullptr;  // column map of C (BatchProblem::cmap)
};

}  // namespace psmma
