// Plain FP64 structured x dense products on gfx950 FP64 MFMA.
//
//   right-structured (psmma_cauchy/toeplitz/hankel/cauchyeig):
//       C(MxN) = alpha * A(MxK) * H(KxN) + beta * C,   H generated from generators
//   left-structured  (pzmdftL):
//       C(MxN) = alpha * F(MxK) * B(KxN) + beta * C,   F(r,c) = w^(r*c mod Nd), complex
//
// Both replace the reference's "build tile into WORK(IPB), then DGEMM/ZGEMM"
// step (SRC/pscauchy_compute.f90:395-407, SRC/pstoeplitz_compute.f90:290-293,
// SRC/pshankel_compute.f90:249-250, SRC/pscauchyeig_compute.f90:399-402,
// SRC/pzmdftL1.f90:490-496) with one kernel in which the structured tile only
// ever exists in LDS.
//
// Tiling: a workgroup of NW = WGM*WGN wave64s owns a BM x BN tile of C.  Each
// wave owns a (BM/WGM) x (BN/WGN) sub-tile held in v_mfma_f64_16x16x4f64
// accumulators.  Per BK-slab of the contraction the dense operand is loaded
// global->registers (prefetched one slab ahead), the structured operand is
// evaluated from its generators, and both are written to a double-buffered LDS
// pair; one barrier per slab.
#include <cstdlib>
#include <type_traits>

#include "psmma_common.h"

namespace psmma {

// MFMA f64 16x16x4 operand/result lane layout used throughout:
//   A fragment: lane l supplies A[l%16][k-slot l/16]
//   B fragment: lane l supplies B[k-slot l/16][l%16]
//   D:          lane l, element r holds D[l/16 + 4*r][l%16]
// (measured on gfx950; the bf16/int8 16x16 forms instead hold D[4*(l/16)+r][l%16]).
// (The k-slot assignment only has to agree between A and B.)

// LDS footprint (doubles) of one right-structured tile configuration.
template <int KIND, int BM, int BN, int BK, bool IDX = false>
struct RightLds {
  static constexpr bool WIN = (KIND == KIND_TOEPLITZ || KIND == KIND_HANKEL) && !IDX;
  static constexpr int ALD = BM + 16, HLD = BN + 16;
  static constexpr int ASZ = BK * ALD;
  static constexpr int HSZ = WIN ? BK + BN : BK * HLD;
};

// One BM x BN output tile at (m0, n0) of C = alpha * A * H(r0:, c0:) + beta * C,
// or with IDX of C = alpha * A * H(ridx, cidx) + beta * C (the rows and columns of
// H given by the 1-based index lists of g; r0, c0 unused).
// As/Hs: LDS areas of 2 * RightLds::ASZ and 2 * RightLds::HSZ doubles.  Ends
// with all waves past a barrier, so the caller may reuse the LDS at once.
// With TAIL (generated kinds, K a multiple of BK) the contraction continues past K
// with the explicit segment tl: C = alpha * (A * H + tl.A * tl.H) + beta * C.
// FAST: interior slabs of the generated kinds skip the bounds selects (more code and
// registers; the single-problem kernels use it).
template <int KIND, int BM, int BN, int BK, int WGM, int WGN, bool IDX = false, bool TAIL = false,
          bool FAST = false>
__device__ __forceinline__ void right_tile(int M, int N, int K, double alpha,
                                           const double* __restrict__ A, int lda, const GenPtrs& g,
                                           int r0, int c0, double beta, double* __restrict__ C,
                                           int ldc, int m0, int n0, double* __restrict__ As_,
                                           double* __restrict__ Hs_, const TailSeg& tl = TailSeg{}) {
  constexpr int NT = WGM * WGN * 64;
  constexpr int WM = BM / WGM, WN = BN / WGN;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int ALD = BM + 16, HLD = BN + 16;
  constexpr int AREG = BM * BK / NT;  // dense elements staged per thread per slab
  // Toeplitz and Hankel tiles are functions of (col - row) / (row + col): one
  // slab of a contiguous block needs only a window of BK+BN-1 generator values.
  // Index-list blocks generate every entry, as the Cauchy kinds do.
  constexpr bool WIN = (KIND == KIND_TOEPLITZ || KIND == KIND_HANKEL) && !IDX;
  // every entry gathered through gen_entry: index-list Toeplitz / Hankel blocks and the
  // internal transposed and explicit kinds
  constexpr bool GATHER = ((KIND == KIND_TOEPLITZ || KIND == KIND_HANKEL) && IDX) ||
                         KIND == KIND_DENSE || KIND == KIND_DENSE_T;
  // transposed Cauchy kinds: H(k, c) = C(c, k), the row generators are C's column ones
  constexpr bool TRANS = KIND == KIND_CAUCHYEIG_T || KIND == KIND_CAUCHY_T;
  static_assert(!(TRANS && IDX), "no index lists for the transposed kinds");
  constexpr int WLEN = BK + BN - 1;
  constexpr int WREG = (WLEN + NT - 1) / NT;
  constexpr int HREG = WIN ? 1 : BK * BN / NT;  // generated entries per thread per slab
  constexpr int ASZ = RightLds<KIND, BM, BN, BK, IDX>::ASZ;
  constexpr int HSZ = RightLds<KIND, BM, BN, BK, IDX>::HSZ;
  static_assert(BM * BK % NT == 0 && BK * BN % NT == 0, "tile/threads mismatch");
  static_assert(WIN || NT % BN == 0, "generated tiles assume a fixed column per thread");
  static_assert(WM % 16 == 0 && WN % 16 == 0 && BK % 4 == 0, "bad wave tile");
  static_assert(!TAIL || (!WIN && !GATHER && !IDX), "the explicit tail follows a generated segment");

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / WGN) * WM, wn0 = (wave % WGN) * WN;
  const int lr = lane & 15, lk = lane >> 4;

  v4d acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};

  double areg[AREG];
  double wreg[WREG];     // window values (Toeplitz/Hankel)
  double rowa[HREG], rowb[HREG];  // row generators (Cauchy kinds)
  double hreg[HREG];
  int growv[HREG];  // global rows of the slab (eigen-Cauchy index lists)
  // transposed eigen-Cauchy: F_{k+1}, DIFL_k, DIFR_k of the slab's rows
  double rowc[KIND == KIND_CAUCHYEIG_T ? HREG : 1], rowd[KIND == KIND_CAUCHYEIG_T ? HREG : 1],
      rowe[KIND == KIND_CAUCHYEIG_T ? HREG : 1];

  // Generator index range of the whole block: window slots outside it are only
  // used by padded (row, col) pairs and are set to 0 so 0 * value stays 0.
  long long glo = 0, ghi = -1;
  if constexpr (KIND == KIND_TOEPLITZ) {
    glo = (long long)g.korder - 1 + c0 - (r0 + K - 1);
    ghi = (long long)g.korder - 1 + (c0 + N - 1) - r0;
  } else if constexpr (KIND == KIND_HANKEL) {
    glo = (long long)r0 + c0;
    ghi = (long long)r0 + K - 1 + c0 + N - 1;
  }

  // Column invariants of the Cauchy kinds: this thread always generates column
  // hcol of every slab.
  const int hcol = tid % BN;
  const bool colok = (n0 + hcol) < N;
  int gcol = c0 + n0 + hcol;
  if constexpr (IDX) gcol = colok ? g.cidx[n0 + hcol] - 1 : 0;
  // (a problem with an empty generated segment, K = 0, is a plain product on its
  // explicit tail: its generators are never read)
  const bool genok = colok && K > 0;
  double cv0 = 0.0, cv1 = 0.0, cv2 = 0.0, cv3 = 0.0, cv4 = 0.0;
  if constexpr (KIND == KIND_CAUCHY) {
    if (genok) {
      cv0 = g.p1[gcol];  // F_c
      cv1 = g.p3[gcol];  // V_c
    }
  } else if constexpr (KIND == KIND_CAUCHY_T) {
    if (genok) {
      cv0 = g.p0[gcol];  // D of C's row c
      cv1 = g.p2[gcol];  // U of C's row c
    }
  } else if constexpr (KIND == KIND_CAUCHYEIG_T) {
    if (genok) {
      cv0 = g.p0[gcol];  // F of C's row c
      cv1 = g.p1[gcol];  // U of C's row c
    }
  } else if constexpr (KIND == KIND_CAUCHYEIG) {
    if (genok) {
      cv0 = g.p0[gcol];                                   // F_c
      cv1 = (gcol + 1 < g.korder) ? g.p0[gcol + 1] : 0.0;  // F_{c+1}
      cv2 = g.p2[gcol];                                   // V_c
      cv3 = g.p3[gcol];                                   // DIFL_c
      cv4 = g.p4[gcol];                                   // DIFR_c
    }
  }

  // Generated kinds: a slab that lies wholly inside the block (every row, column and
  // k index in range) is loaded and evaluated without bounds selects, which the
  // compiler turns into an exec-masked branch per load.
  constexpr bool FASTPATH = FAST && !WIN && !GATHER;
  const bool tile_full = FASTPATH && m0 + BM <= M && n0 + BN <= N;

  auto load_a = [&](int k0) {
    if constexpr (FASTPATH) {
      if (tile_full && k0 + BK <= K) {  // uniform
        const double* Ak = A + (size_t)k0 * lda + m0;
#pragma unroll
        for (int q = 0; q < AREG; ++q) {
          int idx = tid + q * NT;
          areg[q] = Ak[(size_t)(idx / BM) * lda + idx % BM];
        }
        return;
      }
    }
    if constexpr (TAIL) {
      if (k0 >= K) {  // slab of the explicit tail (uniform)
        const int t0 = k0 - K;
#pragma unroll
        for (int q = 0; q < AREG; ++q) {
          int idx = tid + q * NT;
          int row = idx % BM, col = idx / BM;
          int gr = m0 + row, gc = t0 + col;
          areg[q] = (gr < M && gc < tl.K) ? tl.A[(size_t)gc * tl.lda + gr] : 0.0;
        }
        return;
      }
    }
#pragma unroll
    for (int q = 0; q < AREG; ++q) {
      int idx = tid + q * NT;
      int row = idx % BM, col = idx / BM;
      int gr = m0 + row, gc = k0 + col;
      areg[q] = (gr < M && gc < K) ? A[(size_t)gc * lda + gr] : 0.0;
    }
  };
  // Issue the generator loads of slab k0 (consumed after the MFMA loop).
  auto fetch_h = [&](int k0) {
    if constexpr (TAIL) {
      if (k0 >= K) {  // explicit tail: H2 rows k0 - K.., this thread's column
#pragma unroll
        for (int q = 0; q < HREG; ++q) {
          int gk = k0 - K + tid / BN + q * (NT / BN);
          hreg[q] = (gk < tl.K && colok) ? tl.H[(size_t)gk * tl.ldh + (n0 + hcol)] : 0.0;
        }
        return;
      }
    }
    if constexpr (FASTPATH && !IDX) {
      if (tile_full && k0 + BK <= K) {  // uniform: every row of the slab is in range
#pragma unroll
        for (int q = 0; q < HREG; ++q) {
          const int gr = r0 + k0 + tid / BN + q * (NT / BN);
          if constexpr (KIND == KIND_CAUCHY) {
            rowa[q] = g.p0[gr];  // D_r
            rowb[q] = g.p2[gr];  // U_r
          } else if constexpr (KIND == KIND_CAUCHY_T) {
            rowa[q] = g.p1[gr];
            rowb[q] = g.p3[gr];
          } else if constexpr (KIND == KIND_CAUCHYEIG_T) {
            rowa[q] = g.p0[gr];
            rowb[q] = g.p2[gr];
            rowc[q] = (gr + 1 < g.korder) ? g.p0[gr + 1] : 0.0;
            rowd[q] = g.p3[gr];
            rowe[q] = g.p4[gr];
          } else {
            rowa[q] = g.p0[gr];  // F_r
            rowb[q] = g.p1[gr];  // U_r
          }
        }
        return;
      }
    }
    if constexpr (WIN) {
      long long base;
      if constexpr (KIND == KIND_TOEPLITZ)
        base = (long long)g.korder - 1 + (c0 + n0) - (r0 + k0) - (BK - 1);
      else
        base = (long long)(r0 + k0) + (c0 + n0);
#pragma unroll
      for (int q = 0; q < WREG; ++q) {
        int t = tid + q * NT;
        long long gi = base + t;
        wreg[q] = (t < WLEN && gi >= glo && gi <= ghi) ? g.p0[gi] : 0.0;
      }
    } else {
#pragma unroll
      for (int q = 0; q < HREG; ++q) {
        int row = tid / BN + q * (NT / BN);
        int gk = k0 + row;
        if (gk < K) {
          int gr = r0 + gk;
          if constexpr (IDX) gr = g.ridx[gk] - 1;
          if constexpr (GATHER) {
            // gathered kinds: each entry straight from gen_entry
            hreg[q] = colok ? gen_entry<KIND>(g, gr, gcol) : 0.0;
          } else if constexpr (KIND == KIND_CAUCHY) {
            rowa[q] = g.p0[gr];  // D_r
            rowb[q] = g.p2[gr];  // U_r
          } else if constexpr (KIND == KIND_CAUCHY_T) {
            rowa[q] = g.p1[gr];  // F of C's column k
            rowb[q] = g.p3[gr];  // V of C's column k
          } else if constexpr (KIND == KIND_CAUCHYEIG_T) {
            rowa[q] = g.p0[gr];                                  // F_k
            rowb[q] = g.p2[gr];                                  // V_k
            rowc[q] = (gr + 1 < g.korder) ? g.p0[gr + 1] : 0.0;  // F_{k+1}
            rowd[q] = g.p3[gr];                                  // DIFL_k
            rowe[q] = g.p4[gr];                                  // DIFR_k
          } else {
            rowa[q] = g.p0[gr];  // F_r
            rowb[q] = g.p1[gr];  // U_r
          }
          if constexpr (KIND == KIND_CAUCHYEIG) growv[q] = gr;
        } else {
          if constexpr (GATHER) hreg[q] = 0.0;
          rowa[q] = 0.0;
          rowb[q] = 0.0;
        }
      }
    }
  };
  // Evaluate the prefetched slab (the FP64 divisions of the Cauchy kinds).
  auto finish_h = [&](int k0) {
    if constexpr (TAIL) {
      if (k0 >= K) return;  // the tail's entries were loaded as they are
    }
    if constexpr (!WIN && !GATHER) {
      // entry q of this thread's column in the slab (row k0 + gk_off)
      auto eval = [&](int q, int gk) -> double {
        if constexpr (KIND == KIND_CAUCHY) {
          return (rowb[q] * cv1) / (rowa[q] - cv0);
        } else if constexpr (KIND == KIND_CAUCHY_T) {
          return (cv1 * rowb[q]) / (cv0 - rowa[q]);  // C(c, k) = U_c V_k / (D_c - F_k)
        } else if constexpr (KIND == KIND_CAUCHYEIG_T) {
          // C(c, k), c = gcol, k = gr: the eigen-Cauchy form with its roles swapped
          const int gr = r0 + gk;
          const double den = (gcol <= gr) ? ((cv0 - rowa[q]) - rowd[q])
                                          : ((cv0 - rowc[q]) - rowe[q]);
          return (cv1 * rowb[q]) / den;
        } else {
          // SRC/CauchyHssEig.f90:1764-1768: d_r - lambda_c from stored differences
          int gr = IDX ? growv[q] : r0 + gk;
          double den = (gr <= gcol) ? ((rowa[q] - cv0) - cv3) : ((rowa[q] - cv1) - cv4);
          return (rowb[q] * cv2) / den;
        }
      };
      if (tile_full && k0 + BK <= K) {  // uniform: no bounds selects
#pragma unroll
        for (int q = 0; q < HREG; ++q) hreg[q] = eval(q, k0 + tid / BN + q * (NT / BN));
        return;
      }
#pragma unroll
      for (int q = 0; q < HREG; ++q) {
        int row = tid / BN + q * (NT / BN);
        int gk = k0 + row;
        hreg[q] = (gk < K && colok) ? eval(q, gk) : 0.0;
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < AREG; ++q) {
      int idx = tid + q * NT;
      As_[buf * ASZ + (idx / BM) * ALD + idx % BM] = areg[q];
    }
    if constexpr (WIN) {
#pragma unroll
      for (int q = 0; q < WREG; ++q) {
        int t = tid + q * NT;
        if (t < WLEN) Hs_[buf * HSZ + t] = wreg[q];
      }
    } else {
#pragma unroll
      for (int q = 0; q < HREG; ++q) {
        int row = tid / BN + q * (NT / BN);
        Hs_[buf * HSZ + row * HLD + hcol] = hreg[q];
      }
    }
  };

  const int nkt = (K + BK - 1) / BK + (TAIL ? (tl.K + BK - 1) / BK : 0);
  if (nkt > 0) {
    load_a(0);
    fetch_h(0);
    finish_h(0);
    store(0);
  }
  __syncthreads();

  // MFMAs over k rows [KK0, KK1) of the current slab.  Called once over the whole
  // slab; kept as a lambda because the compiler schedules it differently from the
  // same loop written inline, and this form measured faster (Toeplitz N = 8192:
  // 70.3 against 69.7 TF/s).
  auto mfma_rows = [&](const double* as, const double* hs, auto k0c, auto k1c) {
    constexpr int KK0 = decltype(k0c)::value, KK1 = decltype(k1c)::value;
#pragma unroll
    for (int kk = KK0; kk < KK1; kk += 4) {
      double af[TM], bf[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) af[i] = as[(kk + lk) * ALD + wm0 + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        int jj = wn0 + j * 16 + lr, kr = kk + lk;
        if constexpr (WIN && KIND == KIND_TOEPLITZ)
          bf[j] = hs[(BK - 1) - kr + jj];
        else if constexpr (WIN && KIND == KIND_HANKEL)
          bf[j] = hs[kr + jj];
        else
          bf[j] = hs[kr * HLD + jj];
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
  };
  for (int kt = 0; kt < nkt; ++kt) {
    const int cur = kt & 1;
    const bool more = (kt + 1) < nkt;
    if (more) {
      load_a((kt + 1) * BK);
      fetch_h((kt + 1) * BK);
    }
    const double* as = As_ + cur * ASZ;
    const double* hs = Hs_ + cur * HSZ;
    mfma_rows(as, hs, std::integral_constant<int, 0>{}, std::integral_constant<int, BK>{});
    // Window kinds: the next slab is stored branch-free (after the last slab it
    // rewrites the idle buffer, whose last reads finished before the previous
    // barrier, with values nobody reads) and a scheduling fence keeps the stores
    // ahead of the barrier: with BK = 32, Toeplitz N = 8192 70.5 TF/s on MI355X
    // (70.1 with a second fence between two halves of the slab's MFMAs, which
    // helped at BK = 16).  The Cauchy kinds keep the plain order (no gain there).
    if (WIN || more) {
      finish_h((kt + 1) * BK);
      store(cur ^ 1);
    }
    if constexpr (WIN) __builtin_amdgcn_sched_barrier(0);
    __syncthreads();
  }

  // Epilogue: C = alpha*acc + beta*C.  C is not read when beta == 0 (BLAS rule).
  // (The tail form may store output column c in column cmap[c] of C.)
  int oc[TN];
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int gc = n0 + wn0 + j * 16 + lr;
    oc[j] = gc;
    if constexpr (TAIL) {
      if (tl.cmap && gc < N) oc[j] = tl.cmap[gc];
    }
  }
#pragma unroll
  for (int i = 0; i < TM; ++i) {
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      int gc = n0 + wn0 + j * 16 + lr;
      if (gc >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int gr = m0 + wm0 + i * 16 + lk + 4 * r;
        if (gr < M) {
          double* p = C + (size_t)oc[j] * ldc + gr;
          double v = alpha * acc[i][j][r];
          if (beta != 0.0) v += beta * (*p);
          *p = v;
        }
      }
    }
  }
}

template <int KIND, int BM, int BN, int BK, int WGM, int WGN, int MINB = 2>
__global__ __launch_bounds__(WGM* WGN * 64, MINB) void sgemm_f64_right_kernel(
    int M, int N, int K, double alpha, const double* __restrict__ A, int lda, GenPtrs g, int r0,
    int c0, double beta, double* __restrict__ C, int ldc) {
  using L = RightLds<KIND, BM, BN, BK>;
  __shared__ double As[2 * L::ASZ];
  __shared__ double Hs[2 * L::HSZ];
  const int nbn = (N + BN - 1) / BN;
  const int nbm = (M + BM - 1) / BM;
  const int bid = xcd_remap(blockIdx.x, nbn * nbm);
  // the guard-free interior path measured 0.409 -> 0.385 ms for Cauchy N = 2048 and
  // 166.0 -> 163.6 ms for the eigen-Cauchy N = 16384 product on MI355X; the 16-wave
  // shapes (128 registers) spill with it and keep the guarded form
  right_tile<KIND, BM, BN, BK, WGM, WGN, false, false, (WGM * WGN <= 8)>(
      M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc, (bid / nbn) * BM, (bid % nbn) * BN, As, Hs);
}

// One ring step of PS*_COMPUTE on the rank's own index sets: C = alpha * A *
// H(RIndex, CIndex) + beta * C with the 1-based global index lists of g (ConstIndex,
// SRC/Auxil_2p5D.f90:2-66, consumed by ConstCauchy / ConstToeplitz / ConstHankel /
// ConstCauchyEig, SRC/Auxil_2p5D.f90:68-322, :532-637).  The tile body of the
// contiguous kernel with every entry of the slab generated through the lists.
template <int KIND, int BM, int BN, int BK, int WGM, int WGN>
__global__ __launch_bounds__(WGM* WGN * 64, 2) void sgemm_f64_indexed_kernel(
    int M, int N, int K, double alpha, const double* __restrict__ A, int lda, GenPtrs g,
    double beta, double* __restrict__ C, int ldc) {
  using L = RightLds<KIND, BM, BN, BK, true>;
  __shared__ double As[2 * L::ASZ];
  __shared__ double Hs[2 * L::HSZ];
  const int nbn = (N + BN - 1) / BN;
  const int nbm = (M + BM - 1) / BM;
  const int bid = xcd_remap(blockIdx.x, nbn * nbm);
  right_tile<KIND, BM, BN, BK, WGM, WGN, true>(M, N, K, alpha, A, lda, g, 0, 0, beta, C, ldc,
                                               (bid / nbn) * BM, (bid % nbn) * BN, As, Hs);
}

// ---------------------------------------------------------------------------
// Left-structured complex DFT product.  F(r,c) = w^((r0+r)*(c0+c) mod Nd) with
// w = exp(-2*pi*i/Nd) (SRC/pzmdftL1.f90:329-337, unnormalised); the twiddles
// come from an FP64 table tw[e] = (cos(2*pi*e/Nd), -sin(2*pi*e/Nd)) so every
// entry is correctly rounded instead of accumulating the error of Wn**PowK.
// B and C are complex128, interleaved (re, im), column-major, ld in complex
// elements.  alpha and beta are real, as in the reference's argument list.
// The complex product runs in the three-product form (Gauss): P1 = sum Fr Br,
// P2 = sum Fi Bi, P3 = sum (Fr + Fi)(Br + Bi), Re = P1 - P2, Im = P3 - P1 - P2:
// three real MFMAs per complex multiply-add instead of four.  Measured on MI355X,
// F(4096 x 4096) B(4096 x 512): 1.185 -> 1.051 ms, 3.03e-15 -> 3.04e-15 against
// the FFT of the same columns.
template <int BM, int BN, int BK, int WGM, int WGN>
__global__ __launch_bounds__(WGM* WGN * 64, 2) void dft_f64_left_kernel(
    int M, int N, int K, int Nd, double alpha, const double2* __restrict__ tw,
    const double2* __restrict__ B, int ldb, int r0, int c0, double beta, double2* __restrict__ C,
    int ldc) {
  constexpr int NT = WGM * WGN * 64;
  constexpr int WM = BM / WGM, WN = BN / WGN;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int PAD = 16;
  constexpr int FLD = BM + PAD, BLD = BN + PAD;
  constexpr int FREG = BM * BK / NT;
  constexpr int BREG = BK * BN / NT;
  static_assert(BM * BK % NT == 0 && BK * BN % NT == 0, "tile/threads mismatch");
  static_assert(NT % BM == 0, "F generator assumes fixed row per thread");
  static_assert(BK % 8 == 0, "the slab runs as two halves of whole 4-deep MFMA steps");

  __shared__ double Fr[2][BK * FLD];
  __shared__ double Fi[2][BK * FLD];
  __shared__ double Br[2][BK * BLD];
  __shared__ double Bi[2][BK * BLD];

  const int nbn = (N + BN - 1) / BN;
  const int nbm = (M + BM - 1) / BM;
  const int bid = xcd_remap(blockIdx.x, nbn * nbm);
  const int bm = bid / nbn, bn = bid % nbn;
  const int m0 = bm * BM, n0 = bn * BN;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / WGN) * WM, wn0 = (wave % WGN) * WN;
  const int lr = lane & 15, lk = lane >> 4;

  // P1, P2, P3 of the three-product form
  v4d accr[TM][TN], acci[TM][TN], accs[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      accr[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
      acci[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
      accs[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
    }

  // Each thread generates F for one fixed row (tid % BM) and FREG columns
  // (tid / BM + q * NT/BM).  It keeps the exponents e_q = row*col mod Nd and
  // advances them by row*BK mod Nd per slab with one add and one compare.
  const int frow = tid % BM;
  const int grow = m0 + frow;
  const unsigned nd = (unsigned)Nd;
  const unsigned rmod = (unsigned)(((long long)(r0 + grow) % Nd + Nd) % Nd);
  const unsigned step = (unsigned)(((unsigned long long)rmod * BK) % nd);
  unsigned ex[FREG];
#pragma unroll
  for (int q = 0; q < FREG; ++q) {
    int col = tid / BM + q * (NT / BM);
    ex[q] = (unsigned)(((unsigned long long)rmod * (unsigned long long)(c0 + col)) % nd);
  }

  double2 freg[FREG];
  double2 breg[BREG];
  auto gen_f = [&](int k0) {
#pragma unroll
    for (int q = 0; q < FREG; ++q) {
      int col = tid / BM + q * (NT / BM);
      bool ok = (grow < M) && (k0 + col < K);
      freg[q] = ok ? tw[ex[q]] : make_double2(0.0, 0.0);
    }
  };
  auto advance_f = [&]() {
#pragma unroll
    for (int q = 0; q < FREG; ++q) {
      unsigned e = ex[q] + step;
      ex[q] = (e >= nd) ? e - nd : e;
    }
  };
  auto load_b = [&](int k0) {
#pragma unroll
    for (int q = 0; q < BREG; ++q) {
      int idx = tid + q * NT;
      int row = idx % BK, col = idx / BK;  // consecutive threads walk down a column
      int gk = k0 + row, gn = n0 + col;
      breg[q] = (gk < K && gn < N) ? B[(size_t)gn * ldb + gk] : make_double2(0.0, 0.0);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < FREG; ++q) {
      int col = tid / BM + q * (NT / BM);
      Fr[buf][col * FLD + frow] = freg[q].x;
      Fi[buf][col * FLD + frow] = freg[q].y;
    }
#pragma unroll
    for (int q = 0; q < BREG; ++q) {
      int idx = tid + q * NT;
      int row = idx % BK, col = idx / BK;
      Br[buf][row * BLD + col] = breg[q].x;
      Bi[buf][row * BLD + col] = breg[q].y;
    }
  };

  const int nkt = (K + BK - 1) / BK;
  if (nkt > 0) {
    gen_f(0);
    load_b(0);
    store(0);
  }
  __syncthreads();

  for (int kt = 0; kt < nkt; ++kt) {
    const int cur = kt & 1;
    const bool more = (kt + 1) < nkt;
    if (more) {
      advance_f();
      gen_f((kt + 1) * BK);
      load_b((kt + 1) * BK);
    }
    // MFMAs over k rows [KK0, KK1) of the current slab
    auto rows = [&](auto k0c, auto k1c) {
      constexpr int KK0 = decltype(k0c)::value, KK1 = decltype(k1c)::value;
#pragma unroll
      for (int kk = KK0; kk < KK1; kk += 4) {
        double ar[TM], ai[TM], br[TN], bi[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          ar[i] = Fr[cur][(kk + lk) * FLD + wm0 + i * 16 + lr];
          ai[i] = Fi[cur][(kk + lk) * FLD + wm0 + i * 16 + lr];
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          br[j] = Br[cur][(kk + lk) * BLD + wn0 + j * 16 + lr];
          bi[j] = Bi[cur][(kk + lk) * BLD + wn0 + j * 16 + lr];
        }
        double as[TM], bs[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) as[i] = ar[i] + ai[i];
#pragma unroll
        for (int j = 0; j < TN; ++j) bs[j] = br[j] + bi[j];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            accr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], accr[i][j], 0, 0, 0);
            acci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], bi[j], acci[i][j], 0, 0, 0);
            accs[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[i], bs[j], accs[i][j], 0, 0, 0);
          }
      }
    };
    rows(std::integral_constant<int, 0>{}, std::integral_constant<int, BK / 2>{});
    // A scheduling fence between the two halves and a branch-free store of the
    // next slab (after the last slab it rewrites the idle buffer, whose last
    // reads finished before the previous barrier), measured on MI355X in
    // isolation: M = K = 4096, N = 512 1.227 -> 1.183 ms; M = N = 4096, K = 512
    // 1.31 -> 1.23 ms.  The fence is what matters.
    __builtin_amdgcn_sched_barrier(0);
    rows(std::integral_constant<int, BK / 2>{}, std::integral_constant<int, BK>{});
    store(cur ^ 1);
    __builtin_amdgcn_sched_barrier(0);
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < TM; ++i) {
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      int gc = n0 + wn0 + j * 16 + lr;
      if (gc >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int gr = m0 + wm0 + i * 16 + lk + 4 * r;
        if (gr < M) {
          double2* p = C + (size_t)gc * ldc + gr;
          const double re = accr[i][j][r] - acci[i][j][r];
          const double im = accs[i][j][r] - accr[i][j][r] - acci[i][j][r];
          double2 v = make_double2(alpha * re, alpha * im);
          if (beta != 0.0) {
            double2 o = *p;
            v.x += beta * o.x;
            v.y += beta * o.y;
          }
          *p = v;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// One persistent launch for a batch of uneven structured products of every
// kind.  The tiles of all problems form one list: problems ordered by the cost
// of one of their tiles (K x tile area, longest first), each problem's tiles
// contiguous.  Workgroups take tiles from an atomic counter until the list is
// exhausted, find the tile's problem by binary search over the tile prefix
// array and run the single-problem tile body for its (kind, tile class); the
// last workgroup to leave resets the counters, so the next launch (a graph
// replay) starts from zero without a memset.
//
// Tile classes: 0 small (64 x 64, BK = 16, 4 waves of 32 x 32); 1 large, by kind:
// 64 x 256 for Toeplitz / Hankel (1 x 4 waves), 128 x 128 for the Cauchy kinds
// (2 x 2 waves), both with BK = 16 / 8 so that all eight tile bodies inline into
// one kernel of 256 VGPRs with 10 spilled (the single-problem depths, 32 / 16,
// spill 110 there; calling the large bodies instead of inlining them passes
// their uniform arguments in VGPRs).  Measured on MI355X, 512-problem batch:
// 19.4 ms against 20.2 ms for one launch per (kind, class) and 22.5 ms with the
// large bodies called.  All classes use 256 threads.
// Tile shapes of the HSS kind set (the Cauchy kinds, their transposes and
// explicit matrices; short contractions of K = leaf or a rank): class 1 for
// problems of at least 256 rows and columns, class 0 below.
struct HssTiles {
  static constexpr int BM1 = 128, BN1 = 128, BK1 = 8, WM1 = 2, WN1 = 2;
  static constexpr int BM0 = 64, BN0 = 64, BK0 = 16, WM0 = 2, WN0 = 2;
  static constexpr int MIN_BLOCKS = 2;  // workgroups per CU the register budget allows
};

// Tile shape of the leaf kind set (the HSS product's last stage: generated diagonal
// blocks with the explicit F V^T tail): the single-problem kernel's eigen-Cauchy
// shape, 256 x 128 with 8 waves of 64 x 64 (one workgroup per CU), for every problem
// of the set.
struct LeafTiles {
  static constexpr int BM = 256, BN = 128, BK = 16, WM = 4, WN = 2;
};

void batch_tile_shape(int kset, int kind, int cls, int* bm, int* bn) {
  if (kset == 2) {
    *bm = LeafTiles::BM;
    *bn = LeafTiles::BN;
  } else if (kset == 1) {
    *bm = cls == 1 ? HssTiles::BM1 : HssTiles::BM0;
    *bn = cls == 1 ? HssTiles::BN1 : HssTiles::BN0;
  } else if (cls == 1) {
    const bool win = kind == KIND_TOEPLITZ || kind == KIND_HANKEL;
    *bm = win ? 64 : 128;
    *bn = win ? 256 : 128;
  } else {
    *bm = 64;
    *bn = 64;
  }
}

template <int KIND, int BM, int BN, int BK>
constexpr int tile_lds() {
  return 2 * (RightLds<KIND, BM, BN, BK>::ASZ + RightLds<KIND, BM, BN, BK>::HSZ);
}
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int kBatchLds =
    cmax(cmax(cmax(tile_lds<KIND_TOEPLITZ, 64, 256, 16>(), tile_lds<KIND_HANKEL, 64, 256, 16>()),
              cmax(tile_lds<KIND_CAUCHY, 128, 128, 8>(), tile_lds<KIND_CAUCHYEIG, 128, 128, 8>())),
         cmax(cmax(tile_lds<KIND_CAUCHY, 64, 64, 16>(), tile_lds<KIND_CAUCHYEIG, 64, 64, 16>()),
              cmax(tile_lds<KIND_TOEPLITZ, 64, 64, 16>(), tile_lds<KIND_HANKEL, 64, 64, 16>())));
constexpr int kHssLds = cmax(
    cmax(cmax(tile_lds<KIND_DENSE, HssTiles::BM1, HssTiles::BN1, HssTiles::BK1>(),
              tile_lds<KIND_CAUCHYEIG, HssTiles::BM1, HssTiles::BN1, HssTiles::BK1>()),
         tile_lds<KIND_CAUCHY, HssTiles::BM1, HssTiles::BN1, HssTiles::BK1>()),
    cmax(cmax(tile_lds<KIND_DENSE, HssTiles::BM0, HssTiles::BN0, HssTiles::BK0>(),
              tile_lds<KIND_CAUCHYEIG, HssTiles::BM0, HssTiles::BN0, HssTiles::BK0>()),
         tile_lds<KIND_CAUCHY, HssTiles::BM0, HssTiles::BN0, HssTiles::BK0>()));

template <int KIND, int BM, int BN, int BK, int WGM, int WGN, bool TAIL = false>
__device__ __forceinline__ void batch_tile(const BatchProblem& p, int local, double* lds) {
  using L = RightLds<KIND, BM, BN, BK>;
  TailSeg tl;
  if constexpr (TAIL) tl = TailSeg{p.A2, p.lda2, p.H2, p.ldh2, p.k2, p.cmap};
  right_tile<KIND, BM, BN, BK, WGM, WGN, false, TAIL>(
      p.m, p.n, p.k, p.alpha, p.A, p.lda, p.g, p.r0, p.c0, p.beta, p.C, p.ldc,
      (local / p.ntn) * BM, (local % p.ntn) * BN, lds, lds + 2 * L::ASZ, tl);
}

template <int KIND>
__device__ __forceinline__ void batch_tile_kind(const BatchProblem& p, int local, double* lds) {
  if (p.cls == 1) {
    if constexpr (KIND == KIND_TOEPLITZ || KIND == KIND_HANKEL)
      batch_tile<KIND, 64, 256, 16, 1, 4>(p, local, lds);
    else
      batch_tile<KIND, 128, 128, 8, 2, 2>(p, local, lds);
  } else {
    batch_tile<KIND, 64, 64, 16, 2, 2>(p, local, lds);
  }
}

// (The tail form serves k2 = 0 too: it adds no slabs.)
template <int KIND>
__device__ __forceinline__ void leaf_tile_kind(const BatchProblem& p, int local, double* lds) {
  using T = LeafTiles;
  batch_tile<KIND, T::BM, T::BN, T::BK, T::WM, T::WN, true>(p, local, lds);
}

template <int KIND>
__device__ __forceinline__ void hss_tile_kind(const BatchProblem& p, int local, double* lds) {
  using T = HssTiles;
  if constexpr (KIND == KIND_CAUCHY || KIND == KIND_CAUCHYEIG) {
    if (p.k2 > 0) {  // generated diagonal block plus an explicit tail
      if (p.cls == 1)
        batch_tile<KIND, T::BM1, T::BN1, T::BK1, T::WM1, T::WN1, true>(p, local, lds);
      else
        batch_tile<KIND, T::BM0, T::BN0, T::BK0, T::WM0, T::WN0, true>(p, local, lds);
      return;
    }
  }
  if (p.cls == 1)
    batch_tile<KIND, T::BM1, T::BN1, T::BK1, T::WM1, T::WN1>(p, local, lds);
  else
    batch_tile<KIND, T::BM0, T::BN0, T::BK0, T::WM0, T::WN0>(p, local, lds);
}

// Kind sets of the persistent batch kernel: the public kinds (batched API), and
// the kinds of the HSS construction and product (Cauchy kinds, their transposes
// and explicit matrices).  One kernel per set keeps each set's register budget.
enum BatchKindSet : int { KSET_PUBLIC = 0, KSET_HSS = 1, KSET_LEAF = 2 };

constexpr int kLeafLds = tile_lds<KIND_CAUCHYEIG, LeafTiles::BM, LeafTiles::BN, LeafTiles::BK>();

template <int KSET>
__global__ __launch_bounds__(KSET == KSET_LEAF ? 512 : 256,
                             KSET == KSET_PUBLIC ? 2 : (KSET == KSET_HSS ? HssTiles::MIN_BLOCKS : 1)) void
sgemm_f64_batch_kernel(
    const BatchProblem* __restrict__ probs, const int* __restrict__ tbeg, int nprob, int ntiles,
    unsigned* __restrict__ ctr) {
  __shared__ double lds[KSET == KSET_PUBLIC ? kBatchLds : (KSET == KSET_HSS ? kHssLds : kLeafLds)];
  __shared__ int tile_s;
  for (;;) {
    if (threadIdx.x == 0) tile_s = (int)atomicAdd(&ctr[0], 1u);
    __syncthreads();
    const int t = tile_s;  // every thread reads it before the tile body's closing barrier
    if (t >= ntiles) break;
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {  // last problem whose first tile is <= t
      int mid = (lo + hi + 1) >> 1;
      if (tbeg[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const BatchProblem& p = probs[lo];
    const int local = t - tbeg[lo];
    if constexpr (KSET == KSET_PUBLIC) {
      switch (p.kind) {
        case KIND_CAUCHY: batch_tile_kind<KIND_CAUCHY>(p, local, lds); break;
        case KIND_TOEPLITZ: batch_tile_kind<KIND_TOEPLITZ>(p, local, lds); break;
        case KIND_HANKEL: batch_tile_kind<KIND_HANKEL>(p, local, lds); break;
        default: batch_tile_kind<KIND_CAUCHYEIG>(p, local, lds); break;
      }
    } else if constexpr (KSET == KSET_LEAF) {
      if (p.kind == KIND_CAUCHY)
        leaf_tile_kind<KIND_CAUCHY>(p, local, lds);
      else
        leaf_tile_kind<KIND_CAUCHYEIG>(p, local, lds);
    } else {
      switch (p.kind) {
        case KIND_CAUCHY: hss_tile_kind<KIND_CAUCHY>(p, local, lds); break;
        case KIND_CAUCHYEIG: hss_tile_kind<KIND_CAUCHYEIG>(p, local, lds); break;
        case KIND_CAUCHYEIG_T: hss_tile_kind<KIND_CAUCHYEIG_T>(p, local, lds); break;
        case KIND_CAUCHY_T: hss_tile_kind<KIND_CAUCHY_T>(p, local, lds); break;
        case KIND_DENSE: hss_tile_kind<KIND_DENSE>(p, local, lds); break;
        default: hss_tile_kind<KIND_DENSE_T>(p, local, lds); break;
      }
    }
  }
  if (threadIdx.x == 0) {
    // every other workgroup has taken its last tile index before counting itself done
    const unsigned done = atomicAdd(&ctr[1], 1u);
    if (done == gridDim.x - 1) {
      atomicExch(&ctr[0], 0u);
      atomicExch(&ctr[1], 0u);
    }
  }
}

// One launch over the whole tile list; ctr: two zeroed unsigned counters; kset:
// 0 for the public kinds, 1 for the HSS kinds, 2 for the HSS product's leaves.
hipError_t launch_batch_f64(const BatchProblem* dprobs, const int* dtbeg, int nprob, int ntiles,
                            unsigned* ctr, int nworkgroups, hipStream_t s, int kset) {
  if (ntiles <= 0) return hipSuccess;
  const int g = ntiles < nworkgroups ? ntiles : nworkgroups;
  if (kset == KSET_LEAF)
    hipLaunchKernelGGL(sgemm_f64_batch_kernel<KSET_LEAF>, dim3(g), dim3(512), 0, s, dprobs, dtbeg,
                       nprob, ntiles, ctr);
  else if (kset == KSET_HSS)
    hipLaunchKernelGGL(sgemm_f64_batch_kernel<KSET_HSS>, dim3(g), dim3(256), 0, s, dprobs, dtbeg,
                       nprob, ntiles, ctr);
  else
    hipLaunchKernelGGL(sgemm_f64_batch_kernel<KSET_PUBLIC>, dim3(g), dim3(256), 0, s, dprobs,
                       dtbeg, nprob, ntiles, ctr);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Host-side launchers (called from the C ABI in psmma_capi.hip).

struct RightTile {
  int bm, bn;
};

template <int KIND, int BM, int BN, int BK, int WGM, int WGN, int MINB = 2>
static hipError_t launch_right_t(int M, int N, int K, double alpha, const double* A, int lda,
                                 const GenPtrs& g, int r0, int c0, double beta, double* C, int ldc,
                                 hipStream_t s) {
  int nb = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  hipLaunchKernelGGL((sgemm_f64_right_kernel<KIND, BM, BN, BK, WGM, WGN, MINB>), dim3(nb),
                     dim3(WGM * WGN * 64), 0, s, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc);
  return hipGetLastError();
}

template <int KIND>
static hipError_t launch_right_kind(int cfg, int M, int N, int K, double alpha, const double* A,
                                    int lda, const GenPtrs& g, int r0, int c0, double beta,
                                    double* C, int ldc, hipStream_t s) {
  switch (cfg) {
    case 0:  // large problems: 128x128 workgroup tile, 4 waves of 64x64, BK = 16
      return launch_right_t<KIND, 128, 128, 16, 2, 2>(M, N, K, alpha, A, lda, g, r0, c0, beta, C,
                                                      ldc, s);
    case 1:  // as 0 with BK = 32: half the barriers per flop
      return launch_right_t<KIND, 128, 128, 32, 2, 2>(M, N, K, alpha, A, lda, g, r0, c0, beta, C,
                                                      ldc, s);
    case 3:  // 64x256: fewer dense-operand re-reads per flop
      // Window kinds with BK = 32: half the barriers per flop, still two workgroups per
      // CU (45 KB of LDS, 217 registers): Toeplitz N = 8192 68.3 -> 69.6 TF/s on MI355X
      // (BK = 48 the same, BK = 64 62 TF/s at one workgroup per CU).
      if constexpr (KIND == KIND_TOEPLITZ || KIND == KIND_HANKEL)
        return launch_right_t<KIND, 64, 256, 32, 1, 4>(M, N, K, alpha, A, lda, g, r0, c0, beta, C,
                                                       ldc, s);
      else
        return launch_right_t<KIND, 64, 256, 16, 1, 4>(M, N, K, alpha, A, lda, g, r0, c0, beta, C,
                                                       ldc, s);
    case 4:  // 256x128, 8 waves of 64x64: half the generated entries per flop (Cauchy kinds)
      return launch_right_t<KIND, 256, 128, 16, 4, 2, 1>(M, N, K, alpha, A, lda, g, r0, c0, beta,
                                                         C, ldc, s);
    case 5: {  // 128x128, 8 waves of 32x64, BK = 32: two waves per SIMD, one workgroup per CU
      // Chosen from one tile per CU up.  The 147 KB of LDS of the Cauchy kinds keep a
      // second workgroup off the CU (at 74 KB with BK = 16 two could share a CU while
      // others idled): Cauchy N = 2048 0.41 -> 0.37-0.39 ms on MI355X.
      static_assert(KIND == KIND_TOEPLITZ || KIND == KIND_HANKEL ||
                        2 * sizeof(double) * 2 * (RightLds<KIND, 128, 128, 32>::ASZ +
                                                  RightLds<KIND, 128, 128, 32>::HSZ) >
                            160 * 1024,
                    "one workgroup per CU");
      return launch_right_t<KIND, 128, 128, 32, 4, 2, 1>(M, N, K, alpha, A, lda, g, r0, c0, beta,
                                                         C, ldc, s);
    }
    case 6:  // 256x128, 16 waves of 32x64 (plain Cauchy from 1024 tiles: 157 ms against 162 for
             // config 4 at N = 16384 on MI355X; the eigen-Cauchy kind spills more and loses 10%)
      return launch_right_t<KIND, 256, 128, 16, 8, 2, 1>(M, N, K, alpha, A, lda, g, r0, c0, beta,
                                                         C, ldc, s);
    default:  // small problems: 64x64 workgroup tile, 4 waves of 32x32
      return launch_right_t<KIND, 64, 64, 16, 2, 2>(M, N, K, alpha, A, lda, g, r0, c0, beta, C,
                                                    ldc, s);
  }
}

int tuned_right_cfg(int kind, int M, int N, int K);

// Picks a tile shape: PSMMA_F64_CFG (0..6) if set; the autotuner's choice for the
// (kind, size bucket) if one is cached; else the measured rule below.
int choose_right_cfg(int kind, int M, int N, int K) {
  static int forced = [] {
    const char* e = getenv("PSMMA_F64_CFG");
    return e ? atoi(e) : -1;
  }();
  if (forced >= 0) return forced;
  const int tuned = tuned_right_cfg(kind, M, N, K);
  if (tuned >= 0) return tuned;
  long long big = (long long)((M + 127) / 128) * ((N + 127) / 128);
  const bool win = kind == KIND_TOEPLITZ || kind == KIND_HANKEL;
  // Cauchy kinds (measured on MI355X, N x N x N products): the 8-wave 256x128 tile
  // from 1024 128x128 tiles up (N = 4096: 2.56 ms against 2.77 for 128x128; +5% at
  // N = 16384), the 8-wave 128x128 tile from one 128x128 tile per CU (N = 2048:
  // 0.37-0.39 ms with BK = 32 against 0.41 with BK = 16, 0.45 for 4 waves, 0.49 for 64x64),
  // 64x64 below
  if (!win) return big >= 1024 ? (kind == KIND_CAUCHY ? 6 : 4) : big >= 256 ? 5 : 2;
  // Toeplitz / Hankel: 64x256 from 256 128x128 tiles (Toeplitz N = 2048: 0.32 ms against
  // 0.35 for 64x64; best of 0, 2, 3 and two 8-wave 128x128 layouts at N = 2048, 4096, 8192)
  return big >= 256 ? 3 : 2;
}

hipError_t launch_right_f64(int kind, int cfg, int M, int N, int K, double alpha, const double* A,
                            int lda, const GenPtrs& g, int r0, int c0, double beta, double* C,
                            int ldc, hipStream_t s) {
  if (cfg < 0) cfg = choose_right_cfg(kind, M, N, K);
  switch (kind) {
    case KIND_CAUCHY:
      return launch_right_kind<KIND_CAUCHY>(cfg, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc, s);
    case KIND_TOEPLITZ:
      return launch_right_kind<KIND_TOEPLITZ>(cfg, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc,
                                              s);
    case KIND_HANKEL:
      return launch_right_kind<KIND_HANKEL>(cfg, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc, s);
    case KIND_CAUCHYEIG:
      return launch_right_kind<KIND_CAUCHYEIG>(cfg, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc,
                                               s);
  }
  return hipErrorInvalidValue;
}

template <int KIND>
static hipError_t launch_indexed_kind(int M, int N, int K, double alpha, const double* A, int lda,
                                      const GenPtrs& g, double beta, double* C, int ldc,
                                      hipStream_t s) {
  const long long big = (long long)((M + 127) / 128) * ((N + 127) / 128);
  if (big >= 256) {  // 128 x 128 tiles, 4 waves, once they fill the chip
    int nb = ((M + 127) / 128) * ((N + 127) / 128);
    hipLaunchKernelGGL((sgemm_f64_indexed_kernel<KIND, 128, 128, 16, 2, 2>), dim3(nb), dim3(256),
                       0, s, M, N, K, alpha, A, lda, g, beta, C, ldc);
  } else {
    int nb = ((M + 63) / 64) * ((N + 63) / 64);
    hipLaunchKernelGGL((sgemm_f64_indexed_kernel<KIND, 64, 64, 16, 2, 2>), dim3(nb), dim3(256), 0,
                       s, M, N, K, alpha, A, lda, g, beta, C, ldc);
  }
  return hipGetLastError();
}

// C = alpha * A * H(ridx, cidx) + beta * C on the FP64 kernel (g.ridx / g.cidx set).
hipError_t launch_indexed_f64(int kind, int M, int N, int K, double alpha, const double* A,
                              int lda, const GenPtrs& g, double beta, double* C, int ldc,
                              hipStream_t s) {
  switch (kind) {
    case KIND_CAUCHY: return launch_indexed_kind<KIND_CAUCHY>(M, N, K, alpha, A, lda, g, beta, C, ldc, s);
    case KIND_TOEPLITZ: return launch_indexed_kind<KIND_TOEPLITZ>(M, N, K, alpha, A, lda, g, beta, C, ldc, s);
    case KIND_HANKEL: return launch_indexed_kind<KIND_HANKEL>(M, N, K, alpha, A, lda, g, beta, C, ldc, s);
    case KIND_CAUCHYEIG: return launch_indexed_kind<KIND_CAUCHYEIG>(M, N, K, alpha, A, lda, g, beta, C, ldc, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_dft_f64(int M, int N, int K, int Nd, double alpha, const double2* tw,
                          const double2* B, int ldb, int r0, int c0, double beta, double2* C, int ldc,
                          hipStream_t s) {
  // 128 x 64 tiles of 8 waves with 8-deep slabs (57 KB of LDS): F(4096 x 4096) B(4096 x 512)
  // 0.967 ms on MI355X, against 0.975 for 64 x 64 with BK = 8, 1.051 for 64 x 64 with
  // BK = 16, 1.007 for 64 x 128 with BK = 8, 1.051 for 128 x 64 with BK = 16
  constexpr int BM = 128, BN = 64, BK = 8, WGM = 4, WGN = 2;
  int nb = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  hipLaunchKernelGGL((dft_f64_left_kernel<BM, BN, BK, WGM, WGN>), dim3(nb), dim3(WGM * WGN * 64), 0,
                     s, M, N, K, Nd, alpha, tw, B, ldb, r0, c0, beta, C, ldc);
  return hipGetLastError();
}

}  // namespace psmma
