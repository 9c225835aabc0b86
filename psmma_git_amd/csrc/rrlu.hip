// Batched generator-only RRLU of Cauchy-like blocks.  Eigen-Cauchy blocks: the reference's
// Compress_cauchyEig2 (SRC/CauchyHssEig.f90:2640-2829) over rrluCauchyEigrow
// (:7-212) with CauchyPivtEig (:571-717), CauchyPivtEig_CP (:1004-1134),
// searchMax2Eig (:1137-1203), CauchyMaxEigm / CauchyMaxEigmc (:1419-1589),
// computfd / computdd (:470-568) and CauchylikeEigQ (:1882-1952).
//
// A block is the set of global rows PL (1-based) and columns PU of the matrix
// U_i V_j / (F_i - lambda_j), never formed: every F - D and D - D difference is
// rebuilt from the old poles FF and the gaps DIFL / DIFR through the global
// indices, as computfd / computdd do.  The factorisation is a rook-pivoted LU
// on the Cauchy-like generators (O(m + n) work per step), complete pivoting at
// the first convergence test, and every 5th step a swap pass that exchanges a
// chosen row for a Schur-complement row while that grows the chosen block's
// determinant by more than rho (the default-real 1.2E+0).  Output: the rank and, below full rank,
// the row ID H ~= Q Y with Q = P^-1 [I; Z_i W_j / (F_i - F_j)] (m x rank) and
// Y the chosen rows of the block (rank x n).
//
// One 256-thread workgroup per block (m, n <= 256).  The block's permuted
// generators live in LDS; the maximum searches are workgroup reductions that
// return the first index of the largest value (the reference's scan order);
// the generator updates of a step run one entry per thread; the two running
// products of a step (W(k), and U(jj) / W(ii) of a swap) run in the
// reference's order on one thread.  Every entry is computed from the same
// operands in the same order as the restatement (subtractions, quotients,
// products; no expression the compiler could contract into a fused
// multiply-add, and contraction is switched off), so ranks, permutations and
// outputs are bitwise those of the CPU restatement.
//
// Plain Cauchy blocks U_i V_j / (D_i - F_j) (rrluCauchy / compress_cauchy,
// SRC/Cauchylowrank.f90:7-134, :387-491, with CauchyPivtG :191-243,
// CauchyPivtG_CP :246-296, searchMax2 :137-163, CauchyMax :166-188): the same
// scheme with the block's own poles (rho = 1.12, swap pass every 10th step).
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../include/psmma.h"
#include "psmma_common.h"

#pragma clang fp contract(off)

namespace psmma {

constexpr int kRrluMaxDim = 256;
constexpr int kRrluThreads = 256;
constexpr int kRrluWaves = kRrluThreads / 64;
constexpr double kRrluRho = (double)1.2f;  // rho = 1.2E+0, default real (SRC/CauchyHssEig.f90:98)
constexpr int kRrluPeriod = 5;             // swap pass every 5th step (:99)

struct RrluEigState {
  // row positions (permuted together): row pole FF(PL_i), original weight A_i,
  // Schur generator U_i, global index PL_i, row permutation PQ_i (1-based)
  double F[kRrluMaxDim], A[kRrluMaxDim], U[kRrluMaxDim];
  int PL[kRrluMaxDim], PQ[kRrluMaxDim];
  // column positions (permuted together): generator V_j, global index PU_j and
  // the values computfd / computdd read through it: FF(PU_j), FF(PU_j + 1),
  // DIFL(PU_j), DIFR(PU_j)
  double V[kRrluMaxDim], fc[kRrluMaxDim], fc1[kRrluMaxDim], dl[kRrluMaxDim], dr[kRrluMaxDim];
  int PU[kRrluMaxDim];
  double W[kRrluMaxDim];  // interpolation generators of the chosen rows
  int IP[kRrluMaxDim];    // inverse row permutation (output)
  double red_v[kRrluWaves];
  int red_i[kRrluWaves];
  int bc_i;  // a row index published by one thread
};

// computfd: F(row i) - D(col j) for row / column positions.
__device__ __forceinline__ double fd(const RrluEigState& s, int i, int j) {
  return (s.PL[i] <= s.PU[j]) ? ((s.F[i] - s.fc[j]) - s.dl[j]) : ((s.F[i] - s.fc1[j]) - s.dr[j]);
}
// computdd: D(col j) - D(col k).
__device__ __forceinline__ double dd(const RrluEigState& s, int j, int k) {
  return (s.PU[j] < s.PU[k]) ? ((s.fc1[j] - s.fc[k]) + s.dr[j] - s.dl[k])
                             : ((s.fc[j] - s.fc1[k]) + s.dl[j] - s.dr[k]);
}

// Workgroup maximum of x over i in [0, n) with the first index attaining it
// (n = 0: -DBL_MAX and index -1, Fortran's maxval of an empty array).  Uniform
// call: contains barriers; every thread returns the same pair.
template <class S, class Fn>
__device__ void first_max(S& s, int n, Fn x_of, double& best, int& at) {
  const int t = threadIdx.x;
  double b = -1.79769313486231570815e+308;
  int bi = -1;
  for (int i = t; i < n; i += kRrluThreads) {
    const double x = x_of(i);
    if (bi < 0 || x > b) {  // per thread, i increases: the first index is kept
      b = x;
      bi = i;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double b2 = __shfl_xor(b, o);
    const int i2 = __shfl_xor(bi, o);
    if (i2 >= 0 && (bi < 0 || b2 > b || (b2 == b && i2 < bi))) {
      b = b2;
      bi = i2;
    }
  }
  if ((t & 63) == 0) {
    s.red_v[t >> 6] = b;
    s.red_i[t >> 6] = bi;
  }
  __syncthreads();
  b = s.red_v[0];
  bi = s.red_i[0];
  for (int w = 1; w < kRrluWaves; ++w) {
    const double b2 = s.red_v[w];
    const int i2 = s.red_i[w];
    if (i2 >= 0 && (bi < 0 || b2 > b || (b2 == b && i2 < bi))) {
      b = b2;
      bi = i2;
    }
  }
  __syncthreads();  // red_* may be reused at once
  best = b;
  at = bi;
}

__device__ __forceinline__ void swap_rows(RrluEigState& s, int i, int j) {
  if (threadIdx.x == 0 && i != j) {
    double d;
    int n;
    d = s.F[i]; s.F[i] = s.F[j]; s.F[j] = d;
    d = s.U[i]; s.U[i] = s.U[j]; s.U[j] = d;
    d = s.A[i]; s.A[i] = s.A[j]; s.A[j] = d;
    n = s.PL[i]; s.PL[i] = s.PL[j]; s.PL[j] = n;
    n = s.PQ[i]; s.PQ[i] = s.PQ[j]; s.PQ[j] = n;
  }
  __syncthreads();
}
__device__ __forceinline__ void swap_cols(RrluEigState& s, int i, int j) {
  if (threadIdx.x == 0 && i != j) {
    double d;
    int n;
    d = s.V[i]; s.V[i] = s.V[j]; s.V[j] = d;
    d = s.fc[i]; s.fc[i] = s.fc[j]; s.fc[j] = d;
    d = s.fc1[i]; s.fc1[i] = s.fc1[j]; s.fc1[j] = d;
    d = s.dl[i]; s.dl[i] = s.dl[j]; s.dl[j] = d;
    d = s.dr[i]; s.dr[i] = s.dr[j]; s.dr[j] = d;
    n = s.PU[i]; s.PU[i] = s.PU[j]; s.PU[j] = n;
  }
  __syncthreads();
}

// CauchyMaxEigmc down column position c over rows [k, M): |U_i / fd(i, c)|.
__device__ void max_col(RrluEigState& s, int k, int c, int M, double& best, int& at) {
  first_max(s, M - k, [&](int i) { return fabs(s.U[k + i] / fd(s, k + i, c)); }, best, at);
}
// CauchyMaxEigm along row position r over columns [k, N): |V_j / fd(r, j)|.
__device__ void max_row(RrluEigState& s, int k, int r, int N, double& best, int& at) {
  first_max(s, N - k, [&](int j) { return fabs(s.V[k + j] / fd(s, r, k + j)); }, best, at);
}

// CauchyPivtEig: alternating row / column search from position k; a row swap
// falls through to a column swap (the reference's CONTINUE).  The early-return
// pivot magnitude is the pivot entry itself (the reference writes F(PL(k)) with
// a global index into the local pole array; the CPU restatement makes the same
// reading).  Iteration cap as in the restatement.
__device__ void pivt(RrluEigState& s, int k, int M, int N) {
  double junkL, junkU;
  int jjL, jjU;
  max_col(s, k, k, M, junkL, jjL);
  max_row(s, k, k, N, junkU, jjU);
  junkL = junkL * fabs(s.V[k]);
  junkU = junkU * fabs(s.U[k]);
  const double piv = fabs(s.U[k] * s.V[k] / fd(s, k, k));
  __syncthreads();  // every thread has read U(k), V(k) before thread 0 swaps them
  if (junkL <= piv && junkU <= piv) return;
  int flg = (junkL > junkU) ? 1 : 0;
  double junk;
  for (long it = 0; it < 4L * (M + N) + 16; ++it) {
    if (flg == 1) {
      swap_rows(s, k, jjL + k);
      max_row(s, k, k, N, junk, jjU);
      if (jjU == 0) return;
      flg = 0;
    }
    swap_cols(s, k, jjU + k);
    max_col(s, k, k, M, junk, jjL);
    if (jjL == 0) return;
    flg = 1;
  }
}

// CauchyPivtEig_CP: complete pivoting on the k-th Schur complement.  Thread c
// scans column position k + c down the rows (first index of its maximum), then
// the columns' maxima times |V| are compared in column order.
__device__ void pivt_cp(RrluEigState& s, int k, int M, int N) {
  const int t = threadIdx.x;
  double colmax = -1.79769313486231570815e+308;
  int colat = -1;
  if (t < N - k) {
    for (int i = 0; i < M - k; ++i) {
      const double x = fabs(s.U[k + i] / fd(s, k + i, k + t));
      if (colat < 0 || x > colmax) {
        colmax = x;
        colat = i;
      }
    }
  }
  double junk;
  int jj;
  // N - k <= 256 columns: first_max evaluates column c on thread c
  first_max(s, N - k, [&](int c) { return fabs(colmax * s.V[k + c]); }, junk, jj);
  if (t == jj) s.bc_i = colat;  // the row of the winning column
  __syncthreads();
  const int ii = s.bc_i;
  __syncthreads();
  swap_rows(s, k, ii + k);
  swap_cols(s, k, jj + k);
}

__global__ __launch_bounds__(kRrluThreads) void rrlu_eig_kernel(
    const double* __restrict__ ff, const double* __restrict__ difl,
    const double* __restrict__ difr, const double* __restrict__ u, const double* __restrict__ v,
    int korder, const int* __restrict__ plg, const int* __restrict__ pug, int M, int N,
    double tol, int kmax, int* __restrict__ rank_out, int* __restrict__ nswap_out,
    double* __restrict__ q, int ldq, long long qstride, double* __restrict__ y, int ldy,
    long long ystride) {
  __shared__ RrluEigState s;
  const int b = blockIdx.x, t = threadIdx.x;
  const int* pl0 = plg + (size_t)b * M;
  const int* pu0 = pug + (size_t)b * N;
  // global indices outside [1, korder]: rank -2, nothing read through them
  const bool bad = (t < M && (pl0[t] < 1 || pl0[t] > korder)) ||
                   (t < N && (pu0[t] < 1 || pu0[t] > korder));
  if (__syncthreads_or(bad)) {
    if (t == 0) {
      rank_out[b] = -2;
      if (nswap_out) nswap_out[b] = 0;
    }
    return;
  }
  if (t < M) {
    const int p = pl0[t];
    s.PL[t] = p;
    s.F[t] = ff[p - 1];
    s.A[t] = u[p - 1];
    s.U[t] = s.A[t];
    s.PQ[t] = t + 1;
  }
  if (t < N) {
    const int p = pu0[t];
    s.PU[t] = p;
    s.V[t] = v[p - 1];
    s.fc[t] = ff[p - 1];
    s.fc1[t] = (p < korder) ? ff[p] : 0.0;  // read only where PU_j < korder
    s.dl[t] = difl[p - 1];
    s.dr[t] = difr[p - 1];
  }
  s.W[t] = 0.0;
  __syncthreads();

  const int mn = M < N ? M : N;
  int rk = mn, swaps = 0;
  double amax = 0.0;
  for (int k = 0; k < mn; ++k) {  // the reference's k = this k + 1
    pivt(s, k, M, N);
    double ukk = s.U[k] * s.V[k] / fd(s, k, k);
    amax = fmax(amax, fabs(ukk));
    if (fabs(ukk) < tol * amax) {
      pivt_cp(s, k, M, N);
      ukk = s.U[k] * s.V[k] / fd(s, k, k);
      amax = fmax(amax, fabs(ukk));
      if (fabs(ukk) < tol * amax) {
        rk = k;
        break;
      }
    }
    // generator updates of step k (independent entries), then the running
    // product of W(k) on thread 0
    double nu = 0.0, nv = 0.0, nw = 0.0;
    const int ju = k + 1 + t;
    if (ju < M) {
      double temp = fd(s, ju, k);
      temp = (s.F[ju] - s.F[k]) / temp;
      nu = s.U[ju] * temp;
    }
    if (ju < N) {
      double temp = dd(s, ju, k);
      const double temp2 = -fd(s, k, ju);
      temp = temp / temp2;
      nv = s.V[ju] * temp;
    }
    if (t < k) {
      const double temp = -fd(s, t, k);
      nw = s.W[t] * (temp / (s.F[k] - s.F[t]));
    }
    double wk = 0.0;
    if (t == 0) {
      wk = fd(s, k, k) / s.A[k];
      for (int j = 0; j < k; ++j) {
        const double temp = -fd(s, k, j);
        wk = wk * (temp / (s.F[j] - s.F[k]));
      }
    }
    __syncthreads();
    if (ju < M) s.U[ju] = nu;
    if (ju < N) s.V[ju] = nv;
    if (t < k) s.W[t] = nw;
    if (t == 0) s.W[k] = wk;
    __syncthreads();

    // swap pass every kRrluPeriod-th step, repeated while it pays (the cap only
    // guarantees an end; each pass multiplies the chosen block's determinant
    // by more than rho)
    bool again = ((k + 1) % kRrluPeriod) < 1;
    for (int pass = 0; again && pass < 64 * (M + N); ++pass) {
      again = false;
      // searchMax2Eig over the k+1 chosen rows: thread j scans the Schur rows
      const int ns = M - k - 1;
      double cm = -1.79769313486231570815e+308;
      int cat = -1;
      if (t <= k) {
        for (int i = 0; i < ns; ++i) {
          const double x = fabs(s.U[k + 1 + i] / (s.F[k + 1 + i] - s.F[t]));
          if (cat < 0 || x > cm) {
            cm = x;
            cat = i;
          }
        }
        cm = cm * fabs(s.W[t]);
      }
      double nmax;
      int ii;
      // k + 1 <= 256 chosen rows: first_max evaluates row j on thread j
      first_max(s, k + 1, [&](int) { return cm; }, nmax, ii);
      if (t == ii) s.bc_i = cat;
      __syncthreads();
      const int jj = s.bc_i + k + 1;  // Schur-complement row
      __syncthreads();
      if (!(nmax > kRrluRho)) break;
      ++swaps;
      again = true;
      double nv2 = 0.0, nu2 = 0.0, nw2 = 0.0;
      if (ju < N) {
        double temp = -fd(s, ii, ju);
        const double temp2 = -fd(s, jj, ju);
        temp = temp / temp2;
        nv2 = s.V[ju] * temp;
      }
      const int iu = k + 1 + t;
      const bool du = iu < M && iu != jj;
      if (du) nu2 = s.U[iu] * ((s.F[iu] - s.F[jj]) / (s.F[iu] - s.F[ii]));
      const bool dw = t <= k && t != ii;
      if (dw) nw2 = s.W[t] * ((s.F[t] - s.F[ii]) / (s.F[t] - s.F[jj]));
      double ujj = 0.0, wii = 0.0;
      if (t == 0) {
        ujj = s.A[ii] * ((s.F[ii] - s.F[jj]) / fd(s, ii, ii));
        wii = fd(s, jj, ii) / s.A[jj];
        for (int j = 0; j <= k; ++j) {
          if (j == ii) continue;
          double temp = fd(s, ii, j);
          ujj = ujj * ((s.F[ii] - s.F[j]) / temp);
          temp = -fd(s, jj, j);
          wii = wii * (temp / (s.F[j] - s.F[jj]));
        }
      }
      __syncthreads();
      if (ju < N) s.V[ju] = nv2;
      if (du) s.U[iu] = nu2;
      if (dw) s.W[t] = nw2;
      if (t == 0) {
        s.U[jj] = ujj;
        s.W[ii] = wii;
        int n;
        double d;
        n = s.PL[ii]; s.PL[ii] = s.PL[jj]; s.PL[jj] = n;
        n = s.PQ[ii]; s.PQ[ii] = s.PQ[jj]; s.PQ[jj] = n;
        d = s.F[ii]; s.F[ii] = s.F[jj]; s.F[jj] = d;
        d = s.A[ii]; s.A[ii] = s.A[jj]; s.A[jj] = d;
      }
      __syncthreads();
    }
  }

  if (t == 0) {
    rank_out[b] = rk;
    if (nswap_out) nswap_out[b] = swaps;
  }
  if (rk <= 0 || rk >= mn || rk > kmax) return;  // uniform: nothing more to write

  // CauchylikeEigQ: Q(i, j) for original row i = (P^-1 [I; Z W / (F - F)])
  if (t < M) s.IP[s.PQ[t] - 1] = t;
  __syncthreads();
  double* qb = q + (size_t)b * qstride;
  if (t < M) {
    const int p = s.IP[t];
    for (int j = 0; j < rk; ++j) {
      double x;
      if (p < rk)
        x = (p == j) ? 1.0 : 0.0;
      else
        x = s.U[p] * s.W[j] / (s.F[p] - s.F[j]);
      qb[(size_t)j * ldq + t] = x;
    }
  }
  // Y: the chosen rows of the block in the original column order (CauchylikeEig
  // with the original weights)
  double* yb = y + (size_t)b * ystride;
  if (t < N) {
    const int pj = pu0[t];
    const double vj = v[pj - 1];
    const double fpj = ff[pj - 1], fpj1 = (pj < korder) ? ff[pj] : 0.0;
    const double dlj = difl[pj - 1], drj = difr[pj - 1];
    for (int i = 0; i < rk; ++i) {
      const double den = (s.PL[i] <= pj) ? ((s.F[i] - fpj) - dlj) : ((s.F[i] - fpj1) - drj);
      yb[(size_t)t * ldy + i] = s.A[i] * vj / den;
    }
  }
}

// ---------------------------------------------------------------------------
// Plain Cauchy blocks.

constexpr double kRrluRhoC = 1.12;  // rho = 1.12D0 (SRC/Cauchylowrank.f90:70)
constexpr int kRrluPeriodC = 10;    // swap pass every 10th step (:71)

struct RrluCauchyState {
  // row positions: pole D_i, original weight A_i, Schur generator U_i, local
  // permutation PL_i (1-based); column positions: pole F_j, generator V_j, PU_j
  double D[kRrluMaxDim], A[kRrluMaxDim], U[kRrluMaxDim];
  int PL[kRrluMaxDim];
  double F[kRrluMaxDim], V[kRrluMaxDim];
  int PU[kRrluMaxDim];
  double W[kRrluMaxDim];
  int IP[kRrluMaxDim];
  double red_v[kRrluWaves];
  int red_i[kRrluWaves];
  int bc_i;
};

__device__ __forceinline__ void swap_rows(RrluCauchyState& s, int i, int j) {
  if (threadIdx.x == 0 && i != j) {
    double d;
    int n;
    d = s.D[i]; s.D[i] = s.D[j]; s.D[j] = d;
    d = s.U[i]; s.U[i] = s.U[j]; s.U[j] = d;
    d = s.A[i]; s.A[i] = s.A[j]; s.A[j] = d;
    n = s.PL[i]; s.PL[i] = s.PL[j]; s.PL[j] = n;
  }
  __syncthreads();
}
__device__ __forceinline__ void swap_cols(RrluCauchyState& s, int i, int j) {
  if (threadIdx.x == 0 && i != j) {
    double d;
    int n;
    d = s.F[i]; s.F[i] = s.F[j]; s.F[j] = d;
    d = s.V[i]; s.V[i] = s.V[j]; s.V[j] = d;
    n = s.PU[i]; s.PU[i] = s.PU[j]; s.PU[j] = n;
  }
  __syncthreads();
}

// CauchyMax(D(k:M), F(c), U(k:M)): |U_i / (D_i - F_c)| down column position c.
__device__ void max_col(RrluCauchyState& s, int k, int c, int M, double& best, int& at) {
  first_max(s, M - k, [&](int i) { return fabs(s.U[k + i] / (s.D[k + i] - s.F[c])); }, best, at);
}
// CauchyMax(F(k:N), D(r), V(k:N)): |V_j / (F_j - D_r)| along row position r.
__device__ void max_row(RrluCauchyState& s, int k, int r, int N, double& best, int& at) {
  first_max(s, N - k, [&](int j) { return fabs(s.V[k + j] / (s.F[k + j] - s.D[r])); }, best, at);
}

// CauchyPivtG: alternating row / column search from position k (a row swap
// falls through to a column swap, the reference's CONTINUE); iteration cap as
// in the restatement.
__device__ void pivt(RrluCauchyState& s, int k, int M, int N) {
  double junkL, junkU;
  int jjL, jjU;
  max_col(s, k, k, M, junkL, jjL);
  max_row(s, k, k, N, junkU, jjU);
  junkL = junkL * fabs(s.V[k]);
  junkU = junkU * fabs(s.U[k]);
  const double piv = fabs(s.U[k] * s.V[k] / (s.D[k] - s.F[k]));
  __syncthreads();  // every thread has read U(k), V(k) before thread 0 swaps them
  if (junkL <= piv && junkU <= piv) return;
  int flg = (junkL > junkU) ? 1 : 0;
  double junk;
  for (long it = 0; it < 4L * (M + N) + 16; ++it) {
    if (flg == 1) {
      swap_rows(s, k, jjL + k);
      max_row(s, k, k, N, junk, jjU);
      if (jjU == 0) return;
      flg = 0;
    }
    swap_cols(s, k, jjU + k);
    max_col(s, k, k, M, junk, jjL);
    if (jjL == 0) return;
    flg = 1;
  }
}

// CauchyPivtG_CP: complete pivoting on the k-th Schur complement (thread c
// scans column position k + c).
__device__ void pivt_cp(RrluCauchyState& s, int k, int M, int N) {
  const int t = threadIdx.x;
  double colmax = -1.79769313486231570815e+308;
  int colat = -1;
  if (t < N - k) {
    for (int i = 0; i < M - k; ++i) {
      const double x = fabs(s.U[k + i] / (s.D[k + i] - s.F[k + t]));
      if (colat < 0 || x > colmax) {
        colmax = x;
        colat = i;
      }
    }
  }
  double junk;
  int jj;
  // N - k <= 256 columns: first_max evaluates column c on thread c
  first_max(s, N - k, [&](int c) { return fabs(colmax * s.V[k + c]); }, junk, jj);
  if (t == jj) s.bc_i = colat;  // the row of the winning column
  __syncthreads();
  const int ii = s.bc_i;
  __syncthreads();
  swap_rows(s, k, ii + k);
  swap_cols(s, k, jj + k);
}

__global__ __launch_bounds__(kRrluThreads) void rrlu_cauchy_kernel(
    const double* __restrict__ dg, const double* __restrict__ fg, const double* __restrict__ ug,
    const double* __restrict__ vg, int nrow, int ncol, const int* __restrict__ plg,
    const int* __restrict__ pug, int M, int N, double tol, int kmax, int* __restrict__ rank_out,
    double* __restrict__ xo, int ldx, long long xstride, double* __restrict__ yo, int ldy,
    long long ystride) {
  __shared__ RrluCauchyState s;
  const int b = blockIdx.x, t = threadIdx.x;
  const int* pl0 = plg + (size_t)b * M;
  const int* pu0 = pug + (size_t)b * N;
  const bool bad = (t < M && (pl0[t] < 1 || pl0[t] > nrow)) ||
                   (t < N && (pu0[t] < 1 || pu0[t] > ncol));
  if (__syncthreads_or(bad)) {
    if (t == 0) rank_out[b] = -2;
    return;
  }
  if (t < M) {
    const int p = pl0[t] - 1;
    s.D[t] = dg[p];
    s.A[t] = ug[p];
    s.U[t] = s.A[t];
    s.PL[t] = t + 1;
  }
  if (t < N) {
    const int p = pu0[t] - 1;
    s.F[t] = fg[p];
    s.V[t] = vg[p];
    s.PU[t] = t + 1;
  }
  s.W[t] = 0.0;
  __syncthreads();

  const int mn = M < N ? M : N;
  int rk = mn;
  double amax = 0.0;
  for (int k = 0; k < mn; ++k) {
    pivt(s, k, M, N);
    double ukk = s.U[k] * s.V[k] / (s.D[k] - s.F[k]);
    amax = fmax(amax, fabs(ukk));
    if (fabs(ukk) < tol * amax) {
      pivt_cp(s, k, M, N);
      ukk = s.U[k] * s.V[k] / (s.D[k] - s.F[k]);
      amax = fmax(amax, fabs(ukk));
      if (fabs(ukk) < tol * amax) {
        rk = k;
        break;
      }
    }
    double nu = 0.0, nv = 0.0, nw = 0.0, wk = 0.0;
    const int ju = k + 1 + t;
    if (ju < M) nu = s.U[ju] * (s.D[ju] - s.D[k]) / (s.D[ju] - s.F[k]);
    if (ju < N) nv = s.V[ju] * (s.F[ju] - s.F[k]) / (s.F[ju] - s.D[k]);
    if (t < k) nw = s.W[t] * ((s.F[k] - s.D[t]) / (s.D[k] - s.D[t]));
    if (t == 0) {
      wk = (s.D[k] - s.F[k]) / s.A[k];
      for (int j = 0; j < k; ++j) wk = wk * (s.F[j] - s.D[k]) / (s.D[j] - s.D[k]);
    }
    __syncthreads();
    if (ju < M) s.U[ju] = nu;
    if (ju < N) s.V[ju] = nv;
    if (t < k) s.W[t] = nw;
    if (t == 0) s.W[k] = wk;
    __syncthreads();

    // searchMax2 swap pass every kRrluPeriodC-th step (bound shared with the restatement)
    bool again = ((k + 1) % kRrluPeriodC) < 1;
    for (int pass = 0; again && pass < 64 * (M + N); ++pass) {
      again = false;
      const int ns = M - k - 1;
      double cm = -1.79769313486231570815e+308;
      int cat = -1;
      if (t <= k) {
        for (int i = 0; i < ns; ++i) {
          const double x = fabs(s.U[k + 1 + i] / (s.D[k + 1 + i] - s.D[t]));
          if (cat < 0 || x > cm) {
            cm = x;
            cat = i;
          }
        }
        cm = cm * fabs(s.W[t]);
      }
      double nmax;
      int ii;
      // k + 1 <= 256 chosen rows: first_max evaluates row j on thread j
      first_max(s, k + 1, [&](int) { return cm; }, nmax, ii);
      if (t == ii) s.bc_i = cat;
      __syncthreads();
      const int jj = s.bc_i + k + 1;  // Schur-complement row
      __syncthreads();
      if (!(nmax > kRrluRhoC)) break;
      again = true;
      double nv2 = 0.0, nu2 = 0.0, nw2 = 0.0, ujj = 0.0, wii = 0.0;
      if (ju < N) nv2 = s.V[ju] * ((s.F[ju] - s.D[ii]) / (s.F[ju] - s.D[jj]));
      const int iu = k + 1 + t;
      const bool du = iu < M && iu != jj;
      if (du) nu2 = s.U[iu] * ((s.D[iu] - s.D[jj]) / (s.D[iu] - s.D[ii]));
      const bool dw = t <= k && t != ii;
      if (dw) nw2 = s.W[t] * ((s.D[t] - s.D[ii]) / (s.D[t] - s.D[jj]));
      if (t == 0) {
        ujj = s.A[ii] * ((s.D[ii] - s.D[jj]) / (s.D[ii] - s.F[ii]));
        wii = (s.D[jj] - s.F[ii]) / s.A[jj];
        for (int j = 0; j <= k; ++j) {
          if (j == ii) continue;
          ujj = ujj * ((s.D[ii] - s.D[j]) / (s.D[ii] - s.F[j]));
          wii = wii * ((s.F[j] - s.D[jj]) / (s.D[j] - s.D[jj]));
        }
      }
      __syncthreads();
      if (ju < N) s.V[ju] = nv2;
      if (du) s.U[iu] = nu2;
      if (dw) s.W[t] = nw2;
      if (t == 0) {
        s.U[jj] = ujj;
        s.W[ii] = wii;
        int n;
        double d;
        n = s.PL[ii]; s.PL[ii] = s.PL[jj]; s.PL[jj] = n;
        d = s.D[ii]; s.D[ii] = s.D[jj]; s.D[jj] = d;
        d = s.A[ii]; s.A[ii] = s.A[jj]; s.A[jj] = d;
      }
      __syncthreads();
    }
  }

  if (t == 0) rank_out[b] = rk;
  if (rk <= 0 || rk >= M || rk > kmax) return;  // uniform: the reference writes H when rk < M

  // Cauchylike2: X(i, j) for original row i = (P^-1 [I; Z W / (D - D)]); Y: the chosen
  // rows from the original column poles and weights
  if (t < M) s.IP[s.PL[t] - 1] = t;
  __syncthreads();
  double* xb = xo + (size_t)b * xstride;
  if (t < M) {
    const int p = s.IP[t];
    for (int j = 0; j < rk; ++j) {
      double x;
      if (p < rk)
        x = (p == j) ? 1.0 : 0.0;
      else
        x = s.U[p] * s.W[j] / (s.D[p] - s.D[j]);
      xb[(size_t)j * ldx + t] = x;
    }
  }
  double* yb = yo + (size_t)b * ystride;
  if (t < N) {
    const int p = pu0[t] - 1;
    const double fj = fg[p], vj = vg[p];
    for (int i = 0; i < rk; ++i) yb[(size_t)t * ldy + i] = s.A[i] * vj / (s.D[i] - fj);
  }
}

}  // namespace psmma

using namespace psmma;

extern "C" int psmma_rrlu_cauchyeig_batched_dev(int nblk, int m, int n, const int* pl,
                                                const int* pu, const double* fpole,
                                                const double* u, const double* v,
                                                const double* difl, const double* difr,
                                                int korder, double tol, int kmax, int* rank,
                                                int* nswap, double* q, int ldq,
                                                long long stride_q, double* y, int ldy,
                                                long long stride_y, void* stream) {
  if (nblk < 0) return -1;
  if (m < 1 || m > kRrluMaxDim) return -2;
  if (n < 1 || n > kRrluMaxDim) return -3;
  const bool any = nblk > 0;
  if (any && !pl) return -4;
  if (any && !pu) return -5;
  if (any && !fpole) return -6;
  if (any && !u) return -7;
  if (any && !v) return -8;
  if (any && !difl) return -9;
  if (any && !difr) return -10;
  if (korder < 1) return -11;
  if (!(tol >= 0.0)) return -12;
  if (kmax < 1) return -13;
  if (any && !rank) return -14;
  const int mn = m < n ? m : n;
  const int kq = kmax < mn ? kmax : mn;  // Q columns / Y rows that can be written
  if (any && !q) return -16;
  if (ldq < m) return -17;
  if (stride_q < (long long)ldq * kq) return -18;
  if (any && !y) return -19;
  if (ldy < kq) return -20;
  if (stride_y < (long long)ldy * n) return -21;
  if (nblk == 0) return 0;
  rrlu_eig_kernel<<<nblk, kRrluThreads, 0, (hipStream_t)stream>>>(
      fpole, difl, difr, u, v, korder, pl, pu, m, n, tol, kmax, rank, nswap, q, ldq, stride_q, y,
      ldy, stride_y);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "psmma_rrlu_cauchyeig_batched_dev: %s\n", hipGetErrorString(e));
    return 1000 + (int)e;
  }
  return 0;
}

extern "C" int psmma_rrlu_cauchy_batched_dev(int nblk, int m, int n, const int* pl, const int* pu,
                                             const double* d, const double* f, const double* u,
                                             const double* v, int nrow, int ncol, double tol,
                                             int kmax, int* rank, double* x, int ldx,
                                             long long stride_x, double* y, int ldy,
                                             long long stride_y, void* stream) {
  if (nblk < 0) return -1;
  if (m < 1 || m > kRrluMaxDim) return -2;
  if (n < 1 || n > kRrluMaxDim) return -3;
  const bool any = nblk > 0;
  if (any && !pl) return -4;
  if (any && !pu) return -5;
  if (any && !d) return -6;
  if (any && !f) return -7;
  if (any && !u) return -8;
  if (any && !v) return -9;
  if (nrow < 1) return -10;
  if (ncol < 1) return -11;
  if (!(tol >= 0.0)) return -12;
  if (kmax < 1) return -13;
  if (any && !rank) return -14;
  const int mn = m < n ? m : n;
  const int kq = kmax < mn ? kmax : mn;  // X columns / Y rows that can be written
  if (any && !x) return -15;
  if (ldx < m) return -16;
  if (stride_x < (long long)ldx * kq) return -17;
  if (any && !y) return -18;
  if (ldy < kq) return -19;
  if (stride_y < (long long)ldy * n) return -20;
  if (nblk == 0) return 0;
  rrlu_cauchy_kernel<<<nblk, kRrluThreads, 0, (hipStream_t)stream>>>(
      d, f, u, v, nrow, ncol, pl, pu, m, n, tol, kmax, rank, x, ldx, stride_x, y, ldy, stride_y);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "psmma_rrlu_cauchy_batched_dev: %s\n", hipGetErrorString(e));
    return 1000 + (int)e;
  }
  return 0;
}
