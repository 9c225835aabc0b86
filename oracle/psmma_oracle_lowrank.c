/* psmma_oracle_lowrank.c - CPU restatement of the reference's low-rank branch:
 * the truncated pivoted QR (SRC/mdgeqp3.f:153-370), Compress_RRQR
 * (SRC/Auxil_2p5D.f90:1060-1189) and ConstToeplitzlowrank
 * (SRC/Auxil_2p5D.f90:968-1057).
 * TEST INFRASTRUCTURE ONLY: used by tests/ and by bench.py's CPU baseline.
 * Also the plain Cauchy generator RRLU: rrluCauchy, CauchyPivtG(_CP),
 * searchMax2, CauchyMax, compress_cauchy, invp, Cauchylike2
 * (SRC/Cauchylowrank.f90:7-559) behind ConstCauchylowrank.
 *
 * The factorisation is Householder QR with column pivoting on the largest
 * partial column norm, with the partial norms downdated from the new row of R
 * and recomputed when the downdate loses more than half the digits (the
 * published LAPACK xGEQP3 scheme).  The reference's blocked panels (DLAQPS)
 * and this unblocked form choose the same pivots and produce the same R in
 * exact arithmetic; they differ in rounding only.  The reference's early exit
 * is kept at panel granularity: after each panel of nb columns MDGEQP3 stops
 * when |R(J-1,J-1)| <= TOL |R(1,1)| and reports J-1 as an upper bound of the
 * rank; Compress_RRQR then takes the rank as the number of leading diagonal
 * entries above TOL |R(1,1)|.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "psmma_oracle.h"

static double col_norm(int len, const double* x) {
  double s = 0.0;
  for (int i = 0; i < len; ++i) s += x[i] * x[i];
  return sqrt(s);
}

/* Householder reflector H = I - tau v v^T with v(0) = 1 mapping (alpha, x) to
 * (beta, 0), beta = -sign(alpha) ||(alpha, x)||; tau = 0 when x is zero. */
static void householder(int len, double* alpha, double* x, double* tau) {
  double xn = col_norm(len, x);
  if (xn == 0.0) {
    *tau = 0.0;
    return;
  }
  double beta = -copysign(hypot(*alpha, xn), *alpha);
  *tau = (beta - *alpha) / beta;
  double sc = 1.0 / (*alpha - beta);
  for (int i = 0; i < len; ++i) x[i] *= sc;
  *alpha = beta;
}

/* MDGEQP3 (SRC/mdgeqp3.f:153-370) on column-major A (lda).  jpvt: on entry
 * nonzero marks leading (fixed) columns, on exit jpvt[j] = 1-based original
 * column of A*P's column j.  tau: min(m,n) reflector scalars.  *rk: min(m,n),
 * or the upper bound J-1 of the rank at the panel exit. */
int ora_mdgeqp3(int m, int n, double* a, int lda, int* jpvt, double* tau, double tol, int nb,
                int* rk) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < (m > 1 ? m : 1)) return -4;
  int minmn = m < n ? m : n;
  *rk = minmn;
  if (minmn == 0) return 0;
  /* leading columns first (SRC/mdgeqp3.f:247-263) */
  int nfxd = 0;
  for (int j = 0; j < n; ++j) {
    if (jpvt[j] != 0) {
      if (j != nfxd) {
        for (int i = 0; i < m; ++i) {
          double t = a[(size_t)j * lda + i];
          a[(size_t)j * lda + i] = a[(size_t)nfxd * lda + i];
          a[(size_t)nfxd * lda + i] = t;
        }
        jpvt[j] = jpvt[nfxd];
        jpvt[nfxd] = j + 1;
      } else {
        jpvt[j] = j + 1;
      }
      ++nfxd;
    } else {
      jpvt[j] = j + 1;
    }
  }
  double* vn1 = (double*)malloc(sizeof(double) * (size_t)n);
  double* vn2 = (double*)malloc(sizeof(double) * (size_t)n);
  if (!vn1 || !vn2) {
    free(vn1);
    free(vn2);
    return -100;
  }
  const double tol3z = sqrt(ldexp(1.0, -53)); /* sqrt of the unit roundoff */
  double r11 = 0.0;
  for (int i = 0; i < minmn; ++i) {
    /* partial norms of the free columns, once the fixed ones are factored
       (SRC/mdgeqp3.f:310-313) */
    if (i == nfxd)
      for (int j = i; j < n; ++j) vn1[j] = vn2[j] = col_norm(m - i, a + (size_t)j * lda + i);
    int pvt = i;
    if (i >= nfxd) {
      for (int j = i + 1; j < n; ++j)
        if (vn1[j] > vn1[pvt]) pvt = j; /* first largest, as IDAMAX */
    }
    if (pvt != i) {
      for (int r = 0; r < m; ++r) {
        double t = a[(size_t)pvt * lda + r];
        a[(size_t)pvt * lda + r] = a[(size_t)i * lda + r];
        a[(size_t)i * lda + r] = t;
      }
      int t = jpvt[pvt];
      jpvt[pvt] = jpvt[i];
      jpvt[i] = t;
      vn1[pvt] = vn1[i];
      vn2[pvt] = vn2[i];
    }
    double* ci = a + (size_t)i * lda;
    if (i < m - 1)
      householder(m - i - 1, &ci[i], &ci[i + 1], &tau[i]);
    else
      tau[i] = 0.0;
    if (i == 0) r11 = fabs(ci[0]);
    /* apply H(i) to A(i:m, i+1:n) from the left */
    if (tau[i] != 0.0) {
      for (int j = i + 1; j < n; ++j) {
        double* cj = a + (size_t)j * lda;
        double s = cj[i];
        for (int r = i + 1; r < m; ++r) s += ci[r] * cj[r];
        s *= tau[i];
        cj[i] -= s;
        for (int r = i + 1; r < m; ++r) cj[r] -= s * ci[r];
      }
    }
    /* downdate the partial norms from row i of R */
    if (i >= nfxd) {
      for (int j = i + 1; j < n; ++j) {
        if (vn1[j] == 0.0) continue;
        double t = fabs(a[(size_t)j * lda + i]) / vn1[j];
        t = 1.0 - t * t;
        if (t < 0.0) t = 0.0;
        double t2 = t * (vn1[j] / vn2[j]) * (vn1[j] / vn2[j]);
        if (t2 <= tol3z) {
          if (i < m - 1) {
            vn1[j] = col_norm(m - i - 1, a + (size_t)j * lda + i + 1);
            vn2[j] = vn1[j];
          } else {
            vn1[j] = vn2[j] = 0.0;
          }
        } else {
          vn1[j] *= sqrt(t);
        }
      }
    }
    /* the reference's exit test after each full panel of nb columns
       (SRC/mdgeqp3.f:333-346) */
    int done = i + 1;
    if (nb > 1 && nb < minmn && done % nb == 0 && done < minmn &&
        fabs(a[(size_t)(done - 1) * lda + done - 1]) <= tol * r11) {
      *rk = done;
      break;
    }
  }
  free(vn1);
  free(vn2);
  return 0;
}

/* Compress_RRQR (SRC/Auxil_2p5D.f90:1110-1189): A (m x n, destroyed) ~ Q * (R P^T);
 * H = [Q (m x rk) | R P^T (rk x n)] when rk < min(m,n).  Returns rk; at full
 * rank H is left unwritten, as in the reference. */
int ora_compress_rrqr(int m, int n, double* a, int lda, double* h, double tol) {
  int minmn = m < n ? m : n;
  if (minmn == 0) return 0;
  int* jpvt = (int*)calloc((size_t)n, sizeof(int));
  double* tau = (double*)malloc(sizeof(double) * (size_t)minmn);
  int rkk = 0;
  if (!jpvt || !tau || ora_mdgeqp3(m, n, a, lda, jpvt, tau, tol, 32, &rkk) != 0) {
    free(jpvt);
    free(tau);
    return -1;
  }
  /* rank: leading diagonal entries above TOL |R(1,1)| (SRC/Auxil_2p5D.f90:1150-1156) */
  int rk = 0;
  while (rk < rkk && !(fabs(a[(size_t)rk * lda + rk]) <= tol * fabs(a[0]))) ++rk;
  if (rk < minmn) {
    double* r = h + (size_t)m * rk; /* rk x n, leading dimension rk */
    for (int j = 0; j < n; ++j) {
      double* dst = r + (size_t)(jpvt[j] - 1) * rk; /* DLAPMT backward: column j -> jpvt(j) */
      for (int i = 0; i < rk; ++i) dst[i] = (i <= j) ? a[(size_t)j * lda + i] : 0.0;
    }
    /* Q = H(0) ... H(rk-1) [I; 0] (DORGQR with rk reflectors) */
    double* q = h;
    for (int j = 0; j < rk; ++j)
      for (int i = 0; i < m; ++i) q[(size_t)j * m + i] = (i == j) ? 1.0 : 0.0;
    for (int l = rk - 1; l >= 0; --l) {
      const double* v = a + (size_t)l * lda; /* v(l) = 1, v(l+1:m) below the diagonal */
      if (tau[l] == 0.0) continue;
      for (int j = l; j < rk; ++j) {
        double* qj = q + (size_t)j * m;
        double s = qj[l];
        for (int i = l + 1; i < m; ++i) s += v[i] * qj[i];
        s *= tau[l];
        qj[l] -= s;
        for (int i = l + 1; i < m; ++i) qj[i] -= s * v[i];
      }
    }
  }
  free(jpvt);
  free(tau);
  return rk;
}

/* ------------------------------------------------------------------------- */
/* Generator-only structured RRLU of a Cauchy-like block (SRC/Cauchylowrank.f90),
 * restated with 0-based indices; each expression keeps the reference's
 * evaluation order. */

/* CauchyMax (SRC/Cauchylowrank.f90:166-188): largest |u_i / (D_i - f)| over n
 * entries and its first index; n = 0 gives -DBL_MAX (Fortran's maxval of an
 * empty array) and index -1. */
static double cauchy_max(const double* d, double f, const double* u, int n, int* at) {
  double best = -1.79769313486231570815e+308;
  *at = -1;
  for (int i = 0; i < n; ++i) {
    double x = fabs(u[i] / (d[i] - f));
    if (*at < 0 || x > best) {
      best = x;
      *at = i;
    }
  }
  return best;
}

static void dswp(double* x, int i, int j) {
  double t = x[i];
  x[i] = x[j];
  x[j] = t;
}
static void iswp(int* x, int i, int j) {
  int t = x[i];
  x[i] = x[j];
  x[j] = t;
}

/* CauchyPivtG (SRC/Cauchylowrank.f90:191-243): alternating row/column pivot
 * search from position k.  The reference's CONTINUE after a row swap is a
 * no-op, so a row swap falls through to a column swap; that order is kept.
 * The loop ends when the current pivot is the largest of its column or row
 * (an iteration cap guards the restatement against cycling on ties). */
static void cauchy_pivt_g(double* D, double* F, double* u, double* v, int* PL, int* PU, int k,
                          double* A, int M, int N) {
  int jjL, jjU;
  double junkL = cauchy_max(D + k, F[k], u + k, M - k, &jjL);
  double junkU = cauchy_max(F + k, D[k], v + k, N - k, &jjU);
  junkL = junkL * fabs(v[k]);
  junkU = junkU * fabs(u[k]);
  double piv = fabs(u[k] * v[k] / (D[k] - F[k]));
  if (junkL <= piv && junkU <= piv) return;
  int flg = (junkL > junkU) ? 1 : 0;
  for (long it = 0; it < 4L * (M + N) + 16; ++it) {
    if (flg == 1) {
      int jj = jjL;
      dswp(D, k, jj + k);
      dswp(u, k, jj + k);
      dswp(A, k, jj + k);
      iswp(PL, k, jj + k);
      cauchy_max(F + k, D[k], v + k, N - k, &jjU);
      if (jjU == 0) return;
      flg = 0;
    }
    int jj = jjU;
    dswp(F, k, jj + k);
    dswp(v, k, jj + k);
    iswp(PU, k, jj + k);
    cauchy_max(D + k, F[k], u + k, M - k, &jjL);
    if (jjL == 0) return;
    flg = 1;
  }
}

/* CauchyPivtG_CP (SRC/Cauchylowrank.f90:246-296): complete pivoting on the
 * k-th Schur complement. */
static void cauchy_pivt_cp(double* D, double* F, double* u, double* v, int* PL, int* PU, int k,
                           double* A, int M, int N) {
  int ii = 0, jj = 0;
  double junk = 0.0;
  for (int j = 0; j < N - k; ++j) {
    int at;
    double x = cauchy_max(D + k, F[k + j], u + k, M - k, &at);
    x = fabs(x * v[k + j]);
    if (j == 0 || x > junk) {
      junk = x;
      ii = at;
      jj = j;
    }
  }
  dswp(D, k, ii + k);
  dswp(u, k, ii + k);
  dswp(A, k, ii + k);
  iswp(PL, k, ii + k);
  dswp(F, k, jj + k);
  dswp(v, k, jj + k);
  iswp(PU, k, jj + k);
}

/* rrluCauchy (SRC/Cauchylowrank.f90:7-134).  D, F, A, V are modified in place
 * (pivoted; V and the returned U hold the Schur-complement generators, W the
 * interpolation generators).  PL/PU: 1-based permutations.  Returns Rk. */
int ora_rrlu_cauchy(double* D, double* F, double* A, double* V, double tol, double* U,
                    double* W, int* PL, int* PU, int M, int N) {
  for (int i = 0; i < M; ++i) U[i] = A[i];
  int rk = M < N ? M : N;
  const int mn = rk;
  for (int i = 0; i < M; ++i) PL[i] = i + 1;
  for (int j = 0; j < N; ++j) PU[j] = j + 1;
  double amax = 0.0;
  const double rho = 1.12;
  const int prd = 10;
  for (int k = 0; k < mn; ++k) { /* the reference's k = this k + 1 */
    cauchy_pivt_g(D, F, U, V, PL, PU, k, A, M, N);
    double ukk = U[k] * V[k] / (D[k] - F[k]);
    amax = fmax(amax, fabs(ukk));
    if (fabs(ukk) < tol * amax) {
      cauchy_pivt_cp(D, F, U, V, PL, PU, k, A, M, N);
      ukk = U[k] * V[k] / (D[k] - F[k]);
      amax = fmax(amax, fabs(ukk));
      if (fabs(ukk) < tol * amax) {
        rk = k;
        break;
      }
    }
    for (int i = k + 1; i < M; ++i) U[i] = U[i] * (D[i] - D[k]) / (D[i] - F[k]);
    for (int j = k + 1; j < N; ++j) V[j] = V[j] * (F[j] - F[k]) / (F[j] - D[k]);
    for (int j = 0; j < k; ++j) W[j] = W[j] * ((F[k] - D[j]) / (D[k] - D[j]));
    W[k] = (D[k] - F[k]) / A[k];
    for (int j = 0; j < k; ++j) W[k] = W[k] * (F[j] - D[k]) / (D[j] - D[k]);
    /* the swap pass runs every prd-th step (flgl = mod(k, prd) < 1) and repeats while
     * it exchanges a row; the reference's loop has no bound, the restatement stops after
     * 64 (M + N) exchanges (each genuine exchange grows the chosen block's determinant
     * by more than rho; a longer run is rounding cycling between rows) */
    int again = ((k + 1) % prd) < 1;
    for (long pass = 0; again && pass < 64L * (M + N); ++pass) {
      again = 0;
      /* searchMax2 (SRC/Cauchylowrank.f90:137-163) over the k+1 chosen rows */
      int ii = 0, jjl;
      double nmax = cauchy_max(D + k + 1, D[0], U + k + 1, M - k - 1, &jjl) * fabs(W[0]);
      int jj = jjl;
      for (int j = 1; j <= k; ++j) {
        double x = cauchy_max(D + k + 1, D[j], U + k + 1, M - k - 1, &jjl) * fabs(W[j]);
        if (x > nmax) {
          nmax = x;
          ii = j;
          jj = jjl;
        }
      }
      if (nmax > rho) {
        again = 1;
        jj = jj + k + 1; /* 0-based row of the Schur complement */
        for (int j = k + 1; j < N; ++j) V[j] = V[j] * ((F[j] - D[ii]) / (F[j] - D[jj]));
        for (int i = k + 1; i < jj; ++i) U[i] = U[i] * ((D[i] - D[jj]) / (D[i] - D[ii]));
        for (int i = jj + 1; i < M; ++i) U[i] = U[i] * ((D[i] - D[jj]) / (D[i] - D[ii]));
        for (int j = 0; j < ii; ++j) W[j] = W[j] * ((D[j] - D[ii]) / (D[j] - D[jj]));
        for (int j = ii + 1; j <= k; ++j) W[j] = W[j] * ((D[j] - D[ii]) / (D[j] - D[jj]));
        U[jj] = A[ii] * ((D[ii] - D[jj]) / (D[ii] - F[ii]));
        W[ii] = (D[jj] - F[ii]) / A[jj];
        for (int j = 0; j < ii; ++j) {
          U[jj] = U[jj] * ((D[ii] - D[j]) / (D[ii] - F[j]));
          W[ii] = W[ii] * ((F[j] - D[jj]) / (D[j] - D[jj]));
        }
        for (int j = ii + 1; j <= k; ++j) {
          U[jj] = U[jj] * ((D[ii] - D[j]) / (D[ii] - F[j]));
          W[ii] = W[ii] * ((F[j] - D[jj]) / (D[j] - D[jj]));
        }
        iswp(PL, ii, jj);
        dswp(D, ii, jj);
        dswp(A, ii, jj);
      }
    }
  }
  return rk;
}

/* compress_cauchy('R', ...) (SRC/Cauchylowrank.f90:387-491): row ID of the
 * Cauchy-like block U_i V_j / (D_i - F_j) (M x N).  When Rk < M:
 * H = [X (M x Rk) | Y (Rk x N)], X = P^-1 [I; Z_i W_j / (D_i - D_j)],
 * Y = the chosen rows of the block from the original F and V; otherwise H is
 * left unwritten (the reference prints a full-rank warning).  Returns Rk. */
int ora_compress_cauchy(const double* D, const double* F, const double* U, const double* V,
                        double tol, int M, int N, double* H) {
  int mn = M < N ? M : N;
  double* LD = (double*)malloc(sizeof(double) * (size_t)(M + 1));
  double* LU = (double*)malloc(sizeof(double) * (size_t)(M + 1));
  double* LF = (double*)malloc(sizeof(double) * (size_t)(N + 1));
  double* LV = (double*)malloc(sizeof(double) * (size_t)(N + 1));
  double* Z = (double*)calloc((size_t)(M + 1), sizeof(double));
  double* W = (double*)calloc((size_t)(mn + 1), sizeof(double));
  int* PL = (int*)malloc(sizeof(int) * (size_t)(M + 1));
  int* PU = (int*)malloc(sizeof(int) * (size_t)(N + 1));
  int* IP = (int*)malloc(sizeof(int) * (size_t)(M + 1));
  int rk = -1;
  if (LD && LU && LF && LV && Z && W && PL && PU && IP) {
    memcpy(LD, D, sizeof(double) * (size_t)M);
    memcpy(LU, U, sizeof(double) * (size_t)M);
    memcpy(LF, F, sizeof(double) * (size_t)N);
    memcpy(LV, V, sizeof(double) * (size_t)N);
    rk = ora_rrlu_cauchy(LD, LF, LU, LV, tol, Z, W, PL, PU, M, N);
    for (int i = 0; i < M; ++i) IP[PL[i] - 1] = i; /* invp */
    if (rk < M) {
      double* X = H;
      for (int j = 0; j < rk; ++j)
        for (int i = 0; i < M; ++i) {
          const int p = IP[i]; /* row of the pivoted order that is original row i */
          double x;
          if (p < rk)
            x = (p == j) ? 1.0 : 0.0;
          else
            x = Z[p] * W[j] / (LD[p] - LD[j]); /* Cauchylike2 */
          X[(size_t)j * M + i] = x;
        }
      double* Y = H + (size_t)M * rk;
      for (int j = 0; j < N; ++j)
        for (int i = 0; i < rk; ++i) Y[(size_t)j * rk + i] = LU[i] * V[j] / (LD[i] - F[j]);
    }
  }
  free(LD);
  free(LU);
  free(LF);
  free(LV);
  free(Z);
  free(W);
  free(PL);
  free(PU);
  free(IP);
  return rk;
}

/* ------------------------------------------------------------------------- */
/* Generator-only RRLU of an eigenvector (eigen-Cauchy) block
 * (SRC/CauchyHssEig.f90), restated with 0-based positions.  The row and column
 * poles are never formed as differences of stored values: every F(i) - D(j) and
 * D(i) - D(j) is rebuilt from the global old poles FF and the DIFL / DIFR gaps
 * through the global indices PL(i) / PU(j) (1-based, as in the reference), so
 * the block entry is U_i V_j / fd(PL_i, PU_j).  The column poles D themselves
 * are only permuted by the reference, never read, and are not carried here. */

/* computfd (SRC/CauchyHssEig.f90:519-568): F(row) - D(col) from global indices. */
static double eig_fd(const double* ff, const double* difl, const double* difr, int pl, int pu) {
  if (pl <= pu) return (ff[pl - 1] - ff[pu - 1]) - difl[pu - 1];
  return (ff[pl - 1] - ff[pu]) - difr[pu - 1];
}

/* computdd (SRC/CauchyHssEig.f90:470-516): D(j) - D(k) from global indices. */
static double eig_dd(const double* ff, const double* difl, const double* difr, int pj, int pk) {
  if (pj < pk) return (ff[pj] - ff[pk - 1]) + difr[pj - 1] - difl[pk - 1];
  return (ff[pj - 1] - ff[pk]) + difl[pj - 1] - difr[pk - 1];
}

/* CauchyMaxEigmc (SRC/CauchyHssEig.f90:1507-1589): largest |u_i / fd(PL_i, pu)|
 * down one column over n rows, first index of the maximum. */
static double eig_max_col(const double* ff, const double* difl, const double* difr,
                          const int* pl, int pu, const double* u, int n, int* at) {
  double best = -1.79769313486231570815e+308;
  *at = -1;
  for (int i = 0; i < n; ++i) {
    double x = fabs(u[i] / eig_fd(ff, difl, difr, pl[i], pu));
    if (*at < 0 || x > best) {
      best = x;
      *at = i;
    }
  }
  return best;
}

/* CauchyMaxEigm (SRC/CauchyHssEig.f90:1419-1503): largest |v_j / fd(pl, PU_j)|
 * along one row over n columns. */
static double eig_max_row(const double* ff, const double* difl, const double* difr, int pl,
                          const int* pu, const double* v, int n, int* at) {
  double best = -1.79769313486231570815e+308;
  *at = -1;
  for (int j = 0; j < n; ++j) {
    double x = fabs(v[j] / eig_fd(ff, difl, difr, pl, pu[j]));
    if (*at < 0 || x > best) {
      best = x;
      *at = j;
    }
  }
  return best;
}

typedef struct {
  const double *ff, *difl, *difr;
  double *F, *U, *V, *A;
  int *PL, *PU, *PQ;
  int M, N;
} eig_rrlu;

static void eig_swap_rows(eig_rrlu* e, int i, int j) {
  dswp(e->F, i, j);
  dswp(e->U, i, j);
  dswp(e->A, i, j);
  iswp(e->PL, i, j);
  iswp(e->PQ, i, j);
}
static void eig_swap_cols(eig_rrlu* e, int i, int j) {
  dswp(e->V, i, j);
  iswp(e->PU, i, j);
}

/* CauchyPivtEig (SRC/CauchyHssEig.f90:571-717): alternating row/column pivot
 * search from position k; as in CauchyPivtG the CONTINUE after a row swap
 * falls through to a column swap.  The early-return test's pivot magnitude is
 * written F(PL(k)) in the reference (:675-678), i.e. the local row-pole array indexed by
 * a global row index (past its end for every block row after the first); the
 * restatement uses the value of the surrounding terms' convention, FF(PL(k)),
 * which is the pivot entry itself.  An iteration cap guards against cycling
 * on ties. */
static void eig_pivt(eig_rrlu* e, int k) {
  const int M = e->M, N = e->N;
  int jjL, jjU;
  double junkL = eig_max_col(e->ff, e->difl, e->difr, e->PL + k, e->PU[k], e->U + k, M - k, &jjL);
  double junkU = eig_max_row(e->ff, e->difl, e->difr, e->PL[k], e->PU + k, e->V + k, N - k, &jjU);
  junkL = junkL * fabs(e->V[k]);
  junkU = junkU * fabs(e->U[k]);
  double piv = fabs(e->U[k] * e->V[k] / eig_fd(e->ff, e->difl, e->difr, e->PL[k], e->PU[k]));
  if (junkL <= piv && junkU <= piv) return;
  int flg = (junkL > junkU) ? 1 : 0;
  for (long it = 0; it < 4L * (M + N) + 16; ++it) {
    if (flg == 1) {
      eig_swap_rows(e, k, jjL + k);
      eig_max_row(e->ff, e->difl, e->difr, e->PL[k], e->PU + k, e->V + k, N - k, &jjU);
      if (jjU == 0) return;
      flg = 0;
    }
    eig_swap_cols(e, k, jjU + k);
    eig_max_col(e->ff, e->difl, e->difr, e->PL + k, e->PU[k], e->U + k, M - k, &jjL);
    if (jjL == 0) return;
    flg = 1;
  }
}

/* CauchyPivtEig_CP (SRC/CauchyHssEig.f90:1004-1134): complete pivoting on the
 * k-th Schur complement. */
static void eig_pivt_cp(eig_rrlu* e, int k) {
  int ii = 0, jj = 0;
  double junk = 0.0;
  for (int j = 0; j < e->N - k; ++j) {
    int at;
    double x = eig_max_col(e->ff, e->difl, e->difr, e->PL + k, e->PU[k + j], e->U + k, e->M - k,
                           &at);
    x = fabs(x * e->V[k + j]);
    if (j == 0 || x > junk) {
      junk = x;
      ii = at;
      jj = j;
    }
  }
  eig_swap_rows(e, k, ii + k);
  eig_swap_cols(e, k, jj + k);
}

/* rrluCauchyEigrow (SRC/CauchyHssEig.f90:7-212).  F: the block's row poles
 * FF(PL_i); A: its row generators (both permuted with the rows); V: column
 * generators (permuted and turned into Schur-complement generators); U, W:
 * Schur-complement and interpolation generators.  PL / PU: global 1-based
 * row / column indices, permuted.  PQ: the row permutation (1-based).
 * Returns Rk; *nswap counts the swap passes that moved a row. */
int ora_rrlu_cauchyeig_row(const double* ff, const double* difl, const double* difr, double* F,
                           double* A, double* V, double tol, double* U, double* W, int* PL,
                           int* PU, int* PQ, int M, int N, int* nswap) {
  eig_rrlu e = {ff, difl, difr, F, U, V, A, PL, PU, PQ, M, N};
#define FD(j, k) eig_fd(ff, difl, difr, PL[j], PU[k])
  for (int i = 0; i < M; ++i) U[i] = A[i];
  for (int i = 0; i < M; ++i) PQ[i] = i + 1;
  int rk = M < N ? M : N;
  const int mn = rk;
  int swaps = 0;
  double amax = 0.0;
  const double rho = (double)1.2f; /* rho = 1.2E+0, a default-real literal (SRC/CauchyHssEig.f90:98) */
  const int prd = 5;
  for (int k = 0; k < mn; ++k) { /* the reference's k = this k + 1 */
    eig_pivt(&e, k);
    double temp = FD(k, k);
    double ukk = U[k] * V[k] / temp;
    amax = fmax(amax, fabs(ukk));
    if (fabs(ukk) < tol * amax) {
      eig_pivt_cp(&e, k);
      temp = FD(k, k);
      ukk = U[k] * V[k] / temp;
      amax = fmax(amax, fabs(ukk));
      if (fabs(ukk) < tol * amax) {
        rk = k;
        break;
      }
    }
    for (int j = k + 1; j < M; ++j) {
      temp = FD(j, k);
      temp = (F[j] - F[k]) / temp;
      U[j] = U[j] * temp;
    }
    for (int j = k + 1; j < N; ++j) {
      temp = eig_dd(ff, difl, difr, PU[j], PU[k]);
      double temp2 = -FD(k, j);
      temp = temp / temp2;
      V[j] = V[j] * temp;
    }
    for (int j = 0; j < k; ++j) {
      temp = -FD(j, k);
      W[j] = W[j] * (temp / (F[k] - F[j]));
    }
    temp = FD(k, k);
    W[k] = temp / A[k];
    for (int j = 0; j < k; ++j) {
      temp = -FD(k, j);
      W[k] = W[k] * (temp / (F[j] - F[k]));
    }
    /* the swap pass runs every prd-th step (flgl = mod(k, prd) < 1) and repeats while
     * it exchanges a row; the reference's loop has no bound, the restatement stops after
     * 64 (M + N) exchanges (each genuine exchange grows the chosen block's determinant
     * by more than rho; a longer run is rounding cycling between rows) */
    int again = ((k + 1) % prd) < 1;
    for (long pass = 0; again && pass < 64L * (M + N); ++pass) {
      again = 0;
      /* searchMax2Eig (SRC/CauchyHssEig.f90:1137-1203) over the k+1 chosen rows */
      int ii = 0, jjl;
      double nmax = cauchy_max(F + k + 1, F[0], U + k + 1, M - k - 1, &jjl) * fabs(W[0]);
      int jj = jjl;
      for (int j = 1; j <= k; ++j) {
        double x = cauchy_max(F + k + 1, F[j], U + k + 1, M - k - 1, &jjl) * fabs(W[j]);
        if (x > nmax) {
          nmax = x;
          ii = j;
          jj = jjl;
        }
      }
      if (nmax > rho) {
        swaps++;
        again = 1;
        jj = jj + k + 1; /* 0-based row of the Schur complement */
        for (int j = k + 1; j < N; ++j) {
          temp = -FD(ii, j);
          double temp2 = -FD(jj, j);
          temp = temp / temp2;
          V[j] = V[j] * temp;
        }
        for (int i = k + 1; i < jj; ++i) U[i] = U[i] * ((F[i] - F[jj]) / (F[i] - F[ii]));
        for (int i = jj + 1; i < M; ++i) U[i] = U[i] * ((F[i] - F[jj]) / (F[i] - F[ii]));
        for (int j = 0; j < ii; ++j) W[j] = W[j] * ((F[j] - F[ii]) / (F[j] - F[jj]));
        for (int j = ii + 1; j <= k; ++j) W[j] = W[j] * ((F[j] - F[ii]) / (F[j] - F[jj]));
        temp = FD(ii, ii);
        U[jj] = A[ii] * ((F[ii] - F[jj]) / temp);
        temp = FD(jj, ii);
        W[ii] = temp / A[jj];
        for (int j = 0; j <= k; ++j) {
          if (j == ii) continue;
          temp = FD(ii, j);
          U[jj] = U[jj] * ((F[ii] - F[j]) / temp);
          temp = -FD(jj, j);
          W[ii] = W[ii] * (temp / (F[j] - F[jj]));
        }
        iswp(PL, ii, jj);
        iswp(PQ, ii, jj);
        dswp(F, ii, jj);
        dswp(A, ii, jj);
      }
    }
  }
#undef FD
  if (nswap) *nswap = swaps;
  return rk;
}

/* Compress_cauchyEig2 (SRC/CauchyHssEig.f90:2640-2829): row ID of the M x N
 * eigen-Cauchy block with global rows PL / columns PU and gathered generators
 * u (M) / v (N).  When Rk < min(M, N): H = [Q (M x Rk) | Y (Rk x N)],
 * Q = P^-1 [I; Z_i W_j / (F_i - F_j)] (CauchylikeEigQ, :1882-1952) and Y the
 * chosen rows of the block rebuilt from the original u / v (CauchylikeEig,
 * :1693-1774); otherwise H is the explicit block.  Returns Rk. */
int ora_compress_cauchyeig2(const double* ff, const double* difl, const double* difr,
                            const double* u, const double* v, const int* pl, const int* pu,
                            double tol, int M, int N, double* H, int* nswap) {
  int mn = M < N ? M : N;
  double* F = (double*)malloc(sizeof(double) * (size_t)(M + 1));
  double* A = (double*)malloc(sizeof(double) * (size_t)(M + 1));
  double* V = (double*)malloc(sizeof(double) * (size_t)(N + 1));
  double* Z = (double*)calloc((size_t)(M + 1), sizeof(double));
  double* W = (double*)calloc((size_t)(mn + 1), sizeof(double));
  int* PL = (int*)malloc(sizeof(int) * (size_t)(M + 1));
  int* PU = (int*)malloc(sizeof(int) * (size_t)(N + 1));
  int* PQ = (int*)malloc(sizeof(int) * (size_t)(M + 1));
  int* IP = (int*)malloc(sizeof(int) * (size_t)(M + 1));
  int rk = -1;
  if (F && A && V && Z && W && PL && PU && PQ && IP) {
    for (int i = 0; i < M; ++i) {
      F[i] = ff[pl[i] - 1];
      A[i] = u[i];
      PL[i] = pl[i];
    }
    for (int j = 0; j < N; ++j) {
      V[j] = v[j];
      PU[j] = pu[j];
    }
    rk = ora_rrlu_cauchyeig_row(ff, difl, difr, F, A, V, tol, Z, W, PL, PU, PQ, M, N, nswap);
    for (int i = 0; i < M; ++i) IP[PQ[i] - 1] = i; /* invp */
    if (rk < mn) {
      double* Q = H;
      for (int j = 0; j < rk; ++j)
        for (int i = 0; i < M; ++i) {
          const int p = IP[i]; /* row of the pivoted order that is original row i */
          double x;
          if (p < rk)
            x = (p == j) ? 1.0 : 0.0;
          else
            x = Z[p] * W[j] / (F[p] - F[j]);
          Q[(size_t)j * M + i] = x;
        }
      double* Y = H + (size_t)M * rk;
      for (int j = 0; j < N; ++j)
        for (int i = 0; i < rk; ++i) {
          const int pj = pu[j], pi = PL[i];
          Y[(size_t)j * rk + i] = A[i] * v[j] / eig_fd(ff, difl, difr, pi, pj);
        }
    } else {
      for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i)
          H[(size_t)j * M + i] = u[i] * v[j] / eig_fd(ff, difl, difr, pl[i], pu[j]);
    }
  }
  free(F);
  free(A);
  free(V);
  free(Z);
  free(W);
  free(PL);
  free(PU);
  free(PQ);
  free(IP);
  return rk;
}

/* Compress_RRQR's TOL is the default-real literal 1.0E-13 stored into a DOUBLE
 * PRECISION variable (SRC/Auxil_2p5D.f90:1140): its value is the float nearest
 * to 1e-13. */
#define ORA_RRQR_TOL ((double)1.0e-13f)

/* Off-diagonal ring step of the faithful mode.  Cauchy: the generator RRLU
 * above.  Toeplitz: ConstToeplitzlowrank (explicit tile -> Compress_RRQR,
 * SRC/Auxil_2p5D.f90:1050-1056; full rank -> the explicit tile again).
 * Eigen-Cauchy: ConstCauchyEiglowrank (SRC/Auxil_2p5D.f90:642-748) ->
 * Compress_cauchyEig2, which rebuilds the explicit tile at full rank. */
int ora_lowrank_tile(const ora_problem* p, int kbp, int nq, const int* ridx, const int* cidx,
                     double* hbuf, unsigned long hcap, int* rk) {
  if (p->kind == ORA_CAUCHYEIG) {
    /* generators gathered by the index sets (:731-735), TOL = the default-real
     * literal 1.0E-17 (:729); gen = [F D U V DIFL DIFR] with stride ldgen */
    if ((unsigned long)kbp * (unsigned long)nq * 2ul > hcap) return -3;
    const size_t L = (size_t)p->ldgen;
    const double* g = p->gen;
    double* uv = (double*)malloc(sizeof(double) * (size_t)(kbp + nq + 2));
    if (!uv) return -1;
    for (int i = 0; i < kbp; ++i) uv[i] = g[2 * L + ridx[i] - 1];
    for (int j = 0; j < nq; ++j) uv[kbp + j] = g[3 * L + cidx[j] - 1];
    int r = ora_compress_cauchyeig2(g, g + 4 * L, g + 5 * L, uv, uv + kbp, ridx, cidx,
                                    (double)1.0e-17f, kbp, nq, hbuf, NULL);
    free(uv);
    if (r < 0) return -1;
    *rk = r;
    return ORA_TILE_OK;
  }
  if (p->kind == ORA_CAUCHY) {
    /* ConstCauchylowrank (SRC/Auxil_2p5D.f90:876-963): generators gathered by
     * the index sets, compress_cauchy('R') at the default-real literal
     * TOL = 1.0E-17 (:946).  At full rank H keeps whatever the previous step
     * left there and the ring multiplies that stale tile
     * (SRC/pscauchy_compute.f90:379-381). */
    if ((unsigned long)kbp * (unsigned long)nq * 2ul > hcap) return -3;
    const size_t L = (size_t)p->ldgen;
    double* g = (double*)malloc(sizeof(double) * (size_t)(2 * kbp + 2 * nq + 4));
    if (!g) return -1;
    double *d = g, *u = g + kbp, *f = g + 2 * kbp, *v = g + 2 * kbp + nq;
    for (int i = 0; i < kbp; ++i) {
      d[i] = p->gen[ridx[i] - 1];
      u[i] = p->gen[2 * L + ridx[i] - 1];
    }
    for (int j = 0; j < nq; ++j) {
      f[j] = p->gen[L + cidx[j] - 1];
      v[j] = p->gen[3 * L + cidx[j] - 1];
    }
    int r = ora_compress_cauchy(d, f, u, v, (double)1.0e-17f, kbp, nq, hbuf);
    free(g);
    if (r < 0) return -1;
    *rk = r;
    return (r >= kbp) ? ORA_TILE_STALE : ORA_TILE_OK;
  }
  if (p->kind != ORA_TOEPLITZ) return -2;
  if ((unsigned long)kbp * (unsigned long)nq * 2ul > hcap) return -3;
  double* t = (double*)malloc(sizeof(double) * (size_t)kbp * (size_t)(nq > 0 ? nq : 1));
  if (!t) return -1;
  ora_const_toeplitz(kbp, nq, p->gen, p->k, t, ridx, cidx);
  int r = ora_compress_rrqr(kbp, nq, t, kbp, hbuf, ORA_RRQR_TOL);
  free(t);
  if (r < 0) return -1;
  int minmn = kbp < nq ? kbp : nq;
  if (r == minmn) ora_const_toeplitz(kbp, nq, p->gen, p->k, hbuf, ridx, cidx);
  *rk = r;
  return ORA_TILE_OK;
}
