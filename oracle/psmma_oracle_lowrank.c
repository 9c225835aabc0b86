/* psmma_oracle_lowrank.c - CPU restatement of the reference's low-rank branch:
 * the truncated pivoted QR (SRC/mdgeqp3.f:153-370), Compress_RRQR
 * (SRC/Auxil_2p5D.f90:1060-1189) and ConstToeplitzlowrank
 * (SRC/Auxil_2p5D.f90:968-1057).
 * TEST INFRASTRUCTURE ONLY: used by tests/ and by bench.py's CPU baseline.
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

/* Compress_RRQR's TOL is the default-real literal 1.0E-13 stored into a DOUBLE
 * PRECISION variable (SRC/Auxil_2p5D.f90:1140): its value is the float nearest
 * to 1e-13. */
#define ORA_RRQR_TOL ((double)1.0e-13f)

/* Off-diagonal ring step of the faithful mode.  Toeplitz: ConstToeplitzlowrank
 * (explicit tile -> Compress_RRQR, SRC/Auxil_2p5D.f90:1050-1056; full rank ->
 * the explicit tile again).  The Cauchy kinds' generator
 * RRLU is not restated: those kinds report -2. */
int ora_lowrank_tile(const ora_problem* p, int kbp, int nq, const int* ridx, const int* cidx,
                     double* hbuf, unsigned long hcap, int* rk) {
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
