/*
 * psmma_oracle.c - CPU restatement of the reference PSMMA algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP kernels and
 * the CPU baseline that bench.py times; the product library never links it.
 *
 * Each function cites the reference routine it restates (paths relative to the
 * reference repository).  Arrays are column-major, indices handed to and from
 * the index-map routines are 1-based as in the Fortran.
 *
 * The emulated process grid: a P x Q BLACS grid is reproduced by running the
 * per-rank body of PS*_COMPUTE for every (myrow, mycol).  The reference's ring
 * shift (DGESD2D/DGERV2D, SRC/pscauchy_compute.f90:414-423) delivers to rank
 * (myrow, mycol) at step K1 the local A of process column (mycol+K1) mod Q; the
 * emulation reads that block column of the global A directly, which is the
 * same data.  Ranks own disjoint parts of C, so they may run in parallel
 * threads (one per rank, as the reference runs one MPI rank per core).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "psmma_oracle.h"

/* ScaLAPACK NUMROC: number of rows/cols of an n-vector distributed in blocks
 * of nb over nprocs processes owned by iproc (isrcproc = 0). */
int ora_numroc(int n, int nb, int iproc, int nprocs) {
  int nblocks = n / nb;
  int num = (nblocks / nprocs) * nb;
  int extrablks = nblocks % nprocs;
  if (iproc < extrablks)
    num += nb;
  else if (iproc == extrablks)
    num += n % nb;
  return num;
}

/* SRC/Auxil_2p5D.f90:2-66 (ConstIndex): the 1-based global indices owned by
 * process column `mycol` of `npcol` under a block-cyclic(nb) layout of n
 * indices, in local order.  Returns the count. */
int ora_constindex(int n, int nb, int mycol, int npcol, int* idx) {
  int knb = n / nb;
  int nfseg = knb / npcol;
  int cstart = mycol * nb + 1;
  int gap = npcol * nb;
  int tmp = 0;
  for (int i = 0; i < nfseg; ++i) {
    for (int j = 0; j < nb; ++j) idx[tmp + j] = cstart + j;
    cstart += gap;
    tmp += nb;
  }
  int left = n - nfseg * gap - mycol * nb;
  if (left > 0) {
    int len = left < nb ? left : nb;
    for (int j = 0; j < len; ++j) idx[tmp + j] = cstart + j;
    tmp += len;
  }
  return tmp;
}

/* SRC/Auxil_2p5D.f90:752-792 (is_intersect).  NOTE the reference's sense: the
 * flag it returns is TRUE when the sets share NO element (it is then used as
 * "is low rank", SRC/pscauchy_compute.f90:368-371).  Only the first min(m,n)
 * elements of the shorter set are probed, as in the reference. */
int ora_is_intersect(int m, const int* seta, int n, const int* setb) {
  int mn = m < n ? m : n;
  if (m > n) {
    for (int i = 0; i < mn; ++i)
      for (int j = 0; j < m; ++j)
        if (seta[j] == setb[i]) return 0;
  } else {
    for (int i = 0; i < mn; ++i)
      for (int j = 0; j < n; ++j)
        if (setb[j] == seta[i]) return 0;
  }
  return 1;
}

/* SRC/Auxil_2p5D.f90:68-151 (ConstCauchy) + SRC/Cauchylowrank.f90:509-521
 * (Cauchylike): T(i,j) = U(R_i) V(C_j) / (D(R_i) - F(C_j)), B(LDB,4)=[D F U V]. */
void ora_const_cauchy(int lm, int ln, const double* b, int ldb, double* t, const int* ridx,
                      const int* cidx) {
  const double *D = b, *F = b + ldb, *U = b + 2 * (size_t)ldb, *V = b + 3 * (size_t)ldb;
  for (int j = 0; j < ln; ++j) {
    int cj = cidx[j] - 1;
    for (int i = 0; i < lm; ++i) {
      int ri = ridx[i] - 1;
      t[(size_t)j * lm + i] = U[ri] * V[cj] / (D[ri] - F[cj]);
    }
  }
}

/* SRC/Auxil_2p5D.f90:153-236 (ConstToeplitz): T(i,j) = G(K + C_j - R_i). */
void ora_const_toeplitz(int lm, int ln, const double* g, int k, double* t, const int* ridx,
                        const int* cidx) {
  for (int j = 0; j < ln; ++j)
    for (int i = 0; i < lm; ++i) t[(size_t)j * lm + i] = g[k + cidx[j] - ridx[i] - 1];
}

/* SRC/Auxil_2p5D.f90:238-322 (ConstHankel): T(i,j) = G(R_i + C_j - 1). */
void ora_const_hankel(int lm, int ln, const double* g, double* t, const int* ridx,
                      const int* cidx) {
  for (int j = 0; j < ln; ++j)
    for (int i = 0; i < lm; ++i) t[(size_t)j * lm + i] = g[ridx[i] + cidx[j] - 2];
}

/* SRC/Auxil_2p5D.f90:532-637 (ConstCauchyEig) + SRC/CauchyHssEig.f90:1693-1774
 * (CauchylikeEig): with PL = R_i, PU = C_j,
 *   A(i,j) = U(PL) V(PU) / ((F(PL) - F(PU)) - DIFL(PU))      if PL <= PU
 *          = U(PL) V(PU) / ((F(PL) - F(PU+1)) - DIFR(PU))    otherwise. */
void ora_const_cauchyeig(int lm, int ln, double* t, const double* f, const double* u,
                         const double* v, const double* difl, const double* difr,
                         const int* ridx, const int* cidx) {
  for (int j = 0; j < ln; ++j) {
    int pu = cidx[j];
    double fj = f[pu - 1], vj = v[pu - 1];
    for (int i = 0; i < lm; ++i) {
      int pl = ridx[i];
      double den = (pl <= pu) ? ((f[pl - 1] - fj) - difl[pu - 1])
                              : ((f[pl - 1] - f[pu]) - difr[pu - 1]);
      t[(size_t)j * lm + i] = u[pl - 1] * vj / den;
    }
  }
}

/* The ring's DGEMM: the host BLAS's Fortran dgemm_ when one is set (the reference
 * links MKL, SLMakes/SLmake.inc.intel:24-35), else ora_dgemm below. */
typedef void (*ora_blas_dgemm_fn)(const char*, const char*, const int*, const int*, const int*,
                                  const double*, const double*, const int*, const double*,
                                  const int*, const double*, double*, const int*);
static ora_blas_dgemm_fn g_blas_dgemm = NULL;
void ora_dgemm(int m, int n, int k, double alpha, const double* a, int lda, const double* b,
               int ldb, double beta, double* c, int ldc);

void ora_set_blas_dgemm(void* fn) { g_blas_dgemm = (ora_blas_dgemm_fn)fn; }

static void ring_dgemm(int m, int n, int k, double alpha, const double* a, int lda,
                       const double* b, int ldb, double beta, double* c, int ldc) {
  if (g_blas_dgemm && m > 0 && n > 0) {
    const char nn = 'N';
    int ka = lda > 0 ? lda : 1, kb = ldb > 0 ? ldb : 1, kc = ldc > 0 ? ldc : 1;
    g_blas_dgemm(&nn, &nn, &m, &n, &k, &alpha, a, &ka, b, &kb, &beta, c, &kc);
    return;
  }
  ora_dgemm(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

/* DGEMM('N','N') semantics: C = alpha*A*B + beta*C (beta == 0 => C not read).
 * Blocked for cache; FP64 accumulation in natural k order within a block. */
void ora_dgemm(int m, int n, int k, double alpha, const double* a, int lda, const double* b,
               int ldb, double beta, double* c, int ldc) {
  for (int j = 0; j < n; ++j) {
    double* cj = c + (size_t)j * ldc;
    if (beta == 0.0)
      for (int i = 0; i < m; ++i) cj[i] = 0.0;
    else if (beta != 1.0)
      for (int i = 0; i < m; ++i) cj[i] *= beta;
  }
  if (alpha == 0.0 || k == 0) return;
  enum { MB = 128, KB = 256 };
  for (int k0 = 0; k0 < k; k0 += KB) {
    int kb = (k - k0) < KB ? (k - k0) : KB;
    for (int i0 = 0; i0 < m; i0 += MB) {
      int mb = (m - i0) < MB ? (m - i0) : MB;
      for (int j = 0; j < n; ++j) {
        double* cj = c + (size_t)j * ldc + i0;
        const double* bj = b + (size_t)j * ldb + k0;
        double acc[MB];
        for (int i = 0; i < mb; ++i) acc[i] = 0.0;
        for (int kk = 0; kk < kb; ++kk) {
          double s = bj[kk];
          const double* ak = a + (size_t)(k0 + kk) * lda + i0;
          for (int i = 0; i < mb; ++i) acc[i] += ak[i] * s;
        }
        for (int i = 0; i < mb; ++i) cj[i] += alpha * acc[i];
      }
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Ring emulation of PS*_COMPUTE (dense branch at every step, or the
 * reference's low-rank branch for disjoint index sets when mode != 0). */

typedef struct {
  const ora_problem* p;
  int myrow, mycol;
  int status;
  long long lowrank_steps, dense_steps, fullrank_steps, stale_steps;
} rank_job;

static void build_tile(const ora_problem* p, int lm, int ln, const int* ridx, const int* cidx,
                       double* t) {
  switch (p->kind) {
    case ORA_CAUCHY: ora_const_cauchy(lm, ln, p->gen, p->ldgen, t, ridx, cidx); break;
    case ORA_TOEPLITZ: ora_const_toeplitz(lm, ln, p->gen, p->k, t, ridx, cidx); break;
    case ORA_HANKEL: ora_const_hankel(lm, ln, p->gen, t, ridx, cidx); break;
    default: {
      const double* g = p->gen;
      size_t L = (size_t)p->ldgen; /* [F D U V DIFL DIFR] stacked with stride ldgen */
      ora_const_cauchyeig(lm, ln, t, g, g + 2 * L, g + 3 * L, g + 4 * L, g + 5 * L, ridx, cidx);
    }
  }
}

static void* rank_body(void* arg) {
  rank_job* jb = (rank_job*)arg;
  const ora_problem* p = jb->p;
  int P = p->nprow, Q = p->npcol, nb = p->nb;
  int myrow = jb->myrow, mycol = jb->mycol;
  int mp = ora_numroc(p->n, nb, myrow, P);
  int nq = ora_numroc(p->k, nb, mycol, Q);
  if (mp == 0) return NULL;
  int* rows = (int*)malloc(sizeof(int) * (size_t)(mp + nb));
  int* cidx = (int*)malloc(sizeof(int) * (size_t)(nq + nb));
  int* ridx = (int*)malloc(sizeof(int) * (size_t)(p->k / Q + 2 * nb + 1));
  ora_constindex(p->n, nb, myrow, P, rows);
  int lc = ora_constindex(p->k, nb, mycol, Q, cidx);
  size_t maxkbp = (size_t)(p->k / Q + nb);
  double* aloc = (double*)malloc(sizeof(double) * (size_t)mp * (maxkbp + 1));
  double* cloc = (double*)malloc(sizeof(double) * (size_t)mp * (nq + 1));
  /* WORK(IPB) persists across steps, as the reference's workspace does. */
  size_t hcap = (size_t)(maxkbp + 1) * (nq + nb + 1) * 2;
  double* hbuf = (double*)calloc(hcap, sizeof(double));
  double* tmp = (double*)malloc(sizeof(double) * (size_t)mp * (maxkbp + 1));
  if (!rows || !cidx || !ridx || !aloc || !cloc || !hbuf || !tmp) {
    jb->status = -1;
    goto done;
  }
  /* local C <- global C (only read when beta != 0) */
  for (int j = 0; j < nq; ++j)
    for (int i = 0; i < mp; ++i)
      cloc[(size_t)j * mp + i] =
          (p->beta != 0.0) ? p->c[(size_t)(cidx[j] - 1) * p->ldc + rows[i] - 1] : 0.0;
  double tbeta = p->beta;
  for (int k1 = 0; k1 < Q; ++k1) {
    int nroll = (mycol + k1) % Q;
    int lr = ora_constindex(p->k, nb, nroll, Q, ridx);
    int kbp = lr;
    if (lc == 0 || lr == 0) continue;
    /* the A block column this rank holds after k1 ring shifts */
    for (int kk = 0; kk < kbp; ++kk)
      for (int i = 0; i < mp; ++i)
        aloc[(size_t)kk * mp + i] = p->a[(size_t)(ridx[kk] - 1) * p->lda + rows[i] - 1];
    int lowrank = ora_is_intersect(lc, cidx, lr, ridx);
    /* PSHANKEL_COMPUTE has no low-rank branch (SRC/pshankel_compute.f90) */
    if (lowrank && p->mode != ORA_MODE_DENSE && p->kind != ORA_HANKEL) {
      int rk = -1;
      int st = ora_lowrank_tile(p, kbp, nq, ridx, cidx, hbuf, hcap, &rk);
      if (st < 0) {
        jb->status = st;
        goto done;
      }
      if (st == ORA_TILE_STALE) jb->stale_steps++;
      if (rk == (kbp < nq ? kbp : nq)) {
        jb->fullrank_steps++;
        ring_dgemm(mp, nq, kbp, p->alpha, aloc, mp, hbuf, kbp, tbeta, cloc, mp);
      } else {
        jb->lowrank_steps++;
        ring_dgemm(mp, rk, kbp, 1.0, aloc, mp, hbuf, kbp, 0.0, tmp, mp);
        ring_dgemm(mp, nq, rk, p->alpha, tmp, mp, hbuf + (size_t)kbp * rk, rk, tbeta, cloc, mp);
      }
    } else {
      jb->dense_steps++;
      build_tile(p, kbp, nq, ridx, cidx, hbuf);
      ring_dgemm(mp, nq, kbp, p->alpha, aloc, mp, hbuf, kbp, tbeta, cloc, mp);
    }
    tbeta = 1.0;
  }
  for (int j = 0; j < nq; ++j)
    for (int i = 0; i < mp; ++i)
      p->c[(size_t)(cidx[j] - 1) * p->ldc + rows[i] - 1] = cloc[(size_t)j * mp + i];
done:
  free(rows);
  free(cidx);
  free(ridx);
  free(aloc);
  free(cloc);
  free(hbuf);
  free(tmp);
  return NULL;
}

int ora_ps_compute(const ora_problem* p, int nthreads, long long* stats) {
  int P = p->nprow, Q = p->npcol;
  if (P < 1 || Q < 1 || p->nb < 1 || p->n < 0 || p->k < 0) return -1;
  if (p->n == 0 || ((p->alpha == 0.0 || p->k == 0) && p->beta == 1.0)) return 0;
  int nr = P * Q;
  rank_job* jobs = (rank_job*)calloc((size_t)nr, sizeof(rank_job));
  for (int r = 0; r < nr; ++r) {
    jobs[r].p = p;
    jobs[r].myrow = r / Q;
    jobs[r].mycol = r % Q;
  }
  if (nthreads <= 1) {
    for (int r = 0; r < nr; ++r) rank_body(&jobs[r]);
  } else {
    for (int r0 = 0; r0 < nr; r0 += nthreads) {
      int cnt = (nr - r0) < nthreads ? (nr - r0) : nthreads;
      pthread_t th[256];
      if (cnt > 256) cnt = 256;
      for (int t = 0; t < cnt; ++t) pthread_create(&th[t], NULL, rank_body, &jobs[r0 + t]);
      for (int t = 0; t < cnt; ++t) pthread_join(th[t], NULL);
    }
  }
  int st = 0;
  long long s[4] = {0, 0, 0, 0};
  for (int r = 0; r < nr; ++r) {
    if (jobs[r].status) st = jobs[r].status;
    s[0] += jobs[r].dense_steps;
    s[1] += jobs[r].lowrank_steps;
    s[2] += jobs[r].fullrank_steps;
    s[3] += jobs[r].stale_steps;
  }
  if (stats)
    for (int i = 0; i < 4; ++i) stats[i] = s[i];
  free(jobs);
  return st;
}

/* Plain structured product on one process: C = alpha*A*H(r0:,c0:) + beta*C with the
 * tile built explicitly, i.e. the dense step of PS*_COMPUTE on a 1 x 1 grid. */
int ora_structured_gemm(int kind, int m, int n, int k, double alpha, const double* a, int lda,
                        const double* gen, int ldgen, int korder, int r0, int c0, double beta,
                        double* c, int ldc) {
  int* ridx = (int*)malloc(sizeof(int) * (size_t)(k + 1));
  int* cidx = (int*)malloc(sizeof(int) * (size_t)(n + 1));
  double* t = (double*)malloc(sizeof(double) * ((size_t)k * n + 1));
  if (!ridx || !cidx || !t) {
    free(ridx);
    free(cidx);
    free(t);
    return -1;
  }
  for (int i = 0; i < k; ++i) ridx[i] = r0 + i + 1;
  for (int j = 0; j < n; ++j) cidx[j] = c0 + j + 1;
  ora_problem p;
  memset(&p, 0, sizeof p);
  p.kind = kind;
  p.k = korder;
  p.gen = gen;
  p.ldgen = ldgen;
  build_tile(&p, k, n, ridx, cidx, t);
  ora_dgemm(m, n, k, alpha, a, lda, t, k > 0 ? k : 1, beta, c, ldc);
  free(ridx);
  free(cidx);
  free(t);
  return 0;
}
