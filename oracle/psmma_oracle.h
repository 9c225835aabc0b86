/* psmma_oracle.h - CPU restatement of the reference PSMMA algorithm.
 * TEST INFRASTRUCTURE ONLY: used by tests/ and by bench.py's CPU baseline. */
#ifndef PSMMA_ORACLE_H
#define PSMMA_ORACLE_H

enum { ORA_CAUCHY = 0, ORA_TOEPLITZ = 1, ORA_HANKEL = 2, ORA_CAUCHYEIG = 3 };
enum { ORA_MODE_DENSE = 0, ORA_MODE_FAITHFUL = 1 };
enum { ORA_TILE_OK = 0, ORA_TILE_STALE = 1 };

/* One PS*_COMPUTE call on an emulated nprow x npcol grid with global arrays. */
typedef struct {
  int kind;           /* ORA_* */
  int nprow, npcol;   /* emulated BLACS grid */
  int nb;             /* block size (MB = NB = KB, as the reference assumes) */
  int n, k;           /* A is n x k, H is k x k, C is n x k */
  double alpha, beta;
  const double* a;    /* global A, leading dimension lda */
  int lda;
  const double* gen;  /* Cauchy: B(ldgen,4)=[D F U V]; Toeplitz/Hankel: G;
                         CauchyEig: [F D U V DIFL DIFR] stacked with stride ldgen */
  int ldgen;
  double* c;          /* global C, leading dimension ldc */
  int ldc;
  int mode;           /* ORA_MODE_DENSE or ORA_MODE_FAITHFUL */
} ora_problem;

int ora_numroc(int n, int nb, int iproc, int nprocs);
int ora_constindex(int n, int nb, int mycol, int npcol, int* idx);
int ora_is_intersect(int m, const int* seta, int n, const int* setb);
void ora_const_cauchy(int lm, int ln, const double* b, int ldb, double* t, const int* ridx,
                      const int* cidx);
void ora_const_toeplitz(int lm, int ln, const double* g, int k, double* t, const int* ridx,
                        const int* cidx);
void ora_const_hankel(int lm, int ln, const double* g, double* t, const int* ridx,
                      const int* cidx);
void ora_const_cauchyeig(int lm, int ln, double* t, const double* f, const double* u,
                         const double* v, const double* difl, const double* difr,
                         const int* ridx, const int* cidx);
/* Host BLAS Fortran dgemm_ for the ring restatement (NULL: the C DGEMM). */
void ora_set_blas_dgemm(void* fn);
void ora_dgemm(int m, int n, int k, double alpha, const double* a, int lda, const double* b,
               int ldb, double beta, double* c, int ldc);
int ora_ps_compute(const ora_problem* p, int nthreads, long long* stats);
int ora_structured_gemm(int kind, int m, int n, int k, double alpha, const double* a, int lda,
                        const double* gen, int ldgen, int korder, int r0, int c0, double beta,
                        double* c, int ldc);

/* Low-rank branch of one off-diagonal ring step (faithful mode): fills
 * hbuf with [X(kbp x rk) | Y(rk x nq)] or, at full rank, with the tile the
 * reference would multiply.  Returns ORA_TILE_OK / ORA_TILE_STALE or < 0. */
int ora_lowrank_tile(const ora_problem* p, int kbp, int nq, const int* ridx, const int* cidx,
                     double* hbuf, unsigned long hcap, int* rk);

/* Truncated pivoted QR (SRC/mdgeqp3.f): see psmma_oracle_lowrank.c. */
int ora_mdgeqp3(int m, int n, double* a, int lda, int* jpvt, double* tau, double tol, int nb,
                int* rk);
/* Compress_RRQR (SRC/Auxil_2p5D.f90:1060-1189): returns the rank; h = [Q | R P^T]. */
int ora_rrlu_cauchy(double* D, double* F, double* A, double* V, double tol, double* U,
                    double* W, int* PL, int* PU, int M, int N);
/* compress_cauchy('R') (SRC/Cauchylowrank.f90:387-491): returns Rk; h = [X | Y] when Rk < M. */
int ora_compress_cauchy(const double* D, const double* F, const double* U, const double* V,
                        double tol, int M, int N, double* H);
int ora_compress_rrqr(int m, int n, double* a, int lda, double* h, double tol);
/* rrluCauchyEigrow (SRC/CauchyHssEig.f90:7-212): returns Rk. */
int ora_rrlu_cauchyeig_row(const double* ff, const double* difl, const double* difr, double* F,
                           double* A, double* V, double tol, double* U, double* W, int* PL,
                           int* PU, int* PQ, int M, int N, int* nswap);
/* Compress_cauchyEig2 (SRC/CauchyHssEig.f90:2640-2829): returns Rk; h = [Q | Y]
 * when Rk < min(M, N), the explicit block otherwise. */
int ora_compress_cauchyeig2(const double* ff, const double* difl, const double* difr,
                            const double* u, const double* v, const int* pl, const int* pu,
                            double tol, int M, int N, double* H, int* nswap);
#endif
