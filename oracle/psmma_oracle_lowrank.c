/* Low-rank branches of the reference ring step - placeholder until restated. */
#include "psmma_oracle.h"
int ora_lowrank_tile(const ora_problem* p, int kbp, int nq, const int* ridx, const int* cidx,
                     double* hbuf, unsigned long hcap, int* rk) {
  (void)p; (void)kbp; (void)nq; (void)ridx; (void)cidx; (void)hbuf; (void)hcap; (void)rk;
  return -2;
}
int ora_mdgeqp3(int m, int n, double* a, int lda, int* jpvt, double* tau, double tol, int nb,
                int* rk) {
  (void)m; (void)n; (void)a; (void)lda; (void)jpvt; (void)tau; (void)tol; (void)nb; (void)rk;
  return -2;
}
