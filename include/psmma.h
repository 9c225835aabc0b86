/*
 * psmma.h - C ABI of libpsmma_hip.so, the MI355X (gfx950) implementation of the
 * PSMMA structured x dense products.
 *
 * Two layers:
 *
 *  1. Fortran drop-ins.  Lower-case names with one trailing underscore and every
 *     argument by reference, exactly as gfortran/flang/ifort emit calls to the
 *     reference's implicit-interface externals.  A Fortran program that calls
 *     PSMMA_TOEPLITZ(...) links against this library unchanged.  They take host
 *     arrays, run on the current HIP device and return when the result is back
 *     in host memory.  This build serves one GPU's share: the BLACS grid must be
 *     1 x 1 (if a BLACS library is linked, BLACS_GRIDINFO is consulted and a
 *     larger grid is rejected).  INTEGER is 4 bytes, LOGICAL is 4 bytes
 *     (nonzero = .TRUE.).  Errors are reported the reference's way - a message
 *     on stderr and a plain return - and additionally through psmma_last_status().
 *
 *  2. Device API.  Device pointers, a hipStream_t passed as void*, asynchronous
 *     on that stream, status returned (0 = success, -i = argument i invalid,
 *     > 0 = HIP/runtime error; see psmma_last_error()).  It exposes the offsets
 *     (r0, c0) of the block of the global structured matrix that is multiplied,
 *     so a caller can drive one block of a larger product.
 *
 * Matrices are column-major with explicit leading dimensions.  Complex data is
 * interleaved complex128 (re, im), leading dimensions in complex elements.
 */
#ifndef PSMMA_H
#define PSMMA_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Numeric forms.
 *   PSMMA_PREC_FP64     plain FP64 on FP64 MFMA (v_mfma_f64_16x16x4f64).
 *   PSMMA_PREC_SPLIT_I8 error-free modular split: rows of A and columns of H are
 *                       scaled by powers of two and rounded to integers, the
 *                       integer product is computed exactly modulo `slices`
 *                       (2..16) pairwise-coprime moduli <= 256 on int8 MFMA
 *                       (exact int32 accumulation) and rebuilt by the Chinese
 *                       remainder theorem in double-double FP64.  13 moduli keep
 *                       43 bits per operand at K = 8192, 14 keep 47; the Cauchy
 *                       kinds give the structured operand two of A's bits (its
 *                       columns carry their norm in the entries near the pole).
 *   PSMMA_PREC_SPLIT_BF16 the same with bf16 slices on bf16 MFMA: 1..3 slices
 *                       run every kept slice pair into one FP32 accumulator
 *                       (3 slices: ~1e-7); 4..7 slices sum each class exactly
 *                       in FP32 over K-chunks and recombine in FP64.
 * The pole differences of the Cauchy kinds are always formed in FP64.
 *
 * slices = 0 asks for the default of the form: the fewest moduli that keep 45
 * bits per operand at the call's K (int8), or 6 slices (bf16).  The int8 form
 * accumulates products of residues |r| <= 128 in int32, exact for
 * K <= PSMMA_SPLIT_I8_KMAX; a longer contraction is rejected as argument K.
 *
 * Split-precision calls (and the compressed products of section 6) keep their
 * scratch memory in a grow-only workspace per (device, stream): calls on
 * different streams never share it, and calls on one stream are ordered by it.
 * The first call of a given size on a stream allocates (device-synchronising
 * when it grows), so such calls are not capturable until one of that size has
 * run on that stream.
 */
#define PSMMA_PREC_FP64 0
#define PSMMA_PREC_SPLIT_I8 1
#define PSMMA_PREC_SPLIT_BF16 2
/* Split-FP32 form of the DFT product (pzmdft_dev): F and the column-scaled B cut
 * into `slices` (1..6, 0 = 5) fixed-point slices of 8 bits, exact in FP32; all
 * slice products of a class p+q are summed exactly in FP32 over K-slabs of 8 on
 * v_mfma_f32_16x16x4f32 and recombined in FP64.  Error about 2^(-8 slices)
 * normwise; S(S+1)/2 slice pairs cost (S(S+1)/2)/2 times the plain FP64 form's
 * MFMA time, since FP32 MFMA runs at twice the FP64 rate on MI355X. */
#define PSMMA_PREC_SPLIT_FP32 3
#define PSMMA_SPLIT_I8_KMAX 131071

/* Kinds of structured operand (psmma_structured_dev, batched API). */
#define PSMMA_KIND_CAUCHY 0
#define PSMMA_KIND_TOEPLITZ 1
#define PSMMA_KIND_HANKEL 2
#define PSMMA_KIND_CAUCHYEIG 3

/* ===========================================================================
 * 1. Fortran drop-in entry points (host pointers).
 * ======================================================================== */

/* Q = alpha*A*H + beta*Q, H(i,j) = U_i V_j / (D_i - F_j), generators B(LDB,4)=[D F U V].
 * Replaces PSMMA_CAUCHY, SRC/psmma_cauchy.f90:2-3. */
void psmma_cauchy_(const int* n, const int* k, const double* alpha, const double* a,
                   const int* ia, const int* ja, const int* desca, const double* b,
                   const int* ldb, const double* beta, double* q, const int* iq, const int* jq,
                   const int* descq, double* work, const int* redist);

/* Q = alpha*A*H + beta*Q, H(i,j) = B(K+j-i), B of length 2K-1.
 * Replaces PSMMA_TOEPLITZ, SRC/psmma_toeplitz.f90:2-3. */
void psmma_toeplitz_(const int* n, const int* k, const double* alpha, const double* a,
                     const int* ia, const int* ja, const int* desca, const double* b,
                     const int* ldb, const double* beta, double* q, const int* iq, const int* jq,
                     const int* descq, double* work, const int* redist);

/* Q = alpha*A*H + beta*Q, H(i,j) = B(i+j-1).
 * Replaces PSMMA_HANKEL, SRC/psmma_hankel.f90:2-3. */
void psmma_hankel_(const int* n, const int* k, const double* alpha, const double* a,
                   const int* ia, const int* ja, const int* desca, const double* b,
                   const int* ldb, const double* beta, double* q, const int* iq, const int* jq,
                   const int* descq, double* work, const int* redist);

/* Q(IQ:,JQ:) = alpha*A*H + beta*Q(IQ:,JQ:), H(i,j) = U_i V_j / (F_i - D_j) with the
 * pole differences formed from DIFL/DIFR (DIFL_j = D_j - F_j, DIFR_j = D_j - F_{j+1}).
 * Replaces PSMMA_CAUCHYEIG, SRC/psmma_cauchyeig.f90:2-3. */
void psmma_cauchyeig_(const int* n, const int* k, const int* id, const double* alpha,
                      const double* a, const int* ia, const int* ja, const int* desca,
                      const double* f, const double* d, const double* u, const double* v,
                      const double* difl, const double* difr, const double* beta, double* q,
                      const int* iq, const int* jq, const int* descq, double* work,
                      const int* redist);

/* C = alpha*F*B + beta*C, F the unnormalised N x N DFT matrix, complex128 B and C.
 * Replaces PZMDFTL1, SRC/pzmdftL1.f90:1-2 (RRQR variant), and PZMDFTL,
 * SRC/pzmdftL.f90 (SVD variant); on one GPU both reduce to the same product. */
void pzmdftl1_(const int* n, const int* k, const int* nb, const int* kb, const double* alpha,
               double* b, const int* ldb, const double* beta, double* c, const int* ldc,
               const int* imrow, const int* imcol, double* work, double* rwork);
void pzmdftl_(const int* n, const int* k, const int* nb, const int* kb, const double* alpha,
              double* b, const int* ldb, const double* beta, double* c, const int* ldc,
              const int* imrow, const int* imcol, double* work, double* rwork);

/* Compute-level routines (reference internals, also external symbols there).
 * Replace PSCAUCHY_COMPUTE (SRC/pscauchy_compute.f90:2-3), PSTOEPLITZ_COMPUTE
 * (SRC/pstoeplitz_compute.f90:2-3), PSHANKEL_COMPUTE (SRC/pshankel_compute.f90:2-3)
 * and PSCAUCHYEIG_COMPUTE (SRC/pscauchyeig_compute.f90:2-3). */
void pscauchy_compute_(const int* n, const int* k, const double* alpha, const double* a,
                       const int* lda, const int* desca, const double* b, const int* ldb,
                       const double* beta, double* c, const int* ldc, const int* imrow,
                       const int* imcol, double* work);
void pstoeplitz_compute_(const int* n, const int* k, const double* alpha, const double* a,
                         const int* lda, const int* desca, const double* b, const int* ldb,
                         const double* beta, double* c, const int* ldc, const int* imrow,
                         const int* imcol, double* work);
void pshankel_compute_(const int* n, const int* k, const double* alpha, const double* a,
                       const int* lda, const int* desca, const double* b, const int* ldb,
                       const double* beta, double* c, const int* ldc, const int* imrow,
                       const int* imcol, double* work);
void pscauchyeig_compute_(const int* n, const int* k, const int* mb, const int* nb,
                          const double* alpha, const double* a, const int* lda,
                          const int* desca, const double* f, const double* d, const double* u,
                          const double* v, const double* difl, const double* difr,
                          const double* beta, double* c, const int* ldc, const int* imrow,
                          const int* imcol, double* work);

/* Numeric form used by the Fortran entry points (default PSMMA_PREC_FP64).  The
 * real kinds use PSMMA_PREC_SPLIT_I8 / _BF16 when selected (FP64 otherwise), the
 * DFT entries PSMMA_PREC_SPLIT_FP32 when selected (FP64 otherwise).  slices = 0
 * picks each form's default.  Returns 0, or -1 / -2 for an invalid prec / slice
 * count. */
int psmma_set_precision(int prec, int slices);
/* Status of the last Fortran entry call on this thread: 0 or the reference's
 * INFO code (argument number), or a negative runtime error. */
int psmma_last_status(void);
/* Text of the last error on this thread ("" if none). */
const char* psmma_last_error(void);

/* ===========================================================================
 * 2. Device API (device pointers, asynchronous on `stream`).
 *
 * C(m x n) = alpha * A(m x k) * H(r0:r0+k, c0:c0+n) + beta * C
 * ======================================================================== */

/* Toeplitz block: H(r,c) = g[korder - 1 + c - r] (0-based), lg = length of g.
 * Requires 0 <= korder-1+c-r < lg over the block. */
int psmma_toeplitz_dev(int m, int n, int k, double alpha, const double* a, int lda,
                       const double* g, int lg, int korder, int r0, int c0, double beta, double* c,
                       int ldc, int prec, int slices, void* stream);

/* Allocation-free variant of psmma_toeplitz_dev.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, serving
 * loops or harnesses that forbid hipMalloc).  Same arguments and results as
 * psmma_toeplitz_dev for prec = PSMMA_PREC_FP64 (which needs no scratch) and
 * PSMMA_PREC_SPLIT_I8, whose scratch is a static array of
 * PSMMA_NOALLOC_SCRATCH_BYTES per GPU (held from the library's first use on that
 * GPU until the process ends) instead of the grow-only workspace; the call
 * performs no allocation.  A product whose int8 layout exceeds the array
 * (about 2 t M K + 2 t M N bytes for t moduli, e.g. M = N = K = 2048 at t = 14)
 * is rejected with -1 and a message naming the bytes it needs.
 * PSMMA_PREC_SPLIT_BF16 is rejected (-15).  All calls on one GPU share the
 * array, which keeps the last call's contents: calls that overlap on one GPU
 * (other streams or threads) race; order them on one stream. */
#define PSMMA_NOALLOC_SCRATCH_BYTES (128u << 20)
int psmma_toeplitz_noalloc_dev(int m, int n, int k, double alpha, const double* a, int lda,
                               const double* g, int lg, int korder, int r0, int c0, double beta,
                               double* c, int ldc, int prec, int slices, void* stream);

/* Hankel block: H(r,c) = g[r + c] (0-based).  Requires r0+k-1 + c0+n-1 < lg. */
int psmma_hankel_dev(int m, int n, int k, double alpha, const double* a, int lda, const double* g,
                     int lg, int r0, int c0, double beta, double* c, int ldc, int prec, int slices,
                     void* stream);

/* Cauchy-like block: H(r,c) = u[r] v[c] / (d[r] - f[c]).  d,u need r0+k entries,
 * f,v need c0+n entries. */
int psmma_cauchy_dev(int m, int n, int k, double alpha, const double* a, int lda, const double* d,
                     const double* f, const double* u, const double* v, int r0, int c0,
                     double beta, double* c, int ldc, int prec, int slices, void* stream);

/* Eigenvector-Cauchy block: H(r,c) = u[r] v[c] / (fpole[r] - lambda_c) with the
 * difference formed as (fpole[r]-fpole[c]) - difl[c] for r <= c and
 * (fpole[r]-fpole[c+1]) - difr[c] otherwise.  nglob = global order (length of
 * fpole, difl, difr, u, v). */
int psmma_cauchyeig_dev(int m, int n, int k, double alpha, const double* a, int lda,
                        const double* fpole, const double* u, const double* v, const double* difl,
                        const double* difr, int nglob, int r0, int c0, double beta, double* c,
                        int ldc, int prec, int slices, void* stream);

/* DFT block (left-structured, complex128):
 * C(m x n) = alpha * F(r0:r0+m, c0:c0+k) * B(k x n) + beta * C,
 * F(r,c) = exp(-2*pi*i * (r*c mod nd) / nd).  prec: PSMMA_PREC_FP64 or
 * PSMMA_PREC_SPLIT_FP32 (slices 1..6, 0 = 5; ignored for FP64).  The Fortran
 * entries pzmdftl1_/pzmdftl_ use the split-FP32 form when psmma_set_precision
 * selected it, the FP64 form otherwise. */
int pzmdft_dev(int m, int n, int k, int nd, double alpha, const double* b, int ldb, int r0, int c0,
               double beta, double* c, int ldc, int prec, int slices, void* stream);

/* ===========================================================================
 * 3. Batched uneven products (device pointers).
 *
 * One entry per problem: C(m x n) = alpha * A(m x k) * H(r0:r0+k, c0:c0+n) + beta * C.
 * Generators per kind, as in the single-problem calls above:
 *   CAUCHY:    g0 = d, g1 = f, g2 = u, g3 = v
 *   TOEPLITZ:  g0 = g, lg = its length, korder = K
 *   HANKEL:    g0 = g, lg = its length
 *   CAUCHYEIG: g0 = fpole, g1 = u, g2 = v, g3 = difl, g4 = difr, lg = global order
 * ======================================================================== */
typedef struct psmma_problem {
  int kind;
  int m, n, k;
  double alpha, beta;
  const double* a;
  int lda;
  const double *g0, *g1, *g2, *g3, *g4;
  int lg;
  int korder;
  int r0, c0;
  double* c;
  int ldc;
} psmma_problem;

/* All problems in one go on `stream`.  prec: PSMMA_PREC_FP64 (one persistent
 * launch over the tiles of every problem, all kinds and tile classes, longest
 * tiles first, taken from an atomic work counter) or PSMMA_PREC_SPLIT_I8 (each
 * per-element phase of the int8 modular pipeline is one grouped launch over all
 * problems, then one launch runs every 128 x 128 tile's residue GEMMs for all
 * moduli with the recombination in registers; slices = number of moduli, 0 =
 * the fewest that keep 45 bits per operand for the longest contraction of the
 * batch).  The problem table is copied to the
 * device before the call returns; split calls on one device share a workspace
 * and must not overlap. */
int psmma_batched_dev(int nprob, const psmma_problem* probs, int prec, int slices, void* stream);

/* A reusable plan: validates the problems, uploads the table, picks the tile
 * classes (FP64: the size from which problems take the large tiles is timed on
 * the batch when every beta is 0), and captures the whole batch in a hipGraph
 * (FP64: one kernel node).  Launching the plan replays the graph on `stream`.
 * No allocation happens at launch time.  A plan's launches must not overlap
 * each other (they share its work counters and workspace). */
int psmma_batched_plan_create(int nprob, const psmma_problem* probs, int prec, int slices,
                              void** plan);
int psmma_batched_plan_launch(void* plan, void* stream);
/* FP64 plans: the min(m,n) from which problems use the large tiles; split plans: the
 * number of moduli.  nodes: kernel launches captured in the graph. */
int psmma_batched_plan_info(void* plan, int* tile_cfg, int* nodes);
int psmma_batched_plan_destroy(void* plan);

/* ===========================================================================
 * 3b. The reference's P x Q process grid, one rank's share at a time.
 *
 * psmma_constindex: ConstIndex (SRC/Auxil_2p5D.f90:2-66), the 1-based global
 * indices owned by process column `mycol` of `npcol` in a block-cyclic layout
 * of n with block nb (index may be NULL to get the count).  Returns the number
 * of indices, or -1 for invalid arguments.
 *
 * psmma_is_intersect: is_intersect (SRC/Auxil_2p5D.f90:752-792) with the
 * reference's sense: 1 when the two increasing index lists share NO element
 * (the ring step is then off-diagonal: PS*_COMPUTE's low-rank branch), else 0.
 *
 * psmma_ring_step_dev: one ring step of PS*_COMPUTE (the dense branch of
 * SRC/pscauchy_compute.f90:391-407, pstoeplitz_compute.f90:290-293,
 * pshankel_compute.f90:249-250, pscauchyeig_compute.f90:399-402):
 *   C(m x n) = alpha * A(m x k) * H(rindex, cindex) + beta * C,
 * with H the global structured matrix of problem p (kind, generators, lg,
 * korder as in psmma_problem; r0, c0 ignored) and rindex (k entries), cindex
 * (n entries) the 1-based global rows and columns of the step's tile, HOST
 * arrays (as ConstIndex fills them; uploaded on `stream`).  Every index is
 * checked against the generators first: -2 / -3 for a row / column out of
 * range, -1 for other invalid fields.  Asynchronous; returns 0 or > 0 on a
 * runtime error.
 *
 * psmma_redistribute_dev: PDGEMR2D between two block-cyclic layouts of an m x n
 * matrix on an nprow x npcol grid (RSRC = CSRC = 0), e.g. the NB-cyclic and
 * the ceil(N/P)-block layouts of REDIST (SRC/psmma_cauchy.f90:113-162), for the
 * whole grid at once: a holds the P*Q source local arrays back to back in rank
 * order r = prow * npcol + pcol, each column-major with leading dimension its
 * local row count, at the offsets psmma_local_offsets gives; b receives the
 * destination local arrays likewise.  A device permutation (each element read
 * once and written once); asynchronous.  Returns 0 or -(argument number).
 * psmma_local_offsets writes the P*Q + 1 offsets (elements) of such a layout
 * and returns the total element count (-1 for invalid arguments). */
int psmma_constindex(int n, int nb, int mycol, int npcol, int* index);
int psmma_is_intersect(int m, const int* seta, int n, const int* setb);
int psmma_ring_step_dev(const psmma_problem* p, const int* rindex, const int* cindex,
                        void* stream);
long long psmma_local_offsets(int m, int n, int nprow, int npcol, int mb, int nb,
                              long long* offsets);
int psmma_redistribute_dev(int m, int n, int nprow, int npcol, int src_mb, int src_nb,
                           const double* a, int dst_mb, int dst_nb, double* b, void* stream);

/* ===========================================================================
 * 3c. Autotuner.
 *
 * psmma_autotune_f64: times every tile configuration of the FP64 structured
 * kernel on an m x n x k product of `kind` (synthetic operands in memory it
 * allocates and frees; the device is synchronised) and caches the fastest for
 * the (device, kind, ceil log2 of m, n, k) bucket; every later FP64 launch of
 * that bucket uses it.  With PSMMA_AUTOTUNE=1 in the environment a bucket is
 * tuned at its first launch (which then allocates and synchronises).  Returns
 * 0 and the configuration / its time, or -(argument number), > 0 on errors.
 *
 * psmma_split_slices_dev: the fewest int8 moduli (prec PSMMA_PREC_SPLIT_I8) or
 * bf16 slices (PSMMA_PREC_SPLIT_BF16) whose relative Frobenius error, measured
 * on a block of up to 64 of problem p's own columns against the FP64 kernel,
 * is at most `target`; p's output is not touched (the sample goes to the
 * stream's workspace).  Synchronises `stream`.  Returns 0, 1 when no depth
 * reaches the target (the largest is returned), -(argument number).  Batched
 * split plans created with slices = 0 take the largest depth this chooses at
 * 4e-13 (the rest of the 1e-12 goal is a margin for the columns not sampled,
 * shorter problems with larger operand rounding and the fused form's fraction
 * rounding) over the longest-K problem of each kind, whose contraction keeps
 * the fewest bits per operand. */
int psmma_autotune_f64(int kind, int m, int n, int k, int* cfg, float* ms);
int psmma_split_slices_dev(const psmma_problem* p, int prec, double target, int* slices,
                           double* err, void* stream);

/* ===========================================================================
 * 4. Generators of the eigen-Cauchy operand (device pointers).
 *
 * Replaces the merge step that produces psmma_cauchyeig's inputs
 * (SRC/BasicMM.f90:538-590: a DLASD4 call per root, the Loewner recomputation
 * of Z, DNRM2 column scaling).  d (n, strictly increasing), z (n, no zero
 * entry), rho > 0 describe D + rho z z^T.  Outputs (device, length n):
 *   difl[j] = lambda_j - d_j > 0, difr[j] = lambda_j - d_{j+1} < 0 (last root:
 *   d_{n+1} = d_n + rho z^T z), found relative to the nearer pole;
 *   zhat (may be NULL): Loewner's z, for which the lambda_j are exact
 *   eigenvalues (sign of z); v (may be NULL): column normalisers
 *   v_j = 1 / || u ./ (d - lambda_j) || with u = zhat if given, else z.
 * Then psmma_cauchyeig_dev(..., fpole = d, u, v, difl, difr, ...) multiplies
 * by the orthonormal eigenvector matrix.  Returns 0 or -(argument number). */
int psmma_secular_dev(int n, const double* d, const double* z, double rho, double* difl,
                      double* difr, double* zhat, double* v, void* stream);

/* ===========================================================================
 * 5. Batched truncated RRQR (device pointers).
 *
 * Compress_RRQR (SRC/Auxil_2p5D.f90:1060-1189, over MDGEQP3, SRC/mdgeqp3.f:153-370)
 * for nblk column-major m x n blocks, m, n <= 256: block b is a + b*stride_a
 * (leading dimension lda).  Column pivoting on the largest partial norm;
 * rank[b] = number of leading |R(i,i)| > tol |R(0,0)|, with the reference's
 * conventions: rank = min(m,n) at full rank (outputs left unwritten), and here
 * also rank = -1 when kmax (1..32) or more columns stay above the tolerance
 * (outputs left unwritten).  For 0 < rank < min(m,n):
 *   q   + b*stride_q (ldq):  Q, m x rank, orthonormal columns;
 *   rpt + b*stride_r (ldr):  R P^T, rank x n, so that block ~ Q * (R P^T).
 * jpvt[b*n + j] = 1-based source column of pivot position j.  One workgroup
 * per block; asynchronous on `stream`.  Returns 0 or -(argument number). */
int psmma_rrqr_batched_dev(int nblk, int m, int n, const double* a, int lda, long long stride_a,
                           double tol, int kmax, int* rank, int* jpvt, double* q, int ldq,
                           long long stride_q, double* rpt, int ldr, long long stride_r,
                           void* stream);

/* ===========================================================================
 * 5b. Batched generator-only RRLU of eigen-Cauchy blocks (device pointers).
 *
 * Compress_cauchyEig2 (SRC/CauchyHssEig.f90:2640-2829) over rrluCauchyEigrow
 * (:7-212), the compression ConstCauchyEiglowrank applies to off-diagonal
 * eigen-Cauchy tiles (SRC/Auxil_2p5D.f90:642-748).  Block b has the 1-based
 * global rows pl[b*m + i] and columns pu[b*n + j] (m, n <= 256) of the matrix
 * u_i v_j / (fpole_i - lambda_j) of psmma_cauchyeig_dev (order korder, every
 * difference rebuilt from fpole, difl, difr through the global indices); the
 * block itself is never formed.  Rook pivoting, complete pivoting at the first
 * convergence test, a swap pass every 5th step (rho = the default-real 1.2E+0);
 * rank[b] = the first
 * step whose pivot falls below tol times the largest pivot so far (the
 * reference's TOL is the default-real 1.0E-17), min(m,n) at full rank, -2 when
 * an index lies outside [1, korder].  For 0 < rank < min(m,n) and rank <= kmax:
 *   q + b*stride_q (ldq):  Q (m x rank) = P^-1 [I; Z_i W_j / (F_i - F_j)];
 *   y + b*stride_y (ldy):  Y (rank x n), the chosen rows of the block,
 * so that block ~= Q * Y; otherwise q and y are left unwritten.  nswap (may be
 * NULL): swap-pass exchanges per block.  One workgroup per block; results are
 * bitwise those of the CPU restatement.  Asynchronous on `stream`.  Returns 0
 * or -(argument number). */
int psmma_rrlu_cauchyeig_batched_dev(int nblk, int m, int n, const int* pl, const int* pu,
                                     const double* fpole, const double* u, const double* v,
                                     const double* difl, const double* difr, int korder,
                                     double tol, int kmax, int* rank, int* nswap, double* q,
                                     int ldq, long long stride_q, double* y, int ldy,
                                     long long stride_y, void* stream);

/* Plain Cauchy blocks: compress_cauchy('R') over rrluCauchy (SRC/Cauchylowrank.f90:
 * 387-491, :7-134), the compression ConstCauchylowrank applies to off-diagonal Cauchy
 * tiles (SRC/Auxil_2p5D.f90:876-963).  Block b has the 1-based global rows pl[b*m + i]
 * and columns pu[b*n + j] (m, n <= 256) of u_i v_j / (d_i - f_j) with the global
 * generator vectors d, u (length nrow) and f, v (length ncol), gathered as
 * ConstCauchylowrank gathers them.  Rook pivoting, complete pivoting at the first
 * convergence test, a swap pass every 10th step (rho = 1.12); rank[b] as above (-2
 * for an index out of range).  For 0 < rank < m and rank <= kmax (the reference
 * writes H exactly when rank < m):
 *   x + b*stride_x (ldx):  X (m x rank) = P^-1 [I; Z_i W_j / (D_i - D_j)];
 *   y + b*stride_y (ldy):  Y (rank x n), the chosen rows of the block;
 * otherwise x and y are left unwritten.  Bitwise the CPU restatement.  Returns 0
 * or -(argument number). */
int psmma_rrlu_cauchy_batched_dev(int nblk, int m, int n, const int* pl, const int* pu,
                                  const double* d, const double* f, const double* u,
                                  const double* v, int nrow, int ncol, double tol, int kmax,
                                  int* rank, double* x, int ldx, long long stride_x, double* y,
                                  int ldy, long long stride_y, void* stream);

/* ===========================================================================
 * 6. Compressed eigenvector-merge product (device pointers).
 *
 * X(m x n) = alpha * A(m x n) * C + beta * X with C the n x n eigen-Cauchy
 * matrix of psmma_cauchyeig_dev (generators fpole, u, v, difl, difr, order n):
 * C is split into leaf x leaf blocks (leaf <= 256, n % leaf == 0); every
 * off-diagonal block is compressed by the batched truncated RRQR at relative
 * tolerance tol (psmma_rrqr_batched_dev semantics, rank cap 32) from the
 * generators, and multiplied in compressed form (two launches of the persistent
 * FP64 batch kernel: A's block rows times the Q factors, then per block column
 * the diagonal block generated in LDS with the low-rank columns as the explicit
 * tail of its contraction); incompressible blocks use the structured FP64
 * kernel.  The reference's off-diagonal ring
 * step (SRC/pstoeplitz_compute.f90:263-288: Compress_RRQR per tile, then two
 * DGEMMs) on a grid of leaf tiles.  Error: each block within ~tol of its largest
 * R entry.  stats (host, may be NULL, 5 entries): blocks, low-rank blocks,
 * incompressible blocks, zero blocks, sum of ranks.  Synchronises `stream` once
 * (the ranks decide the GEMM shapes).  Returns 0 or -(argument number). */
int psmma_cauchyeig_blr_dev(int m, int n, double alpha, const double* a, int lda,
                            const double* fpole, const double* u, const double* v,
                            const double* difl, const double* difr, int leaf, double tol,
                            double beta, double* c, int ldc, long long* stats, void* stream);

/* The same product with a choice of compressor for the off-diagonal blocks:
 *   PSMMA_BLR_RRQR: the batched truncated RRQR above (Compress_RRQR);
 *   PSMMA_BLR_RRLU: the generator-only RRLU of psmma_rrlu_cauchyeig_batched_dev
 *     (Compress_cauchyEig2, the reference's compressor of off-diagonal eigen-Cauchy
 *     ring tiles, SRC/Auxil_2p5D.f90:642-748), block ~= Q Y with Q and Y the row
 *     ID; tol is relative to the largest pivot; ranks above 32 keep the block in
 *     the structured FP64 kernel.
 * No counterpart in the reference (its compressed merge fixes the compressor).
 * Returns 0 or -(argument number). */
#define PSMMA_BLR_RRQR 0
#define PSMMA_BLR_RRLU 1
int psmma_cauchyeig_blr_ex_dev(int m, int n, double alpha, const double* a, int lda,
                               const double* fpole, const double* u, const double* v,
                               const double* difl, const double* difr, int leaf, double tol,
                               int compressor, double beta, double* c, int ldc,
                               long long* stats, void* stream);

/* The compressed form of the plain Cauchy product X = alpha * A * C + beta * X,
 * C(i,j) = u_i v_j / (d_i - f_j) (n x n, generators d, f, u, v of length n, the
 * columns of psmma_cauchy's B): off-diagonal leaf blocks compressed by the batched
 * RRQR or by the generator-only RRLU (compress_cauchy, the reference's compressor of
 * off-diagonal Cauchy ring tiles, SRC/pscauchy_compute.f90:355-424), the rest as in
 * psmma_cauchyeig_blr_ex_dev.  No counterpart in the reference as a single call.
 * Returns 0 or -(argument number). */
int psmma_cauchy_blr_dev(int m, int n, double alpha, const double* a, int lda, const double* d,
                         const double* f, const double* u, const double* v, int leaf, double tol,
                         int compressor, double beta, double* c, int ldc, long long* stats,
                         void* stream);

/* Compressed Toeplitz and Hankel products X = alpha * A * H + beta * X, H the n x n
 * block of psmma_toeplitz_dev (H(r,c) = g[korder - 1 + c - r]) or psmma_hankel_dev
 * (H(r,c) = g[r + c]), 0 <= r, c < n: every off-diagonal leaf block is compressed by
 * the batched truncated RRQR straight from the generator (ConstToeplitzlowrank ->
 * Compress_RRQR, SRC/Auxil_2p5D.f90:968-1189; tol as there, the reference uses
 * 1e-13) and multiplied in compressed form, blocks of rank above 32 or full rank
 * take the structured FP64 kernel (the reference regenerates the tile,
 * SRC/Auxil_2p5D.f90:1054-1056); the diagonal blocks run on the window tiles.  The
 * reference's off-diagonal Toeplitz ring step (SRC/pstoeplitz_compute.f90:263-295)
 * on a grid of leaf tiles; the reference's Hankel ring never compresses
 * (SRC/pshankel_compute.f90:238-265), this is its analogue.  stats and
 * synchronisation as psmma_cauchyeig_blr_dev.  Returns 0 or -(argument number). */
int psmma_toeplitz_blr_dev(int m, int n, double alpha, const double* a, int lda, const double* g,
                           int lg, int korder, int leaf, double tol, double beta, double* c,
                           int ldc, long long* stats, void* stream);
int psmma_hankel_blr_dev(int m, int n, double alpha, const double* a, int lda, const double* g,
                         int lg, int leaf, double tol, double beta, double* c, int ldc,
                         long long* stats, void* stream);

/* ===========================================================================
 * 6b. Nested HSS representation and HSS x dense product (device pointers).
 *
 * psmma_hss_build_dev: an HSS representation of the n x n eigen-Cauchy matrix
 * (kind PSMMA_KIND_CAUCHYEIG, g0..g4 = fpole, u, v, difl, difr as in
 * psmma_cauchyeig_dev) or plain Cauchy matrix (PSMMA_KIND_CAUCHY, g0..g3 = d, f,
 * u, v), the counterpart of Cauchy2hssEig_omp (SRC/CauchyHssEig.f90:1955-2259):
 * a complete binary tree whose node ranges halve (left half = floor) until they
 * hold at most `leaf` indices (2 <= leaf <= 256, any n; all leaves at one depth),
 * nested row and column bases U, V and skeleton couplings B per node (the
 * HSSMM generators, SRC/aux_hss.f90:27-38), each node's bases from an
 * interpolative decomposition of `samples` (1..128) random samples of its
 * block row / column, truncated at |R_kk| <= tol |R_00|; samples = 0 chooses
 * the count: 64 (96 from n = 16384), then 96, then 128 while some node's rank comes within 10 of
 * it (the oversampling margin).  `seed` fixes the samples.  The generators stay
 * on the device (the generator arrays must outlive the handle); the device is
 * synchronised.  stats (may be NULL, 6 entries): levels, largest rank, sum of
 * ranks, IDs whose rank came within 10 of the sample count below full width
 * (accuracy not guaranteed there), generator storage (doubles), samples.
 * Returns 0 or -(argument number), > 0 on runtime errors.
 *
 * psmma_hss_product_dev: X(m x n) = alpha * A(m x n) * C + beta * X with C in
 * HSS form (up pass, sibling couplings, down pass, leaf diagonal blocks
 * generated from the generators; one persistent launch per stage).  Synchronises
 * `stream` once (after the stages).
 *
 * psmma_hss_layout: the HSSMM-like layout, nodes in post order (the reference's
 * tree order): parent (-1 for the root), row / column range [lo, lo + size),
 * dims of U (w x r), V (w' x r'), B (r x r'_sibling), their offsets (3 per node,
 * doubles) into the device array *h; any output may be NULL.  For siblings a, b:
 * C(I_a, I_b) ~ Ut_a B_a Vt_b^T, Ut = U for leaves and diag(Ut_left, Ut_right) U
 * for parents (likewise Vt).
 *
 * These have no single counterpart among the reference's externals (its HSS
 * construction is a module routine, not on a PSMMA entry path). */
int psmma_hss_build_dev(int kind, int n, const double* g0, const double* g1, const double* g2,
                        const double* g3, const double* g4, int leaf, double tol, int samples,
                        unsigned long long seed, void** hss, long long* stats, void* stream);
int psmma_hss_product_dev(void* hss, int m, double alpha, const double* a, int lda, double beta,
                          double* c, int ldc, void* stream);
int psmma_hss_layout(void* hss, int* nnodes, int* parent, int* lo, int* size, int* udim,
                     int* vdim, int* bdim, long long* offsets, const double** h);
int psmma_hss_destroy(void* hss);

/* Library version string. */
const char* psmma_version(void);

#ifdef __cplusplus
}
#endif
#endif /* PSMMA_H */
