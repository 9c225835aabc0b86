// C ABI of libpsmma_hip.so: the Fortran drop-in entry points (host arrays) and
// the device-pointer API.  Declarations and contracts: include/psmma.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/psmma.h"
#include "psmma_common.h"

namespace psmma {
hipError_t launch_right_f64(int kind, int cfg, int M, int N, int K, double alpha, const double* A,
                            int lda, const GenPtrs& g, int r0, int c0, double beta, double* C,
                            int ldc, hipStream_t s);
hipError_t launch_dft_f64(int M, int N, int K, int Nd, double alpha, const double2* tw,
                          const double2* B, int ldb, int r0, int c0, double beta, double2* C, int ldc,
                          hipStream_t s);
int launch_dft_split_f32(int M, int N, int K, int Nd, int slices, double alpha, const double2* tw,
                         const double2* B, int ldb, int r0, int c0, double beta, double2* C,
                         int ldc, hipStream_t s);
int launch_right_split(int kind, int prec, int slices, int M, int N, int K, double alpha,
                       const double* A, int lda, const GenPtrs& g, int r0, int c0, double beta,
                       double* C, int ldc, hipStream_t s);
int launch_right_split_i8_noalloc(int kind, int slices, int M, int N, int K, double alpha,
                                  const double* A, int lda, const GenPtrs& g, int r0, int c0,
                                  double beta, double* C, int ldc, hipStream_t s, size_t* need);
size_t split_noalloc_capacity();
}  // namespace psmma

using namespace psmma;

// ---------------------------------------------------------------------------
// Error state (per thread).
static thread_local std::string g_err;
static thread_local int g_status = 0;
static int g_prec = PSMMA_PREC_FP64;
static int g_slices = 0;  // 0: chosen per call from K (psmma.h)

static void set_err(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}

// PXERBLA's convention: report the routine name and the argument number, then
// return without touching the outputs.
static void xerbla(const char* name, int info) {
  fprintf(stderr, "{    0,    0}:  On entry to %s parameter number %d had an illegal value\n", name,
          info);
  set_err("On entry to %s parameter number %d had an illegal value", name, info);
  g_status = info;
}

static int hip_fail(hipError_t e, const char* where) {
  set_err("%s: %s", where, hipGetErrorString(e));
  fprintf(stderr, "psmma: %s: %s\n", where, hipGetErrorString(e));
  g_status = -1000 - (int)e;
  return 1000 + (int)e;
}

// ---------------------------------------------------------------------------
// BLACS grid.  The Fortran entry points serve one GPU's share, so they need a
// 1 x 1 grid.  If the host program links a BLACS, its BLACS_GRIDINFO is used to
// check that; otherwise (the weak reference is null) the grid is 1 x 1.
extern "C" void blacs_gridinfo_(const int*, int*, int*, int*, int*) __attribute__((weak));

static bool grid_ok_ctxt(const int* ictxt, const char* name) {
  if (blacs_gridinfo_ == nullptr || ictxt == nullptr) return true;
  int nprow = 1, npcol = 1, myrow = 0, mycol = 0;
  blacs_gridinfo_(ictxt, &nprow, &npcol, &myrow, &mycol);
  if (nprow == 1 && npcol == 1) return true;
  set_err("%s: process grid %d x %d; this library computes one GPU's share and needs a 1 x 1 grid",
          name, nprow, npcol);
  fprintf(stderr, "psmma: %s\n", g_err.c_str());
  g_status = -2;
  return false;
}

static bool grid_ok(const int* desc, const char* name) {
  return desc == nullptr || grid_ok_ctxt(&desc[1], name);
}

// PZMDFTL1 / PZMDFTL take no descriptor: they read the BLACS context from
// COMMON /CONTEXT/ ICTXT (SRC/pzmdftL1.f90:178), the global symbol context_ of a
// Fortran caller that declares the block (weak: null when nobody does).
extern "C" int context_ __attribute__((weak));

// ---------------------------------------------------------------------------
// Device staging buffers for the host-pointer entry points: grow-only, one set
// per device, guarded by a mutex (the entry points are synchronous).
struct Staging {
  void* p[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t cap[4] = {0, 0, 0, 0};
};
static std::mutex g_mu;
static std::unordered_map<int, Staging> g_stage;

static void* stage_buf(int slot, size_t bytes) {
  int dev = 0;
  hipGetDevice(&dev);
  Staging& s = g_stage[dev];
  if (bytes == 0) bytes = 8;
  if (s.cap[slot] < bytes) {
    if (s.p[slot]) hipFree(s.p[slot]);
    s.p[slot] = nullptr;
    s.cap[slot] = 0;
    if (hipMalloc(&s.p[slot], bytes) != hipSuccess) return nullptr;
    s.cap[slot] = bytes;
  }
  return s.p[slot];
}

// FP64 twiddle tables, one per (device, Nd), built on the host in long double
// and rounded once: tw[e] = (cos(2 pi e/Nd), -sin(2 pi e/Nd)).
static std::mutex g_tw_mu;
static std::unordered_map<long long, double2*> g_tw;

static const double2* twiddles(int nd, hipStream_t s) {
  int dev = 0;
  hipGetDevice(&dev);
  long long key = ((long long)dev << 32) | (unsigned)nd;
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto it = g_tw.find(key);
  if (it != g_tw.end()) return it->second;
  std::vector<double2> h(nd);
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int e = 0; e < nd; ++e) {
    // Reduce by symmetry to the first octant so cos/sin see a small argument.
    long double ang = twopi * (long double)e / (long double)nd;
    h[e] = make_double2((double)cosl(ang), (double)(-sinl(ang)));
  }
  double2* d = nullptr;
  if (hipMalloc(&d, sizeof(double2) * nd) != hipSuccess) return nullptr;
  if (hipMemcpyAsync(d, h.data(), sizeof(double2) * nd, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    hipFree(d);
    return nullptr;
  }
  g_tw[key] = d;
  return d;
}

// ---------------------------------------------------------------------------
// Device API.

static int run_right(int kind, int m, int n, int k, double alpha, const double* a, int lda,
                     const GenPtrs& g, int r0, int c0, double beta, double* c, int ldc, int prec,
                     int slices, hipStream_t s) {
  if (m == 0 || n == 0) return 0;
  if (prec == PSMMA_PREC_FP64) {
    hipError_t e = launch_right_f64(kind, -1, m, n, k, alpha, a, lda, g, r0, c0, beta, c, ldc, s);
    if (e != hipSuccess) return hip_fail(e, "launch_right_f64");
    return 0;
  }
  return launch_right_split(kind, prec, slices, m, n, k, alpha, a, lda, g, r0, c0, beta, c, ldc, s);
}

static int check_common(const char* name, int m, int n, int k, const double* a, int lda,
                        const double* c, int ldc, int prec, int slices) {
  int st = 0;
  if (m < 0) st = -1;
  else if (n < 0) st = -2;
  else if (k < 0) st = -3;
  else if (lda < (m > 1 ? m : 1)) st = -6;
  else if (ldc < (m > 1 ? m : 1)) st = -14;
  else if (prec != PSMMA_PREC_FP64 && prec != PSMMA_PREC_SPLIT_I8 && prec != PSMMA_PREC_SPLIT_BF16)
    st = -15;
  else if (prec != PSMMA_PREC_FP64 && (slices < 0 || slices > 16)) st = -16;
  else if (m > 0 && n > 0 && (c == nullptr || (k > 0 && a == nullptr))) st = -5;
  if (st) {
    set_err("%s: invalid argument %d", name, -st);
    return st;
  }
  if (prec == PSMMA_PREC_SPLIT_I8 && k > PSMMA_SPLIT_I8_KMAX) {
    set_err("%s: K = %d exceeds the int8 split form's exact-accumulation limit %d", name, k,
            PSMMA_SPLIT_I8_KMAX);
    return -3;
  }
  return 0;
}

// Toeplitz block arguments shared by psmma_toeplitz_dev and its allocation-free variant.
static int check_toeplitz(const char* name, int m, int n, int k, const double* a, int lda,
                          const double* g, int lg, int korder, int r0, int c0, const double* c,
                          int ldc, int prec, int slices) {
  int st = check_common(name, m, n, k, a, lda, c, ldc, prec, slices);
  if (st) return st;
  if (r0 < 0 || c0 < 0) return -10;
  if (m > 0 && n > 0 && k > 0) {
    long long lo = (long long)korder - 1 + c0 - (r0 + k - 1);
    long long hi = (long long)korder - 1 + (c0 + n - 1) - r0;
    if (g == nullptr || lo < 0 || hi >= lg) {
      set_err("%s: generator index range [%lld,%lld] outside [0,%d)", name, lo, hi, lg);
      return -7;
    }
  }
  return 0;
}

extern "C" int psmma_toeplitz_dev(int m, int n, int k, double alpha, const double* a, int lda,
                                  const double* g, int lg, int korder, int r0, int c0, double beta,
                                  double* c, int ldc, int prec, int slices, void* stream) {
  int st = check_toeplitz("psmma_toeplitz_dev", m, n, k, a, lda, g, lg, korder, r0, c0, c, ldc,
                          prec, slices);
  if (st) return st;
  GenPtrs gp{g, nullptr, nullptr, nullptr, nullptr, korder, r0 + k, c0 + n};
  return run_right(KIND_TOEPLITZ, m, n, k, alpha, a, lda, gp, r0, c0, beta, c, ldc, prec, slices,
                   (hipStream_t)stream);
}

// No counterpart in the reference: psmma_toeplitz_dev for callers that cannot
// allocate.  The int8 form's scratch is the library's static per-GPU array.
extern "C" int psmma_toeplitz_noalloc_dev(int m, int n, int k, double alpha, const double* a,
                                          int lda, const double* g, int lg, int korder, int r0,
                                          int c0, double beta, double* c, int ldc, int prec,
                                          int slices, void* stream) {
  const char* name = "psmma_toeplitz_noalloc_dev";
  int st = check_toeplitz(name, m, n, k, a, lda, g, lg, korder, r0, c0, c, ldc, prec, slices);
  if (st) return st;
  if (prec == PSMMA_PREC_SPLIT_BF16) {
    set_err("%s: the bf16 form is not available without allocation (use FP64 or SPLIT_I8)", name);
    return -15;
  }
  if (m == 0 || n == 0) return 0;
  GenPtrs gp{g, nullptr, nullptr, nullptr, nullptr, korder, r0 + k, c0 + n};
  hipStream_t s = (hipStream_t)stream;
  if (prec == PSMMA_PREC_FP64 || k == 0) {
    hipError_t e = launch_right_f64(KIND_TOEPLITZ, -1, m, n, k, alpha, a, lda, gp, r0, c0, beta, c,
                                    ldc, s);
    return e == hipSuccess ? 0 : hip_fail(e, name);
  }
  size_t need = 0;
  st = launch_right_split_i8_noalloc(KIND_TOEPLITZ, slices, m, n, k, alpha, a, lda, gp, r0, c0,
                                     beta, c, ldc, s, &need);
  if (st == -1)
    set_err("%s: the product needs %zu bytes of scratch; the static scratch holds %zu", name, need,
            split_noalloc_capacity());
  else if (st == -16)
    set_err("%s: invalid slice count %d for K = %d", name, slices, k);
  else if (st > 0)
    set_err("%s: runtime error %d", name, st);
  return st;
}

extern "C" int psmma_hankel_dev(int m, int n, int k, double alpha, const double* a, int lda,
                                const double* g, int lg, int r0, int c0, double beta, double* c,
                                int ldc, int prec, int slices, void* stream) {
  int st = check_common("psmma_hankel_dev", m, n, k, a, lda, c, ldc, prec, slices);
  if (st) return st;
  if (r0 < 0 || c0 < 0) return -9;
  if (m > 0 && n > 0 && k > 0) {
    long long hi = (long long)(r0 + k - 1) + (c0 + n - 1);
    if (g == nullptr || hi >= lg) {
      set_err("psmma_hankel_dev: generator index %lld outside [0,%d)", hi, lg);
      return -7;
    }
  }
  GenPtrs gp{g, nullptr, nullptr, nullptr, nullptr, 0, r0 + k, c0 + n};
  return run_right(KIND_HANKEL, m, n, k, alpha, a, lda, gp, r0, c0, beta, c, ldc, prec, slices,
                   (hipStream_t)stream);
}

extern "C" int psmma_cauchy_dev(int m, int n, int k, double alpha, const double* a, int lda,
                                const double* d, const double* f, const double* u, const double* v,
                                int r0, int c0, double beta, double* c, int ldc, int prec,
                                int slices, void* stream) {
  int st = check_common("psmma_cauchy_dev", m, n, k, a, lda, c, ldc, prec, slices);
  if (st) return st;
  if (r0 < 0 || c0 < 0) return -11;
  if (m > 0 && n > 0 && k > 0 && (!d || !f || !u || !v)) return -7;
  GenPtrs gp{d, f, u, v, nullptr, 0, r0 + k, c0 + n};
  return run_right(KIND_CAUCHY, m, n, k, alpha, a, lda, gp, r0, c0, beta, c, ldc, prec, slices,
                   (hipStream_t)stream);
}

extern "C" int psmma_cauchyeig_dev(int m, int n, int k, double alpha, const double* a, int lda,
                                   const double* fpole, const double* u, const double* v,
                                   const double* difl, const double* difr, int nglob, int r0,
                                   int c0, double beta, double* c, int ldc, int prec, int slices,
                                   void* stream) {
  int st = check_common("psmma_cauchyeig_dev", m, n, k, a, lda, c, ldc, prec, slices);
  if (st) return st;
  if (r0 < 0 || c0 < 0 || r0 + k > nglob || c0 + n > nglob) {
    set_err("psmma_cauchyeig_dev: block [%d,%d)x[%d,%d) outside order %d", r0, r0 + k, c0, c0 + n,
            nglob);
    return -13;
  }
  if (m > 0 && n > 0 && k > 0 && (!fpole || !u || !v || !difl || !difr)) return -7;
  GenPtrs gp{fpole, u, v, difl, difr, nglob, r0 + k, c0 + n};
  return run_right(KIND_CAUCHYEIG, m, n, k, alpha, a, lda, gp, r0, c0, beta, c, ldc, prec, slices,
                   (hipStream_t)stream);
}

extern "C" int pzmdft_dev(int m, int n, int k, int nd, double alpha, const double* b, int ldb,
                          int r0, int c0, double beta, double* c, int ldc, int prec, int slices,
                          void* stream) {
  int st = 0;
  if (m < 0) st = -1;
  else if (n < 0) st = -2;
  else if (k < 0) st = -3;
  else if (nd < 1) st = -4;
  else if (ldb < (k > 1 ? k : 1)) st = -7;
  else if (r0 < 0) st = -8;
  else if (c0 < 0) st = -9;
  else if (ldc < (m > 1 ? m : 1)) st = -12;
  else if (prec != PSMMA_PREC_FP64 && prec != PSMMA_PREC_SPLIT_FP32) st = -13;
  else if (prec == PSMMA_PREC_SPLIT_FP32 && (slices < 0 || slices > 6)) st = -14;
  else if (m > 0 && n > 0 && (c == nullptr || (k > 0 && b == nullptr))) st = -6;
  if (st) {
    set_err(st == -13 ? "pzmdft_dev: the DFT product takes PSMMA_PREC_FP64 or PSMMA_PREC_SPLIT_FP32"
                      : "pzmdft_dev: invalid argument %d",
            -st);
    return st;
  }
  if (m == 0 || n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const double2* tw = twiddles(nd, s);
  if (!tw) return hip_fail(hipErrorOutOfMemory, "pzmdft_dev twiddle table");
  if (prec == PSMMA_PREC_SPLIT_FP32) {
    st = launch_dft_split_f32(m, n, k, nd, slices, alpha, tw, (const double2*)b, ldb, r0, c0, beta,
                              (double2*)c, ldc, s);
    if (st > 0) set_err("pzmdft_dev: runtime error %d", st);
    return st;
  }
  hipError_t e = launch_dft_f64(m, n, k, nd, alpha, tw, (const double2*)b, ldb, r0, c0, beta,
                                (double2*)c, ldc, s);
  if (e != hipSuccess) return hip_fail(e, "launch_dft_f64");
  return 0;
}

extern "C" const char* psmma_version(void) { return "psmma-hip 0.1 (gfx950)"; }
extern "C" const char* psmma_last_error(void) { return g_err.c_str(); }
extern "C" int psmma_last_status(void) { return g_status; }
extern "C" int psmma_set_precision(int prec, int slices) {
  if (prec != PSMMA_PREC_FP64 && prec != PSMMA_PREC_SPLIT_I8 && prec != PSMMA_PREC_SPLIT_BF16 &&
      prec != PSMMA_PREC_SPLIT_FP32)
    return -1;
  if (prec != PSMMA_PREC_FP64 && (slices < 0 || slices > 16)) return -2;
  g_prec = prec;
  g_slices = slices;
  return 0;
}

// ---------------------------------------------------------------------------
// Fortran drop-ins.
//
// On a 1 x 1 grid every process-column ring of the reference collapses to one
// step whose row and column index sets are both {1..K}; is_intersect reports
// them as intersecting, so the step is the dense branch: generate the full
// K x K tile and DGEMM (SRC/pscauchy_compute.f90:391-407 and siblings).  That
// step is exactly what the kernels compute.

// Numeric form of the real-kind drop-ins (the DFT's split-FP32 form does not apply to them).
static int real_prec() { return g_prec == PSMMA_PREC_SPLIT_FP32 ? PSMMA_PREC_FP64 : g_prec; }

namespace {

// Host -> device -> host driver for C(m x n) = alpha*A(m x k)*H + beta*C.
template <class Launch>
void host_right(const char* name, int m, int n, int k, double alpha, const double* a, int lda,
                double beta, double* c, int ldc, Launch&& launch) {
  std::lock_guard<std::mutex> lk(g_mu);
  hipStream_t s = nullptr;
  double* dA = (double*)stage_buf(0, sizeof(double) * (size_t)m * k);
  double* dC = (double*)stage_buf(1, sizeof(double) * (size_t)m * n);
  if (!dA || !dC) {
    hip_fail(hipErrorOutOfMemory, name);
    return;
  }
  hipError_t e = hipSuccess;
  if (k > 0)
    e = hipMemcpy2DAsync(dA, sizeof(double) * m, a, sizeof(double) * lda, sizeof(double) * m, k,
                         hipMemcpyHostToDevice, s);
  if (e == hipSuccess && beta != 0.0)
    e = hipMemcpy2DAsync(dC, sizeof(double) * m, c, sizeof(double) * ldc, sizeof(double) * m, n,
                         hipMemcpyHostToDevice, s);
  if (e != hipSuccess) {
    hip_fail(e, name);
    return;
  }
  int st = launch(dA, m, dC, m, s);
  if (st != 0) {
    if (g_status == 0) g_status = -3;
    fprintf(stderr, "psmma: %s: %s\n", name, g_err.c_str());
    return;
  }
  e = hipMemcpy2DAsync(c, sizeof(double) * ldc, dC, sizeof(double) * m, sizeof(double) * m, n,
                       hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) hip_fail(e, name);
}

double* upload(int slot, const double* h, size_t count) {
  double* d = (double*)stage_buf(slot, sizeof(double) * count);
  if (!d) return nullptr;
  if (count && hipMemcpy(d, h, sizeof(double) * count, hipMemcpyHostToDevice) != hipSuccess)
    return nullptr;
  return d;
}

// Shared body of PSCAUCHY_COMPUTE / PSTOEPLITZ_COMPUTE / PSHANKEL_COMPUTE on a
// 1 x 1 grid: argument checks with the reference's INFO numbers
// (SRC/pscauchy_compute.f90:234-257, :307-319), then one dense step.
void compute_right(int kind, const char* name, int N, int K, double alpha, const double* A, int LDA,
                   const int* desca, const double* B, int LDB, double beta, double* C, int LDC,
                   int imrow, int imcol) {
  g_status = 0;
  g_err.clear();
  int nb = desca ? desca[5] : 1;
  int info = 0;
  if (N < 0) info = 2;
  else if (K < 0) info = 3;
  else if (nb < 1) info = 5;
  else if (nb < 1) info = 6;
  else if (LDA < 1) info = 9;
  else if (LDC < 1) info = 14;
  else if (imrow != 0) info = 15;
  else if (imcol != 0) info = 16;
  if (info) return xerbla(name, info);
  if (N == 0) return;  // MP == 0
  if (LDA < N) return xerbla(name, 9);
  if (LDC < N) return xerbla(name, 14);
  if ((alpha == 0.0 || K == 0) && beta == 1.0) return;

  std::vector<double> sym;  // symmetric Toeplitz storage (LDB = K), expanded to 2K-1
  const double* gen = B;
  int lgen = LDB;
  if (kind == KIND_TOEPLITZ && K > 0 && LDB < 2 * K - 1) {
    if (LDB < K) return xerbla(name, 8);
    // B(1:K) holds the first column bottom-to-top with B(K) the diagonal; a
    // symmetric matrix's first row mirrors it: B(K+j) = B(K-j).
    sym.resize(2 * K - 1);
    for (int i = 0; i < K; ++i) sym[i] = B[i];
    for (int j = 1; j < K; ++j) sym[K - 1 + j] = B[K - 1 - j];
    gen = sym.data();
    lgen = 2 * K - 1;
  }
  if (kind == KIND_HANKEL && K > 0 && LDB < 2 * K - 1) return xerbla(name, 8);
  if (kind == KIND_CAUCHY && LDB < K) return xerbla(name, 8);

  int prec = real_prec(), slices = g_slices;
  host_right(name, N, K, K, alpha, A, LDA, beta, C, LDC,
             [&](const double* dA, int lda, double* dC, int ldc, hipStream_t s) -> int {
               if (kind == KIND_CAUCHY) {
                 // B(LDB,4) = [D F U V], all four columns of length >= K.
                 double* dg = (double*)stage_buf(2, sizeof(double) * (size_t)LDB * 4);
                 if (!dg || hipMemcpy(dg, B, sizeof(double) * (size_t)LDB * 4,
                                      hipMemcpyHostToDevice) != hipSuccess)
                   return hip_fail(hipErrorOutOfMemory, name);
                 return psmma_cauchy_dev(N, K, K, alpha, dA, lda, dg, dg + LDB, dg + 2 * LDB,
                                         dg + 3 * LDB, 0, 0, beta, dC, ldc, prec, slices, s);
               }
               double* dg = upload(2, gen, (size_t)lgen);
               if (!dg) return hip_fail(hipErrorOutOfMemory, name);
               if (kind == KIND_TOEPLITZ)
                 return psmma_toeplitz_dev(N, K, K, alpha, dA, lda, dg, lgen, K, 0, 0, beta, dC,
                                           ldc, prec, slices, s);
               return psmma_hankel_dev(N, K, K, alpha, dA, lda, dg, lgen, 0, 0, beta, dC, ldc,
                                       prec, slices, s);
             });
}

// PSMMA_CAUCHY / _TOEPLITZ / _HANKEL on a 1 x 1 grid.  REDIST=.false. passes A
// and Q straight to the compute routine and ignores IA/JA/IQ/JQ
// (SRC/psmma_toeplitz.f90:99-110).  REDIST=.true. redistributes A from (1,1)
// and Q back to (IQ,JQ) with PDGEMR2D (SRC/psmma_toeplitz.f90:139,157); on one
// process the block-cyclic and block layouts coincide, so the redistribution is
// the identity apart from that (IQ,JQ) offset.
void entry_right(int kind, const char* name, const int* n, const int* k, const double* alpha,
                 const double* a, const int* iq, const int* jq, const int* desca, const double* b,
                 const int* ldb, const double* beta, double* q, const int* descq,
                 const int* redist) {
  if (!grid_ok(descq, name)) return;
  int lda = desca[8], ldq = descq[8];
  double* qq = q;
  if (*redist) qq = q + (size_t)(*iq - 1) + (size_t)(*jq - 1) * ldq;
  compute_right(kind, name, *n, *k, *alpha, a, lda, desca, b, *ldb, *beta, qq, ldq, 0, 0);
}

}  // namespace

extern "C" void pscauchy_compute_(const int* n, const int* k, const double* alpha, const double* a,
                                  const int* lda, const int* desca, const double* b,
                                  const int* ldb, const double* beta, double* c, const int* ldc,
                                  const int* imrow, const int* imcol, double* /*work*/) {
  if (!grid_ok(desca, "PSCAUCHY_COMPUTE")) return;
  compute_right(KIND_CAUCHY, "PSCAUCHY_COMPUTE ", *n, *k, *alpha, a, *lda, desca, b, *ldb, *beta, c,
                *ldc, *imrow, *imcol);
}
extern "C" void pstoeplitz_compute_(const int* n, const int* k, const double* alpha,
                                    const double* a, const int* lda, const int* desca,
                                    const double* b, const int* ldb, const double* beta, double* c,
                                    const int* ldc, const int* imrow, const int* imcol,
                                    double* /*work*/) {
  if (!grid_ok(desca, "PSTOEPLITZ_COMPUTE")) return;
  compute_right(KIND_TOEPLITZ, "PSTOEPLITZ_COMPUTE ", *n, *k, *alpha, a, *lda, desca, b, *ldb,
                *beta, c, *ldc, *imrow, *imcol);
}
extern "C" void pshankel_compute_(const int* n, const int* k, const double* alpha, const double* a,
                                  const int* lda, const int* desca, const double* b,
                                  const int* ldb, const double* beta, double* c, const int* ldc,
                                  const int* imrow, const int* imcol, double* /*work*/) {
  if (!grid_ok(desca, "PSHANKEL_COMPUTE")) return;
  compute_right(KIND_HANKEL, "PSHANKEL_COMPUTE ", *n, *k, *alpha, a, *lda, desca, b, *ldb, *beta, c,
                *ldc, *imrow, *imcol);
}

extern "C" void psmma_cauchy_(const int* n, const int* k, const double* alpha, const double* a,
                              const int* /*ia*/, const int* /*ja*/, const int* desca,
                              const double* b, const int* ldb, const double* beta, double* q,
                              const int* iq, const int* jq, const int* descq, double* /*work*/,
                              const int* redist) {
  entry_right(KIND_CAUCHY, "PSCAUCHY_COMPUTE ", n, k, alpha, a, iq, jq, desca, b, ldb, beta, q,
              descq, redist);
}
extern "C" void psmma_toeplitz_(const int* n, const int* k, const double* alpha, const double* a,
                                const int* /*ia*/, const int* /*ja*/, const int* desca,
                                const double* b, const int* ldb, const double* beta, double* q,
                                const int* iq, const int* jq, const int* descq, double* /*work*/,
                                const int* redist) {
  entry_right(KIND_TOEPLITZ, "PSTOEPLITZ_COMPUTE ", n, k, alpha, a, iq, jq, desca, b, ldb, beta, q,
              descq, redist);
}
extern "C" void psmma_hankel_(const int* n, const int* k, const double* alpha, const double* a,
                              const int* /*ia*/, const int* /*ja*/, const int* desca,
                              const double* b, const int* ldb, const double* beta, double* q,
                              const int* iq, const int* jq, const int* descq, double* /*work*/,
                              const int* redist) {
  entry_right(KIND_HANKEL, "PSHANKEL_COMPUTE ", n, k, alpha, a, iq, jq, desca, b, ldb, beta, q,
              descq, redist);
}

// PSCAUCHYEIG_COMPUTE on a 1 x 1 grid (checks: SRC/pscauchyeig_compute.f90:232-254, :305-323).
static void compute_eig(int N, int K, int NB, int KB, double alpha, const double* A, int LDA,
                        const double* F, const double* U, const double* V, const double* DIFL,
                        const double* DIFR, double beta, double* C, int LDC, int imrow, int imcol) {
  const char* name = "PSCAUCHYEIG_COMPUTE ";
  g_status = 0;
  g_err.clear();
  int info = 0;
  if (N < 0) info = 1;
  else if (K < 0) info = 2;
  else if (NB < 1) info = 3;
  else if (KB < 1) info = 4;
  else if (LDA < 1) info = 7;
  else if (LDC < 1) info = 17;
  else if (imrow != 0) info = 18;
  else if (imcol != 0) info = 19;
  if (info) return xerbla(name, info);
  if (N == 0) return;
  if (LDA < N) return xerbla(name, 7);
  if (LDC < N) return xerbla(name, 17);
  if ((alpha == 0.0 || K == 0) && beta == 1.0) return;
  int prec = real_prec(), slices = g_slices;
  host_right(name, N, K, K, alpha, A, LDA, beta, C, LDC,
             [&](const double* dA, int lda, double* dC, int ldc, hipStream_t s) -> int {
               double* dg = (double*)stage_buf(2, sizeof(double) * (size_t)K * 5);
               if (!dg) return hip_fail(hipErrorOutOfMemory, name);
               const double* src[5] = {F, U, V, DIFL, DIFR};
               for (int i = 0; i < 5; ++i)
                 if (hipMemcpy(dg + (size_t)i * K, src[i], sizeof(double) * K,
                               hipMemcpyHostToDevice) != hipSuccess)
                   return hip_fail(hipErrorInvalidValue, name);
               return psmma_cauchyeig_dev(N, K, K, alpha, dA, lda, dg, dg + K, dg + 2 * K,
                                          dg + 3 * K, dg + 4 * K, K, 0, 0, beta, dC, ldc, prec,
                                          slices, s);
             });
}

extern "C" void pscauchyeig_compute_(const int* n, const int* k, const int* mb, const int* nb,
                                     const double* alpha, const double* a, const int* lda,
                                     const int* desca, const double* f, const double* /*d*/,
                                     const double* u, const double* v, const double* difl,
                                     const double* difr, const double* beta, double* c,
                                     const int* ldc, const int* imrow, const int* imcol,
                                     double* /*work*/) {
  if (!grid_ok(desca, "PSCAUCHYEIG_COMPUTE")) return;
  compute_eig(*n, *k, *mb, *nb, *alpha, a, *lda, f, u, v, difl, difr, *beta, c, *ldc, *imrow,
              *imcol);
}

// PSMMA_CAUCHYEIG: both branches compute into a work copy of Q and then copy it
// to Q(IQ:IQ+N-1, JQ:JQ+K-1) (PDLACPY at SRC/psmma_cauchyeig.f90:126, PDGEMR2D at
// :199), so the result always lands at (IQ,JQ).  D (the new eigenvalues) is only
// read by the reference's low-rank compression, never by the dense step.
extern "C" void psmma_cauchyeig_(const int* n, const int* k, const int* /*id*/, const double* alpha,
                                 const double* a, const int* /*ia*/, const int* /*ja*/,
                                 const int* desca, const double* f, const double* /*d*/,
                                 const double* u, const double* v, const double* difl,
                                 const double* difr, const double* beta, double* q, const int* iq,
                                 const int* jq, const int* descq, double* /*work*/,
                                 const int* /*redist*/) {
  if (!grid_ok(descq, "PSMMA_CAUCHYEIG")) return;
  int ldq = descq[8];
  double* qq = q + (size_t)(*iq - 1) + (size_t)(*jq - 1) * ldq;
  compute_eig(*n, *k, descq[4], descq[5], *alpha, a, desca[8], f, u, v, difl, difr, *beta, qq, ldq,
              0, 0);
}

// PZMDFTL1 / PZMDFTL on a 1 x 1 grid (checks: SRC/pzmdftL1.f90:189-276).  With
// one process row F11 is the whole N x N DFT matrix, which has full rank, so
// the RRQR (or SVD) factorisation reproduces F and the product is F*B.  The
// reference sizes the local columns of B and C from N (NQ = NUMROC(N,...)) and
// contracts over K, so it needs K == N; that is checked here.
static void entry_dft(const char* name, const int* n, const int* k, const int* nb, const int* kb,
                      const double* alpha, const double* b, const int* ldb, const double* beta,
                      double* c, const int* ldc, const int* imrow, const int* imcol) {
  g_status = 0;
  g_err.clear();
  if (!grid_ok_ctxt(&context_, name)) return;
  int N = *n, K = *k;
  int info = 0;
  if (N < 0) info = 1;
  else if (K < 0) info = 2;
  else if (*nb < 1) info = 3;
  else if (*kb < 1) info = 4;
  else if (*ldb < 1) info = 7;
  else if (*ldc < 1) info = 10;
  else if (*imrow != 0) info = 11;
  else if (*imcol != 0) info = 12;
  if (info) return xerbla(name, info);
  if (*ldb < K) return xerbla(name, 7);
  if (*ldc < N) return xerbla(name, 10);
  if (N == 0 || ((*alpha == 0.0 || K == 0) && *beta == 1.0)) return;
  if (K != N) {
    set_err("%s: the DFT product contracts over K and sizes columns by N; it needs K == N", name);
    fprintf(stderr, "psmma: %s\n", g_err.c_str());
    g_status = 2;
    return;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  hipStream_t s = nullptr;
  size_t cplx = 2 * sizeof(double);
  double* dB = (double*)stage_buf(0, cplx * (size_t)K * N);
  double* dC = (double*)stage_buf(1, cplx * (size_t)N * N);
  if (!dB || !dC) return (void)hip_fail(hipErrorOutOfMemory, name);
  hipError_t e = hipMemcpy2DAsync(dB, cplx * K, b, cplx * (*ldb), cplx * K, N,
                                  hipMemcpyHostToDevice, s);
  if (e == hipSuccess && *beta != 0.0)
    e = hipMemcpy2DAsync(dC, cplx * N, c, cplx * (*ldc), cplx * N, N, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return (void)hip_fail(e, name);
  // the staging lock stays held: dB and dC are in use until the copy back ends
  const int prec = g_prec == PSMMA_PREC_SPLIT_FP32 ? PSMMA_PREC_SPLIT_FP32 : PSMMA_PREC_FP64;
  const int slices = (prec == PSMMA_PREC_SPLIT_FP32 && g_slices <= 6) ? g_slices : 0;
  int st = pzmdft_dev(N, N, K, N, *alpha, dB, K, 0, 0, *beta, dC, N, prec, slices, s);
  if (st != 0) {
    if (g_status == 0) g_status = -3;
    fprintf(stderr, "psmma: %s: %s\n", name, g_err.c_str());
    return;
  }
  e = hipMemcpy2DAsync(c, cplx * (*ldc), dC, cplx * N, cplx * N, N, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) hip_fail(e, name);
}

extern "C" void pzmdftl1_(const int* n, const int* k, const int* nb, const int* kb,
                          const double* alpha, double* b, const int* ldb, const double* beta,
                          double* c, const int* ldc, const int* imrow, const int* imcol,
                          double* /*work*/, double* /*rwork*/) {
  entry_dft("PDMDFTL ", n, k, nb, kb, alpha, b, ldb, beta, c, ldc, imrow, imcol);
}
extern "C" void pzmdftl_(const int* n, const int* k, const int* nb, const int* kb,
                         const double* alpha, double* b, const int* ldb, const double* beta,
                         double* c, const int* ldc, const int* imrow, const int* imcol,
                         double* /*work*/, double* /*rwork*/) {
  entry_dft("PDMDFTL ", n, k, nb, kb, alpha, b, ldb, beta, c, ldc, imrow, imcol);
}
