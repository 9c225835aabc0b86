// Batched truncated rank-revealing QR (column-pivoted Householder QR stopped
// at a relative tolerance): the compression of off-diagonal blocks, with the
// semantics of the reference's Compress_RRQR (SRC/Auxil_2p5D.f90:1060-1189)
// over MDGEQP3 (SRC/mdgeqp3.f:153-370):
//   pivot = the column of largest partial norm (first on ties), norms
//   downdated from each new row of R and recomputed when the downdate loses
//   half the digits; rank = number of leading |R(i,i)| above tol |R(0,0)|;
//   output H = [Q (m x rank) | R P^T (rank x n)] when rank < min(m,n).
//
// One 256-thread workgroup per block (m, n <= 256), left-looking: the block
// is never updated in place.  After k reflectors the updated matrix is
// A - V F^T (V the reflectors, F = A^T V T accumulated one column per step,
// as LAPACK's blocked panel does), so each step reads the original block once
// (A^T v), regenerating it from the generators when the block is a structured
// one.  V lives in LDS; each thread owns one pivot position (its source
// column, F row, R column and partial norms) in registers and, for the
// row-wise work, one row.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>

#include "../../include/psmma.h"
#include "psmma_common.h"

namespace psmma {

constexpr int kRrqrKmax = 32;   // rank cap of the batched kernel
constexpr int kRrqrMaxDim = 256;
constexpr int kVld = kRrqrMaxDim + 1;  // padded LDS pitch: conflict-free row and column walks

struct SrcExplicit {
  const double* a;
  int lda;
  long long stride;
  const double* base;
  __device__ void bind(int b) { base = a + (size_t)b * stride; }
  __device__ double operator()(int i, int j) const { return base[i + (size_t)j * lda]; }
};

template <int KIND>
struct SrcGen {
  GenPtrs g;
  const int* r0;
  const int* c0;
  int rb, cb;
  __device__ void bind(int b) {
    rb = r0[b];
    cb = c0[b];
  }
  __device__ double operator()(int i, int j) const { return gen_entry<KIND>(g, rb + i, cb + j); }
};

// Sum over the workgroup in a fixed order; every thread gets the same value.
__device__ double block_sum(double x, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  return ((red[0] + red[1]) + red[2]) + red[3];
}

template <int KMAX, class Src>
__global__ __launch_bounds__(256, 1) void rrqr_kernel(Src src, int m, int n, double tol, int kmax,
                                                      int* __restrict__ rank_out,
                                                      int* __restrict__ jpvt_out, long long jstride,
                                                      double* __restrict__ q, int ldq,
                                                      long long qstride, double* __restrict__ rpt,
                                                      int ldr, long long rstride) {
  __shared__ double V[KMAX * kVld];
  __shared__ double E[KMAX * kVld];  // Q accumulation; exchange/broadcast scratch before that
  __shared__ double s_tau[KMAX];
  __shared__ double red[4];
  __shared__ double s_key[4];
  __shared__ int s_idx[4];

  const int b = blockIdx.x;
  const int t = threadIdx.x;
  src.bind(b);
  const int minmn = min(m, n);
  const double tol3z = 1.0536712127723509e-08;  // sqrt(2^-53)

  double F[KMAX], R[KMAX];
#pragma unroll
  for (int l = 0; l < KMAX; ++l) F[l] = R[l] = 0.0;
  int col = t;
  double vn1 = -1.0, vn2 = -1.0;
  if (t < n) {
    double s = 0.0;
    for (int i = 0; i < m; ++i) {
      double x = src(i, t);
      s += x * x;
    }
    vn1 = vn2 = sqrt(s);
  }

  int rank = 0;
  double r00 = 0.0;
  for (int k = 0; k < minmn; ++k) {
    // (1) pivot: largest partial norm among positions k..n-1, first on ties
    double key = (t >= k && t < n) ? vn1 : -1.0;
    int idx = t;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      double k2 = __shfl_xor(key, o);
      int i2 = __shfl_xor(idx, o);
      if (k2 > key || (k2 == key && i2 < idx)) {
        key = k2;
        idx = i2;
      }
    }
    if ((t & 63) == 0) {
      s_key[t >> 6] = key;
      s_idx[t >> 6] = idx;
    }
    __syncthreads();
    int p = s_idx[0];
    double pk = s_key[0];
    for (int w = 1; w < 4; ++w)
      if (s_key[w] > pk) {
        pk = s_key[w];
        p = s_idx[w];
      }
    // (2) exchange positions k and p (source column, F row, R column, norms)
    double* xa = E;                 // written by the thread at p
    double* xb = E + 2 * KMAX + 8;  // written by the thread at k
    if (p != k) {
      if (t == p || t == k) {
        double* dst = (t == p) ? xa : xb;
#pragma unroll
        for (int l = 0; l < KMAX; ++l) {
          dst[l] = F[l];
          dst[KMAX + l] = R[l];
        }
        dst[2 * KMAX] = vn1;
        dst[2 * KMAX + 1] = vn2;
        dst[2 * KMAX + 2] = (double)col;
      }
      __syncthreads();
      if (t == p || t == k) {
        const double* srcx = (t == p) ? xb : xa;
#pragma unroll
        for (int l = 0; l < KMAX; ++l) {
          F[l] = srcx[l];
          R[l] = srcx[KMAX + l];
        }
        vn1 = srcx[2 * KMAX];
        vn2 = srcx[2 * KMAX + 1];
        col = (int)srcx[2 * KMAX + 2];
      }
      __syncthreads();
    }
    // (3) the pivot column, updated: a(:, col_k) - V F_k^T
    double* fk = E + 4 * KMAX + 16;  // F row of position k, then its column index
    if (t == k) {
#pragma unroll
      for (int l = 0; l < KMAX; ++l) fk[l] = F[l];
      fk[KMAX] = (double)col;
    }
    __syncthreads();
    const int ck = (int)fk[KMAX];
    double ahat = 0.0;
    if (t < m) {
      ahat = src(t, ck);
#pragma unroll
      for (int l = 0; l < KMAX; ++l)
        if (l < k) ahat -= V[l * kVld + t] * fk[l];
    }
    // (4) Householder reflector on rows k..m-1 of the pivot column
    const double xn2 = block_sum((t > k && t < m) ? ahat * ahat : 0.0, red);
    double* s_alpha = E + 5 * KMAX + 24;
    if (t == k) *s_alpha = ahat;
    __syncthreads();
    const double alpha = *s_alpha;
    double tau = 0.0, beta = alpha, sc = 0.0;
    if (xn2 != 0.0 && k < m - 1) {
      beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
      tau = (beta - alpha) / beta;
      sc = 1.0 / (alpha - beta);
    }
    V[k * kVld + t] = (t == k) ? 1.0 : ((t > k && t < m) ? ahat * sc : 0.0);
    if (t == 0) s_tau[k] = tau;
    if (k == 0) r00 = fabs(beta);
    if (t == k) {
#pragma unroll
      for (int l = 0; l < KMAX; ++l)
        if (l == k) R[l] = beta;
    }
    __syncthreads();
    // rank decisions (uniform): |R(k,k)| <= tol |R(0,0)| ends the rank at k
    if (fabs(beta) <= tol * r00) {
      rank = k;
      break;
    }
    if (k + 1 == minmn) {
      rank = minmn;
      break;
    }
    if (k + 1 == kmax) {
      rank = -1;  // kmax or more columns above the tolerance
      break;
    }
    // (5) y = V(:, 0:k)^T v_k, then f_k = tau (A^T v_k - F y) and row k of R
    double* y = E + 6 * KMAX + 32;
    if (t < k) {
      double s = 0.0;
      for (int i = k; i < m; ++i) s += V[t * kVld + i] * V[k * kVld + i];
      y[t] = s;
    }
    __syncthreads();
    if (t > k && t < n) {
      double g = 0.0;
      for (int i = k; i < m; ++i) g += src(i, col) * V[k * kVld + i];
      double fy = 0.0, vf = 0.0;
#pragma unroll
      for (int l = 0; l < KMAX; ++l)
        if (l < k) {
          fy += F[l] * y[l];
          vf += V[l * kVld + k] * F[l];
        }
      const double f = tau * (g - fy);
      const double rk = (src(k, col) - vf) - f;
#pragma unroll
      for (int l = 0; l < KMAX; ++l)
        if (l == k) {
          F[l] = f;
          R[l] = rk;
        }
      // (6) partial norm downdate, or recomputation from the updated column
      if (vn1 != 0.0) {
        double tt = fabs(rk) / vn1;
        tt = fmax(1.0 - tt * tt, 0.0);
        const double r12 = vn1 / vn2;
        if (tt * r12 * r12 <= tol3z) {
          double s = 0.0;
          for (int i = k + 1; i < m; ++i) {
            double x = src(i, col);
#pragma unroll
            for (int l = 0; l < KMAX; ++l)
              if (l <= k) x -= V[l * kVld + i] * F[l];
            s += x * x;
          }
          vn1 = vn2 = sqrt(s);
        } else {
          vn1 *= sqrt(tt);
        }
      }
    }
    __syncthreads();
  }

  if (t == 0) rank_out[b] = rank;
  if (t < n) jpvt_out[(size_t)b * jstride + t] = col + 1;
  if (rank <= 0 || rank >= minmn) return;  // uniform: nothing more to write

  // R P^T: position t's column of R goes to column col
  if (t < n) {
    double* dst = rpt + (size_t)b * rstride + (size_t)col * ldr;
#pragma unroll
    for (int l = 0; l < KMAX; ++l)
      if (l < rank) dst[l] = R[l];
  }
  // Q = H(0) ... H(rank-1) [I; 0], accumulated backwards in LDS
  for (int c = 0; c < rank; ++c) E[c * kVld + t] = (t == c) ? 1.0 : 0.0;
  __shared__ double s_w[KMAX];
  __syncthreads();
  for (int l = rank - 1; l >= 0; --l) {
    if (t >= l && t < rank) {
      double s = 0.0;
      for (int i = l; i < m; ++i) s += V[l * kVld + i] * E[t * kVld + i];
      s_w[t] = s_tau[l] * s;
    }
    __syncthreads();
    if (t >= l && t < m) {
      const double vi = V[l * kVld + t];
      for (int c = l; c < rank; ++c) E[c * kVld + t] -= vi * s_w[c];
    }
    __syncthreads();
  }
  if (t < m) {
    double* qb = q + (size_t)b * qstride;
    for (int c = 0; c < rank; ++c) qb[t + (size_t)c * ldq] = E[c * kVld + t];
  }
}

template <class Src>
hipError_t launch_rrqr(const Src& src, int nblk, int m, int n, double tol, int kmax, int* rank,
                       int* jpvt, long long jstride, double* q, int ldq, long long qstride,
                       double* rpt, int ldr, long long rstride, hipStream_t s) {
  if (nblk <= 0) return hipSuccess;
  rrqr_kernel<kRrqrKmax, Src><<<nblk, 256, 0, s>>>(src, m, n, tol, kmax, rank, jpvt, jstride, q,
                                                   ldq, qstride, rpt, ldr, rstride);
  return hipGetLastError();
}

// Blocks (r0[b], c0[b]) of size m x n of a structured matrix (device block lists).
hipError_t launch_rrqr_structured(int kind, const GenPtrs& g, const int* r0, const int* c0, int nblk,
                                  int m, int n, double tol, int kmax, int* rank, int* jpvt,
                                  long long jstride, double* q, int ldq, long long qstride,
                                  double* rpt, int ldr, long long rstride, hipStream_t s) {
  switch (kind) {
    case KIND_CAUCHY:
      return launch_rrqr(SrcGen<KIND_CAUCHY>{g, r0, c0, 0, 0}, nblk, m, n, tol, kmax, rank, jpvt,
                         jstride, q, ldq, qstride, rpt, ldr, rstride, s);
    case KIND_TOEPLITZ:
      return launch_rrqr(SrcGen<KIND_TOEPLITZ>{g, r0, c0, 0, 0}, nblk, m, n, tol, kmax, rank,
                         jpvt, jstride, q, ldq, qstride, rpt, ldr, rstride, s);
    case KIND_HANKEL:
      return launch_rrqr(SrcGen<KIND_HANKEL>{g, r0, c0, 0, 0}, nblk, m, n, tol, kmax, rank, jpvt,
                         jstride, q, ldq, qstride, rpt, ldr, rstride, s);
    default:
      return launch_rrqr(SrcGen<KIND_CAUCHYEIG>{g, r0, c0, 0, 0}, nblk, m, n, tol, kmax, rank,
                         jpvt, jstride, q, ldq, qstride, rpt, ldr, rstride, s);
  }
}

}  // namespace psmma

using namespace psmma;

extern "C" int psmma_rrqr_batched_dev(int nblk, int m, int n, const double* a, int lda,
                                      long long stride_a, double tol, int kmax, int* rank,
                                      int* jpvt, double* q, int ldq, long long stride_q,
                                      double* rpt, int ldr, long long stride_r, void* stream) {
  if (nblk < 0) return -1;
  if (m < 0 || m > kRrqrMaxDim) return -2;
  if (n < 0 || n > kRrqrMaxDim) return -3;
  if (nblk > 0 && m > 0 && n > 0 && !a) return -4;
  if (lda < (m > 1 ? m : 1)) return -5;
  if (stride_a < (long long)lda * n) return -6;
  if (!(tol >= 0.0)) return -7;
  if (kmax < 1 || kmax > kRrqrKmax) return -8;
  if (nblk > 0 && (!rank || !jpvt)) return -9;
  int qcols = kmax < (m < n ? m : n) ? kmax : (m < n ? m : n);
  if (ldq < (m > 1 ? m : 1) || stride_q < (long long)ldq * qcols) return -12;
  if (ldr < (qcols > 1 ? qcols : 1) || stride_r < (long long)ldr * n) return -15;
  if (nblk > 0 && (!q || !rpt)) return -11;
  if (nblk == 0) return 0;
  hipError_t e = launch_rrqr(SrcExplicit{a, lda, stride_a, nullptr}, nblk, m, n, tol, kmax, rank,
                             jpvt, n, q, ldq, stride_q, rpt, ldr, stride_r, (hipStream_t)stream);
  if (e != hipSuccess) {
    fprintf(stderr, "psmma_rrqr_batched_dev: %s\n", hipGetErrorString(e));
    return 1000 + (int)e;
  }
  return 0;
}
