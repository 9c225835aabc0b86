// Batched truncated rank-revealing QR (column-pivoted Householder QR stopped
// at a relative tolerance): the compression of off-diagonal blocks, with the
// semantics of the reference's Compress_RRQR (SRC/Auxil_2p5D.f90:1060-1189)
// over MDGEQP3 (SRC/mdgeqp3.f:153-370):
//   pivot = the column of largest partial norm (first on ties), norms
//   downdated from each new row of R and recomputed when the downdate loses
//   half the digits; rank = number of leading |R(i,i)| above tol |R(0,0)|;
//   output H = [Q (m x rank) | R P^T (rank x n)] when rank < min(m,n).
//
// One 512-thread workgroup per block (m, n <= 256), left-looking: the block
// is never updated in place.  After k reflectors the updated matrix is
// A - V F^T (V the reflectors, F = A^T V T accumulated one column per step,
// as LAPACK's blocked panel does), so each step reads the original block once
// (A^T v), regenerating it from the generators when the block is a structured
// one.  V lives in LDS; each pivot position (its source column, F row, R
// column and partial norms) lives in the registers of two threads, which split
// the row loops; the row-wise work takes one row per thread.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <type_traits>

#include "../../include/psmma.h"
#include "psmma_common.h"

namespace psmma {

constexpr int kRrqrKmax = 32;   // rank cap of the batched kernel
constexpr int kRrqrMaxDim = 256;
constexpr int kVld = kRrqrMaxDim + 1;  // padded LDS pitch: conflict-free row and column walks
// Rows per group in the passes over a column: a group's inputs are all fetched
// before any is used, so one LDS (or memory) latency is paid per group.  Full
// groups carry no row test (a test per row turns each entry into a branch of
// its own); the last, partial group clamps its row indices and multiplies the
// rows past the end by zero.
constexpr int kGrp = 8;

// Block sources.  bind() stages what the block needs in LDS; col(c) returns
// the per-column context a thread keeps for its current source column;
// fetch(i, ctx) loads what entry (i, c) needs, values<G>(raw, rows, ctx, out)
// finishes a group of G = kG entries; at(i, ctx) is one entry.  A source with
// kCaches evaluates the block in the first pass (generate, in groups of kGrp)
// and keeps a copy.  Sources that only load use groups of 16 rows (more loads
// in flight per wave).
struct SrcExplicit {
  const double* a;
  int lda;
  long long stride;
  const double* base;
  using Ctx = const double*;
  __device__ void bind(int b, double*, int, int) { base = a + (size_t)b * stride; }
  __device__ Ctx col(int c) const { return base + (size_t)c * lda; }
  static constexpr bool kCaches = false;
  static constexpr int kG = 2 * kGrp;
  using Raw = double;
  __device__ Raw fetch(int i, Ctx p) const { return p[i]; }
  template <int G>
  __device__ void values(const Raw* r, const int*, Ctx, double* out) const {
#pragma unroll
    for (int j = 0; j < G; ++j) out[j] = r[j];
  }
  __device__ double at(int i, Ctx p) const { return p[i]; }
};

template <int KIND>
struct SrcGen {
  GenPtrs g;
  const int* r0;
  const int* c0;
  int rb, cb;
  using Ctx = int;
  __device__ void bind(int b, double*, int, int) {
    rb = r0[b];
    cb = c0[b];
  }
  __device__ Ctx col(int c) const { return cb + c; }
  static constexpr bool kCaches = false;
  static constexpr int kG = kGrp;
  using Raw = double;
  __device__ Raw fetch(int i, Ctx c) const { return gen_entry<KIND>(g, rb + i, c); }
  template <int G>
  __device__ void values(const Raw* r, const int*, Ctx, double* out) const {
#pragma unroll
    for (int j = 0; j < G; ++j) out[j] = r[j];
  }
  __device__ double at(int i, Ctx c) const { return gen_entry<KIND>(g, rb + i, c); }
};

// n / d by the reciprocal refinement that FP64 division on this target runs
// (v_rcp_f64, two Newton steps, one residual correction) without its range
// scaling (v_div_scale / v_div_fmas / v_div_fixup).  Equal to n / d whenever
// the scaling would not apply: finite, normal operands and quotient, away from
// the ends of the exponent range.  The scaling steps pass a flag through VCC,
// which serialises every division in a thread; this form leaves independent
// quotients free to overlap.
__device__ __forceinline__ double div_refined(double n, double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  const double q = n * r;
  return fma(fma(-d, q, n), r, q);
}

// Eigen-Cauchy blocks: the row poles and weights staged in LDS, the column's
// pole, next pole, DIFL, DIFR and normaliser in registers; the arithmetic of
// gen_entry<KIND_CAUCHYEIG> with the quotient taken by div_refined (the
// entries' denominators are pole differences and their numerators products of
// weights: normal numbers far from the exponent range's ends, where division
// would scale).
struct SrcEigStaged {
  GenPtrs g;
  const int* r0;
  const int* c0;
  int rb, cb;
  const double* s;
  struct Ctx {
    double fc, fc1, dl, dr, vc;
    int cg;
  };
  __device__ void bind(int b, double* st, int m, int) {
    rb = r0[b];
    cb = c0[b];
    const int t = threadIdx.x;
    if (t < m) {
      st[t] = g.p0[rb + t];
      st[kRrqrMaxDim + t] = g.p1[rb + t];
    }
    s = st;
  }
  __device__ Ctx col(int c) const {
    const int cg = cb + c;
    return Ctx{g.p0[cg], g.p0[min(cg + 1, g.korder - 1)], g.p3[cg], g.p4[cg], g.p2[cg], cg};
  }
  static constexpr bool kCaches = false;
  static constexpr int kG = kGrp;
  struct Raw {
    double f, u;
  };
  __device__ Raw fetch(int i, const Ctx&) const { return Raw{s[i], s[kRrqrMaxDim + i]}; }
  __device__ double den(double f, int i, const Ctx& x) const {
    return (rb + i <= x.cg) ? ((f - x.fc) - x.dl) : ((f - x.fc1) - x.dr);
  }
  // div_refined over the group, written stage by stage so that the G
  // independent quotients interleave
  template <int G>
  __device__ void values(const Raw* r, const int* ii, const Ctx& x, double* out) const {
    double d[G], nu[G], rc[G], e[G];
#pragma unroll
    for (int j = 0; j < G; ++j) {
      d[j] = den(r[j].f, ii[j], x);
      nu[j] = r[j].u * x.vc;
    }
#pragma unroll
    for (int j = 0; j < G; ++j) rc[j] = __builtin_amdgcn_rcp(d[j]);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
#pragma unroll
      for (int j = 0; j < G; ++j) e[j] = fma(-d[j], rc[j], 1.0);
#pragma unroll
      for (int j = 0; j < G; ++j) rc[j] = fma(rc[j], e[j], rc[j]);
    }
#pragma unroll
    for (int j = 0; j < G; ++j) e[j] = nu[j] * rc[j];
#pragma unroll
    for (int j = 0; j < G; ++j) out[j] = fma(fma(-d[j], e[j], nu[j]), rc[j], e[j]);
  }
  __device__ double at(int i, const Ctx& x) const {
    return div_refined(s[kRrqrMaxDim + i] * x.vc, den(s[i], i, x));
  }
};

// Eigen-Cauchy blocks evaluated once: the first pass (the initial column
// norms) computes each entry as SrcEigStaged does and stores it in a row-major
// copy of the block in device memory, entry (i, j) at i * n + j, so that the
// threads of a wave (one column each) read a row in one coalesced access;
// every later pass reads the copy instead of dividing again.  The copy is
// written and read by the one workgroup that owns the block, and the first
// pass ends in a workgroup barrier.
struct SrcEigCached {
  SrcEigStaged gen;
  double* cache;  // nblk blocks of m * n entries
  int ld;
  double* base;
  struct Ctx {
    SrcEigStaged::Ctx g;
    int c;
  };
  static constexpr bool kCaches = true;
  static constexpr int kG = 2 * kGrp;
  __device__ void bind(int b, double* st, int m, int n) {
    gen.bind(b, st, m, n);
    ld = n;
    base = cache + (size_t)b * m * n;
  }
  __device__ Ctx col(int c) const { return Ctx{gen.col(c), c}; }
  using Raw = double;
  __device__ Raw fetch(int i, const Ctx& x) const { return base[i * ld + x.c]; }
  template <int G>
  __device__ void values(const Raw* r, const int*, const Ctx&, double* out) const {
#pragma unroll
    for (int j = 0; j < G; ++j) out[j] = r[j];
  }
  __device__ double at(int i, const Ctx& x) const { return base[i * ld + x.c]; }
  __device__ void generate(const int* ii, const Ctx& x, double* out) const {
    SrcEigStaged::Raw r[kGrp];
#pragma unroll
    for (int j = 0; j < kGrp; ++j) r[j] = gen.fetch(ii[j], x.g);
    __builtin_amdgcn_sched_barrier(0);
    gen.values<kGrp>(r, ii, x.g, out);
#pragma unroll
    for (int j = 0; j < kGrp; ++j) base[ii[j] * ld + x.c] = out[j];
  }
};

constexpr int kThreads = 512;  // two threads per pivot position (row halves)
constexpr int kWaves = kThreads / 64;
constexpr int kFld = kRrqrKmax + 1;  // F row pitch in LDS

// Sum over the workgroup in a fixed order; every thread gets the same value.
__device__ double block_sum(double x, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < kWaves; ++w) r += red[w];
  return r;
}

// Thread layout: thread t owns pivot position pos = t % 256 together with its
// twin t ^ 256: the two keep identical copies of the position's source column,
// column context, R column and partial norms, split the row loops in halves
// and exchange the two partial sums through LDS (added in a fixed order, so the
// copies stay identical).  The position's F row lives in LDS (written by the
// first twin).  Row-wise work (the pivot column, the reflector, Q) runs on the
// first 256 threads, one row each.
template <int KMAX, class Src>
__global__ __launch_bounds__(kThreads, 1) void rrqr_kernel(Src src, int m, int n, double tol,
                                                           int kmax, int* __restrict__ rank_out,
                                                           int* __restrict__ jpvt_out,
                                                           long long jstride,
                                                           double* __restrict__ q, int ldq,
                                                           long long qstride,
                                                           double* __restrict__ rpt, int ldr,
                                                           long long rstride) {
  static_assert(KMAX == kRrqrKmax, "F pitch");
  __shared__ double V[KMAX * kVld];
  __shared__ double Fs[kRrqrMaxDim * kFld];  // F rows by position; T and W after the loop
  __shared__ double Y[KMAX * KMAX];          // column l: V(:, 0:l)^T v_l
  __shared__ double part[(kThreads / KMAX) * KMAX];
  __shared__ double hpart[2 * kRrqrMaxDim];  // per-half partial sums of each position
  __shared__ double X[2 * KMAX + 8];         // R columns of the two exchanged positions
  __shared__ double s_alpha;
  __shared__ double s_tau[KMAX];
  __shared__ double red[kWaves];
  __shared__ double s_key[kWaves];
  __shared__ int s_idx[kWaves];
  __shared__ int s_col[2];
  __shared__ double gst[2 * kRrqrMaxDim];  // row generators of a structured block

  const int b = blockIdx.x;
  const int t = threadIdx.x;
  const int pos = t & (kRrqrMaxDim - 1), half = t >> 8;
  src.bind(b, gst, m, n);
  for (int x = t; x < KMAX * kVld; x += kThreads) V[x] = 0.0;
  for (int x = t; x < kRrqrMaxDim * kFld; x += kThreads) Fs[x] = 0.0;
  __syncthreads();
  const int minmn = min(m, n);
  const double tol3z = 1.0536712127723509e-08;  // sqrt(2^-53)
  // this thread's half of rows [lo, m)
  auto split = [&](int lo, int& i0, int& i1) {
    const int mid = lo + ((m - lo) >> 1);
    i0 = half ? mid : lo;
    i1 = half ? m : mid;
  };
  // combine the two halves' partial sums (uniform call: contains barriers)
  auto combine = [&](double mine) {
    hpart[half * kRrqrMaxDim + pos] = mine;
    __syncthreads();
    const double r = hpart[pos] + hpart[kRrqrMaxDim + pos];
    __syncthreads();
    return r;
  };

  double R[KMAX];
#pragma unroll
  for (int l = 0; l < KMAX; ++l) R[l] = 0.0;
  int col = pos;
  typename Src::Ctx cx = src.col(min(pos, max(n - 1, 0)));
  double* Frow = Fs + pos * kFld;
  double vn1 = -1.0, vn2 = -1.0;
  {
    int i0, i1;
    split(0, i0, i1);
    constexpr int G0 = Src::kCaches ? kGrp : Src::kG;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    auto group = [&](int i, auto tail) {
      int ii[G0];
#pragma unroll
      for (int j = 0; j < G0; ++j) ii[j] = decltype(tail)::value ? min(i + j, i1 - 1) : i + j;
      double x[G0];
      if constexpr (Src::kCaches) {
        src.generate(ii, cx, x);
      } else {
        typename Src::Raw r[G0];
#pragma unroll
        for (int j = 0; j < G0; ++j) r[j] = src.fetch(ii[j], cx);
        __builtin_amdgcn_sched_barrier(0);
        src.template values<G0>(r, ii, cx, x);
      }
#pragma unroll
      for (int j = 0; j < G0; ++j) {
        if (decltype(tail)::value) x[j] *= (i + j < i1) ? 1.0 : 0.0;
        acc[j & 3] += x[j] * x[j];
      }
    };
    if (pos < n) {
      int i = i0;
      for (; i + G0 <= i1; i += G0) group(i, std::false_type{});
      if (i < i1) group(i, std::true_type{});
    }
    const double s = combine((acc[0] + acc[1]) + (acc[2] + acc[3]));
    if (pos < n) vn1 = vn2 = sqrt(s);
  }

  int rank = 0;
  double r00 = 0.0;
  for (int k = 0; k < minmn; ++k) {
    // (1) pivot: largest partial norm among positions k..n-1, first on ties
    double key = (half == 0 && pos >= k && pos < n) ? vn1 : -1.0;
    int idx = pos;
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
    for (int w = 1; w < kWaves; ++w)
      if (s_key[w] > pk || (s_key[w] == pk && s_idx[w] < p)) {
        pk = s_key[w];
        p = s_idx[w];
      }
    // (2) exchange positions k and p: source column, norms, R column, F row
    if (p != k) {
      const bool mine = pos == p || pos == k;
      const int other = (pos == p) ? k : p;
      double ovn1 = 0.0, ovn2 = 0.0, fo = 0.0;
      if (half == 0 && mine) {
        hpart[pos] = vn1;
        hpart[kRrqrMaxDim + pos] = vn2;
        s_col[pos == p ? 0 : 1] = col;
#pragma unroll
        for (int l = 0; l < KMAX; ++l) X[(pos == p ? 0 : KMAX + 4) + l] = R[l];
      }
      // F rows: thread l of position k's twin pair swaps entry l
      if (t < KMAX) {
        fo = Fs[k * kFld + t];
        Fs[k * kFld + t] = Fs[p * kFld + t];
      }
      __syncthreads();
      if (t < KMAX) Fs[p * kFld + t] = fo;
      if (mine) {
        ovn1 = hpart[other];
        ovn2 = hpart[kRrqrMaxDim + other];
        col = s_col[pos == p ? 1 : 0];
#pragma unroll
        for (int l = 0; l < KMAX; ++l) R[l] = X[(pos == p ? KMAX + 4 : 0) + l];
        vn1 = ovn1;
        vn2 = ovn2;
        cx = src.col(col);
      }
      __syncthreads();
    }
    // (3) the pivot column, updated: a(:, col_k) - V F_k^T (first 256 threads: rows)
    if (t == k) s_col[0] = col;
    __syncthreads();
    const int ck = s_col[0];
    double ahat = 0.0;
    if (half == 0 && t < m) {
      const typename Src::Ctx ckx = src.col(ck);
      ahat = src.at(t, ckx);
      const double* fk = Fs + k * kFld;
      for (int l = 0; l < k; ++l) ahat -= V[l * kVld + t] * fk[l];
    }
    // (4) Householder reflector on rows k..m-1 of the pivot column
    const double xn2 = block_sum((half == 0 && t > k && t < m) ? ahat * ahat : 0.0, red);
    if (t == k) s_alpha = ahat;
    __syncthreads();
    const double alpha = s_alpha;
    double tau = 0.0, beta = alpha, sc = 0.0;
    if (xn2 != 0.0 && k < m - 1) {
      beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
      tau = (beta - alpha) / beta;
      sc = 1.0 / (alpha - beta);
    }
    if (half == 0) V[k * kVld + t] = (t == k) ? 1.0 : ((t > k && t < m) ? ahat * sc : 0.0);
    if (t == 0) s_tau[k] = tau;
    if (k == 0) r00 = fabs(beta);
    if (pos == k) {
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
    // (5) y = V(:, 0:k)^T v_k, segmented over the rows: thread (segment, l)
    double* y = Y + k * KMAX;
    {
      constexpr int NSEG = kThreads / KMAX, ROWS = kRrqrMaxDim / NSEG;
      const int l = t % KMAX, seg = t / KMAX;
      double sy = 0.0;
      if (l < k) {
        const int i0 = max(k, seg * ROWS), i1 = min(m, seg * ROWS + ROWS);
        for (int i = i0; i < i1; ++i) sy += V[l * kVld + i] * V[k * kVld + i];
      }
      part[seg * KMAX + l] = sy;
      __syncthreads();
      if (t < k) {
        double yy = 0.0;
        for (int g2 = 0; g2 < NSEG; ++g2) yy += part[g2 * KMAX + t];
        y[t] = yy;
      }
      __syncthreads();
    }
    // (6) f_k = tau (A^T v_k - F y) and row k of R, for positions after k
    const bool live = pos > k && pos < n;
    double g = 0.0;
    if (live) {
      int i0, i1;
      split(k, i0, i1);
      constexpr int G = Src::kG;
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      const double* vk = V + k * kVld;
      auto group = [&](int i, auto tail) {
        typename Src::Raw r[G];
        double w[G];
        int ii[G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
          ii[j] = decltype(tail)::value ? min(i + j, i1 - 1) : i + j;
          r[j] = src.fetch(ii[j], cx);
          w[j] = vk[ii[j]];
          if (decltype(tail)::value) w[j] *= (i + j < i1) ? 1.0 : 0.0;
        }
        __builtin_amdgcn_sched_barrier(0);
        double x[G];
        src.template values<G>(r, ii, cx, x);
#pragma unroll
        for (int j = 0; j < G; ++j) acc[j & 3] += x[j] * w[j];
      };
      int i = i0;
      for (; i + G <= i1; i += G) group(i, std::false_type{});
      if (i < i1) group(i, std::true_type{});
      g = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    }
    g = combine(g);
    bool recompute = false;
    if (live) {
      double fy = 0.0, vf = 0.0;
      for (int l = 0; l < k; ++l) {
        const double fl = Frow[l];
        fy += fl * y[l];
        vf += V[l * kVld + k] * fl;
      }
      const double f = tau * (g - fy);
      const double rk = (src.at(k, cx) - vf) - f;
#pragma unroll
      for (int l = 0; l < KMAX; ++l)
        if (l == k) R[l] = rk;
      if (half == 0) Frow[k] = f;
      // (7) partial norm downdate, or recomputation from the updated column
      if (vn1 != 0.0) {
        double tt = fabs(rk) / vn1;
        tt = fmax(1.0 - tt * tt, 0.0);
        const double r12 = vn1 / vn2;
        if (tt * r12 * r12 <= tol3z)
          recompute = true;
        else
          vn1 *= sqrt(tt);
      }
    }
    // (uniform; the barrier also makes F(pos, k) visible to both twins)
    if (!__syncthreads_or(recompute)) continue;
    double sr = 0.0;
    if (recompute) {
      int i0, i1;
      split(k + 1, i0, i1);
      constexpr int G = Src::kG;
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      auto group = [&](int i, auto tail) {
        typename Src::Raw r[G];
        int ii[G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
          ii[j] = decltype(tail)::value ? min(i + j, i1 - 1) : i + j;
          r[j] = src.fetch(ii[j], cx);
        }
        __builtin_amdgcn_sched_barrier(0);
        double x[G];
        src.template values<G>(r, ii, cx, x);
        // two reflectors per LDS round trip; each entry still subtracts them in order
        int l = 0;
        for (; l + 1 <= k; l += 2) {
          const double f0 = Frow[l], f1 = Frow[l + 1];
          const double* v0 = V + l * kVld;
          const double* v1 = v0 + kVld;
          double p0[G], p1[G];
#pragma unroll
          for (int j = 0; j < G; ++j) {
            p0[j] = v0[ii[j]];
            p1[j] = v1[ii[j]];
          }
#pragma unroll
          for (int j = 0; j < G; ++j) x[j] = (x[j] - p0[j] * f0) - p1[j] * f1;
        }
        if (l <= k) {
          const double fl = Frow[l];
          const double* vl = V + l * kVld;
#pragma unroll
          for (int j = 0; j < G; ++j) x[j] -= vl[ii[j]] * fl;
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
          if (decltype(tail)::value) x[j] *= (i + j < i1) ? 1.0 : 0.0;
          acc[j & 3] += x[j] * x[j];
        }
      };
      int i = i0;
      for (; i + G <= i1; i += G) group(i, std::false_type{});
      if (i < i1) group(i, std::true_type{});
      sr = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    }
    sr = combine(sr);
    if (recompute) vn1 = vn2 = sqrt(sr);
  }

  if (t == 0) rank_out[b] = rank;
  if (half == 0 && t < n) jpvt_out[(size_t)b * jstride + t] = col + 1;
  if (rank <= 0 || rank >= minmn) return;  // uniform: nothing more to write

  // R P^T: position t's column of R goes to column col
  if (half == 0 && t < n) {
    double* dst = rpt + (size_t)b * rstride + (size_t)col * ldr;
#pragma unroll
    for (int l = 0; l < KMAX; ++l)
      if (l < rank) dst[l] = R[l];
  }
  // Q = H(0) ... H(rank-1) [I; 0] = [I; 0] - V T V(0:rank, :)^T, with T the
  // upper-triangular factor of the compact WY form (T(l,l) = tau_l,
  // T(0:l, l) = -tau_l T(0:l, 0:l) y_l), then W = T V(0:rank, 0:rank)^T.
  // T and W reuse the F rows' LDS.
  __syncthreads();
  double* Tm = Fs;
  double* Wm = Fs + KMAX * KMAX;
  for (int x = t; x < KMAX * KMAX; x += kThreads) Tm[x] = 0.0;
  __syncthreads();
  if (t == 0) {
    for (int l = 0; l < rank; ++l) {
      Tm[l * KMAX + l] = s_tau[l];
      for (int j = 0; j < l; ++j) {
        double acc = 0.0;
        for (int p2 = j; p2 < l; ++p2) acc += Tm[p2 * KMAX + j] * Y[l * KMAX + p2];
        Tm[l * KMAX + j] = -s_tau[l] * acc;  // stored as Tm[col * KMAX + row]
      }
    }
  }
  __syncthreads();
  for (int e2 = t; e2 < rank * rank; e2 += kThreads) {
    const int l = e2 / rank, c = e2 % rank;
    double acc = 0.0;
    for (int j = l; j <= c; ++j) acc += Tm[j * KMAX + l] * V[j * kVld + c];  // V(c,j) = 0 for j > c
    Wm[c * KMAX + l] = acc;
  }
  __syncthreads();
  // rows pos, columns split between the halves
  if (pos < m) {
    double* qb = q + (size_t)b * qstride;
    for (int c = half; c < rank; c += 2) {
      double acc = (pos == c) ? 1.0 : 0.0;
      for (int l = 0; l < rank; ++l) acc -= V[l * kVld + pos] * Wm[c * KMAX + l];
      qb[pos + (size_t)c * ldq] = acc;
    }
  }
}

template <class Src>
hipError_t launch_rrqr(const Src& src, int nblk, int m, int n, double tol, int kmax, int* rank,
                       int* jpvt, long long jstride, double* q, int ldq, long long qstride,
                       double* rpt, int ldr, long long rstride, hipStream_t s) {
  if (nblk <= 0) return hipSuccess;
  rrqr_kernel<kRrqrKmax, Src><<<nblk, kThreads, 0, s>>>(src, m, n, tol, kmax, rank, jpvt, jstride,
                                                        q, ldq, qstride, rpt, ldr, rstride);
  return hipGetLastError();
}

// Blocks (r0[b], c0[b]) of size m x n of a structured matrix (device block lists).
// cache: optional workspace of nblk * m * n doubles; eigen-Cauchy blocks are then
// evaluated once and read back from it.
hipError_t launch_rrqr_structured(int kind, const GenPtrs& g, const int* r0, const int* c0, int nblk,
                                  int m, int n, double tol, int kmax, int* rank, int* jpvt,
                                  long long jstride, double* q, int ldq, long long qstride,
                                  double* rpt, int ldr, long long rstride, double* cache,
                                  hipStream_t s) {
  if (kind == KIND_CAUCHYEIG && cache)
    return launch_rrqr(SrcEigCached{SrcEigStaged{g, r0, c0, 0, 0, nullptr}, cache, 0, nullptr}, nblk,
                       m, n, tol, kmax, rank, jpvt, jstride, q, ldq, qstride, rpt, ldr, rstride, s);
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
      return launch_rrqr(SrcEigStaged{g, r0, c0, 0, 0, nullptr}, nblk, m, n, tol, kmax, rank, jpvt,
                         jstride, q, ldq, qstride, rpt, ldr, rstride, s);
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
