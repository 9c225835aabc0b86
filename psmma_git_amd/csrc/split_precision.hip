// Split-precision structured x dense products on gfx950: an int8 modular form
// (the fast one) and a bf16-slice form.
//
// ---- int8 modular form (PSMMA_PREC_SPLIT_I8) ----
//
// C = alpha * A * H + beta * C with FP64 inputs and an FP64 result, computed
// from exact integer products (an error-free modular decomposition, the
// "Ozaki scheme II" family):
//
//  1. scale each row i of A by 2^sa_i and each column j of H by 2^sh_j so the
//     largest entry is below 2^a (rows of A) or 2^h (columns of H), and round
//     to integers A', H' (this rounding is the only approximation; a = h =
//     abits, except for the Cauchy kinds, whose H takes two bits from A);
//  2. for t pairwise-coprime moduli m_l <= 256 form the symmetric residues
//     A'_l = A' mod m_l, H'_l = H' mod m_l as int8 matrices;
//  3. compute C'_l = A'_l * H'_l exactly on v_mfma_i32_16x16x64_i8 (int32
//     accumulation cannot overflow: |residue| <= 128, K <= 2^17) and keep
//     C'_l mod m_l;
//  4. rebuild X = A' * H' from its residues (Chinese remainder theorem, Garner's
//     mixed-radix form with signed digits, evaluated in double-double FP64); X is
//     determined exactly because M = prod m_l > 2 K 2^(a + h) > 2 |X|; scale
//     back: C = alpha * X * 2^-(sa_i + sh_j) + beta * C.
//
// The structured operand is evaluated from its generators (FP64, divisions of
// the Cauchy kinds included) once per entry in the residue pass; only the
// residues travel to the int8 GEMM.  Error: |C - AH| <= 2^-abits-ish relative
// to row/column maxima; t = 13 gives abits = 43 at K = 8192, t = 14 gives 47.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <unordered_map>
#include <type_traits>
#include <vector>

#include "psmma_common.h"
#include "workspace.h"

typedef long v2l __attribute__((ext_vector_type(2)));

namespace psmma {

constexpr int kMaxModuli = 16;
// Longest contraction of the int8 form: sum_k |r_a r_b| <= K * 2^14 < 2^31.
constexpr int kSplitI8Kmax = 131071;
// slices = 0: the fewest moduli that keep this many bits per operand (int8
// form), or this many bf16 slices (the count that stays <= 1e-13 in the tests).
constexpr int kAutoSplitBits = 45;
constexpr int kAutoBf16Slices = 6;
// bf16 forms with at most this many slices run every kept slice pair into one
// FP32 accumulator (their ~2^(-8s) slicing error is above FP32 accumulation
// rounding); more slices sum each class exactly per K-chunk and flush to FP64.
constexpr int kBf16Fp32AccSlices = 3;
// Exponent of an all-zero row/column (below any FP64 exponent; sums stay in int).
constexpr int kNoExp = -10000;
// Pairwise coprime (distinct prime factors), all <= 256 so a symmetric
// residue fits int8.
constexpr int kModuli[kMaxModuli] = {256, 255, 253, 251, 247, 241, 239, 233,
                                     229, 227, 223, 217, 211, 199, 197, 193};

// y_l / m_l = wa + wb + wc for t = T moduli (y_l = (M/m_l)^-1 mod m_l): wa on
// the 2^-41 grid and wb on the 2^-81 grid, so that a residue (< 2^8) times
// either is exact and sums of up to 16 such products stay exact in FP64; wc is
// the rounded rest (< 2^-81).  Evaluated at compile time, so the CRT kernels
// take the constants as literals (t-length arrays in a kernel argument do not
// fit the scalar registers next to everything else).
template <int T>
struct CrtFrac {
  double wa[T], wb[T], wc[T];
};
constexpr long long cx_inv_mod(long long a, long long m) {
  long long t = 0, nt = 1, r = m, nr = ((a % m) + m) % m;
  while (nr) {
    long long q = r / nr, tmp = t - q * nt;
    t = nt;
    nt = tmp;
    tmp = r - q * nr;
    r = nr;
    nr = tmp;
  }
  return (t % m + m) % m;
}
template <int T>
constexpr CrtFrac<T> make_crt_frac() {
  CrtFrac<T> f{};
  for (int l = 0; l < T; ++l) {
    const long long m = kModuli[l];
    long long ml = 1;  // (M / m_l) mod m_l
    for (int j = 0; j < T; ++j)
      if (j != l) ml = ml * (kModuli[j] % m) % m;
    const unsigned long long y = (unsigned long long)cx_inv_mod(ml, m);
    const unsigned long long num = y << 41, wa = num / m, r1 = num % m;
    const unsigned long long num2 = r1 << 40, wb = num2 / m, r2 = num2 % m;
    f.wa[l] = (double)wa * 0x1p-41;
    f.wb[l] = (double)wb * 0x1p-81;
    f.wc[l] = ((double)r2 / (double)m) * 0x1p-81;
  }
  return f;
}

struct CrtParams {
  int t;
  int m[kMaxModuli];
  int y[kMaxModuli];         // (M/m_l)^-1 mod m_l
  double inv_hi[kMaxModuli];  // 1/m_l as a double-double
  double inv_lo[kMaxModuli];
  double M_hi, M_lo;          // M = prod m_l as a double-double
  double rm[kMaxModuli];      // 1/m_l rounded (for the residue quotient)
  unsigned char minv[kMaxModuli][kMaxModuli];  // minv[j][l] = m_j^-1 mod m_l (j < l)
};

// ---------------------------------------------------------------------------
// Host: CRT constants.

static long long inv_mod(long long a, long long m) {
  long long t = 0, nt = 1, r = m, nr = ((a % m) + m) % m;
  while (nr) {
    long long q = r / nr, tmp;
    tmp = t - q * nt; t = nt; nt = tmp;
    tmp = r - q * nr; r = nr; nr = tmp;
  }
  return (t % m + m) % m;
}

static CrtParams make_crt(int t) {
  CrtParams p{};
  p.t = t;
  unsigned __int128 M = 1;
  for (int l = 0; l < t; ++l) M *= (unsigned)kModuli[l];
  for (int l = 0; l < t; ++l) {
    int m = kModuli[l];
    p.m[l] = m;
    unsigned __int128 Ml = M / (unsigned)m;
    long long r = (long long)(Ml % (unsigned)m);
    p.y[l] = (int)inv_mod(r, m);
    double hi = 1.0 / m;
    double lo = std::fma(-hi, (double)m, 1.0) / m;  // 1 - hi*m is exact in the fma
    p.inv_hi[l] = hi;
    p.inv_lo[l] = lo;
    p.rm[l] = 1.0 / m;
    for (int j = 0; j < l; ++j) p.minv[j][l] = (unsigned char)inv_mod(kModuli[j] % m, m);
  }
  // M as hi + lo (M < 2^128): split the 128-bit integer exactly.
  unsigned long long top = (unsigned long long)(M >> 64), bot = (unsigned long long)M;
  long double Ml = (long double)top * 18446744073709551616.0L + (long double)bot;
  double hi = (double)Ml;
  // exact remainder M - hi in 128-bit integer arithmetic
  __int128 hi_i = 0;
  {
    int e;
    double f = frexp(hi, &e);  // hi = f * 2^e, f in [0.5,1)
    long long mant = (long long)ldexp(f, 53);
    hi_i = (__int128)mant << (e - 53);
  }
  __int128 rem = (__int128)M - hi_i;
  double lo = (double)(long double)rem;
  p.M_hi = hi;
  p.M_lo = lo;
  return p;
}

// Largest abits with 2 * K * 2^(2 abits) < M, capped so A' stays exact in FP64.
static int choose_abits(int t, int K) {
  double lg = 0.0;
  for (int l = 0; l < t; ++l) lg += std::log2((double)kModuli[l]);
  int lk = 0;
  while ((1LL << lk) < (long long)(K > 1 ? K : 1)) ++lk;
  int ab = ((int)std::floor(lg) - 2 - lk) / 2;
  if (ab > 52) ab = 52;
  return ab;
}

// Bits of the two operands (A rows to a bits, H columns to h bits), a + h <=
// 2 choose_abits, so |X| <= K 2^(a + h) < M / 4 as above.  The Cauchy kinds give H
// kCauchyHShift bits more and A as many fewer: a column of U_k V_j / (D_k - F_j)
// carries its norm in the few entries near the pole, while the rounding error of
// every entry is up to 2^-h of that column's largest, so H's rounding dominates
// (summed over K entries against a norm of a few); A is dense and balanced.
// Measured with an exact model of the rounding on the 512-problem batch at t = 13
// (max over its Cauchy problems 1.24e-12 -> 3.3e-13; Toeplitz / Hankel unchanged
// at 1.3e-13 with no shift) and on the eigenvector merge at N = 16384, t = 14
// (2.6e-13 -> 8.6e-14; shifts of 1 and 3 bits lie between).
constexpr int kCauchyHShift = 2;
struct OpBits {
  int a, h;
};
static OpBits choose_bits(int t, int K, int kind) {
  const int ab = choose_abits(t, K);
  const int d = (kind == KIND_CAUCHY || kind == KIND_CAUCHYEIG) ? kCauchyHShift : 0;
  const int h = std::min(52, ab + d);
  return {ab - (h - ab), h};
}

// ---------------------------------------------------------------------------
// Device helpers.

// Symmetric residue of an integer-valued double |v| < 2^53 modulo m:
// q = rint(v/m) may be off by one (v/m is rounded); r = v - q*m is exact.
__device__ __forceinline__ int sym_residue(double v, int m, double rm) {
  double q = rint(v * rm);
  double r = fma(-q, (double)m, v);
  int ri = (int)r;
  int half = m >> 1;
  if (ri >= half) ri -= m;
  if (ri < -half) ri += m;
  return ri;  // in [-m/2, m/2) subset of [-128, 127]
}

typedef unsigned int v4u __attribute__((ext_vector_type(4)));

// Loads and stores through the global address space.  Pointers read from a
// problem table are generic and would otherwise be accessed with flat
// instructions.
template <class T>
__device__ __forceinline__ T gld(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ void gst(T* p, T v) {
  *(__attribute__((address_space(1))) T*)p = v;
}

// Sixteen residues of integer-valued doubles |v| <= 2^52, packed four per word.
// q = rint(v * fl(1/m)) has an absolute error of at most |v| / m * 2^-52 <=
// 1/193, so r = v - q m satisfies |r| <= 0.5052 m: below 128 for every m <= 253
// (integer r), which then needs no correction into [-m/2, m/2); the moduli 255
// and 256 take sym_residue's correction.
template <bool ADJ>
__device__ __forceinline__ v4u residue_words(const double* v, int m, double rm) {
  v4u w;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    unsigned word = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      int r;
      if constexpr (ADJ) {
        r = sym_residue(v[c * 4 + b], m, rm);
      } else {
        const double x = v[c * 4 + b];
        r = (int)fma(-rint(x * rm), (double)m, x);
      }
      word |= ((unsigned)(r & 0xff)) << (8 * b);
    }
    w[c] = word;
  }
  return w;
}

__device__ __forceinline__ int exponent_of(double x) {
  int e;
  frexp(fabs(x), &e);
  return x == 0.0 ? kNoExp : e;
}

__device__ __forceinline__ void fill_int_kernel_body(int* p, int n, int v, int bx_, int by_) {
  int i = bx_ * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void fill_int_kernel(int* p, int n, int v) {
  fill_int_kernel_body(p, n, v, blockIdx.x, blockIdx.y);
}

// Max exponent per row of A (column-major): threads walk rows, blocks split K.
__device__ __forceinline__ void row_exp_kernel_body(int M, int K, const double* __restrict__ A, int lda, int kchunk,
                               int* __restrict__ ea, int bx_, int by_) {
  int i = bx_ * blockDim.x + threadIdx.x;
  int k0 = by_ * kchunk;
  if (i >= M) return;
  int k1 = min(K, k0 + kchunk);
  // eight independent maxima: eight loads in flight per thread
  double m8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double* a = A + (size_t)k0 * lda + i;
  int k = k0;
  for (; k + 8 <= k1; k += 8, a += (size_t)8 * lda) {
#pragma unroll
    for (int q = 0; q < 8; ++q) m8[q] = fmax(m8[q], fabs(a[(size_t)q * lda]));
  }
  for (; k < k1; ++k, a += lda) m8[0] = fmax(m8[0], fabs(*a));
  double mx = fmax(fmax(fmax(m8[0], m8[1]), fmax(m8[2], m8[3])),
                   fmax(fmax(m8[4], m8[5]), fmax(m8[6], m8[7])));
  if (mx > 0.0) atomicMax(&ea[i], exponent_of(mx));
}

__global__ void row_exp_kernel(int M, int K, const double* __restrict__ A, int lda, int kchunk,
                               int* __restrict__ ea) {
  row_exp_kernel_body(M, K, A, lda, kchunk, ea, blockIdx.x, blockIdx.y);
}

// Max exponent per column of H: threads walk columns, blocks split K.
template <int KIND>
__device__ __forceinline__ void col_exp_kernel_body(int K, int N, GenPtrs g, int r0, int c0, int kchunk,
                               int* __restrict__ eh, int bx_, int by_) {
  int j = bx_ * blockDim.x + threadIdx.x;
  int k0 = by_ * kchunk;
  if (j >= N) return;
  int k1 = min(K, k0 + kchunk);
  double mx = 0.0;
  for (int k = k0; k < k1; ++k) mx = fmax(mx, fabs(gen_entry<KIND>(g, r0 + k, c0 + j)));
  if (mx > 0.0) atomicMax(&eh[j], exponent_of(mx));
}

template <int KIND>
__global__ void col_exp_kernel(int K, int N, GenPtrs g, int r0, int c0, int kchunk,
                               int* __restrict__ eh) {
  col_exp_kernel_body<KIND>(K, N, g, r0, c0, kchunk, eh, blockIdx.x, blockIdx.y);
}

// Residues of A: a unit is 64 rows x kSplitAK k of column-major A, taken as four
// 16 x 256 sub-tiles transposed through LDS (each wave then stores runs of 256
// bytes of four rows per modulus: 275 us against 362 us for 64 x 64 sub-tiles,
// whose 64-byte runs cost the stores, at M = K = 8192, t = 13 on MI355X); the
// next sub-tile's global loads are in flight while the current one's residues
// are formed and stored.  Writes t row-major int8 planes Ares[l][i * kp + k]
// (zero for k >= K; kp a multiple of kSplitAK).
constexpr int kSplitAK = 256;
__device__ __forceinline__ void split_a_kernel_body(int M, int K, const double* __restrict__ A,
                                                      int lda, const int* __restrict__ ea,
                                                      int abits, CrtParams cp,
                                                      signed char* __restrict__ res, int kp,
                                                      size_t plane, int bx_, int by_) {
  __shared__ double tile[16 * 273];  // 16 rows x 256 k: groups of 16 k padded to 17
  const int i0 = bx_ * 64, kbase = by_ * kSplitAK;
  const int t = threadIdx.x;
  // thread -> row t / 16 of each 16-row sub-tile, 16 consecutive k from (t % 16) * 16
  const int ir = t / 16, kg = t % 16, kb = kg * 16;
  double pre[16];
  auto load = [&](int r0) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int iq = t % 16, kk = t / 16 + 16 * q;
      const int gr = i0 + r0 + iq, gk = kbase + kk;
      pre[q] = (gr < M && gk < K) ? gld(A + (size_t)gk * lda + gr) : 0.0;
    }
  };
  load(0);
#pragma unroll 1
  for (int st = 0; st < 4; ++st) {
    const int r0 = st * 16;
#pragma unroll
    for (int q = 0; q < 16; ++q) tile[(t % 16) * 273 + q * 17 + t / 16] = pre[q];
    __syncthreads();
    if (st + 1 < 4) load(r0 + 16);
    const int gi = i0 + r0 + ir;
    if (gi < M && kbase + kb < kp) {  // kp may end half-way into the last unit
      const int sh = abits - gld(ea + gi);
      double v[16];
#pragma unroll
      for (int e = 0; e < 16; ++e) v[e] = rint(ldexp(tile[ir * 273 + kg * 17 + e], sh));
      for (int l = 0; l < cp.t; ++l) {
        const int m = cp.m[l];
        const double rm = cp.rm[l];
        v4u* dst = (v4u*)(res + l * plane + (size_t)gi * kp + kbase + kb);
        gst(dst, m > 253 ? residue_words<true>(v, m, rm) : residue_words<false>(v, m, rm));
      }
    }
    __syncthreads();  // every row view is read before the next sub-tile is stored
  }
}

__global__ __launch_bounds__(256) void split_a_kernel(int M, int K, const double* __restrict__ A,
                                                      int lda, const int* __restrict__ ea,
                                                      int abits, CrtParams cp,
                                                      signed char* __restrict__ res, int kp,
                                                      size_t plane) {
  split_a_kernel_body(M, K, A, lda, ea, abits, cp, res, kp, plane, blockIdx.x, blockIdx.y);
}

// Residues of H: column-major int8 planes Hres[l][j * kp + k] (k contiguous),
// entries generated from the generators.  Thread -> (column, 16 consecutive k).
template <int KIND>
__device__ __forceinline__ void split_h_kernel_body(int K, int N, GenPtrs g, int r0, int c0,
                                                      const int* __restrict__ eh, int hbits,
                                                      CrtParams cp, signed char* __restrict__ res,
                                                      int kp, size_t plane, int bx_, int by_) {
  const int t = threadIdx.x;
  const int j = bx_ * 16 + t / 16;
  const int kb = by_ * 256 + (t % 16) * 16;
  if (j >= N || kb >= kp) return;
  const int sh = hbits - eh[j];
  double v[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    int k = kb + e;
    v[e] = (k < K) ? rint(ldexp(gen_entry<KIND>(g, r0 + k, c0 + j), sh)) : 0.0;
  }
  for (int l = 0; l < cp.t; ++l) {
    const int m = cp.m[l];
    const double rm = cp.rm[l];
    v4u* dst = (v4u*)(res + l * plane + (size_t)j * kp + kb);
    gst(dst, m > 253 ? residue_words<true>(v, m, rm) : residue_words<false>(v, m, rm));
  }
}

template <int KIND>
__global__ __launch_bounds__(256) void split_h_kernel(int K, int N, GenPtrs g, int r0, int c0,
                                                      const int* __restrict__ eh, int hbits,
                                                      CrtParams cp, signed char* __restrict__ res,
                                                      int kp, size_t plane) {
  split_h_kernel_body<KIND>(K, N, g, r0, c0, eh, hbits, cp, res, kp, plane, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Window form for Toeplitz/Hankel.  With one power-of-two scale for the whole
// block of H, H' is built from an integer generator, so its residues are
// residues of that generator: R_l[u], u = k + (N-1-j) (Toeplitz, generator
// read backwards) or u = k + j (Hankel), u in [0, K+N-1).  The GEMM then only
// streams the residues of A; each workgroup stages the K+BN-1 bytes of its
// window in LDS once and assembles every B fragment from five dword reads and
// four v_alignbyte_b32 (the 16 bytes a lane supplies are contiguous in u).

// max exponent of |g| over [lo, hi] -> *e
__device__ __forceinline__ void gen_range_exp_kernel_body(const double* __restrict__ g, long long lo, long long hi,
                                     int* __restrict__ e, int bx_, int by_) {
  long long i = lo + (long long)bx_ * blockDim.x + threadIdx.x;
  double v = (i <= hi) ? fabs(g[i]) : 0.0;
  // wave-level max, then one atomic per wave
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0 && v > 0.0) atomicMax(e, exponent_of(v));
}

__global__ void gen_range_exp_kernel(const double* __restrict__ g, long long lo, long long hi,
                                     int* __restrict__ e) {
  gen_range_exp_kernel_body(g, lo, hi, e, blockIdx.x, blockIdx.y);
}

__device__ __forceinline__ void fill_from_kernel_body(int* __restrict__ dst, int n, const int* __restrict__ src, int bx_, int by_) {
  int i = bx_ * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = *src;
}

__global__ void fill_from_kernel(int* __restrict__ dst, int n, const int* __restrict__ src) {
  fill_from_kernel_body(dst, n, src, blockIdx.x, blockIdx.y);
}

// Residue planes of the scaled generator: R_l[u] = sym(rint(g[src(u)] 2^sh)) mod m_l,
// src(u) = hi - u when reversed (Toeplitz) else lo + u; zero for u >= L.
__device__ __forceinline__ void gen_residue_kernel_body(const double* __restrict__ g, long long lo, long long hi,
                                   int reversed, const int* __restrict__ e, int hbits, CrtParams cp,
                                   signed char* __restrict__ res, size_t gplane, int lpad, int bx_, int by_) {
  int u = bx_ * blockDim.x + threadIdx.x;
  if (u >= lpad) return;
  long long L = hi - lo + 1;
  double v = 0.0;
  if (u < L) v = rint(ldexp(g[reversed ? hi - u : lo + u], hbits - *e));
  for (int l = 0; l < cp.t; ++l) res[l * gplane + u] = (signed char)sym_residue(v, cp.m[l], cp.rm[l]);
}

__global__ void gen_residue_kernel(const double* __restrict__ g, long long lo, long long hi,
                                   int reversed, const int* __restrict__ e, int hbits, CrtParams cp,
                                   signed char* __restrict__ res, size_t gplane, int lpad) {
  gen_residue_kernel_body(g, lo, hi, reversed, e, hbits, cp, res, gplane, lpad, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// int8 modular GEMM, one modulus per blockIdx.y:
//   R_l = (Ares_l * B_l) mod m_l, stored as uint8, column-major with leading
//   dimension ldo = M rounded up to 16 (whole dwords / 16-byte chunks).
// Ares_l: M x kp row-major int8.  B_l is either streamed (BMODE_DENSE: kp x N
// column-major residues) or assembled from a generator window (BMODE_TOE /
// BMODE_HANK: residues R_l[u] of the scaled generator, u = k + (N-1-j) or
// u = k + j, staged once per workgroup in LDS; every B fragment is five dword
// reads and four v_alignbyte_b32).
//
// Pipeline: global loads run two stages ahead in registers (the loads for
// stage s+2 are issued when stage s starts and reach LDS at the end of stage
// s+1); LDS fragments run one MFMA k-step ahead; one barrier per stage.
// v_mfma_i32_16x16x64_i8 lane layout (measured): A lane l supplies
// A[l%16][16*(l/16) .. +15], B lane l supplies B[16*(l/16) .. +15][l%16],
// D lane l element r holds D[4*(l/16)+r][l%16].
enum { BMODE_DENSE = 0, BMODE_TOE = 1, BMODE_HANK = 2 };
constexpr int kFastModK = 16384;  // longest padded contraction of the FP32 reduction mod m

// acc mod m in [0, m) for |acc| <= 2^28.  t' = fl(fl(acc) * fl(1/m) - 1/4) (one
// FMA) is within 3 * 2^-24 * |acc| / m <= 1/4 of acc / m - 1/4, so
// q = floor(t') is floor(acc / m) or one less (|q| < 2^21: the 24-bit multiply
// is exact), r = acc - q m lies in [0, 2m) and one unsigned min finishes.
__device__ __forceinline__ unsigned mod_acc(int x, int m, float rmf) {
  const int q = (int)floorf(fmaf((float)x, rmf, -0.25f));
  const unsigned r = (unsigned)(x + __mul24(q, -m));  // one 24-bit multiply-add
  return min(r, r - (unsigned)m);
}

template <int BMODE, int BM, int BN, int BK, int WGM, int WGN, int KMAX, bool STAGE2 = true>
__device__ __forceinline__ void gemm_i8_body(
    int M, int N, int kp, const signed char* __restrict__ Ares, const signed char* __restrict__ Bres,
    size_t aplane, size_t bplane, int lpad, const CrtParams& cp, unsigned char* __restrict__ out,
    size_t oplane, int ldo, int bmi, int bni, int l) {
  constexpr bool WIN = BMODE != BMODE_DENSE;
  constexpr int NT = WGM * WGN * 64;
  constexpr int WM = BM / WGM, WN = BN / WGN;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int LDB = BK + 16;  // staged row pitch: conflict-free 16-byte fragment reads
  constexpr int CH = BK / 16;
  constexpr int AQ = BM * CH / NT;
  constexpr int BQ = WIN ? 1 : BN * CH / NT;
  constexpr int KS = BK / 64;
  constexpr int WBYTES = WIN ? KMAX + BN + 32 : 16;
  static_assert(BM * CH % NT == 0 && (WIN || BN * CH % NT == 0), "tile/threads mismatch");
  static_assert(BK % 64 == 0, "BK must be a multiple of the MFMA depth");

  __shared__ __align__(16) signed char As[2][BM * LDB];
  __shared__ __align__(16) signed char Bs[2][WIN ? 16 : BN * LDB];
  __shared__ __align__(16) unsigned int Win[WBYTES / 4];

  const int m = cp.m[l];
  const signed char* Al = Ares + l * aplane;
  const signed char* Bl = Bres + l * bplane;
  unsigned char* Ol = out + l * oplane;
  const int m0 = bmi * BM, n0 = bni * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / WGN) * WM, wn0 = (wave % WGN) * WN;
  const int lr = lane & 15, lg = lane >> 4;

  if constexpr (WIN) {
    // window u in [ubase, ubase + kp + BN + 16), zero outside [0, lpad)
    const int ubase = (BMODE == BMODE_TOE) ? (N - n0 - BN) : n0;
    unsigned char* wb = (unsigned char*)Win;
    for (int x = tid; x < kp + BN + 16; x += NT) {
      int u = ubase + x;
      wb[x] = (u >= 0 && u < lpad) ? (unsigned char)Bl[u] : 0;
    }
  }

  v4i acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = (v4i){0, 0, 0, 0};

  int cj[TN];
#pragma unroll
  for (int j = 0; j < TN; ++j)
    cj[j] = (BMODE == BMODE_TOE) ? (BN - 1 - wn0 - 16 * j - lr) : (wn0 + 16 * j + lr);

  // Branch-free staging loads: out-of-range rows/columns read a clamped valid
  // address and are replaced by zero with a select.
  int arow[AQ], apart[AQ], brow[BQ], bpart[BQ];
  bool aok[AQ], bok[BQ];
#pragma unroll
  for (int q = 0; q < AQ; ++q) {
    int c = tid + q * NT;
    int row = c / CH;
    apart[q] = c % CH;
    aok[q] = (m0 + row) < M;
    arow[q] = min(m0 + row, M - 1);
  }
#pragma unroll
  for (int q = 0; q < BQ; ++q) {
    int c = tid + q * NT;
    int col = c / CH;
    bpart[q] = c % CH;
    bok[q] = (n0 + col) < N;
    brow[q] = min(n0 + col, N - 1);
  }
  auto gload = [&](int k0, uint4* ra, uint4* rb) {
#pragma unroll
    for (int q = 0; q < AQ; ++q) {
      uint4 v = *(const uint4*)(Al + (size_t)arow[q] * kp + k0 + apart[q] * 16);
      ra[q] = aok[q] ? v : make_uint4(0, 0, 0, 0);
    }
    if constexpr (!WIN) {
#pragma unroll
      for (int q = 0; q < BQ; ++q) {
        uint4 v = *(const uint4*)(Bl + (size_t)brow[q] * kp + k0 + bpart[q] * 16);
        rb[q] = bok[q] ? v : make_uint4(0, 0, 0, 0);
      }
    }
  };
  auto lstore = [&](int buf, const uint4* ra, const uint4* rb) {
#pragma unroll
    for (int q = 0; q < AQ; ++q) {
      int row = (tid + q * NT) / CH;
      *(uint4*)(&As[buf][row * LDB + apart[q] * 16]) = ra[q];
    }
    if constexpr (!WIN) {
#pragma unroll
      for (int q = 0; q < BQ; ++q) {
        int col = (tid + q * NT) / CH;
        *(uint4*)(&Bs[buf][col * LDB + bpart[q] * 16]) = rb[q];
      }
    }
  };

  v2l af[2][TM];
  v2l bfd[2][WIN ? 1 : TN];           // streamed B fragments
  unsigned bw[2][WIN ? TN : 1][5];    // window dwords
  auto frags = [&](int buf, int kglob, int ks, int slot) {
#pragma unroll
    for (int i = 0; i < TM; ++i)
      af[slot][i] = *(const v2l*)(&As[buf][(wm0 + i * 16 + lr) * LDB + ks * 64 + lg * 16]);
    if constexpr (WIN) {
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        int w0 = (kglob + ks * 64 + lg * 16 + cj[j]) >> 2;
#pragma unroll
        for (int d = 0; d < 5; ++d) bw[slot][j][d] = Win[w0 + d];
      }
    } else {
#pragma unroll
      for (int j = 0; j < TN; ++j)
        bfd[slot][j] = *(const v2l*)(&Bs[buf][(wn0 + j * 16 + lr) * LDB + ks * 64 + lg * 16]);
    }
  };
  auto bfrag = [&](int slot, int j, int kglob, int ks) -> v2l {
    if constexpr (WIN) {
      int sh = (kglob + ks * 64 + lg * 16 + cj[j]) & 3;
      unsigned o0 = __builtin_amdgcn_alignbyte(bw[slot][j][1], bw[slot][j][0], sh);
      unsigned o1 = __builtin_amdgcn_alignbyte(bw[slot][j][2], bw[slot][j][1], sh);
      unsigned o2 = __builtin_amdgcn_alignbyte(bw[slot][j][3], bw[slot][j][2], sh);
      unsigned o3 = __builtin_amdgcn_alignbyte(bw[slot][j][4], bw[slot][j][3], sh);
      v2l r;
      r[0] = (long)(((unsigned long)o1 << 32) | o0);
      r[1] = (long)(((unsigned long)o3 << 32) | o2);
      return r;
    } else {
      return bfd[slot][j];
    }
  };

  const int nkt = kp / BK;
  uint4 ra0[AQ], rb0[BQ], ra1[STAGE2 ? AQ : 1], rb1[STAGE2 ? BQ : 1];

  if constexpr (!STAGE2) {
    // One register stage (large tiles: the registers go to the accumulators):
    // the loads of tile kt+1 are issued at the start of tile kt and stored to
    // LDS after its MFMAs were issued, which cover their latency.
    gload(0, ra0, rb0);
    lstore(0, ra0, rb0);
    __syncthreads();
    frags(0, 0, 0, 0);
    for (int kt = 0; kt < nkt; ++kt) {
      const int cur = kt & 1;
      const int kglob = kt * BK;
      const int slot0 = (kt * KS) & 1;
      if (kt + 1 < nkt) gload((kt + 1) * BK, ra0, rb0);
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int slot = (slot0 + ks) & 1;
        if (ks + 1 < KS) frags(cur, kglob, ks + 1, slot ^ 1);
        v2l bf[TN];
#pragma unroll
        for (int j = 0; j < TN; ++j) bf[j] = bfrag(slot, j, kglob, ks);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
            acc[i][j] =
                __builtin_amdgcn_mfma_i32_16x16x64_i8(af[slot][i], bf[j], acc[i][j], 0, 0, 0);
        if (ks + 1 == KS && kt + 1 < nkt) {
          lstore(cur ^ 1, ra0, rb0);
          __syncthreads();
          frags(cur ^ 1, kglob + BK, 0, slot ^ 1);
        }
      }
    }
  } else {

  // One stage: hold = registers with stage kt+1, into = registers for stage kt+2.
  auto stage = [&](int kt, int slot0, uint4* hold_a, uint4* hold_b, uint4* into_a,
                   uint4* into_b) {
    const int cur = kt & 1;
    const int kglob = kt * BK;
    if (kt + 2 < nkt) gload((kt + 2) * BK, into_a, into_b);
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int slot = (slot0 + ks) & 1;
      if (ks + 1 < KS) {
        frags(cur, kglob, ks + 1, slot ^ 1);
      } else if (kt + 1 < nkt) {
        lstore(cur ^ 1, hold_a, hold_b);
        __syncthreads();
        frags(cur ^ 1, kglob + BK, 0, slot ^ 1);
      }
      v2l bf[TN];
#pragma unroll
      for (int j = 0; j < TN; ++j) bf[j] = bfrag(slot, j, kglob, ks);
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[slot][i], bf[j], acc[i][j], 0, 0, 0);
    }
  };

  gload(0, ra0, rb0);
  lstore(0, ra0, rb0);
  if (nkt > 1) gload(BK, ra1, rb1);
  __syncthreads();
  frags(0, 0, 0, 0);
  constexpr int SLOT1 = KS & 1;  // slot in which an odd stage starts
  for (int kt = 0; kt < nkt; kt += 2) {
    stage(kt, 0, ra1, rb1, ra0, rb0);
    if (kt + 1 < nkt) stage(kt + 1, SLOT1, ra0, rb0, ra1, rb1);
  }
  }

  // Reduction mod m.  Residues lie in [-128, 127], so |acc| <= kp * 2^14; for
  // kp <= kFastModK that is <= 2^28 (mod_acc).  Longer contractions use %.
  const float rmf = 1.0f / (float)m;
  // the uniform choice of reduction is taken once, outside the element loops
  auto reduce_store = [&](auto fastc) {
    constexpr bool FAST = decltype(fastc)::value;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        int gc = n0 + wn0 + j * 16 + lr;
        int gr = m0 + wm0 + i * 16 + lg * 4;  // a multiple of 4: gr + 3 < ldo when gr < M
        if (gc >= N || gr >= M) continue;
        unsigned word = 0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int x = acc[i][j][r];
          unsigned y;
          if constexpr (FAST) {
            y = mod_acc(x, m, rmf);
          } else {
            int z = x % m;
            y = (unsigned)(z < 0 ? z + m : z);
          }
          word |= y << (8 * r);
        }
        *(unsigned*)(Ol + (size_t)gc * ldo + gr) = word;
      }
  };
  if (kp <= kFastModK) reduce_store(std::integral_constant<bool, true>{});
  else reduce_store(std::integral_constant<bool, false>{});
}


// ---------------------------------------------------------------------------
// int8 modular GEMM with direct-to-LDS staging: 256 x 256 workgroup tiles,
// eight waves of 128 x 64 (32 accumulators each, two waves per SIMD), k steps
// of 64 through a ring of kDmaStages LDS stages filled by global_load_lds
// (no staging registers, no LDS store instructions).  The fragments of step
// k+1 are read from LDS while the MFMAs of step k run; one barrier per step.
// (Measured on MI355X at N = 8192: 17% faster end to end than the register-
// staged 256 x 256 form; four waves of 128 x 128 are slower, one wave per
// SIMD cannot keep the MFMA pipe busy.)
//
// LDS layout of a stage: 64 bytes (one k step) per row (A) or column (B); the
// four 16-byte chunks of row r are stored in the order c ^ g(r), with
// g(r) = (r >> 1) & 3: the fragment reads (lane l: row l % 16, chunk l / 16)
// then run without bank conflicts (measured with SQ_LDS_BANK_CONFLICT on
// gfx950; the unswizzled layout and other XOR patterns spend half of their
// LDS cycles in conflicts).
// One global_load_lds_dwordx4 fills 16 rows: lane i writes LDS bytes
// [16 i, 16 i + 16) of the group, i.e. row i / 4, stored chunk i % 4, so it
// loads the global chunk (i % 4) ^ g(i / 4).  Rows past M (columns past N)
// load the last valid row; their products only reach outputs that are not
// stored.  Residues are zero for k >= K, so every k step is read whole.
constexpr int kDmaStages = 5;  // 5 x 32 KB: DMA issued four k steps ahead
#pragma clang diagnostic ignored "-Winline-asm"  // M0 is listed as clobbered on purpose

__device__ __forceinline__ int dma_swz(int r) { return (r >> 1) & 3; }

// s_waitcnt with only the vector-memory count bounded (expcnt, lgkmcnt left free).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// BN = 256: one workgroup per CU, five stages, eight waves.  BN = 128: 256 x 128
// tiles, four waves, three stages (72 KB): two workgroups per CU, so one
// workgroup's prologue and epilogue overlap the other's MFMAs (short
// contractions of a batch).  Waves are 128 x 64 in both.
template <int BMODE, int KMAX, int BN>
__device__ __forceinline__ void gemm_i8_dma_body(
    int M, int N, int kp, const signed char* __restrict__ Ares, const signed char* __restrict__ Bres,
    size_t aplane, size_t bplane, int lpad, const CrtParams& cp, unsigned char* __restrict__ out,
    size_t oplane, int ldo, int bmi, int bni, int l) {
  static_assert(BN == 256 || BN == 128, "tile width");
  constexpr bool WIN = BMODE != BMODE_DENSE;
  constexpr int BM = 256, BK = 64, NS = BN == 256 ? kDmaStages : 3;
  constexpr int WGN = BN / 64, NW = 2 * WGN, WM = 128, WN = 64;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int NT = NW * 64;
  constexpr int GPWA = 16 / NW, GPWB = (BN / 16) / NW;  // 16-row DMA groups per wave
  constexpr int SA = BM * BK, SB = WIN ? 16 : BN * BK;
  constexpr int DPS = WIN ? GPWA : GPWA + GPWB;  // DMA instructions per wave per k step
  constexpr int WBYTES = WIN ? KMAX + BN + 32 : 16;

  // one LDS allocation: NS stages of A, then NS stages of B or the window
  constexpr int BYTES_B = WIN ? WBYTES : NS * SB;
  static_assert(NS * SA + BYTES_B <= 160 * 1024, "LDS of one CU");
  __shared__ __align__(1024) signed char lds[NS * SA + BYTES_B];
  signed char(*Ast)[SA] = (signed char(*)[SA])lds;
  signed char(*Bst)[SB] = (signed char(*)[SB])(lds + NS * SA);
  unsigned int* Win = (unsigned int*)(lds + NS * SA);

  const int m = cp.m[l];
  const signed char* Al = Ares + l * aplane;
  const signed char* Bl = Bres + l * bplane;
  unsigned char* Ol = out + l * oplane;
  const int m0 = bmi * BM, n0 = bni * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / WGN) * WM, wn0 = (wave % WGN) * WN;
  const int lr = lane & 15, lg = lane >> 4;

  if constexpr (WIN) {
    const int ubase = (BMODE == BMODE_TOE) ? (N - n0 - BN) : n0;
    unsigned char* wb = (unsigned char*)Win;
    for (int x = tid; x < kp + BN + 16; x += NT) {
      int u = ubase + x;
      wb[x] = (u >= 0 && u < lpad) ? (unsigned char)Bl[u] : 0;
    }
  }

  // DMA sources: wave w fills row groups GPWA*w .. GPWA*w+GPWA-1 (16 rows each)
  // of A and column groups GPWB*w .. of B.  Offsets fit 32 bits (checked on the host).
  const int rg = lane >> 2;                                   // row within a group
  const int gch = (lane & 3) ^ dma_swz(rg);                   // global chunk this lane loads
  unsigned aoff[GPWA], boff[GPWB];
#pragma unroll
  for (int q = 0; q < GPWA; ++q) {
    const int r = 16 * (GPWA * wave + q) + rg;
    aoff[q] = (unsigned)min(m0 + r, M - 1) * (unsigned)kp + 16u * gch;
  }
#pragma unroll
  for (int q = 0; q < GPWB; ++q) {
    const int r = 16 * (GPWB * wave + q) + rg;
    boff[q] = WIN ? 0u : (unsigned)min(n0 + r, N - 1) * (unsigned)kp + 16u * gch;
  }
  // DMA of k step kt into stage kt % NS.  Written as inline assembly: the
  // compiler's builtin makes every later LDS read wait for all outstanding
  // DMA (it cannot tell the stages apart), which would serialise the ring.
  // The instruction has no register results; its completion is tracked by
  // the explicit vmcnt waits below.  Source: uniform base (SGPRs) + 32-bit
  // lane offset; destination: M0 + 16 * lane.
  auto dma16 = [&](const signed char* gbase, unsigned voff, const signed char* ldst) {
    const unsigned la = __builtin_amdgcn_readfirstlane((unsigned)(size_t)ldst);
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, %2"
                 :
                 : "s"(la), "v"(voff), "s"(gbase)
                 : "memory", "m0");
  };
  auto issue = [&](int kt, int st) {
    const signed char* ak = Al + (size_t)kt * BK;
#pragma unroll
    for (int q = 0; q < GPWA; ++q) dma16(ak, aoff[q], &Ast[st][(GPWA * wave + q) * 1024]);
    if constexpr (!WIN) {
      const signed char* bk = Bl + (size_t)kt * BK;
#pragma unroll
      for (int q = 0; q < GPWB; ++q) dma16(bk, boff[q], &Bst[st][(GPWB * wave + q) * 1024]);
    }
  };

  v4i acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = (v4i){0, 0, 0, 0};

  int cj[TN];
#pragma unroll
  for (int j = 0; j < TN; ++j)
    cj[j] = (BMODE == BMODE_TOE) ? (BN - 1 - wn0 - 16 * j - lr) : (wn0 + 16 * j + lr);

  const int lofs = lr * 64 + ((lg ^ dma_swz(lr)) * 16);  // this lane's fragment bytes in a block
  // A and streamed B fragments are read one step ahead into two register
  // slots; window B words are read at the start of their own step.
  v2l af[2][TM];
  v2l bfd[2][WIN ? 1 : TN];
  unsigned bw[WIN ? TN : 1][5];
  auto frags = [&](int kt, int slot) {
    const signed char* sa = Ast[kt % NS];
#pragma unroll
    for (int i = 0; i < TM; ++i) af[slot][i] = *(const v2l*)(sa + (wm0 + 16 * i) * 64 + lofs);
    if constexpr (!WIN) {
      const signed char* sb = Bst[kt % NS];
#pragma unroll
      for (int j = 0; j < TN; ++j)
        bfd[slot][j] = *(const v2l*)(sb + (wn0 + 16 * j) * 64 + lofs);
    }
  };
  auto window_words = [&](int kt) {
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      int w0 = (kt * BK + lg * 16 + cj[j]) >> 2;
#pragma unroll
      for (int d = 0; d < 5; ++d) bw[j][d] = Win[w0 + d];
    }
  };
  auto bfrag = [&](int slot, int j, int kt) -> v2l {
    if constexpr (WIN) {
      int sh = (kt * BK + lg * 16 + cj[j]) & 3;
      unsigned o0 = __builtin_amdgcn_alignbyte(bw[j][1], bw[j][0], sh);
      unsigned o1 = __builtin_amdgcn_alignbyte(bw[j][2], bw[j][1], sh);
      unsigned o2 = __builtin_amdgcn_alignbyte(bw[j][3], bw[j][2], sh);
      unsigned o3 = __builtin_amdgcn_alignbyte(bw[j][4], bw[j][3], sh);
      v2l r;
      r[0] = (long)(((unsigned long)o1 << 32) | o0);
      r[1] = (long)(((unsigned long)o3 << 32) | o2);
      return r;
    } else {
      return bfd[slot][j];
    }
  };

  const int nkt = kp / BK;
  // prologue: steps 0 .. NS-2 in flight (nkt >= 4 >= NS - 1), step 0 landed everywhere
#pragma unroll
  for (int q = 0; q < NS - 1; ++q) issue(q, q);
  wait_vmcnt<(NS - 2) * DPS>();
  if constexpr (WIN) __syncthreads();  // the window was written with LDS stores
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
  frags(0, 0);

  // one k step; the register slot (kt parity) is a compile-time constant
  auto step = [&](int kt, auto slotc) {
    constexpr int slot = decltype(slotc)::value;
    // One barrier per two steps.  At an even step, steps kt+1 and kt+2 have
    // landed for this wave (only kt+3 may be in flight), then for every wave,
    // and every wave has read its fragments of steps <= kt-1.  The DMA of this
    // step and of the next (odd) one write stages (kt-1) % NS and kt % NS,
    // whose fragments were read before this barrier; the prefetches of this
    // step and the next read stages kt+1 and kt+2.
    // The loop body has no conditions: past the last step, the DMA re-reads
    // step nkt-1 into the free stage and the fragment prefetch reads a landed
    // stage; neither result is used.
    // With three stages (DMA two steps ahead) the barrier is taken every step,
    // once step kt+1 has landed (nothing later is in flight at that point).
    static_assert(NS == 5 || NS == 3, "barrier schedules for five or three stages");
    if constexpr (NS == 5) {
      if constexpr (slot == 0) {
        wait_vmcnt<DPS>();
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
      }
    } else {
      wait_vmcnt<0>();
      asm volatile("" ::: "memory");
      __builtin_amdgcn_s_barrier();
      asm volatile("" ::: "memory");
    }
    if constexpr (WIN) window_words(kt);
    issue(min(kt + NS - 1, nkt - 1), (kt + NS - 1) % NS);
    frags(kt + 1, slot ^ 1);
    // keep the prefetch ahead of the MFMAs (the scheduler would otherwise sink
    // it below them and fold the two register slots into one)
    __builtin_amdgcn_sched_barrier(0);
    v2l bf[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) bf[j] = bfrag(slot, j, kt);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[slot][i], bf[j], acc[i][j], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
  };
  for (int kt = 0; kt < nkt; kt += 2) {  // kp is a multiple of 256: nkt is even, >= 4
    step(kt, std::integral_constant<int, 0>{});
    step(kt + 1, std::integral_constant<int, 1>{});
  }
  wait_vmcnt<0>();  // no DMA outlives the loop

  const float rmf = 1.0f / (float)m;
  // The residue tile goes through LDS (column-major, pitch OP bytes) and
  // leaves as 16-byte column chunks: 4-byte stores scattered over 16 columns
  // per instruction made the epilogue store-bound.  OP / 4 = 68 = 4 (mod 32)
  // spreads a wave's dword writes over all banks.  ldo is a multiple of 16,
  // so a chunk that starts inside a column's ldo rows ends inside them.
  constexpr int OP = 272;
  static_assert(BN * OP <= NS * SA + BYTES_B, "output tile fits the stage memory");
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();  // every wave is done with the stages
  asm volatile("" ::: "memory");
  // the uniform choice of reduction is taken once, outside the element loops
  auto reduce_store = [&](auto fastc) {
    constexpr bool FAST = decltype(fastc)::value;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        unsigned word = 0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int x = acc[i][j][r];
          unsigned y;
          if constexpr (FAST) {
            y = mod_acc(x, m, rmf);
          } else {
            int z = x % m;
            y = (unsigned)(z < 0 ? z + m : z);
          }
          word |= y << (8 * r);
        }
        *(unsigned*)(lds + (wn0 + j * 16 + lr) * OP + wm0 + i * 16 + lg * 4) = word;
      }
  };
  if (kp <= kFastModK) reduce_store(std::integral_constant<bool, true>{});
  else reduce_store(std::integral_constant<bool, false>{});
  __syncthreads();
  constexpr int CHUNKS = BN * (BM / 16);
  static_assert(CHUNKS % NT == 0, "chunks per thread");
#pragma unroll
  for (int it = 0; it < CHUNKS / NT; ++it) {
    const int x = it * NT + tid;
    const int c = x >> 4, q = x & 15;
    const int gc = n0 + c, gr = m0 + 16 * q;
    if (gc < N && gr < ldo)
      *(uint4*)(Ol + (size_t)gc * ldo + gr) = *(const uint4*)(lds + c * OP + 16 * q);
  }
}

template <int BMODE, int KMAX, int BN>
__global__ __launch_bounds__(BN / 2 * 4, BN == 256 ? 1 : 2) void gemm_i8_dma_kernel(
    int M, int N, int kp, const signed char* __restrict__ Ares, const signed char* __restrict__ Bres,
    size_t aplane, size_t bplane, int lpad, CrtParams cp, unsigned char* __restrict__ out,
    size_t oplane, int ldo) {
  const int nbn = (N + BN - 1) / BN, nbm = (M + 255) / 256;
  int bmi, bni;
  patch_remap<4, 8>(blockIdx.x, nbm, nbn, bmi, bni);  // one XCD's 32 resident tiles share panels
  gemm_i8_dma_body<BMODE, KMAX, BN>(M, N, kp, Ares, Bres, aplane, bplane, lpad, cp, out, oplane, ldo,
                                bmi, bni, blockIdx.y);
}

template <int BMODE, int BM, int BN, int BK, int WGM, int WGN, int MINB, int KMAX,
          bool STAGE2 = true>
__global__ __launch_bounds__(WGM* WGN * 64, MINB) void gemm_i8_kernel(
    int M, int N, int kp, const signed char* __restrict__ Ares, const signed char* __restrict__ Bres,
    size_t aplane, size_t bplane, int lpad, CrtParams cp, unsigned char* __restrict__ out,
    size_t oplane, int ldo) {
  const int nbn = (N + BN - 1) / BN, nbm = (M + BM - 1) / BM;
  int bmi, bni;
  patch_remap<4, 8>(blockIdx.x, nbm, nbn, bmi, bni);
  gemm_i8_body<BMODE, BM, BN, BK, WGM, WGN, KMAX, STAGE2>(M, N, kp, Ares, Bres, aplane, bplane,
                                                          lpad, cp, out, oplane, ldo, bmi, bni,
                                                          blockIdx.y);
}

// The error-free transformations below rely on every rounding being the one
// written: FMA contraction (hipcc's default) would fuse a product into an
// addition and break them, so it is off in these functions.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// CRT recombination + scaling: C(i,j) = alpha * X * 2^-(sa_i + sh_j) + beta * C(i,j).
//
// Fast form: X/M = frac(sum_l r_l y_l / m_l) with r_l the stored residue and
// y_l = (M/m_l)^-1 mod m_l.  With y_l / m_l = wa_l + wb_l + wc_l (CrtParams),
// sum r_l wa_l and sum r_l wb_l are formed exactly by FMAs and sum r_l wc_l
// (< 2^-69) in plain FP64, so the fraction is known to ~2^-106 absolute and
// gives X to full FP64 precision whenever |X/M| >= 2^-48.
// Exact form (taken only for the rare |X/M| < 2^-48): Garner's mixed-radix
// conversion with signed digits d_l in [-m_l/2, m_l/2) (exact small-integer
// arithmetic, moduli known at compile time), X = d_0 + m_0 (d_1 + m_1 (...)),
// evaluated by Horner in double-double from the most significant digit, so a
// small |X| loses nothing to the size of M.
template <int T>
__device__ __forceinline__ double crt_garner(const unsigned char* __restrict__ res, size_t oplane,
                                          size_t off, const CrtParams& cp) {
#pragma clang fp contract(off)
  int d[T];
#pragma unroll
  for (int l = 0; l < T; ++l) {
    const int m = kModuli[l];
    int v = res[l * oplane + off];  // in [0, m)
#pragma unroll
    for (int q = 0; q < l; ++q) v = ((v - d[q] + 2 * m) * (int)cp.minv[q][l]) % m;
    d[l] = (v >= (m + 1) / 2) ? v - m : v;
  }
  double yh = (double)d[T - 1], yl = 0.0;
#pragma unroll
  for (int l = T - 2; l >= 0; --l) {
    const double m = (double)kModuli[l];
    double ph = yh * m;
    double pl = fma(yh, m, -ph) + yl * m;
    double sh, se;
    two_sum(ph, (double)d[l], sh, se);
    double sl = pl + se;
    yh = sh + sl;
    yl = sl - (yh - sh);
  }
  return yh + yl;
}

// Four consecutive rows i0 .. i0+3 (i0 a multiple of 4) of column j: one dword
// of residues per modulus.  Rows >= M are padding and are not written.
template <int T>
__device__ __forceinline__ void crt_quad(int M, const unsigned char* __restrict__ res,
                                         size_t oplane, int ldo, const CrtParams& cp,
                                         const int* __restrict__ ea, const int* __restrict__ eh,
                                         int abits, int hbits, double alpha, double beta,
                                         double* __restrict__ C, int ldc, int i0, int j) {
#pragma clang fp contract(off)
  const size_t off = (size_t)j * ldo + i0;
  unsigned w[T];
#pragma unroll
  for (int l = 0; l < T; ++l) w[l] = *(const unsigned*)(res + l * oplane + off);
  const int s_h = hbits - eh[j];
  double* cj = C + (size_t)j * ldc;
  constexpr CrtFrac<T> fr = make_crt_frac<T>();
  unsigned slow = 0;  // rows that need the exact form
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = i0 + r;
    if (i >= M) break;
    double sa = 0.0, sb = 0.0, sc = 0.0;
#pragma unroll
    for (int l = 0; l < T; ++l) {
      const double x = (double)((w[l] >> (8 * r)) & 255u);
      sa = fma(x, fr.wa[l], sa);  // exact
      sb = fma(x, fr.wb[l], sb);  // exact
      sc = fma(x, fr.wc[l], sc);
    }
    double f, fl;
    two_sum(sa - rint(sa), sb, f, fl);
    fl += sc;
    if (f > 0.5) f -= 1.0;
    else if (f < -0.5) f += 1.0;
    if (fabs(f) < 0x1p-48) {
      slow |= 1u << r;
      continue;
    }
    const double xh = f * cp.M_hi;
    const double x = xh + (fma(f, cp.M_hi, -xh) + f * cp.M_lo + fl * cp.M_hi);
    double v = alpha * ldexp(x, -(abits - ea[i] + s_h));
    if (beta != 0.0) v += beta * cj[i];
    cj[i] = v;
  }
  // one copy of the exact form, looped over the rows that need it
  while (slow) {
    const int r = __builtin_ctz(slow);
    slow &= slow - 1;
    const int i = i0 + r;
    const double x = crt_garner<T>(res, oplane, off + r, cp);
    double v = alpha * ldexp(x, -(abits - ea[i] + s_h));
    if (beta != 0.0) v += beta * cj[i];
    cj[i] = v;
  }
}

// Thread e of the padded ldo x N residue grid handles rows 4e .. 4e+3 (mod ldo).
template <int T>
__device__ __forceinline__ void crt_unit(int M, int N, const unsigned char* __restrict__ res,
                                         size_t oplane, int ldo, const CrtParams& cp,
                                         const int* __restrict__ ea, const int* __restrict__ eh,
                                         int abits, int hbits, double alpha, double beta,
                                         double* __restrict__ C, int ldc, size_t e) {
  const size_t el = e * 4;
  int i0, j;
  if (el < 0x80000000ull) {  // 32-bit index arithmetic whenever it suffices
    const unsigned o = (unsigned)el, lu = (unsigned)ldo;
    j = (int)(o / lu);
    i0 = (int)(o - (unsigned)j * lu);
  } else {
    j = (int)(el / ldo);
    i0 = (int)(el - (size_t)j * ldo);
  }
  if (j >= N) return;
  crt_quad<T>(M, res, oplane, ldo, cp, ea, eh, abits, hbits, alpha, beta, C, ldc, i0, j);
}

template <int T>
__global__ __launch_bounds__(256) void crt_kernel(int M, int N, const unsigned char* __restrict__ res,
                                                  size_t oplane, int ldo, CrtParams cp,
                                                  const int* __restrict__ ea,
                                                  const int* __restrict__ eh, int abits, int hbits,
                                                  double alpha, double beta, double* __restrict__ C,
                                                  int ldc) {
  crt_unit<T>(M, N, res, oplane, ldo, cp, ea, eh, abits, hbits, alpha, beta, C, ldc,
              (size_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// Launches the int8 modular GEMM for all t moduli.  cfg: PSMMA_I8_CFG or by size.
template <int BMODE>
static void launch_gemm_i8(int t, int M, int N, int kp, const signed char* ares,
                           const signed char* bres, size_t aplane, size_t bplane, int lpad,
                           const CrtParams& cp, unsigned char* ores, size_t oplane, int ldo,
                           hipStream_t s) {
  constexpr int KMAX = 16384;
  static int forced = [] {
    const char* e = getenv("PSMMA_I8_CFG");
    return e ? atoi(e) : -1;
  }();
  int cfg = forced;
  auto nblk = [&](int bm, int bn) { return ((M + bm - 1) / bm) * ((N + bn - 1) / bn); };
  // 256 x 256 direct-to-LDS tiles (one workgroup per CU) once the t planes
  // have enough of them to fill the chip; 128 x 128 below that.
  if (cfg < 0) cfg = (nblk(256, 256) * t >= 256) ? 5 : 0;
  // the direct-to-LDS form addresses a residue plane with 32-bit offsets
  if ((cfg == 5 || cfg == 6) && ((size_t)M * kp >= (1ull << 32) || (size_t)N * kp >= (1ull << 32)))
    cfg = 4;
  switch (cfg) {
    case 1:
      hipLaunchKernelGGL((gemm_i8_kernel<BMODE, 128, 256, 128, 2, 2, 1, KMAX>),
                         dim3(nblk(128, 256), t), dim3(256), 0, s, M, N, kp, ares, bres, aplane,
                         bplane, lpad, cp, ores, oplane, ldo);
      break;
    case 2:
      hipLaunchKernelGGL((gemm_i8_kernel<BMODE, 256, 128, 128, 2, 2, 1, KMAX>),
                         dim3(nblk(256, 128), t), dim3(256), 0, s, M, N, kp, ares, bres, aplane,
                         bplane, lpad, cp, ores, oplane, ldo);
      break;
    case 4:
      hipLaunchKernelGGL((gemm_i8_kernel<BMODE, 256, 256, 128, 2, 2, 1, KMAX, false>),
                         dim3(nblk(256, 256), t), dim3(256), 0, s, M, N, kp, ares, bres, aplane,
                         bplane, lpad, cp, ores, oplane, ldo);
      break;
    case 5:
      hipLaunchKernelGGL((gemm_i8_dma_kernel<BMODE, KMAX, 256>), dim3(nblk(256, 256), t),
                         dim3(512), 0, s, M, N, kp, ares, bres, aplane, bplane, lpad, cp, ores,
                         oplane, ldo);
      break;
    case 6:
      hipLaunchKernelGGL((gemm_i8_dma_kernel<BMODE, KMAX, 128>), dim3(nblk(256, 128), t),
                         dim3(256), 0, s, M, N, kp, ares, bres, aplane, bplane, lpad, cp, ores,
                         oplane, ldo);
      break;

    default:
      hipLaunchKernelGGL((gemm_i8_kernel<BMODE, 128, 128, 128, 2, 2, 2, KMAX>),
                         dim3(nblk(128, 128), t), dim3(256), 0, s, M, N, kp, ares, bres, aplane,
                         bplane, lpad, cp, ores, oplane, ldo);
  }
}

static int finish_crt(int t, int M, int N, const unsigned char* ores, size_t oplane, int ldo,
                      const CrtParams& cp, const int* ea, const int* eh, int abits, int hbits,
                      double alpha, double beta, double* C, int ldc, hipStream_t s) {
  const size_t nquad = (size_t)ldo * N / 4;
  dim3 cg((unsigned)((nquad + 255) / 256));
#define PSMMA_CRT_CASE(TT)                                                                     \
  case TT:                                                                                     \
    hipLaunchKernelGGL(crt_kernel<TT>, cg, dim3(256), 0, s, M, N, ores, oplane, ldo, cp, ea,    \
                       eh, abits, hbits, alpha, beta, C, ldc);                                 \
    break;
  switch (t) {
    PSMMA_CRT_CASE(2) PSMMA_CRT_CASE(3) PSMMA_CRT_CASE(4) PSMMA_CRT_CASE(5) PSMMA_CRT_CASE(6)
    PSMMA_CRT_CASE(7) PSMMA_CRT_CASE(8) PSMMA_CRT_CASE(9) PSMMA_CRT_CASE(10) PSMMA_CRT_CASE(11)
    PSMMA_CRT_CASE(12) PSMMA_CRT_CASE(13) PSMMA_CRT_CASE(14) PSMMA_CRT_CASE(15) PSMMA_CRT_CASE(16)
    default: return -16;
  }
#undef PSMMA_CRT_CASE
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : 1000 + (int)e;
}

// Workspace layout of one int8 modular product: row / column exponents, t
// residue planes of A, of H (dense form) or of the generator window (Toeplitz /
// Hankel with kp <= kWindowKmax), and t residue planes of the result.
constexpr int kWindowKmax = 16384;  // longest contraction whose window one workgroup stages in LDS

struct I8Layout {
  bool window = false;
  int kp = 0, ldo = 0, lpad = 0;
  long long glo = 0, ghi = -1;
  size_t aplane = 0, hplane = 0, oplane = 0;
  size_t off_eh = 0, off_ares = 0, off_hres = 0, off_ores = 0, bytes = 0;
};

static I8Layout i8_layout(int kind, int t, int M, int N, int K, const GenPtrs& g, int r0, int c0) {
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  I8Layout L;
  L.kp = ((K + 255) / 256) * 256;
  L.ldo = (M + 15) & ~15;  // residue column pitch (16-byte chunks)
  L.aplane = (size_t)M * L.kp;
  L.oplane = (size_t)L.ldo * N;
  L.window = (kind == KIND_TOEPLITZ || kind == KIND_HANKEL) && L.kp <= kWindowKmax;
  size_t hbytes;
  if (L.window) {
    const bool toe = kind == KIND_TOEPLITZ;
    L.glo = toe ? (long long)g.korder - 1 + c0 - (r0 + K - 1) : (long long)r0 + c0;
    L.ghi = toe ? (long long)g.korder - 1 + (c0 + N - 1) - r0 : (long long)r0 + K - 1 + c0 + N - 1;
    L.lpad = (int)(((L.ghi - L.glo + 1) + 15) / 16 * 16);
    L.hplane = (size_t)L.lpad;
    hbytes = (size_t)t * L.lpad;
  } else {
    L.hplane = (size_t)L.kp * N;
    hbytes = (size_t)t * L.hplane;
  }
  L.off_eh = al(sizeof(int) * M);
  L.off_ares = L.off_eh + al(sizeof(int) * N);
  L.off_hres = L.off_ares + al((size_t)t * L.aplane);
  L.off_ores = L.off_hres + al(hbytes);
  L.bytes = L.off_ores + al(std::max((size_t)t * L.oplane, sizeof(int)));
  return L;
}

// Enqueues the int8 modular product in the workspace `ws` (i8_layout bytes).
template <int KIND>
static int enqueue_split_i8(int t, int M, int N, int K, double alpha, const double* A, int lda,
                            const GenPtrs& g, int r0, int c0, double beta, double* C, int ldc,
                            const I8Layout& L, char* ws, hipStream_t s) {
  const CrtParams cp = make_crt(t);
  const OpBits bits = choose_bits(t, K, KIND);
  const int abits = bits.a, hbits = bits.h;
  const int kp = L.kp, ldo = L.ldo;
  int* ea = (int*)ws;
  int* eh = (int*)(ws + L.off_eh);
  signed char* ares = (signed char*)(ws + L.off_ares);
  signed char* hres = (signed char*)(ws + L.off_hres);
  unsigned char* ores = (unsigned char*)(ws + L.off_ores);

  hipLaunchKernelGGL(fill_int_kernel, dim3((M + 255) / 256), dim3(256), 0, s, ea, M, kNoExp);
  hipLaunchKernelGGL(fill_int_kernel, dim3((N + 255) / 256), dim3(256), 0, s, eh, N, kNoExp);
  const int kchunk = 256;
  hipLaunchKernelGGL(row_exp_kernel, dim3((M + 255) / 256, (K + kchunk - 1) / kchunk), dim3(256), 0,
                     s, M, K, A, lda, kchunk, ea);
  hipLaunchKernelGGL(split_a_kernel, dim3((M + 63) / 64, kp / kSplitAK), dim3(256), 0, s, M, K, A, lda,
                     ea, abits, cp, ares, kp, L.aplane);
  if constexpr (KIND == KIND_TOEPLITZ || KIND == KIND_HANKEL) {
    if (L.window) {
      // one scale for the block of H; residues of the generator window
      constexpr bool TOE = (KIND == KIND_TOEPLITZ);
      int* e1 = (int*)ores;  // scratch int, consumed before the GEMM writes ores
      hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, e1, 1, kNoExp);
      long long len = L.ghi - L.glo + 1;
      hipLaunchKernelGGL(gen_range_exp_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s,
                         g.p0, L.glo, L.ghi, e1);
      hipLaunchKernelGGL(fill_from_kernel, dim3((N + 255) / 256), dim3(256), 0, s, eh, N, e1);
      hipLaunchKernelGGL(gen_residue_kernel, dim3((L.lpad + 255) / 256), dim3(256), 0, s, g.p0,
                         L.glo, L.ghi, TOE ? 1 : 0, e1, hbits, cp, hres, (size_t)L.lpad, L.lpad);
      launch_gemm_i8<TOE ? BMODE_TOE : BMODE_HANK>(t, M, N, kp, ares, hres, L.aplane,
                                                     (size_t)L.lpad, L.lpad, cp, ores, L.oplane,
                                                     ldo, s);
      return finish_crt(t, M, N, ores, L.oplane, ldo, cp, ea, eh, abits, hbits, alpha, beta, C,
                        ldc, s);
    }
  }
  hipLaunchKernelGGL(col_exp_kernel<KIND>, dim3((N + 255) / 256, (K + kchunk - 1) / kchunk),
                     dim3(256), 0, s, K, N, g, r0, c0, kchunk, eh);
  hipLaunchKernelGGL(split_h_kernel<KIND>, dim3((N + 15) / 16, (kp + 255) / 256), dim3(256), 0, s,
                     K, N, g, r0, c0, eh, hbits, cp, hres, kp, L.hplane);
  launch_gemm_i8<BMODE_DENSE>(t, M, N, kp, ares, hres, L.aplane, L.hplane, 0, cp, ores, L.oplane,
                              ldo, s);
  return finish_crt(t, M, N, ores, L.oplane, ldo, cp, ea, eh, abits, hbits, alpha, beta, C, ldc,
                    s);
}

static int enqueue_split_i8_kind(int kind, int t, int M, int N, int K, double alpha,
                                 const double* A, int lda, const GenPtrs& g, int r0, int c0,
                                 double beta, double* C, int ldc, const I8Layout& L, char* ws,
                                 hipStream_t s) {
  switch (kind) {
    case KIND_CAUCHY:
      return enqueue_split_i8<KIND_CAUCHY>(t, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc, L, ws, s);
    case KIND_TOEPLITZ:
      return enqueue_split_i8<KIND_TOEPLITZ>(t, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc, L, ws, s);
    case KIND_HANKEL:
      return enqueue_split_i8<KIND_HANKEL>(t, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc, L, ws, s);
    case KIND_CAUCHYEIG:
      return enqueue_split_i8<KIND_CAUCHYEIG>(t, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc, L, ws, s);
  }
  return -1;
}

hipError_t launch_right_f64(int kind, int cfg, int M, int N, int K, double alpha, const double* A,
                            int lda, const GenPtrs& g, int r0, int c0, double beta, double* C,
                            int ldc, hipStream_t s);

static int launch_right_split_i8(int kind, int t, int M, int N, int K, double alpha,
                                 const double* A, int lda, const GenPtrs& g, int r0, int c0,
                                 double beta, double* C, int ldc, hipStream_t s) {
  if (t < 2 || t > kMaxModuli) return -16;
  if (choose_abits(t, K) < 8) return -16;
  const I8Layout L = i8_layout(kind, t, M, N, K, g, r0, c0);
  std::unique_lock<std::mutex> lk;
  char* ws = (char*)ws_lease(WS_SPLIT, L.bytes, s, lk);
  if (!ws) return 1002;
  return enqueue_split_i8_kind(kind, t, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc, L, ws, s);
}

// ===========================================================================
// bf16-slice form (PSMMA_PREC_SPLIT_BF16).
//
// With the same power-of-two row/column scales (|x| < 1 after scaling), each
// value is cut into s slices on a fixed-point grid: x_1 = rint(x 2^8) 2^-8,
// x_p = rint(r_{p-1} 2^(8p)) 2^(-8p) with r_p the remainder.  Every slice has
// at most 8 significant bits, so it is exact in bf16, and |x - sum x_p| <= 2^(-8s-1).
// Products x_p y_q are exact in FP32; all pairs with p + q = d sit on the grid
// 2^(-8d) with |x_p y_q| <= 2^16 grid units, so their FP32 sum over a K-chunk of
// 256 stays below 2^24 units and is exact.  Pairs with p + q <= s + 1 are kept,
// one FP32 accumulator per class d, flushed into FP64 accumulators per chunk.

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ __forceinline__ unsigned bf16_bits_exact(double v) {
  // v has <= 8 significant bits: the float conversion and the truncation to
  // the upper 16 bits of the float are both exact.
  float f = (float)v;
  return __float_as_uint(f) >> 16;
}

// A slices: s planes of bf16, row-major M x kp (k contiguous).
__global__ __launch_bounds__(256) void split_a_bf16_kernel(int M, int K, const double* __restrict__ A,
                                                           int lda, const int* __restrict__ ea,
                                                           int s, unsigned short* __restrict__ out,
                                                           int kp, size_t plane) {
  __shared__ double tile[64][65];
  const int i0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    int ii = t % 64, kk = t / 64 + 4 * q;
    int gi = i0 + ii, gk = k0 + kk;
    tile[kk][ii] = (gi < M && gk < K) ? A[(size_t)gk * lda + gi] : 0.0;
  }
  __syncthreads();
  const int ii = t / 4, kb = (t % 4) * 16;
  const int gi = i0 + ii;
  if (gi >= M) return;
  const int sh = -ea[gi];
  double r[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) r[e] = ldexp(tile[kb + e][ii], sh);
  for (int p = 1; p <= s; ++p) {
    unsigned w[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      unsigned word = 0;
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        int e = c * 2 + b;
        double xp = ldexp(rint(ldexp(r[e], 8 * p)), -8 * p);
        r[e] -= xp;
        word |= bf16_bits_exact(xp) << (16 * b);
      }
      w[c] = word;
    }
    uint4* dst = (uint4*)(out + (size_t)(p - 1) * plane + (size_t)gi * kp + k0 + kb);
    dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
    dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
  }
}

// H slices: s planes of bf16, column-major kp x N (k contiguous).
template <int KIND>
__global__ __launch_bounds__(256) void split_h_bf16_kernel(int K, int N, GenPtrs g, int r0, int c0,
                                                           const int* __restrict__ eh, int s,
                                                           unsigned short* __restrict__ out, int kp,
                                                           size_t plane) {
  const int t = threadIdx.x;
  const int j = blockIdx.x * 16 + t / 16;
  const int kb = blockIdx.y * 256 + (t % 16) * 16;
  if (j >= N || kb >= kp) return;
  const int sh = -eh[j];
  double r[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    int k = kb + e;
    r[e] = (k < K) ? ldexp(gen_entry<KIND>(g, r0 + k, c0 + j), sh) : 0.0;
  }
  for (int p = 1; p <= s; ++p) {
    unsigned w[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      unsigned word = 0;
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        int e = c * 2 + b;
        double xp = ldexp(rint(ldexp(r[e], 8 * p)), -8 * p);
        r[e] -= xp;
        word |= bf16_bits_exact(xp) << (16 * b);
      }
      w[c] = word;
    }
    uint4* dst = (uint4*)(out + (size_t)(p - 1) * plane + (size_t)j * kp + kb);
    dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
    dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
  }
}

// bf16 slice GEMM: workgroup 64 x 64 (4 waves of 32 x 32), fragments read from
// global memory (L1/L2), v_mfma_f32_16x16x32_bf16.  Lane layout (measured):
// A lane l supplies A[l%16][8*(l/16) .. +7], B lane l supplies B[8*(l/16) .. +7][l%16],
// D lane l element r holds D[4*(l/16)+r][l%16].
template <int S>
__global__ __launch_bounds__(256) void gemm_bf16_slices_kernel(
    int M, int N, int kp, const unsigned short* __restrict__ As,
    const unsigned short* __restrict__ Hs, size_t aplane, size_t hplane,
    const int* __restrict__ ea, const int* __restrict__ eh, double alpha, double beta,
    double* __restrict__ C, int ldc) {
  constexpr int TM = 2, TN = 2;
  const int nbn = (N + 63) / 64, nbm = (M + 63) / 64;
  const int bid = xcd_remap(blockIdx.x, nbn * nbm);
  const int m0 = (bid / nbn) * 64, n0 = (bid % nbn) * 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / 2) * 32, wn0 = (wave % 2) * 32;
  const int lr = lane & 15, lg = lane >> 4;

  double acc64[TM][TN][4];
  v4f acc[S][TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[i][j][r] = 0.0;

  int arow[TM], hcol[TN];
  bool aok[TM], hok[TN];
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    arow[i] = m0 + wm0 + i * 16 + lr;
    aok[i] = arow[i] < M;
  }
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    hcol[j] = n0 + wn0 + j * 16 + lr;
    hok[j] = hcol[j] < N;
  }

  for (int kc = 0; kc < kp; kc += 256) {
#pragma unroll
    for (int d = 0; d < S; ++d)
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[d][i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
    for (int kk = kc; kk < kc + 256; kk += 32) {
      v8bf af[S][TM], bf[S][TN];
#pragma unroll
      for (int p = 0; p < S; ++p) {
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          uint4 u = aok[i] ? *(const uint4*)(As + p * aplane + (size_t)arow[i] * kp + kk + lg * 8)
                           : make_uint4(0, 0, 0, 0);
          af[p][i] = __builtin_bit_cast(v8bf, u);
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          uint4 u = hok[j] ? *(const uint4*)(Hs + p * hplane + (size_t)hcol[j] * kp + kk + lg * 8)
                           : make_uint4(0, 0, 0, 0);
          bf[p][j] = __builtin_bit_cast(v8bf, u);
        }
      }
#pragma unroll
      for (int p = 0; p < S; ++p)
#pragma unroll
        for (int q = 0; q < S - p; ++q)  // 0-based p + q <= S - 1: class d = p + q
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
              acc[p + q][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[p][i], bf[q][j],
                                                                         acc[p + q][i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          double sum = 0.0;
#pragma unroll
          for (int d = S - 1; d >= 0; --d) sum += (double)acc[d][i][j][r];
          acc64[i][j][r] += sum;
        }
  }

#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      int gc = n0 + wn0 + j * 16 + lr;
      if (gc >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int gr = m0 + wm0 + i * 16 + lg * 4 + r;
        if (gr >= M) continue;
        double v = alpha * ldexp(acc64[i][j][r], ea[gr] + eh[gc]);
        double* p = C + (size_t)gc * ldc + gr;
        if (beta != 0.0) v += beta * (*p);
        *p = v;
      }
    }
}

// bf16 slice GEMM with one FP32 accumulator (slices <= 3): the kept slice pairs
// (p, q), p + q <= s - 1, are concatenated along k into one contraction of
// npair * kp bf16 products, lowest class first, and accumulated in FP32 on
// v_mfma_f32_16x16x32_bf16.  Staging is the int8 direct-to-LDS design: 256 x 256
// tiles, eight waves of 128 x 64, k steps of 64 bytes (32 bf16) through a ring
// of kDmaStages LDS stages filled by global_load_lds, the same chunk swizzle.
// The bf16 operand bytes of a lane (A: row l % 16, bf16 k 8 (l / 16) .. +7)
// are the int8 form's (row l % 16, bytes 16 (l / 16) .. +15), so the stage
// layout and the fragment reads carry over unchanged.  Rows past M (columns
// past N) load the last valid row; their outputs are not stored.
// pairs: pair s of the contraction is (A plane (pairs >> 4s) & 3, H plane
// (pairs >> (4s + 2)) & 3).
__global__ __launch_bounds__(512, 1) void gemm_bf16_dma_kernel(
    int M, int N, int kp, const unsigned short* __restrict__ As,
    const unsigned short* __restrict__ Hs, size_t aplane, size_t hplane, int npair, int pairs,
    const int* __restrict__ ea, const int* __restrict__ eh, double alpha, double beta,
    double* __restrict__ C, int ldc) {
  constexpr int BM = 256, BN = 256, BK = 64, NS = kDmaStages;
  constexpr int WGN = BN / 64, NW = 2 * WGN, WM = 128, WN = 64;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int GPWA = 16 / NW, GPWB = (BN / 16) / NW;
  constexpr int SA = BM * BK, SB = BN * BK;
  constexpr int DPS = GPWA + GPWB;
  static_assert(NS * (SA + SB) <= 160 * 1024, "LDS of one CU");
  __shared__ __align__(1024) signed char lds[NS * (SA + SB)];
  signed char(*Ast)[SA] = (signed char(*)[SA])lds;
  signed char(*Bst)[SB] = (signed char(*)[SB])(lds + NS * SA);

  const int nbn = (N + BN - 1) / BN, nbm = (M + BM - 1) / BM;
  int bmi, bni;
  patch_remap<4, 8>(blockIdx.x, nbm, nbn, bmi, bni);
  const int m0 = bmi * BM, n0 = bni * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / WGN) * WM, wn0 = (wave % WGN) * WN;
  const int lr = lane & 15, lg = lane >> 4;
  const unsigned rowb = 2u * (unsigned)kp;  // bytes per row of a plane

  const int rg = lane >> 2;
  const int gch = (lane & 3) ^ dma_swz(rg);
  unsigned aoff[GPWA], boff[GPWB];
#pragma unroll
  for (int q = 0; q < GPWA; ++q) {
    const int r = 16 * (GPWA * wave + q) + rg;
    aoff[q] = (unsigned)min(m0 + r, M - 1) * rowb + 16u * gch;
  }
#pragma unroll
  for (int q = 0; q < GPWB; ++q) {
    const int r = 16 * (GPWB * wave + q) + rg;
    boff[q] = (unsigned)min(n0 + r, N - 1) * rowb + 16u * gch;
  }
  const char* Ab = (const char*)As;
  const char* Hb = (const char*)Hs;
  auto dma16 = [&](const char* gbase, unsigned voff, const signed char* ldst) {
    const unsigned la = __builtin_amdgcn_readfirstlane((unsigned)(size_t)ldst);
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, %2"
                 :
                 : "s"(la), "v"(voff), "s"(gbase)
                 : "memory", "m0");
  };
  // the issue stream walks the k steps of segment after segment; past the last
  // step it re-reads that step into the free stage (results unused)
  const int nks = kp / 32, nkt = npair * nks;
  int iseg = 0, ik = 0;
  auto issue_next = [&](int st) {
    const int pr = (pairs >> (4 * iseg)) & 15;
    const char* ak = Ab + 2 * (size_t)(pr & 3) * aplane + (size_t)ik * BK;
    const char* bk = Hb + 2 * (size_t)(pr >> 2) * hplane + (size_t)ik * BK;
#pragma unroll
    for (int q = 0; q < GPWA; ++q) dma16(ak, aoff[q], &Ast[st][(GPWA * wave + q) * 1024]);
#pragma unroll
    for (int q = 0; q < GPWB; ++q) dma16(bk, boff[q], &Bst[st][(GPWB * wave + q) * 1024]);
    if (++ik == nks) {
      if (iseg + 1 < npair) {
        ++iseg;
        ik = 0;
      } else {
        ik = nks - 1;
      }
    }
  };

  v4f acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};

  const int lofs = lr * 64 + ((lg ^ dma_swz(lr)) * 16);
  v2l af[2][TM], bfd[2][TN];
  auto frags = [&](int kt, int slot) {
    const signed char* sa = Ast[kt % NS];
    const signed char* sb = Bst[kt % NS];
#pragma unroll
    for (int i = 0; i < TM; ++i) af[slot][i] = *(const v2l*)(sa + (wm0 + 16 * i) * 64 + lofs);
#pragma unroll
    for (int j = 0; j < TN; ++j) bfd[slot][j] = *(const v2l*)(sb + (wn0 + 16 * j) * 64 + lofs);
  };

#pragma unroll
  for (int q = 0; q < NS - 1; ++q) issue_next(q);
  wait_vmcnt<(NS - 2) * DPS>();
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
  frags(0, 0);

  // the int8 kernel's five-stage schedule: one barrier per two k steps
  auto step = [&](int kt, auto slotc) {
    constexpr int slot = decltype(slotc)::value;
    static_assert(NS == 5, "barrier schedule for five stages");
    if constexpr (slot == 0) {
      wait_vmcnt<DPS>();
      asm volatile("" ::: "memory");
      __builtin_amdgcn_s_barrier();
      asm volatile("" ::: "memory");
    }
    issue_next((kt + NS - 1) % NS);
    frags(kt + 1, slot ^ 1);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(v8bf, af[slot][i]), __builtin_bit_cast(v8bf, bfd[slot][j]),
            acc[i][j], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
  };
  for (int kt = 0; kt < nkt; kt += 2) {  // nks is a multiple of 8: nkt is even, >= 8
    step(kt, std::integral_constant<int, 0>{});
    step(kt + 1, std::integral_constant<int, 1>{});
  }
  wait_vmcnt<0>();

  // C = alpha * acc * 2^(ea_i + eh_j) + beta * C: four consecutive rows of one
  // column per lane and block (32 contiguous bytes)
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int gc = n0 + wn0 + 16 * j + lr;
    if (gc >= N) continue;
    const int ec = eh[gc];
    double* cc = C + (size_t)gc * ldc;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int gr0 = m0 + wm0 + 16 * i + 4 * lg;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = gr0 + r;
        if (gr >= M) continue;
        double v = alpha * ldexp((double)acc[i][j][r], ea[gr] + ec);
        if (beta != 0.0) v += beta * cc[gr];
        cc[gr] = v;
      }
    }
  }
}

template <int KIND>
static int run_split_bf16(int s, int M, int N, int K, double alpha, const double* A, int lda,
                          const GenPtrs& g, int r0, int c0, double beta, double* C, int ldc,
                          hipStream_t st) {
  const int kp = ((K + 255) / 256) * 256;
  const size_t aplane = (size_t)M * kp, hplane = (size_t)kp * N;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t bytes = al(sizeof(int) * M) + al(sizeof(int) * N) + al(2 * s * aplane) + al(2 * s * hplane);
  std::unique_lock<std::mutex> lk;
  char* ws = (char*)ws_lease(WS_SPLIT, bytes, st, lk);
  if (!ws) return 1002;
  int* ea = (int*)ws;
  int* eh = (int*)(ws + al(sizeof(int) * M));
  unsigned short* as = (unsigned short*)(ws + al(sizeof(int) * M) + al(sizeof(int) * N));
  unsigned short* hs = (unsigned short*)((char*)as + al(2 * s * aplane));
  hipLaunchKernelGGL(fill_int_kernel, dim3((M + 255) / 256), dim3(256), 0, st, ea, M, kNoExp);
  hipLaunchKernelGGL(fill_int_kernel, dim3((N + 255) / 256), dim3(256), 0, st, eh, N, kNoExp);
  const int kchunk = 256;
  hipLaunchKernelGGL(row_exp_kernel, dim3((M + 255) / 256, (K + kchunk - 1) / kchunk), dim3(256),
                     0, st, M, K, A, lda, kchunk, ea);
  hipLaunchKernelGGL(col_exp_kernel<KIND>, dim3((N + 255) / 256, (K + kchunk - 1) / kchunk),
                     dim3(256), 0, st, K, N, g, r0, c0, kchunk, eh);
  hipLaunchKernelGGL(split_a_bf16_kernel, dim3((M + 63) / 64, kp / 64), dim3(256), 0, st, M, K, A,
                     lda, ea, s, as, kp, aplane);
  hipLaunchKernelGGL(split_h_bf16_kernel<KIND>, dim3((N + 15) / 16, kp / 256), dim3(256), 0, st, K,
                     N, g, r0, c0, eh, s, hs, kp, hplane);
  if (s <= kBf16Fp32AccSlices && 2 * (size_t)M * kp < (1ull << 32) &&
      2 * (size_t)N * kp < (1ull << 32)) {
    // kept pairs p + q <= s - 1, lowest class first
    int npair = 0, pairs = 0;
    for (int d = s - 1; d >= 0; --d)
      for (int p = 0; p <= d; ++p) pairs |= (p | ((d - p) << 2)) << (4 * npair++);
    const int nb256 = ((M + 255) / 256) * ((N + 255) / 256);
    hipLaunchKernelGGL(gemm_bf16_dma_kernel, dim3(nb256), dim3(512), 0, st, M, N, kp, as, hs,
                       aplane, hplane, npair, pairs, ea, eh, alpha, beta, C, ldc);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : 1000 + (int)e;
  }
  int nb = ((M + 63) / 64) * ((N + 63) / 64);
#define PSMMA_BF16_CASE(SS)                                                                \
  case SS:                                                                                    \
    hipLaunchKernelGGL(gemm_bf16_slices_kernel<SS>, dim3(nb), dim3(256), 0, st, M, N, kp, as, \
                       hs, aplane, hplane, ea, eh, alpha, beta, C, ldc);                     \
    break;
  switch (s) {
    PSMMA_BF16_CASE(1)
    PSMMA_BF16_CASE(2)
    PSMMA_BF16_CASE(3)
    PSMMA_BF16_CASE(4)
    PSMMA_BF16_CASE(5)
    PSMMA_BF16_CASE(6)
    PSMMA_BF16_CASE(7)
    default: return -16;
  }
#undef PSMMA_BF16_CASE
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : 1000 + (int)e;
}

static int launch_right_split_bf16(int kind, int s, int M, int N, int K, double alpha,
                                   const double* A, int lda, const GenPtrs& g, int r0, int c0,
                                   double beta, double* C, int ldc, hipStream_t st) {
  if (s < 1 || s > 7) return -16;
  switch (kind) {
    case KIND_CAUCHY: return run_split_bf16<KIND_CAUCHY>(s, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc, st);
    case KIND_TOEPLITZ: return run_split_bf16<KIND_TOEPLITZ>(s, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc, st);
    case KIND_HANKEL: return run_split_bf16<KIND_HANKEL>(s, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc, st);
    case KIND_CAUCHYEIG: return run_split_bf16<KIND_CAUCHYEIG>(s, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc, st);
  }
  return -1;
}

// ---------------------------------------------------------------------------
// Grouped int8 modular pipeline for a batch of problems: every phase is one
// launch over the work units of all problems (a unit is one workgroup of the
// single-problem kernel); a workgroup finds its problem by binary search over
// the phase's unit prefix sums.  All problems use the same moduli (the batch's
// CrtParams); each keeps its own power-of-two scales and abits.
enum {
  PH_FILL = 0, PH_ROWEXP, PH_COLEXP, PH_EHFILL, PH_SPLITA, PH_SPLITB,
  PH_GEMM_D, PH_GEMM_T, PH_GEMM_H,     // 128 x 128 tiles
  PH_GEMM_DB, PH_GEMM_TB, PH_GEMM_HB,  // 256 x 256 tiles (large problems)
  PH_CRT,
  PH_FUSED,  // the fused form: 128 x 128 tiles, all moduli and the recombination
  NPH
};

struct SplitProblem {
  int kind, bmode, M, N, K, kp, lpad, abits, hbits;
  double alpha, beta;
  const double* A;
  int lda;
  GenPtrs g;
  int r0, c0;
  double* C;
  int ldc;
  long long glo, ghi;
  int *ea, *eh, *e1;
  signed char *ares, *bres;
  unsigned char* ores;
  size_t aplane, bplane, oplane;
  int ldo;  // residue leading dimension: M rounded up to 16
  int gx[NPH];  // x-extent of the phase's unit grid
  int ub[NPH];  // first unit of this problem in the phase
};

// The problem owning unit u of a phase and the unit's index within that
// problem: one table entry per unit, filled on the host when the batch is
// bound, so a workgroup finds its problem with a single load (a search over
// the problem list costs a chain of dependent loads per workgroup, comparable
// to the whole work of a small unit).
using UnitOwner = int2;

// One kernel per phase: each carries only its own LDS and registers.
template <int ph>
__global__ __launch_bounds__(256) void gsplit_phase_kernel(const SplitProblem* __restrict__ P,
                                                           const UnitOwner* __restrict__ own,
                                                           CrtParams cp) {
  const UnitOwner o = own[blockIdx.x];
  const SplitProblem& q = P[o.x];
  const int lu = o.y;
  const int bx = lu % q.gx[ph], by = lu / q.gx[ph];
  const bool dense = q.bmode == BMODE_DENSE;
  (void)dense;
  switch (ph) {
    case PH_FILL: {
      int na = (q.M + 255) / 256, nh = (q.N + 255) / 256;
      if (lu < na) fill_int_kernel_body(q.ea, q.M, kNoExp, lu, 0);
      else if (lu < na + nh) fill_int_kernel_body(q.eh, q.N, kNoExp, lu - na, 0);
      else fill_int_kernel_body(q.e1, 1, kNoExp, 0, 0);
      break;
    }
    case PH_ROWEXP:
      row_exp_kernel_body(q.M, q.K, q.A, q.lda, 256, q.ea, bx, by);
      break;
    case PH_COLEXP:
      if (!dense) {
        gen_range_exp_kernel_body(q.g.p0, q.glo, q.ghi, q.e1, bx, 0);
      } else if (q.kind == KIND_CAUCHY) {
        col_exp_kernel_body<KIND_CAUCHY>(q.K, q.N, q.g, q.r0, q.c0, 256, q.eh, bx, by);
      } else if (q.kind == KIND_TOEPLITZ) {
        col_exp_kernel_body<KIND_TOEPLITZ>(q.K, q.N, q.g, q.r0, q.c0, 256, q.eh, bx, by);
      } else if (q.kind == KIND_HANKEL) {
        col_exp_kernel_body<KIND_HANKEL>(q.K, q.N, q.g, q.r0, q.c0, 256, q.eh, bx, by);
      } else {
        col_exp_kernel_body<KIND_CAUCHYEIG>(q.K, q.N, q.g, q.r0, q.c0, 256, q.eh, bx, by);
      }
      break;
    case PH_EHFILL:
      fill_from_kernel_body(q.eh, q.N, q.e1, bx, 0);
      break;
    case PH_SPLITA:
      split_a_kernel_body(q.M, q.K, q.A, q.lda, q.ea, q.abits, cp, q.ares, q.kp, q.aplane, bx, by);
      break;
    case PH_SPLITB:
      if (!dense) {
        gen_residue_kernel_body(q.g.p0, q.glo, q.ghi, q.bmode == BMODE_TOE ? 1 : 0, q.e1, q.hbits,
                                cp, q.bres, q.bplane, q.lpad, bx, 0);
      } else if (q.kind == KIND_CAUCHY) {
        split_h_kernel_body<KIND_CAUCHY>(q.K, q.N, q.g, q.r0, q.c0, q.eh, q.hbits, cp, q.bres,
                                         q.kp, q.bplane, bx, by);
      } else if (q.kind == KIND_TOEPLITZ) {
        split_h_kernel_body<KIND_TOEPLITZ>(q.K, q.N, q.g, q.r0, q.c0, q.eh, q.hbits, cp, q.bres,
                                           q.kp, q.bplane, bx, by);
      } else if (q.kind == KIND_HANKEL) {
        split_h_kernel_body<KIND_HANKEL>(q.K, q.N, q.g, q.r0, q.c0, q.eh, q.hbits, cp, q.bres,
                                         q.kp, q.bplane, bx, by);
      } else {
        split_h_kernel_body<KIND_CAUCHYEIG>(q.K, q.N, q.g, q.r0, q.c0, q.eh, q.hbits, cp, q.bres,
                                            q.kp, q.bplane, bx, by);
      }
      break;
  }
}

constexpr int kBatchKmax = 16384;  // longest contraction of a generator-window problem in a batch

template <int BMODE>
__global__ __launch_bounds__(256, 2) void gsplit_gemm_kernel(const SplitProblem* __restrict__ P,
                                                             const UnitOwner* __restrict__ own,
                                                             CrtParams cp) {
  const UnitOwner o = own[blockIdx.x];
  const SplitProblem& q = P[o.x];
  const int lu = o.y;
  const int ntn = (q.N + 127) / 128, ntm = (q.M + 127) / 128;
  const int l = lu / (ntm * ntn), r = lu % (ntm * ntn);
  gemm_i8_body<BMODE, 128, 128, 128, 2, 2, kBatchKmax>(q.M, q.N, q.kp, q.ares, q.bres, q.aplane,
                                                      q.bplane, q.lpad, cp, q.ores, q.oplane,
                                                      q.ldo, r / ntn, r % ntn, l);
}

// Large problems use the 256 x 256 direct-to-LDS tiles (512 threads).
constexpr int kBigTile = 512;  // smallest min(M, N) and K of the large class (measured best)

template <int BMODE>
__global__ __launch_bounds__(512, 1) void gsplit_gemm_big_kernel(const SplitProblem* __restrict__ P,
                                                                 const UnitOwner* __restrict__ own,
                                                                 CrtParams cp) {
  const UnitOwner o = own[blockIdx.x];
  const SplitProblem& q = P[o.x];
  const int lu = o.y;
  const int ntn = (q.N + 255) / 256, ntm = (q.M + 255) / 256;
  const int l = lu / (ntm * ntn), r = lu % (ntm * ntn);
  gemm_i8_dma_body<BMODE, kBatchKmax, 256>(q.M, q.N, q.kp, q.ares, q.bres, q.aplane, q.bplane,
                                           q.lpad, cp, q.ores, q.oplane, q.ldo, r / ntn, r % ntn, l);
}

// A grouped CRT unit is 256 threads x kCrtQuads groups of four rows.
constexpr int kCrtQuads = 4;
template <int T>
__global__ __launch_bounds__(256) void gsplit_crt_kernel(const SplitProblem* __restrict__ P,
                                                         const UnitOwner* __restrict__ own,
                                                         CrtParams cp) {
  const UnitOwner o = own[blockIdx.x];
  const SplitProblem& q = P[o.x];
  const size_t e0 = (size_t)o.y * (256 * kCrtQuads) + threadIdx.x;
#pragma unroll 1
  for (int h = 0; h < kCrtQuads; ++h)
    crt_unit<T>(q.M, q.N, q.ores, q.oplane, q.ldo, cp, q.ea, q.eh, q.abits, q.hbits, q.alpha,
                q.beta, q.C, q.ldc, e0 + (size_t)h * 256);
}

// ---------------------------------------------------------------------------
// int8 modular GEMM with the recombination in registers (fused form).
//
// One workgroup owns one 128 x 128 output tile and runs the t residue GEMMs of
// that tile back to back as one k loop of t * kp / 64 steps (the DMA ring runs
// on across the modulus boundaries).  When modulus l is done, every int32
// accumulator x is reduced to its symmetric residue r (|r| <= 128) and folded
// into a 64-bit fixed-point fraction
//     S += (r + 128) * W_l  (mod 2^64),   W_l = round(2^65 y_l / m_l),
// y_l = (M / m_l)^-1 mod m_l (M = prod m_l).  By the Chinese remainder theorem
// X / M = sum_l r_l y_l / m_l (mod 1), so after the last modulus
//     S - sum_l 128 W_l = 2^65 X / M  (mod 2^64),
// and that is X / M itself: the operand bits keep |X| <= K 2^(a + h) < M / 4, so
// |2^65 X / M| < 2^63 fits a signed 64-bit integer.  The rounding of the W_l
// adds at most t * 128 * 1/2 units of 2^-65 to X / M: |dX| <= t 2^-59 M, about
// 2^-45 of a typical |X| at K = 2048, below the 2^-abits operand rounding.
// C = alpha * X * 2^-(sa_i + sh_j) + beta * C is then written directly: no
// residue planes of the result and no recombination pass.
//
// Four waves of 64 x 64 (16 accumulators and 32 fraction words each) and two
// workgroups per CU, so one workgroup's residue folds and epilogue overlap the
// other's MFMAs.  k steps of 64 bytes go through a five-stage ring filled by
// global_load_lds three steps ahead (one spare stage: the fragments of a step
// are read at the step itself, and one barrier per two steps keeps the ring
// safe).  Dense B stages 128 columns x 64 bytes like A; the generator-window
// kinds stage the t windows of the tile once, at the start, in 16-byte chunks.
struct FusedCrt {
  int t;
  int m[kMaxModuli];
  float rmf[kMaxModuli];             // fl(1 / m_l), for mod_acc
  unsigned long long w[kMaxModuli];  // round(2^65 y_l / m_l) mod 2^64
  unsigned long long k0;             // sum_l 128 w_l mod 2^64 (the residue offset)
  double scale;                      // M * 2^-65
};

static FusedCrt make_fused(int t) {
  FusedCrt f{};
  f.t = t;
  unsigned __int128 M = 1;
  for (int l = 0; l < t; ++l) M *= (unsigned)kModuli[l];
  for (int l = 0; l < t; ++l) {
    const unsigned m = (unsigned)kModuli[l];
    const long long y = inv_mod((long long)((M / m) % m), m);
    const unsigned __int128 num = ((unsigned __int128)y << 65) + m / 2;
    f.m[l] = (int)m;
    f.rmf[l] = 1.0f / (float)m;
    f.w[l] = (unsigned long long)(num / m);  // the quotient's low 64 bits
    f.k0 += 128ull * f.w[l];
  }
  const unsigned long long top = (unsigned long long)(M >> 64), bot = (unsigned long long)M;
  const long double Ml = (long double)top * 18446744073709551616.0L + (long double)bot;
  f.scale = (double)ldexpl(Ml, -65);
  return f;
}

constexpr int kFusedWinKmax = 2048;  // longest padded contraction of the fused window form
constexpr int kFusedWinStride = kFusedWinKmax + 128 + 32;  // LDS bytes per modulus window
constexpr int kFusedLds = 5 * 128 * 64 * 2;  // five stages of A and B (dense form)
static_assert(5 * 128 * 64 + kMaxModuli * kFusedWinStride <= kFusedLds, "window form fits");

template <int BMODE>
__device__ __forceinline__ void gemm_i8_fused_body(
    signed char* __restrict__ lds, int M, int N, int kp, int lpad,
    const signed char* __restrict__ Ares, const signed char* __restrict__ Bres, size_t aplane,
    size_t bplane, const FusedCrt& fc, const int* __restrict__ ea, const int* __restrict__ eh,
    int abits, int hbits, double alpha, double beta, double* __restrict__ C, int ldc, int bmi,
    int bni) {
  constexpr bool WIN = BMODE != BMODE_DENSE;
  constexpr int BM = 128, BN = 128, BK = 64, NS = 5, D = 3;
  constexpr int WM = 64, WN = 64, TM = 4, TN = 4, NT = 256;
  constexpr int GPW = 2;  // 16-row DMA groups per wave and operand
  constexpr int SA = BM * BK, SB = BN * BK;
  constexpr int DPS = WIN ? GPW : 2 * GPW;  // DMA instructions per wave per k step
  signed char(*Ast)[SA] = (signed char(*)[SA])lds;
  signed char(*Bst)[SB] = (signed char(*)[SB])(lds + NS * SA);
  unsigned* Win = (unsigned*)(lds + NS * SA);

  const int t = fc.t;
  const int m0 = bmi * BM, n0 = bni * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
  const int lr = lane & 15, lg = lane >> 4;

  // window form: window byte x of modulus l holds R_l[ws + x], ws = the tile's
  // first generator position rounded down to 16 (chunks of 16 bytes are either
  // inside [0, lpad) or outside it: lpad is a multiple of 16)
  int cj[TN];
  if constexpr (WIN) {
    const int ubase = (BMODE == BMODE_TOE) ? (N - n0 - BN) : n0;
    const int ws = (ubase >> 4) << 4;
    const int nch = (kp + BN + 32) / 16;
    for (int x = tid; x < t * nch; x += NT) {
      const int l = x / nch, c = x - l * nch;
      const int u = ws + 16 * c;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (u >= 0 && u < lpad) v = *(const uint4*)(Bres + l * bplane + u);
      *(uint4*)((char*)Win + l * kFusedWinStride + 16 * c) = v;
    }
#pragma unroll
    for (int j = 0; j < TN; ++j)
      cj[j] = ubase - ws +
              ((BMODE == BMODE_TOE) ? (BN - 1 - wn0 - 16 * j - lr) : (wn0 + 16 * j + lr));
  }

  const int rg = lane >> 2;
  const int gch = (lane & 3) ^ dma_swz(rg);
  unsigned aoff[GPW], boff[GPW];
#pragma unroll
  for (int q = 0; q < GPW; ++q) {
    const int r = 16 * (GPW * wave + q) + rg;
    aoff[q] = (unsigned)min(m0 + r, M - 1) * (unsigned)kp + 16u * gch;
    boff[q] = WIN ? 0u : (unsigned)min(n0 + r, N - 1) * (unsigned)kp + 16u * gch;
  }
  auto dma16 = [&](const signed char* gbase, unsigned voff, const signed char* ldst) {
    const unsigned la = __builtin_amdgcn_readfirstlane((unsigned)(size_t)ldst);
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, %2"
                 :
                 : "s"(la), "v"(voff), "s"(gbase)
                 : "memory", "m0");
  };
  // the issue stream walks modulus after modulus; past the last step it
  // re-reads that step into the spare stage (never read)
  // (running source pointers and uniform LDS offsets keep the per-step
  // address work in a few scalar instructions)
  const int nks = kp / BK;
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  int il = 0, ik = 0;
  const signed char* ap = Ares;
  const signed char* bp = Bres;
  auto issue_next = [&](int st) {
#pragma unroll
    for (int q = 0; q < GPW; ++q) dma16(ap, aoff[q], &Ast[st][(GPW * wv + q) * 1024]);
    if constexpr (!WIN) {
#pragma unroll
      for (int q = 0; q < GPW; ++q) dma16(bp, boff[q], &Bst[st][(GPW * wv + q) * 1024]);
    }
    if (++ik < nks) {
      ap += BK;
      bp += BK;
    } else if (il + 1 < t) {  // first step of the next modulus's planes
      ++il;
      ik = 0;
      ap += (ptrdiff_t)aplane - (kp - BK);
      bp += (ptrdiff_t)bplane - (kp - BK);
    } else {
      ik = nks - 1;  // past the end: the last step again
    }
  };

  v4i acc[TM][TN];
  unsigned long long S[TM][TN][4];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      acc[i][j] = (v4i){0, 0, 0, 0};
#pragma unroll
      for (int r = 0; r < 4; ++r) S[i][j][r] = 0ull;
    }

  const int lofs = lr * 64 + ((lg ^ dma_swz(lr)) * 16);
  // one k step: fragments of stage st (k step kk of modulus l), then 16 MFMAs
  // (A fragments are read one row block at a time, just before their MFMAs:
  // the 64 fraction words leave no room for all of them at once)
  auto mfma_step = [&](int st, int kk, int l) {
    v2l bf[TN];
    if constexpr (WIN) {
      const unsigned* wl = Win + l * (kFusedWinStride / 4);
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int x0 = kk * BK + lg * 16 + cj[j];
        const unsigned* wp = wl + (x0 >> 2);
        const int sh = x0 & 3;
        unsigned w[5];
#pragma unroll
        for (int d = 0; d < 5; ++d) w[d] = wp[d];
        const unsigned o0 = __builtin_amdgcn_alignbyte(w[1], w[0], sh);
        const unsigned o1 = __builtin_amdgcn_alignbyte(w[2], w[1], sh);
        const unsigned o2 = __builtin_amdgcn_alignbyte(w[3], w[2], sh);
        const unsigned o3 = __builtin_amdgcn_alignbyte(w[4], w[3], sh);
        bf[j][0] = (long)(((unsigned long)o1 << 32) | o0);
        bf[j][1] = (long)(((unsigned long)o3 << 32) | o2);
      }
    } else {
#pragma unroll
      for (int j = 0; j < TN; ++j) bf[j] = *(const v2l*)(Bst[st] + (wn0 + 16 * j) * 64 + lofs);
    }
    // the next row block's fragment is read while this one's MFMAs run
    v2l af = *(const v2l*)(Ast[st] + wm0 * 64 + lofs);
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      v2l an = af;
      if (i + 1 < TM) an = *(const v2l*)(Ast[st] + (wm0 + 16 * (i + 1)) * 64 + lofs);
#pragma unroll
      for (int j = 0; j < TN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af, bf[j], acc[i][j], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
      af = an;
    }
  };
  // end of modulus l: accumulators -> symmetric residues -> fraction words
  // (mod_acc: kp <= kFastModK keeps |x| <= 2^28; the host routes longer
  // contractions elsewhere)
  auto fold = [&](int l) {
    const int m = fc.m[l];
    const float rmf = fc.rmf[l];
    const unsigned long long w = fc.w[l];
    const unsigned hm = (unsigned)(m + 1) >> 1;
    const unsigned neg = 128u - (unsigned)m;  // r - m + 128 (mod 2^32)
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const unsigned rr = mod_acc(acc[i][j][r], m, rmf);
          const unsigned u = rr + (rr >= hm ? neg : 128u);  // in [0, 256)
          S[i][j][r] += (unsigned long long)u * w;
        }
        acc[i][j] = (v4i){0, 0, 0, 0};
      }
  };

  if constexpr (WIN) __syncthreads();  // the windows were written with LDS stores
#pragma unroll
  for (int q = 0; q < D; ++q) issue_next(q);
  int sc = 0;  // stage of the current step: kt % NS
  auto next_stage = [](int s) { return s == NS - 1 ? 0 : s + 1; };
  for (int l = 0; l < t; ++l) {
    for (int kk = 0; kk < nks; kk += 2) {  // nks is even (kp a multiple of 128)
      // steps kt and kt+1 have landed once only step kt+2 may be in flight;
      // after the barrier every wave has read the stages of steps <= kt-1,
      // which the DMA of steps kt+3 and kt+4 overwrite
      wait_vmcnt<DPS>();
      asm volatile("" ::: "memory");
      __builtin_amdgcn_s_barrier();
      asm volatile("" ::: "memory");
      const int s1 = next_stage(sc);
      const int s3 = next_stage(next_stage(s1));
      issue_next(s3);
      mfma_step(sc, kk, l);
      issue_next(next_stage(s3));
      mfma_step(s1, kk + 1, l);
      sc = next_stage(s1);
    }
    fold(l);
  }
  wait_vmcnt<0>();  // no DMA outlives the loop

  // C = alpha * X * 2^-(sa_i + sh_j) + beta * C, X = (S - k0) * M * 2^-65
  const unsigned long long k0 = fc.k0;
  const double scale = fc.scale;
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int gc = n0 + wn0 + 16 * j + lr;
    if (gc >= N) continue;
    const int sj = hbits - eh[gc];
    double* cc = C + (size_t)gc * ldc;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int gr0 = m0 + wm0 + 16 * i + 4 * lg;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = gr0 + r;
        if (gr >= M) continue;
        const double x = (double)(long long)(S[i][j][r] - k0) * scale;
        double v = alpha * ldexp(x, -(abits - ea[gr] + sj));
        if (beta != 0.0) v += beta * cc[gr];
        cc[gr] = v;
      }
    }
  }
}

// Grouped fused GEMM: one launch over the 128 x 128 tiles of every problem,
// all kinds.  The host lists each problem's tiles in kFusedPatch^2 patches;
// each run of that many consecutive units goes to one XCD, so the workgroups
// of a patch run there together and share its row and column panels in L2.
constexpr int kFusedPatch = 4;
__global__ __launch_bounds__(256, 2) void gsplit_fused_kernel(const SplitProblem* __restrict__ P,
                                                              const UnitOwner* __restrict__ own,
                                                              FusedCrt fc, int nunits) {
  __shared__ __align__(1024) signed char lds[kFusedLds];
  constexpr int NXCD = 8, G = kFusedPatch * kFusedPatch;
  int u = blockIdx.x;
  const int full = nunits / (NXCD * G) * (NXCD * G);
  if (u < full) u = (u / (NXCD * G)) * (NXCD * G) + (u % NXCD) * G + (u / NXCD) % G;
  const UnitOwner o = own[u];
  const SplitProblem& q = P[o.x];
  const int ntn = (q.N + 127) / 128;
  const int bmi = o.y / ntn, bni = o.y - bmi * ntn;
  if (q.bmode == BMODE_TOE)
    gemm_i8_fused_body<BMODE_TOE>(lds, q.M, q.N, q.kp, q.lpad, q.ares, q.bres, q.aplane, q.bplane,
                                  fc, q.ea, q.eh, q.abits, q.hbits, q.alpha, q.beta, q.C, q.ldc,
                                  bmi, bni);
  else if (q.bmode == BMODE_HANK)
    gemm_i8_fused_body<BMODE_HANK>(lds, q.M, q.N, q.kp, q.lpad, q.ares, q.bres, q.aplane,
                                   q.bplane, fc, q.ea, q.eh, q.abits, q.hbits, q.alpha, q.beta,
                                   q.C, q.ldc, bmi, bni);
  else
    gemm_i8_fused_body<BMODE_DENSE>(lds, q.M, q.N, q.kp, q.lpad, q.ares, q.bres, q.aplane,
                                    q.bplane, fc, q.ea, q.eh, q.abits, q.hbits, q.alpha, q.beta,
                                    q.C, q.ldc, bmi, bni);
}

// Host plan of a grouped split batch (table, unit totals, workspace size).
struct SplitBatch {
  int t = 0;
  bool fused = false;  // the fused GEMM + recombination form
  CrtParams cp{};
  FusedCrt fc{};
  std::vector<SplitProblem> tab;
  int total[NPH] = {0};
  size_t ws_bytes = 0;
};

// The fused form addresses residue planes with 32-bit offsets: every problem's
// M * kp (and N * kp where B is a dense plane) must stay below 2^32; its
// residue reduction needs kp <= kFastModK.
static bool fused_fits(const BatchProblem& b) {
  if (b.m == 0 || b.n == 0 || b.k == 0) return true;
  const size_t kp = ((size_t)b.k + 255) / 256 * 256;
  if (kp > (size_t)kFastModK) return false;
  const bool win = (b.kind == KIND_TOEPLITZ || b.kind == KIND_HANKEL) && kp <= kFusedWinKmax;
  return (size_t)b.m * kp < (1ull << 32) && (win || (size_t)b.n * kp < (1ull << 32));
}

// Lays out the batch in a workspace starting at `ws` (nullptr: sizing only).
// Problems must be validated; empty ones are skipped.  sb.fused selects the
// fused form (set by the caller before the first layout).
void split_batch_layout(const BatchProblem* probs, int nprob, int t, char* ws, SplitBatch& sb) {
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  sb.t = t;
  sb.cp = make_crt(t);
  sb.fc = make_fused(t);
  sb.tab.clear();
  for (int i = 0; i < NPH; ++i) sb.total[i] = 0;
  size_t off = 0;
  std::vector<int> order(nprob);
  for (int i = 0; i < nprob; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return probs[x].k > probs[y].k; });
  for (int oi = 0; oi < nprob; ++oi) {
    const BatchProblem& b = probs[order[oi]];
    if (b.m == 0 || b.n == 0 || b.k == 0) continue;
    SplitProblem q{};
    q.kind = b.kind;
    q.bmode = (b.kind == KIND_TOEPLITZ) ? BMODE_TOE : (b.kind == KIND_HANKEL) ? BMODE_HANK : BMODE_DENSE;
    q.M = b.m; q.N = b.n; q.K = b.k;
    // the fused GEMM steps in pairs of 64-deep k steps; the residue-plane
    // GEMMs need a multiple of 256
    const int kq = sb.fused ? 128 : 256;
    q.kp = ((b.k + kq - 1) / kq) * kq;
    if (q.bmode != BMODE_DENSE && q.kp > (sb.fused ? kFusedWinKmax : kBatchKmax))
      q.bmode = BMODE_DENSE;
    {
      const OpBits bits = choose_bits(t, b.k, b.kind);
      q.abits = bits.a;
      q.hbits = bits.h;
    }
    q.alpha = b.alpha; q.beta = b.beta; q.A = b.A; q.lda = b.lda; q.g = b.g;
    q.r0 = b.r0; q.c0 = b.c0; q.C = b.C; q.ldc = b.ldc;
    const bool toe = b.kind == KIND_TOEPLITZ;
    q.glo = toe ? (long long)b.g.korder - 1 + b.c0 - (b.r0 + b.k - 1) : (long long)b.r0 + b.c0;
    q.ghi = toe ? (long long)b.g.korder - 1 + (b.c0 + b.n - 1) - b.r0
                : (long long)b.r0 + b.k - 1 + b.c0 + b.n - 1;
    long long L = q.ghi - q.glo + 1;
    q.lpad = (int)((L + 15) / 16 * 16);
    q.aplane = (size_t)q.M * q.kp;
    q.bplane = (q.bmode == BMODE_DENSE) ? (size_t)q.kp * q.N : (size_t)q.lpad;
    q.ldo = (q.M + 15) & ~15;
    q.oplane = (size_t)q.ldo * q.N;
    auto take = [&](size_t bytes) {
      char* p = ws ? ws + off : nullptr;
      off += al(bytes);
      return p;
    };
    q.ea = (int*)take(sizeof(int) * q.M);
    q.eh = (int*)take(sizeof(int) * q.N);
    q.e1 = (int*)take(sizeof(int));
    q.ares = (signed char*)take(t * q.aplane);
    q.bres = (signed char*)take(t * q.bplane);
    q.ores = sb.fused ? nullptr : (unsigned char*)take(t * q.oplane);
    const bool dense = q.bmode == BMODE_DENSE;
    const int kc = (q.K + 255) / 256;
    int gx[NPH], gy[NPH];
    for (int ph = 0; ph < NPH; ++ph) gx[ph] = 1, gy[ph] = 0;
    gx[PH_FILL] = (q.M + 255) / 256 + (q.N + 255) / 256 + 1; gy[PH_FILL] = 1;
    gx[PH_ROWEXP] = (q.M + 255) / 256; gy[PH_ROWEXP] = kc;
    if (dense) { gx[PH_COLEXP] = (q.N + 255) / 256; gy[PH_COLEXP] = kc; }
    else { gx[PH_COLEXP] = (int)((L + 255) / 256); gy[PH_COLEXP] = 1; }
    if (!dense) { gx[PH_EHFILL] = (q.N + 255) / 256; gy[PH_EHFILL] = 1; }
    gx[PH_SPLITA] = (q.M + 63) / 64; gy[PH_SPLITA] = (q.kp + kSplitAK - 1) / kSplitAK;
    if (dense) { gx[PH_SPLITB] = (q.N + 15) / 16; gy[PH_SPLITB] = (q.kp + 255) / 256; }
    else { gx[PH_SPLITB] = (q.lpad + 255) / 256; gy[PH_SPLITB] = 1; }
    // the large class: the direct-to-LDS tiles, whose residue planes must be
    // addressable with 32-bit offsets
    const bool big = std::min(q.M, q.N) >= kBigTile && q.K >= kBigTile &&
                     (size_t)q.M * q.kp < (1ull << 32) &&
                     (q.bmode != BMODE_DENSE || (size_t)q.N * q.kp < (1ull << 32));
    if (sb.fused) {
      gx[PH_FUSED] = ((q.M + 127) / 128) * ((q.N + 127) / 128); gy[PH_FUSED] = 1;
    } else {
      int gph = dense ? PH_GEMM_D : (q.bmode == BMODE_TOE ? PH_GEMM_T : PH_GEMM_H);
      if (big) gph += PH_GEMM_DB - PH_GEMM_D;
      const int tile = big ? 256 : 128;
      gx[gph] = ((q.M + tile - 1) / tile) * ((q.N + tile - 1) / tile) * t; gy[gph] = 1;
      gx[PH_CRT] = (int)(((size_t)q.ldo * q.N / 4 + 256 * kCrtQuads - 1) / (256 * kCrtQuads));
      gy[PH_CRT] = 1;
    }
    for (int ph = 0; ph < NPH; ++ph) {
      q.gx[ph] = gx[ph];
      q.ub[ph] = sb.total[ph];
      sb.total[ph] += gx[ph] * gy[ph];
    }
    sb.tab.push_back(q);
  }
  sb.ws_bytes = off;
}

// Device layout of a bound batch: the problem table, then the unit owners of
// every phase in phase order (UnitOwner is 8 bytes and SplitProblem a multiple
// of 8 bytes, so the owners are aligned).
static_assert(sizeof(SplitProblem) % 8 == 0, "owner table alignment");
size_t owner_count(const SplitBatch& sb) {
  size_t n = 0;
  for (int ph = 0; ph < NPH; ++ph) n += (size_t)sb.total[ph];
  return n;
}
const UnitOwner* phase_owners(const SplitBatch& sb, const SplitProblem* dtab, int ph) {
  size_t off = 0;
  for (int p = 0; p < ph; ++p) off += (size_t)sb.total[p];
  return (const UnitOwner*)(dtab + sb.tab.size()) + off;
}

// Enqueues the grouped pipeline; dtab is the device copy of the bound layout.
int split_batch_enqueue(const SplitBatch& sb, const SplitProblem* dtab, hipStream_t s, int* nodes) {
  const int np = (int)sb.tab.size();
  int n = 0;
  if (np == 0) {
    if (nodes) *nodes = 0;
    return 0;
  }
  for (int ph : {PH_FILL, PH_ROWEXP, PH_COLEXP, PH_EHFILL, PH_SPLITA, PH_SPLITB}) {
    if (sb.total[ph] == 0) continue;
    auto kern = ph == PH_FILL     ? gsplit_phase_kernel<PH_FILL>
                : ph == PH_ROWEXP ? gsplit_phase_kernel<PH_ROWEXP>
                : ph == PH_COLEXP ? gsplit_phase_kernel<PH_COLEXP>
                : ph == PH_EHFILL ? gsplit_phase_kernel<PH_EHFILL>
                : ph == PH_SPLITA ? gsplit_phase_kernel<PH_SPLITA>
                                  : gsplit_phase_kernel<PH_SPLITB>;
    hipLaunchKernelGGL(kern, dim3(sb.total[ph]), dim3(256), 0, s, dtab,
                       phase_owners(sb, dtab, ph), sb.cp);
    ++n;
  }
  if (sb.fused) {
    if (sb.total[PH_FUSED]) {
      hipLaunchKernelGGL(gsplit_fused_kernel, dim3(sb.total[PH_FUSED]), dim3(256), 0, s, dtab,
                         phase_owners(sb, dtab, PH_FUSED), sb.fc, sb.total[PH_FUSED]);
      ++n;
    }
    if (nodes) *nodes = n;
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : 1000 + (int)e;
  }
  if (sb.total[PH_GEMM_D]) {
    hipLaunchKernelGGL(gsplit_gemm_kernel<BMODE_DENSE>, dim3(sb.total[PH_GEMM_D]), dim3(256), 0, s,
                       dtab, phase_owners(sb, dtab, PH_GEMM_D), sb.cp);
    ++n;
  }
  if (sb.total[PH_GEMM_T]) {
    hipLaunchKernelGGL(gsplit_gemm_kernel<BMODE_TOE>, dim3(sb.total[PH_GEMM_T]), dim3(256), 0, s,
                       dtab, phase_owners(sb, dtab, PH_GEMM_T), sb.cp);
    ++n;
  }
  if (sb.total[PH_GEMM_H]) {
    hipLaunchKernelGGL(gsplit_gemm_kernel<BMODE_HANK>, dim3(sb.total[PH_GEMM_H]), dim3(256), 0, s,
                       dtab, phase_owners(sb, dtab, PH_GEMM_H), sb.cp);
    ++n;
  }
  if (sb.total[PH_GEMM_DB]) {
    hipLaunchKernelGGL(gsplit_gemm_big_kernel<BMODE_DENSE>, dim3(sb.total[PH_GEMM_DB]), dim3(512),
                       0, s, dtab, phase_owners(sb, dtab, PH_GEMM_DB), sb.cp);
    ++n;
  }
  if (sb.total[PH_GEMM_TB]) {
    hipLaunchKernelGGL(gsplit_gemm_big_kernel<BMODE_TOE>, dim3(sb.total[PH_GEMM_TB]), dim3(512), 0,
                       s, dtab, phase_owners(sb, dtab, PH_GEMM_TB), sb.cp);
    ++n;
  }
  if (sb.total[PH_GEMM_HB]) {
    hipLaunchKernelGGL(gsplit_gemm_big_kernel<BMODE_HANK>, dim3(sb.total[PH_GEMM_HB]), dim3(512), 0,
                       s, dtab, phase_owners(sb, dtab, PH_GEMM_HB), sb.cp);
    ++n;
  }
  dim3 cg(sb.total[PH_CRT]);
  const UnitOwner* crt_own = phase_owners(sb, dtab, PH_CRT);
#define PSMMA_GCRT_CASE(TT)                                                              \
  case TT:                                                                               \
    hipLaunchKernelGGL(gsplit_crt_kernel<TT>, cg, dim3(256), 0, s, dtab, crt_own, sb.cp); \
    break;
  switch (sb.t) {
    PSMMA_GCRT_CASE(2) PSMMA_GCRT_CASE(3) PSMMA_GCRT_CASE(4) PSMMA_GCRT_CASE(5) PSMMA_GCRT_CASE(6)
    PSMMA_GCRT_CASE(7) PSMMA_GCRT_CASE(8) PSMMA_GCRT_CASE(9) PSMMA_GCRT_CASE(10) PSMMA_GCRT_CASE(11)
    PSMMA_GCRT_CASE(12) PSMMA_GCRT_CASE(13) PSMMA_GCRT_CASE(14) PSMMA_GCRT_CASE(15) PSMMA_GCRT_CASE(16)
    default: return -16;
  }
#undef PSMMA_GCRT_CASE
  ++n;
  if (nodes) *nodes = n;
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : 1000 + (int)e;
}

// Fewest moduli that keep `bits` bits per operand for every problem of the batch.
int split_batch_moduli(const BatchProblem* probs, int nprob, int bits) {
  int kmax = 1;
  for (int i = 0; i < nprob; ++i) kmax = std::max(kmax, probs[i].k);
  for (int t = 2; t <= kMaxModuli; ++t)
    if (choose_abits(t, kmax) >= bits) return t;
  return kMaxModuli;
}

// Opaque grouped split batch for the batched C ABI: create (sizes the layout),
// bind (lays it out in a caller-owned workspace and uploads the table),
// enqueue (all phases on a stream; safe to capture), destroy.
struct GSplit {
  SplitBatch sb;
  std::vector<BatchProblem> probs;
  SplitProblem* dtab = nullptr;
};

void* gsplit_create(const BatchProblem* probs, int nprob, int t, size_t* ws_bytes) {
  GSplit* h = new GSplit;
  h->probs.assign(probs, probs + nprob);
  // PSMMA_GSPLIT_FUSED=0 keeps the residue-plane pipeline (read per plan)
  const char* fe = getenv("PSMMA_GSPLIT_FUSED");
  const bool fused_on = !(fe && atoi(fe) == 0);
  h->sb.fused = fused_on && std::all_of(probs, probs + nprob, fused_fits);
  split_batch_layout(h->probs.data(), nprob, t, nullptr, h->sb);
  if (ws_bytes) *ws_bytes = h->sb.ws_bytes;
  return h;
}

int gsplit_bind(void* hv, void* ws) {
  GSplit* h = (GSplit*)hv;
  split_batch_layout(h->probs.data(), (int)h->probs.size(), h->sb.t, (char*)ws, h->sb);
  if (h->sb.tab.empty()) return 0;
  const SplitBatch& sb = h->sb;
  const size_t np = sb.tab.size();
  const size_t bytes = sizeof(SplitProblem) * np;
  // unit owners, phase by phase: problem i holds units [ub, next problem's ub)
  std::vector<UnitOwner> own(owner_count(sb));
  size_t off = 0;
  for (int ph = 0; ph < NPH; ++ph) {
    for (size_t i = 0; i < np; ++i) {
      const int u0 = sb.tab[i].ub[ph];
      const int u1 = (i + 1 < np) ? sb.tab[i + 1].ub[ph] : sb.total[ph];
      if (ph == PH_FUSED && u1 > u0) {  // (no units unless the batch is fused)
        // tiles in patch order: P x P patches of the tile grid, row-major inside
        // a patch, so the P^2 consecutive units one XCD takes share P row and P
        // column panels in its L2
        constexpr int P = kFusedPatch;
        const int ntm = (sb.tab[i].M + 127) / 128, ntn = (sb.tab[i].N + 127) / 128;
        int u = u0;
        for (int pr = 0; pr < ntm; pr += P)
          for (int pc = 0; pc < ntn; pc += P)
            for (int r = pr; r < std::min(ntm, pr + P); ++r)
              for (int c = pc; c < std::min(ntn, pc + P); ++c)
                own[off + u++] = UnitOwner{(int)i, r * ntn + c};
        continue;
      }
      for (int u = u0; u < u1; ++u) own[off + u] = UnitOwner{(int)i, u - u0};
    }
    off += (size_t)sb.total[ph];
  }
  if (hipMalloc(&h->dtab, bytes + sizeof(UnitOwner) * own.size()) != hipSuccess) return 1002;
  hipError_t e = hipMemcpy(h->dtab, sb.tab.data(), bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess && !own.empty())
    e = hipMemcpy(h->dtab + np, own.data(), sizeof(UnitOwner) * own.size(),
                  hipMemcpyHostToDevice);
  return e == hipSuccess ? 0 : 1000 + (int)e;
}

int gsplit_enqueue(void* hv, hipStream_t s, int* nodes) {
  GSplit* h = (GSplit*)hv;
  return split_batch_enqueue(h->sb, h->dtab, s, nodes);
}

void gsplit_destroy(void* hv) {
  GSplit* h = (GSplit*)hv;
  if (!h) return;
  if (h->dtab) {
    hipDeviceSynchronize();
    hipFree(h->dtab);
  }
  delete h;
}

int gsplit_moduli(const BatchProblem* probs, int nprob, int bits) {
  return split_batch_moduli(probs, nprob, bits);
}

// Fewest int8 moduli that keep `bits` bits per operand for a contraction of
// length K (43 bits give a relative error below 1e-12 in the tests and bench).
int split_moduli_for_bits(int K, int bits) {
  for (int t = 2; t <= kMaxModuli; ++t)
    if (choose_abits(t, K) >= bits) return t;
  return kMaxModuli;
}

// Split-precision dispatcher.  slices = number of int8 moduli (PSMMA_PREC_SPLIT_I8)
// or of bf16 slices per operand (PSMMA_PREC_SPLIT_BF16).  An empty contraction
// (K == 0) has no products to split: it is the beta * C epilogue of the FP64 kernel.
int launch_right_split(int kind, int prec, int slices, int M, int N, int K, double alpha,
                       const double* A, int lda, const GenPtrs& g, int r0, int c0, double beta,
                       double* C, int ldc, hipStream_t s) {
  if (K == 0) {
    hipError_t e = launch_right_f64(kind, -1, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc, s);
    return e == hipSuccess ? 0 : 1000 + (int)e;
  }
  if (prec == 1) {
    // int32 accumulation of products of residues |r| <= 128 is exact for K < 2^17
    if (K > kSplitI8Kmax) return -3;
    const int t = slices ? slices : split_moduli_for_bits(K, kAutoSplitBits);
    return launch_right_split_i8(kind, t, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc, s);
  }
  if (slices == 0) slices = kAutoBf16Slices;
  return launch_right_split_bf16(kind, slices, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc, s);
}

// ===========================================================================
// Allocation-free int8 modular product (psmma_toeplitz_noalloc_dev).
//
// The scratch memory of the int8 form lives in a file-scope device array of
// fixed size instead of the grow-only per-stream workspace, so a call allocates
// nothing (callers under graph capture, serving loops that forbid hipMalloc).
// Every GPU the library runs on holds one copy of the array until the process
// ends.  In host code `split_scratch` names a host-side stand-in; the device
// address of each GPU's copy comes from hipGetSymbolAddress, taken once per
// device with that device current, and only that address is passed to kernels.
// All calls on one GPU share the array: calls that overlap on one GPU race.
constexpr size_t kNoallocScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char split_scratch[kNoallocScratchBytes];

static std::mutex g_scratch_mu;
static std::unordered_map<int, char*> g_scratch_ptr;

// Device address of the current device's scratch array (nullptr on failure).
static char* scratch_address() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) return it->second;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(split_scratch)) != hipSuccess || p == nullptr)
    return nullptr;
  g_scratch_ptr[dev] = (char*)p;
  return (char*)p;
}

size_t split_noalloc_capacity() { return kNoallocScratchBytes; }

// Scratch bytes the int8 form needs for one product (0 when it cannot run).
size_t split_i8_bytes(int kind, int slices, int M, int N, int K, const GenPtrs& g, int r0, int c0) {
  const int t = slices ? slices : split_moduli_for_bits(K, kAutoSplitBits);
  if (t < 2 || t > kMaxModuli || choose_abits(t, K) < 8) return 0;
  return i8_layout(kind, t, M, N, K, g, r0, c0).bytes;
}

// The int8 form in the static scratch array.  Returns 0, -16 (slices), -1 with
// *need set when the product needs more scratch than the array holds, or a
// runtime error (> 0).
int launch_right_split_i8_noalloc(int kind, int slices, int M, int N, int K, double alpha,
                                  const double* A, int lda, const GenPtrs& g, int r0, int c0,
                                  double beta, double* C, int ldc, hipStream_t s, size_t* need) {
  if (K > kSplitI8Kmax) return -3;
  const int t = slices ? slices : split_moduli_for_bits(K, kAutoSplitBits);
  if (t < 2 || t > kMaxModuli || choose_abits(t, K) < 8) return -16;
  const I8Layout L = i8_layout(kind, t, M, N, K, g, r0, c0);
  if (need) *need = L.bytes;
  if (L.bytes > kNoallocScratchBytes) return -1;
  char* ws = scratch_address();
  if (!ws) return 1000 + (int)hipErrorInvalidSymbol;
  return enqueue_split_i8_kind(kind, t, M, N, K, alpha, A, lda, g, r0, c0, beta, C, ldc, L, ws, s);
}

}  // namespace psmma
