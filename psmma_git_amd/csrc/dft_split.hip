// Split-FP32 form of the left-structured DFT product (PSMMA_PREC_SPLIT_FP32):
//
//   C(m x n) = alpha * F(r0:r0+m, c0:c0+k) * B(k x n) + beta * C,
//   F(r,c) = exp(-2 pi i ((r*c) mod nd) / nd), complex128 B and C,
//
// the product that PZMDFTL1 forms with ZGEMM (SRC/pzmdftL1.f90:488-498), here
// on FP32 MFMA (v_mfma_f32_16x16x4f32) with an error-free split:
//
//  * every column j of B is scaled by 2^-e_j so that |Re|, |Im| < 1; F's entries
//    already lie in [-1, 1];
//  * each value x in [-1, 1] is cut into S fixed-point slices of 8 bits:
//    x_1 = rint(x 2^8) 2^-8, x_p = rint(r_{p-1} 2^(8p)) 2^(-8p) with r_p the
//    remainder, so |x - sum_p x_p| <= 2^(-8S-1) and every slice has at most 9
//    significant bits (exact in FP32);
//  * the complex product runs in the three-product form: per class, P1 = sum
//    Fr_p Br_q, P2 = sum Fi_p Bi_q and P3 = sum (Fr_p + Fi_p)(Br_q + Bi_q), then
//    Re = P1 - P2 and Im = P3 - P1 - P2.  Slices on one grid sum exactly (at most
//    10 significant bits), so every product is exact in FP32; all pairs with
//    p + q = d lie on the grid 2^(-8d), and for d <= 7 one k step adds at most
//    2^17 grid units to P1 or P2 and 2^19 to P3, so the FP32 sum of a K-slab of
//    16 stays below 2^24 units and is exact.  Pairs with p + q <= S + 1 are kept,
//    three FP32 accumulators per class, flushed into FP64 accumulators after
//    every slab (FP64 recombination, where P1 - P2 and P3 - P1 - P2 are exact:
//    the result equals the four-product form's bit for bit);
//  * the twiddles come from the FP64 table, so F's slices are those of the
//    correctly rounded entries; they are generated straight into LDS.
//
// Error: about 2^(-8S) relative to the row/column maxima (normwise), plus the
// FP64 recombination.  Cost: S(S+1)/2 slice pairs of 3 real FP32 products,
// against 3 FP64 products for the plain form; FP32 MFMA runs at twice the FP64
// MFMA rate on MI355X, so S = 2 costs 1.5x the plain form and S = 5 (about
// 1e-12) 7.5x.  The form exists for callers that want FP32 MFMA; the plain FP64
// kernel is the fast exact one on this GPU.
#include <hip/hip_runtime.h>

#include <climits>
#include <mutex>

#include "psmma_common.h"
#include "workspace.h"

namespace psmma {

namespace {

constexpr int kDftNoExp = INT_MIN / 4;

// Largest frexp exponent of |Re| and |Im| over column j of B (k x n, ld ldb):
// e_j with max < 2^e_j.  Threads split k; one atomicMax per workgroup.
__global__ __launch_bounds__(256) void dft_col_exp_kernel(int K, const double2* __restrict__ B,
                                                          int ldb, int kchunk, int* __restrict__ e) {
  const int j = blockIdx.x;
  const int k0 = blockIdx.y * kchunk;
  const int k1 = min(K, k0 + kchunk);
  int mx = kDftNoExp;
  for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
    double2 v = B[(size_t)j * ldb + k];
    double a = fmax(fabs(v.x), fabs(v.y));
    if (a != 0.0) {
      int ex;
      frexp(a, &ex);
      mx = max(mx, ex);
    }
  }
  __shared__ int red[256];
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[0] != kDftNoExp) atomicMax(&e[j], red[0]);
}

__global__ void dft_fill_kernel(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// S slices of 8 bits of x in [-1, 1] (exact FP64 steps, FP32 results).
template <int S>
__device__ __forceinline__ void slice8(double x, float* out) {
  double r = x;
#pragma unroll
  for (int p = 0; p < S; ++p) {
    const double sc = (double)(1ll << (8 * (p + 1)));  // 2^(8(p+1)), exact for p < 6
    const double q = __builtin_rint(r * sc) / sc;        // exact: power-of-two scaling
    out[p] = (float)q;
    r = r - q;                                           // exact (Sterbenz-type)
  }
}

}  // namespace

// v_mfma_f32_16x16x4f32 lane layout: A lane l supplies A[l % 16][k = l / 16],
// B lane l supplies B[k = l / 16][l % 16], D lane l element r holds
// D[4 * (l / 16) + r][l % 16].
template <int S, int BM, int BN, int BK, int WGM, int WGN>
__global__ __launch_bounds__(WGM* WGN * 64, 1) void dft_split_f32_kernel(
    int M, int N, int K, int Nd, double alpha, const double2* __restrict__ tw,
    const double2* __restrict__ B, int ldb, const int* __restrict__ eb, int r0, int c0, double beta,
    double2* __restrict__ C, int ldc) {
  constexpr int NT = WGM * WGN * 64;
  constexpr int WM = BM / WGM, WN = BN / WGN;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int FLD = BM + 4, BLD = BN + 4;
  constexpr int FREG = BM * BK / NT;
  constexpr int BREG = BK * BN / NT;
  constexpr int FSZ = BK * FLD, BSZ = BK * BLD;  // one slice, one part
  constexpr int NCLS = S;                         // classes d = 2 .. S + 1
  static_assert(BM * BK % NT == 0 && BK * BN % NT == 0, "tile/threads mismatch");
  static_assert(NT % BM == 0, "F generator assumes a fixed row per thread");
  static_assert(BK <= 16, "exact FP32 class sums need K-slabs of at most 16");
  static_assert(S >= 1 && S <= 6, "1..6 slices");

  // [buf][slice][re, im][k][row|col]
  __shared__ float Fs[2][S][2][FSZ];
  __shared__ float Bs[2][S][2][BSZ];

  const int nbn = (N + BN - 1) / BN;
  const int nbm = (M + BM - 1) / BM;
  const int bid = xcd_remap(blockIdx.x, nbn * nbm);
  const int m0 = (bid / nbn) * BM, n0 = (bid % nbn) * BN;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / WGN) * WM, wn0 = (wave % WGN) * WN;
  const int lr = lane & 15, lk = lane >> 4;

  // P1, P2, P3 of each class (three-product form)
  v4f accr[NCLS][TM][TN], acci[NCLS][TM][TN], accs[NCLS][TM][TN];
  double sr[TM][TN][4], si[TM][TN][4];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
#pragma unroll
      for (int d = 0; d < NCLS; ++d) {
        accr[d][i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
        acci[d][i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
        accs[d][i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) sr[i][j][r] = si[i][j][r] = 0.0;
    }

  // F: each thread generates one fixed row and FREG columns per slab; exponents
  // e_q = row*col mod Nd advance by row*BK mod Nd per slab.
  const int frow = tid % BM;
  const int grow = m0 + frow;
  const unsigned nd = (unsigned)Nd;
  const unsigned rmod = (unsigned)(((long long)(r0 + grow) % Nd + Nd) % Nd);
  const unsigned step = (unsigned)(((unsigned long long)rmod * BK) % nd);
  unsigned ex[FREG];
#pragma unroll
  for (int q = 0; q < FREG; ++q) {
    int col = tid / BM + q * (NT / BM);
    ex[q] = (unsigned)(((unsigned long long)rmod * (unsigned long long)(c0 + col)) % nd);
  }

  double2 freg[FREG], breg[BREG];
  double bscale[BREG];
#pragma unroll
  for (int q = 0; q < BREG; ++q) {
    int col = (tid + q * NT) / BK;
    int gn = n0 + col;
    bscale[q] = (gn < N && eb[gn] != kDftNoExp) ? ldexp(1.0, -eb[gn]) : 0.0;
  }
  auto gen_f = [&](int k0) {
#pragma unroll
    for (int q = 0; q < FREG; ++q) {
      int col = tid / BM + q * (NT / BM);
      bool ok = (grow < M) && (k0 + col < K);
      freg[q] = ok ? tw[ex[q]] : make_double2(0.0, 0.0);
    }
  };
  auto advance_f = [&]() {
#pragma unroll
    for (int q = 0; q < FREG; ++q) {
      unsigned e = ex[q] + step;
      ex[q] = (e >= nd) ? e - nd : e;
    }
  };
  auto load_b = [&](int k0) {
#pragma unroll
    for (int q = 0; q < BREG; ++q) {
      int idx = tid + q * NT;
      int row = idx % BK, col = idx / BK;
      int gk = k0 + row, gn = n0 + col;
      breg[q] = (gk < K && gn < N) ? B[(size_t)gn * ldb + gk] : make_double2(0.0, 0.0);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < FREG; ++q) {
      int col = tid / BM + q * (NT / BM);
      float fr[S], fi[S];
      slice8<S>(freg[q].x, fr);
      slice8<S>(freg[q].y, fi);
#pragma unroll
      for (int p = 0; p < S; ++p) {
        Fs[buf][p][0][col * FLD + frow] = fr[p];
        Fs[buf][p][1][col * FLD + frow] = fi[p];
      }
    }
#pragma unroll
    for (int q = 0; q < BREG; ++q) {
      int idx = tid + q * NT;
      int row = idx % BK, col = idx / BK;
      float br[S], bi[S];
      slice8<S>(breg[q].x * bscale[q], br);
      slice8<S>(breg[q].y * bscale[q], bi);
#pragma unroll
      for (int p = 0; p < S; ++p) {
        Bs[buf][p][0][row * BLD + col] = br[p];
        Bs[buf][p][1][row * BLD + col] = bi[p];
      }
    }
  };

  const int nkt = (K + BK - 1) / BK;
  if (nkt > 0) {
    gen_f(0);
    load_b(0);
    store(0);
  }
  __syncthreads();

  for (int kt = 0; kt < nkt; ++kt) {
    const int cur = kt & 1;
    const bool more = (kt + 1) < nkt;
    if (more) {
      advance_f();
      gen_f((kt + 1) * BK);
      load_b((kt + 1) * BK);
    }
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      // B's slices for every q, then F's slice p at a time (fewer live registers)
      float fbr[S][TN], fbi[S][TN], fbs[S][TN];
#pragma unroll
      for (int q = 0; q < S; ++q)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          fbr[q][j] = Bs[cur][q][0][(kk + lk) * BLD + wn0 + j * 16 + lr];
          fbi[q][j] = Bs[cur][q][1][(kk + lk) * BLD + wn0 + j * 16 + lr];
          fbs[q][j] = fbr[q][j] + fbi[q][j];  // exact: one grid, at most 10 bits
        }
      // class d = p + q (0-based: dd = p + q, kept when dd <= S - 1)
#pragma unroll
      for (int p = 0; p < S; ++p) {
        float far[TM], fai[TM], fas[TM];
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          far[i] = Fs[cur][p][0][(kk + lk) * FLD + wm0 + i * 16 + lr];
          fai[i] = Fs[cur][p][1][(kk + lk) * FLD + wm0 + i * 16 + lr];
          fas[i] = far[i] + fai[i];
        }
#pragma unroll
        for (int q = 0; q + p < S; ++q)
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) {
              v4f& p1 = accr[p + q][i][j];
              v4f& p2 = acci[p + q][i][j];
              v4f& p3 = accs[p + q][i][j];
              p1 = __builtin_amdgcn_mfma_f32_16x16x4f32(far[i], fbr[q][j], p1, 0, 0, 0);
              p2 = __builtin_amdgcn_mfma_f32_16x16x4f32(fai[i], fbi[q][j], p2, 0, 0, 0);
              p3 = __builtin_amdgcn_mfma_f32_16x16x4f32(fas[i], fbs[q][j], p3, 0, 0, 0);
            }
      }
    }
    // flush the exact slab sums into FP64
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int d = 0; d < NCLS; ++d) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double p1 = (double)accr[d][i][j][r], p2 = (double)acci[d][i][j][r];
            sr[i][j][r] += p1 - p2;                                // exact differences
            si[i][j][r] += ((double)accs[d][i][j][r] - p1) - p2;
          }
          accr[d][i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
          acci[d][i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
          accs[d][i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
        }
    if (more) store(cur ^ 1);
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      int gc = n0 + wn0 + j * 16 + lr;
      if (gc >= N) continue;
      const int e = eb[gc];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int gr = m0 + wm0 + i * 16 + 4 * lk + r;
        if (gr >= M) continue;
        double2* p = C + (size_t)gc * ldc + gr;
        double vr = 0.0, vi = 0.0;
        if (e != kDftNoExp) {
          vr = alpha * ldexp(sr[i][j][r], e);
          vi = alpha * ldexp(si[i][j][r], e);
        }
        if (beta != 0.0) {
          double2 o = *p;
          vr += beta * o.x;
          vi += beta * o.y;
        }
        *p = make_double2(vr, vi);
      }
    }
}

template <int S>
static void launch_dft_split_s(int M, int N, int K, int Nd, double alpha, const double2* tw,
                               const double2* B, int ldb, const int* eb, int r0, int c0,
                               double beta, double2* C, int ldc, hipStream_t s) {
  // 8 waves of 32 x 16: the three accumulators per class stay in registers up to 6
  // slices (4 waves of 32 x 32 spill from 4); F(4096 x 4096) B(4096 x 512) on MI355X,
  // 2 / 3 / 5 slices: 1.88 / 3.44 / 7.25 ms (four-product form on 4 waves: 2.14 / 3.97 / 8.65)
  constexpr int BM = 64, BN = 64, BK = 8, WGM = 2, WGN = 4;
  int nb = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  hipLaunchKernelGGL((dft_split_f32_kernel<S, BM, BN, BK, WGM, WGN>), dim3(nb),
                     dim3(WGM * WGN * 64), 0, s, M, N, K, Nd, alpha, tw, B, ldb, eb, r0, c0, beta,
                     C, ldc);
}

// Default slice count of the split-FP32 DFT form (about 2^-40 normwise).
constexpr int kDftSplitDefaultSlices = 5;

// C = alpha * F * B + beta * C in the split-FP32 form with `slices` (0: default)
// 8-bit slices per operand.  Returns 0, -13 (slices out of 1..6) or a runtime
// error (> 0).
int launch_dft_split_f32(int M, int N, int K, int Nd, int slices, double alpha, const double2* tw,
                         const double2* B, int ldb, int r0, int c0, double beta, double2* C,
                         int ldc, hipStream_t s) {
  if (slices == 0) slices = kDftSplitDefaultSlices;
  if (slices < 1 || slices > 6) return -13;
  std::unique_lock<std::mutex> lk;
  int* eb = (int*)ws_lease(WS_SPLIT, sizeof(int) * (size_t)(N > 0 ? N : 1), s, lk);
  if (!eb) return 1002;
  hipLaunchKernelGGL(dft_fill_kernel, dim3((N + 255) / 256), dim3(256), 0, s, eb, N, kDftNoExp);
  const int kchunk = 2048;
  if (K > 0)
    hipLaunchKernelGGL(dft_col_exp_kernel, dim3(N, (K + kchunk - 1) / kchunk), dim3(256), 0, s, K,
                       B, ldb, kchunk, eb);
  switch (slices) {
    case 1: launch_dft_split_s<1>(M, N, K, Nd, alpha, tw, B, ldb, eb, r0, c0, beta, C, ldc, s); break;
    case 2: launch_dft_split_s<2>(M, N, K, Nd, alpha, tw, B, ldb, eb, r0, c0, beta, C, ldc, s); break;
    case 3: launch_dft_split_s<3>(M, N, K, Nd, alpha, tw, B, ldb, eb, r0, c0, beta, C, ldc, s); break;
    case 4: launch_dft_split_s<4>(M, N, K, Nd, alpha, tw, B, ldb, eb, r0, c0, beta, C, ldc, s); break;
    case 5: launch_dft_split_s<5>(M, N, K, Nd, alpha, tw, B, ldb, eb, r0, c0, beta, C, ldc, s); break;
    default: launch_dft_split_s<6>(M, N, K, Nd, alpha, tw, B, ldb, eb, r0, c0, beta, C, ldc, s);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : 1000 + (int)e;
}

}  // namespace psmma
