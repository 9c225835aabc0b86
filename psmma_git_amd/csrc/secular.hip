// Generators of the eigen-Cauchy operand of a divide-and-conquer merge: the
// roots of the secular equation of D + rho z z^T, Loewner's recomputed z and
// the column normalisers.  These replace the per-root zero-finder loop, the
// z recomputation and the column normalisation of the reference's merge step
// (SRC/BasicMM.f90:538-590: a DLASD4 loop, the Loewner product for Z and DNRM2
// column scaling), whose outputs psmma_cauchyeig consumes as F, U, V, DIFL, DIFR
// (SRC/psmma_cauchyeig.f90:2-3,21-22).
//
// One wavefront per root (or per row / column): the O(n) sums of every
// iteration are spread over the 64 lanes and reduced with cross-lane shuffles.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/psmma.h"

namespace {

constexpr int kWaves = 4;      // roots per 256-thread workgroup
constexpr int kMaxIter = 320;  // the safeguard halves the bracket at least every 2 steps

// Sum over the wave; every lane returns lane 0's value so that wave-uniform
// control flow sees one number (lanes would otherwise differ in the last bit).
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return __shfl(x, 0);
}

__device__ __forceinline__ double wave_prod(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x *= __shfl_xor(x, o);
  return __shfl(x, 0);
}

// lambda_c - d_r from the shifted root (never from lambda itself): positive
// terms only for r <= c, non-positive terms only for r > c.  The tile generator
// forms its denominator as the negation of exactly this expression.
__device__ __forceinline__ double lam_minus_d(const double* d, const double* difl,
                                              const double* difr, int r, int c) {
  return (r <= c) ? ((d[c] - d[r]) + difl[c]) : ((d[c + 1] - d[r]) + difr[c]);
}

// Root j of f(lam) = 1 + rho sum_i z_i^2 / (d_i - lam), d strictly increasing,
// rho > 0: lam_j in (d_j, d_{j+1}), the last one in (d_n, d_n + rho z^T z).
// The secular function at the interval midpoint picks the nearer pole as the
// origin s; the root is found as tau = lam - s on the shifted poles d_i - s by
// Newton's method inside a bisection bracket (f is increasing in tau between
// the poles), so DIFL_j = lam_j - d_j and DIFR_j = lam_j - d_{j+1} keep a small
// relative error however close the root is to a pole.
__global__ __launch_bounds__(256) void secular_roots_kernel(int n, const double* __restrict__ d,
                                                            const double* __restrict__ z,
                                                            double rho, double* __restrict__ difl,
                                                            double* __restrict__ difr) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (j >= n) return;  // uniform over the wave
  double zz = 0.0;
  for (int i = lane; i < n; i += 64) zz += z[i] * z[i];
  zz = wave_sum(zz);
  const double lo = d[j];
  const double hi = (j + 1 < n) ? d[j + 1] : d[n - 1] + rho * zz;
  const double mid = 0.5 * (lo + hi);
  double s = 0.0;
  for (int i = lane; i < n; i += 64) s += z[i] * z[i] / (d[i] - mid);
  const bool left = 1.0 + rho * wave_sum(s) >= 0.0;  // root left of the midpoint
  const double org = left ? lo : hi;
  double a = left ? 0.0 : lo - hi;
  double b = left ? mid - lo : 0.0;
  double tau = 0.5 * (a + b);
  double w1 = b - a, w2 = b - a;  // bracket widths one and two steps back
  const double eps = 1.1102230246251565e-16;
  for (int it = 0; it < kMaxIter; ++it) {
    double s1 = 0.0, s2 = 0.0;
    for (int i = lane; i < n; i += 64) {
      const double zi = z[i];
      const double r = 1.0 / ((d[i] - org) - tau);
      const double w = zi * zi * r;
      s1 += w;
      s2 += w * r;
    }
    const double g = 1.0 + rho * wave_sum(s1);
    const double gp = rho * wave_sum(s2);
    if (g == 0.0) break;
    if (g < 0.0) a = tau; else b = tau;
    double nt = tau - g / gp;
    const double w = b - a;
    if (!(nt > a && nt < b) || w > 0.5 * w2) nt = 0.5 * (a + b);
    w2 = w1;
    w1 = w;
    const bool done = fabs(nt - tau) <= 2.0 * eps * fabs(nt) ||
                      w <= 4.0 * eps * fmax(fabs(a), fabs(b));
    tau = nt;
    if (done) break;
  }
  if (lane == 0) {
    difl[j] = left ? tau : tau + (hi - lo);
    difr[j] = left ? tau - (hi - lo) : tau;
  }
}

// Loewner's theorem: the z that makes the computed lam_j exact eigenvalues of
// D + rho zhat zhat^T is zhat_i^2 = prod_j (lam_j - d_i) / prod_{j != i} (d_j - d_i) / rho,
// taken with the sign of z_i.  The factors are paired as ratios
// (lam_j - d_i) / (d_j - d_i), each within a factor ~N of 1.
__global__ __launch_bounds__(256) void loewner_kernel(int n, const double* __restrict__ d,
                                                      const double* __restrict__ z, double rho,
                                                      const double* __restrict__ difl,
                                                      const double* __restrict__ difr,
                                                      double* __restrict__ zhat) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (i >= n) return;
  double p = 1.0;
  for (int j = lane; j < n; j += 64) {
    const double num = lam_minus_d(d, difl, difr, i, j);
    p *= (j == i) ? num : num / (d[j] - d[i]);
  }
  p = wave_prod(p);
  if (lane == 0) zhat[i] = copysign(sqrt(fabs(p) / rho), z[i]);
}

// v_c = 1 / || u ./ (d - lam_c) ||, with the same differences as the tile.
__global__ __launch_bounds__(256) void colnorm_kernel(int n, const double* __restrict__ d,
                                                      const double* __restrict__ u,
                                                      const double* __restrict__ difl,
                                                      const double* __restrict__ difr,
                                                      double* __restrict__ v) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (c >= n) return;
  double s = 0.0;
  for (int r = lane; r < n; r += 64) {
    const double x = u[r] / lam_minus_d(d, difl, difr, r, c);
    s += x * x;
  }
  s = wave_sum(s);
  if (lane == 0) v[c] = 1.0 / sqrt(s);
}

}  // namespace

extern "C" int psmma_secular_dev(int n, const double* d, const double* z, double rho,
                                 double* difl, double* difr, double* zhat, double* v,
                                 void* stream) {
  if (n < 0) return -1;
  if (n == 0) return 0;
  if (!d) return -2;
  if (!z) return -3;
  if (!(rho > 0.0)) return -4;
  if (!difl) return -5;
  if (!difr) return -6;
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((n + kWaves - 1) / kWaves), block(64 * kWaves);
  secular_roots_kernel<<<grid, block, 0, s>>>(n, d, z, rho, difl, difr);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && zhat) {
    loewner_kernel<<<grid, block, 0, s>>>(n, d, z, rho, difl, difr, zhat);
    e = hipGetLastError();
  }
  if (e == hipSuccess && v) {
    colnorm_kernel<<<grid, block, 0, s>>>(n, d, zhat ? zhat : z, difl, difr, v);
    e = hipGetLastError();
  }
  if (e != hipSuccess) {
    fprintf(stderr, "psmma_secular_dev: %s\n", hipGetErrorString(e));
    return 1000 + (int)e;
  }
  return 0;
}
