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
constexpr int kMaxIter = 200;  // 24 model steps, then bisection to full precision

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
// origin s; the root is found as tau = lam - s on the shifted poles d_i - s, so
// DIFL_j = lam_j - d_j and DIFR_j = lam_j - d_{j+1} keep a small relative error
// however close the root is to a pole.
//
// Iteration: the rational model of Bunch, Nielsen and Sorensen (Numer. Math. 31
// (1978) 31-48) in the form LAPACK's DLAED4 uses: with psi / phi the sums over
// the poles left / right of the interval and Dl, Dr the two interval poles minus
// the current point, f is matched by c + s / (Dl - eta) + S / (Dr - eta) with the
// derivatives of psi and phi, and the root of that model is the next step
// (quadratic convergence; one pole for the last root).  f is increasing in tau
// between the poles, so the sign of f keeps a bracket; a step leaving it, or a
// run of 24 model steps without convergence, falls back to bisection, which
// then halves the bracket to full precision.
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
  const bool last = j + 1 >= n;
  const double lo = d[j];
  const double hi = last ? d[n - 1] + rho * zz : d[j + 1];
  const double mid = 0.5 * (lo + hi);
  double s = 0.0;
  for (int i = lane; i < n; i += 64) s += z[i] * z[i] / (d[i] - mid);
  const bool left = 1.0 + rho * wave_sum(s) >= 0.0;  // root left of the midpoint
  const double org = left ? lo : hi;
  const double pl = left ? 0.0 : lo - hi;  // the interval's poles, shifted
  const double pr = left ? hi - lo : 0.0;
  double a = left ? 0.0 : lo - hi;
  double b = left ? mid - lo : 0.0;
  double tau = 0.5 * (a + b);
  const double eps = 1.1102230246251565e-16;
  for (int it = 0; it < kMaxIter; ++it) {
    double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0;
    for (int i = lane; i < n; i += 64) {
      const double zi = z[i];
      const double r = 1.0 / ((d[i] - org) - tau);
      const double w = zi * zi * r;
      if (i <= j) {
        psi += w;
        dpsi += w * r;
      } else {
        phi += w;
        dphi += w * r;
      }
    }
    psi = wave_sum(psi);
    dpsi = wave_sum(dpsi);
    phi = wave_sum(phi);
    dphi = wave_sum(dphi);
    const double g = 1.0 + rho * (psi + phi);
    // f is zero within its rounding error: psi sums negative terms and phi
    // positive ones, so 1 + rho (phi - psi) bounds the magnitudes summed
    if (fabs(g) <= 8.0 * eps * (1.0 + rho * (phi - psi))) break;
    if (g < 0.0) a = tau; else b = tau;
    const double dl = pl - tau, dr = pr - tau;  // interval poles minus the point
    double eta;
    if (last) {  // one pole: g ~ c + s / (dl - eta), s = dl^2 rho dpsi
      const double sl = rho * dpsi * dl * dl;
      const double c = g - rho * dpsi * dl;
      eta = dl + sl / c;
    } else {
      const double A = (dl + dr) * g - dl * dr * rho * (dpsi + dphi);
      const double B = dl * dr * g;
      const double C = g - dl * rho * dpsi - dr * rho * dphi;
      const double disc = sqrt(fabs(A * A - 4.0 * B * C));
      if (C == 0.0)
        eta = B / A;
      else if (A <= 0.0)
        eta = (A - disc) / (2.0 * C);
      else
        eta = 2.0 * B / (A + disc);
    }
    // converged: the model's step is below the resolution of tau (the step may
    // land on a bracket end, which a bisection would throw away)
    if (fabs(eta) <= 2.0 * eps * fabs(tau)) {
      tau += eta;
      break;
    }
    double nt = tau + eta;
    if (!(nt > a && nt < b) || it >= 24) nt = 0.5 * (a + b);
    const double w = b - a;
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
