// Short-range Rys rule on the device: the n-point Gauss rule of the weight
// exp(-x t^2) dt on [lower, 1], in the variable s = t^2, returned as
// u = s / (1 - s) and w (the rule CINTsr_rys_roots returns,
// src/rys_roots.c:149-267, declared at src/rys_roots.h:4).
//
// Method (Gautschi's discretised Stieltjes procedure with a Sturm-bisection
// eigensolver): one wave per rule.  The weight is sampled at 1024 points:
// 16-point Gauss-Legendre in t on 64 pieces of equal width in s over
// [lower^2, min(1, lower^2 + 100/x)] (beyond the upper end the weight has
// fallen below e^-100 of its value at lower; across one piece it changes by
// at most e^1.6), one piece per lane, with weights taken relative to
// exp(-x lower^2) so that no rule underflows before the last step.  The
// wave generates the monic three-term recurrence
// p_{k+1} = (s - a_k) p_k - b_k p_{k-1} of that discrete measure (two wave
// sums per step).  Lane i < n then finds the i-th eigenvalue of the Jacobi
// matrix (diagonal a_k, off-diagonal sqrt(b_k)) by bisection on its Sturm
// sequence, to the last bit, and its weight as the Christoffel number
// 1 / sum_k p_k(s)^2 / (b_0 ... b_k), times exp(-x lower^2).
#include <hip/hip_runtime.h>

#include "eri_launch.h"

namespace mcint {
namespace {

constexpr int SR_PIECES = 64;  // pieces of equal width in s, one per lane
constexpr int SR_GL = 16;      // Gauss-Legendre points per piece
constexpr int SR_PER_LANE = SR_GL;

// node g (0..15) of the 16-point Gauss-Legendre rule on [-1, 1]: Newton's
// method on P_16 from the asymptotic guess cos(pi (g + 3/4) / (16 + 1/2))
__device__ void gauss_legendre16(int g, double& xi, double& wi) {
  double z = cos(3.14159265358979323846 * (g + 0.75) / (SR_GL + 0.5));
  double p0 = 1, p1 = z, dp = 1;
  for (int it = 0; it < 10; ++it) {
    p0 = 1;
    p1 = z;
    for (int k = 2; k <= SR_GL; ++k) {
      const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = SR_GL * (z * p1 - p0) / (z * z - 1);
    z -= p1 / dp;
  }
  p0 = 1;
  p1 = z;
  for (int k = 2; k <= SR_GL; ++k) {
    const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  dp = SR_GL * (z * p1 - p0) / (z * z - 1);
  xi = z;
  wi = 2 / ((1 - z * z) * dp * dp);
}

// sum over the wave, lane 0's total on every lane
__device__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return __shfl(v, 0);
}

__global__ __launch_bounds__(64) void sr_rys_kernel(int n, const double* __restrict__ xs,
                                                    const double* __restrict__ lows, int64_t count,
                                                    double* __restrict__ u, double* __restrict__ w) {
  const int64_t e = blockIdx.x;
  if (e >= count) return;  // uniform over the wave
  const int lane = threadIdx.x;
  const double x = xs[e], lo = lows[e];
  const double s_lo = lo * lo;
  double s_hi = 1.0;
  if (x > 0 && s_lo + 100.0 / x < 1.0) s_hi = s_lo + 100.0 / x;
  const double ds = (s_hi - s_lo) / SR_PIECES;
  const int piece = lane;
  const double t0 = sqrt(s_lo + ds * piece);
  const double t1 = piece == SR_PIECES - 1 ? sqrt(s_hi) : sqrt(s_lo + ds * (piece + 1));
  const double mid = 0.5 * (t0 + t1), half = 0.5 * (t1 - t0);
  double S[SR_PER_LANE], W[SR_PER_LANE], P[SR_PER_LANE], Q[SR_PER_LANE];
#pragma unroll
  for (int j = 0; j < SR_PER_LANE; ++j) {
    double xi, wi;
    gauss_legendre16(j, xi, wi);
    const double t = mid + half * xi;
    S[j] = t * t;
    W[j] = half * wi * exp(-x * (S[j] - s_lo));
    P[j] = 1;
    Q[j] = 0;
  }
  double a[NRYS_MAX], b[NRYS_MAX];
  double nrm_prev = 1;
  for (int k = 0; k < n; ++k) {
    double q0 = 0, q1 = 0;
#pragma unroll
    for (int j = 0; j < SR_PER_LANE; ++j) {
      const double q = W[j] * P[j] * P[j];
      q0 += q;
      q1 += q * S[j];
    }
    q0 = wave_sum(q0);
    q1 = wave_sum(q1);
    a[k] = q1 / q0;
    b[k] = k ? q0 / nrm_prev : q0;
    nrm_prev = q0;
#pragma unroll
    for (int j = 0; j < SR_PER_LANE; ++j) {
      const double pn = (S[j] - a[k]) * P[j] - (k ? b[k] * Q[j] : 0.0);
      Q[j] = P[j];
      P[j] = pn;
    }
  }
  if (lane >= n) return;
  // eigenvalue number `lane` (ascending): the Sturm count of eigenvalues
  // below m is the number of negative pivots of J - m I
  double L = s_lo, H = s_hi;
  for (int it = 0; it < 200; ++it) {
    const double m = 0.5 * (L + H);
    if (m <= L || m >= H) break;
    double d = a[0] - m;
    int cnt = d < 0;
    for (int k = 1; k < n; ++k) {
      if (d == 0) d = 1e-300;
      d = (a[k] - m) - b[k] / d;
      cnt += d < 0;
    }
    if (cnt > lane)
      H = m;
    else
      L = m;
  }
  const double s = 0.5 * (L + H);
  // Christoffel number of the discrete measure at s
  double pk = 1, pkm = 0, nrm = b[0], sum = 1 / b[0];
  for (int k = 0; k + 1 < n; ++k) {
    const double pn = (s - a[k]) * pk - (k ? b[k] * pkm : 0.0);
    pkm = pk;
    pk = pn;
    nrm *= b[k + 1];
    sum += pk * pk / nrm;
  }
  u[e * n + lane] = s / (1 - s);
  w[e * n + lane] = exp(-x * s_lo) / sum;
}

}  // namespace

hipError_t launch_sr_rys_batch(int n, const double* x, const double* lower, int64_t count,
                               double* u, double* w, hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(sr_rys_kernel, dim3((unsigned)count), dim3(64), 0, stream, n, x, lower, count,
                     u, w);
  return hipGetLastError();
}

}  // namespace mcint
