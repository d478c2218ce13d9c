// Rys roots and weights on the device, from the generated piecewise
// polynomials (rys_tables.inc) below RYS_XMAX(n) and the large-x limit of
// the rule above it.  Included by exactly one translation unit.
#pragma once
#include "mcint_internal.h"

#define RYS_TABLE_QUAL static __device__ const
#include "rys_tables.inc"
#undef RYS_TABLE_QUAL

namespace mcint {

// Root/weight number f (0 <= f < 2n: roots first, then weights) of the
// n-point rule at x < RYS_XMAX(n).
__device__ __forceinline__ double rys_poly_eval(int n, int f, double x) {
  const int it = (int)x;
  const double z = 2.0 * (x - (double)it) - 1.0;
  const double* c = rys_coef + rys_table_offset[n] + (size_t)it * RYS_NCOEF * 2 * n + f;
  double acc = c[(RYS_NCOEF - 1) * 2 * n];
#pragma unroll
  for (int j = RYS_NCOEF - 2; j >= 0; --j) acc = fma(acc, z, c[j * 2 * n]);
  return acc;
}

// One root r of the n-point rule (runtime n): u = t^2/(1-t^2), w.
__device__ __forceinline__ void rys_root_one(int n, int r, double x, double& u, double& w) {
  if (x >= (double)RYS_XMAX(n)) {
    const double rr = rys_lag_r[rys_lag_offset[n] + r];
    u = rr / (x - rr);
    w = rys_lag_w[rys_lag_offset[n] + r] / sqrt(x);
  } else {
    u = rys_poly_eval(n, r, x);
    w = rys_poly_eval(n, n + r, x);
  }
}

// All NR roots (compile-time NR): one Horner chain per function, the
// coefficients of one interval are contiguous per lane.
template <int NR>
__device__ __forceinline__ void rys_roots(double x, double* u, double* w) {
  if (x >= (double)RYS_XMAX(NR)) {
    const double rsx = 1.0 / sqrt(x);
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const double rr = rys_lag_r[rys_lag_offset[NR] + i];
      u[i] = rr / (x - rr);
      w[i] = rys_lag_w[rys_lag_offset[NR] + i] * rsx;
    }
  } else {
    const int it = (int)x;
    const double z = 2.0 * (x - (double)it) - 1.0;
    const double* c = rys_coef + rys_table_offset[NR] + (size_t)it * RYS_NCOEF * 2 * NR;
    double acc[2 * NR];
#pragma unroll
    for (int f = 0; f < 2 * NR; ++f) acc[f] = c[(RYS_NCOEF - 1) * 2 * NR + f];
#pragma unroll
    for (int j = RYS_NCOEF - 2; j >= 0; --j) {
#pragma unroll
      for (int f = 0; f < 2 * NR; ++f) acc[f] = fma(acc[f], z, c[j * 2 * NR + f]);
    }
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      u[i] = acc[i];
      w[i] = acc[NR + i];
    }
  }
}

}  // namespace mcint
