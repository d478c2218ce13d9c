// Rys roots and weights on the device, from the generated piecewise
// polynomials (rys_tables.inc) below RYS_XMAX(n) and the large-x limit of
// the rule above it.  Included by exactly one translation unit.
#pragma once
#include "mcint_internal.h"

#define RYS_TABLE_QUAL static __device__ const
#include "rys_tables.inc"
#undef RYS_TABLE_QUAL

namespace mcint {

// 1/d for positive normal d: hardware reciprocal refined by two Newton
// steps (error below 1 ulp; the IEEE division sequence costs about twice
// the instructions and handles cases that cannot occur here).
__device__ __forceinline__ double rcp_pos(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

// 1/sqrt(x) and sqrt(x) for positive normal x: the hardware reciprocal
// square root refined by two Newton-Raphson steps y <- y (3/2 - x y^2 / 2),
// and for sqrt one more residual correction s <- s + (x - s^2) y / 2 (error
// about 1 ulp; the library sqrt also rescales denormal and special inputs,
// which cannot occur here).
__device__ __forceinline__ double rsqrt_pos(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = fma(y, fma(-h * y, y, 0.5), y);
  return fma(y, fma(-h * y, y, 0.5), y);
}
__device__ __forceinline__ double sqrt_pos(double x) {
  const double y = rsqrt_pos(x);
  const double s = x * y;
  return fma(fma(-s, s, x), 0.5 * y, s);
}

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
    u = rr * rcp_pos(x - rr);
    w = rys_lag_w[rys_lag_offset[n] + r] * rsqrt_pos(x);
  } else {
    u = rys_poly_eval(n, r, x);
    w = rys_poly_eval(n, n + r, x);
  }
}

// Root tables staged in LDS (n <= 4: at most 55 KB, which leaves room for
// two workgroups of 256 threads per CU; the 75-98 KB tables of n = 5, 6
// would allow only one and are read from the global copy).  In LDS each
// interval's block of RYS_NCOEF * 2n coefficients is padded by two doubles
// (keeping 16-byte alignment): lanes reading different intervals then start
// on different LDS banks (an unpadded block is a multiple of 8 banks long).
constexpr int RYS_LDS_NMAX = 4;
template <int NR>
struct RysLds {
  static constexpr bool IN_LDS = NR <= RYS_LDS_NMAX;
  static constexpr int BLOCK = RYS_NCOEF * 2 * NR;            // one interval, global layout
  static constexpr int STRIDE = IN_LDS ? BLOCK + 2 : BLOCK;   // one interval, as evaluated
  static constexpr int N = IN_LDS ? RYS_XMAX(NR) * STRIDE : 1;
};

// Copies the n = NR table into s (all threads of the block; ends with a
// barrier) and returns the table to evaluate from: s, or the global copy
// when the table does not fit.
template <int NR>
__device__ __forceinline__ const double* rys_stage_table(double* s) {
  using L = RysLds<NR>;
  if constexpr (L::IN_LDS) {
    const double* g = rys_coef + rys_table_offset[NR];
    for (int i = threadIdx.x; i < RYS_XMAX(NR) * L::BLOCK; i += blockDim.x) {
      const int iv = i / L::BLOCK;
      s[i + 2 * iv] = g[i];  // = iv * STRIDE + (i - iv * BLOCK)
    }
    __syncthreads();
    return s;
  } else {
    return rys_coef + rys_table_offset[NR];
  }
}

// All NR roots (compile-time NR): one Horner chain per function, the
// coefficients of one interval are contiguous per lane.  tab: the n = NR
// table (rys_coef + rys_table_offset[NR], or its LDS copy).
template <int NR>
__device__ __forceinline__ void rys_roots(double x, double* u, double* w,
                                          const double* __restrict__ tab) {
  if (x >= (double)RYS_XMAX(NR)) {
    const double rsx = rsqrt_pos(x);
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const double rr = rys_lag_r[rys_lag_offset[NR] + i];
      u[i] = rr * rcp_pos(x - rr);
      w[i] = rys_lag_w[rys_lag_offset[NR] + i] * rsx;
    }
  } else {
    const int it = (int)x;
    const double z = 2.0 * (x - (double)it) - 1.0;
    const double* c = tab + it * RysLds<NR>::STRIDE;
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
