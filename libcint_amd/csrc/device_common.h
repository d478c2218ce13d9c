// Device-side helpers shared by every device translation unit.
#pragma once
#include "eri_launch.h"
#include "rys_device.h"

#define C2S_QUAL static __device__ const
#include "c2s_tables.inc"
#undef C2S_QUAL

namespace mcint {

__device__ __forceinline__ double common_fac_sp(int l) {
  // 1/(2 sqrt(pi)) and sqrt(3/(4 pi)); src/g1e.c:565-572
  return l == 0 ? 0.282094791773878143 : (l == 1 ? 0.488602511902919921 : 1.0);
}

constexpr double PI3_2_SQRTPI = 34.986836655249725;  // 2 pi^3 / sqrt(pi)

__device__ __forceinline__ int cart_idx(int l, int lx, int lz) {
  const int m = l - lx;
  return m * (m + 1) / 2 + lz;
}

// Gradient digest: adds the per-team partials g[k][x] (atom[k] < 0: none)
// of one block to slot (block % nslot) of the accumulators, one atomic per
// run of equal atoms and per component.  Call after a barrier that made g
// and atom visible.
__device__ __forceinline__ void grad_block_flush(const BucketArgs& A, const double (*g)[3],
                                                 const int* atom, int nteam) {
  if (threadIdx.x < 3) {
    const int c = threadIdx.x;
    double* base = A.grad + (int64_t)(blockIdx.x % A.nslot) * A.grad_stride + c;
    int cur = -1;
    double s = 0.0;
    for (int k = 0; k < nteam; ++k) {
      const int at = atom[k];
      if (at < 0) continue;
      if (at != cur) {
        if (cur >= 0) atomicAdd(base + 3 * cur, s);
        cur = at;
        s = 0.0;
      }
      s += g[k][c];
    }
    if (cur >= 0) atomicAdd(base + 3 * cur, s);
  }
}

}  // namespace mcint
