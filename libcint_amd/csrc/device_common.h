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

// J/K digest of unique quartets.  With the degeneracy factor
// f = (i==j ? 1/2 : 1) (k==l ? 1/2 : 1) (ij==kl ? 1/2 : 1) of a quartet of
// real shells (i>=j, k>=l, ij>=kl in any order), the sum over its 8 index
// maps reduces to six block updates of half matrices
//     J'_ij += 2f (ij|kl) D_kl      J'_kl += 2f (ij|kl) D_ij
//     K'_ik += f (ij|kl) D_jl       K'_il += f (ij|kl) D_jk
//     K'_jk += f (ij|kl) D_il       K'_jl += f (ij|kl) D_ik
// and J = J' + J'^T, K = K' + K'^T.  The set of updates is the same for
// every order of the four shells within the symmetry group, so kernels
// may digest a quartet in their canonical shell order.
__host__ __device__ inline double jk_degeneracy(int i, int j, int k, int l) {
  return (i == j ? 0.5 : 1.0) * (k == l ? 0.5 : 1.0) *
         (((i == k && j == l) || (i == l && j == k)) ? 0.5 : 1.0);
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
