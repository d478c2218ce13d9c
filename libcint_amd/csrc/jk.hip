// Coulomb and exchange matrices from blocks of unique shell quartets.
//
// For every quartet (i>=j, k>=l, ij>=kl in any order) of a computed ERI
// list, each distinct shell permutation (pq|rs) of the 8-fold symmetry
// group adds
//     J[p,q] += (pq|rs) D[r,s]        K[p,r] += (pq|rs) D[q,s]
// over the AO functions of the block.  This is the caller-side layer of
// libcint (PySCF's _vhf does it there); here it runs on the device next to
// the integrals.
#include "device_common.h"

namespace mcint {

__global__ void jk_digest_kernel(const int* __restrict__ shls, const int64_t* __restrict__ off,
                                 int64_t nq, const double* __restrict__ eri,
                                 const int* __restrict__ ao_loc, const double* __restrict__ D,
                                 int nao, double* __restrict__ J, double* __restrict__ K) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const int sh[4] = {shls[4 * q], shls[4 * q + 1], shls[4 * q + 2], shls[4 * q + 3]};
  int a0[4], dn[4];
  for (int s = 0; s < 4; ++s) {
    a0[s] = ao_loc[sh[s]];
    dn[s] = ao_loc[sh[s] + 1] - a0[s];
  }
  const double* blk = eri + off[q];
  // the 8 index maps of the (ij|kl) symmetry group, original axis per slot
  const int perms[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2}, {1, 0, 3, 2},
                           {2, 3, 0, 1}, {3, 2, 0, 1}, {2, 3, 1, 0}, {3, 2, 1, 0}};
  const int64_t N = nao;
  for (int p = 0; p < 8; ++p) {
    // skip permutations that repeat an earlier shell tuple
    bool dup = false;
    for (int r = 0; r < p && !dup; ++r) {
      dup = sh[perms[p][0]] == sh[perms[r][0]] && sh[perms[p][1]] == sh[perms[r][1]] &&
            sh[perms[p][2]] == sh[perms[r][2]] && sh[perms[p][3]] == sh[perms[r][3]];
    }
    if (dup) continue;
    const int* pm = perms[p];
    for (int l = 0; l < dn[3]; ++l)
      for (int k = 0; k < dn[2]; ++k)
        for (int j = 0; j < dn[1]; ++j)
          for (int i = 0; i < dn[0]; ++i) {
            const double v = blk[i + dn[0] * (j + dn[1] * (k + (int64_t)dn[2] * l))];
            if (v == 0.0) continue;
            const int idx[4] = {a0[0] + i, a0[1] + j, a0[2] + k, a0[3] + l};
            const int64_t P = idx[pm[0]], Q = idx[pm[1]], R = idx[pm[2]], S = idx[pm[3]];
            if (J) atomicAdd(J + P * N + Q, v * D[R * N + S]);
            if (K) atomicAdd(K + P * N + R, v * D[Q * N + S]);
          }
  }
}

__global__ void grad_reduce_kernel(const double* __restrict__ acc, int nslot, int n,
                                   double* __restrict__ grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = 0; k < nslot; ++k) s += acc[(int64_t)k * n + i];
  grad[i] += s;
}

hipError_t launch_grad_reduce(const double* acc, int nslot, int n, double* grad, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(grad_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                     acc, nslot, n, grad);
  return hipGetLastError();
}

hipError_t launch_jk_digest(const int* shls, const int64_t* off, int64_t nq, const double* eri,
                            const int* ao_loc, const double* D, int nao, double* J, double* K,
                            hipStream_t stream) {
  if (nq <= 0) return hipSuccess;
  hipLaunchKernelGGL(jk_digest_kernel, dim3((unsigned)((nq + 127) / 128)), dim3(128), 0, stream,
                     shls, off, nq, eri, ao_loc, D, nao, J, K);
  return hipGetLastError();
}

}  // namespace mcint
