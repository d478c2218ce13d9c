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

// ---------------------------------------------------------------------
// Grouped digest of a stored ERI buffer: the six block updates of
// jk_degeneracy's half matrices (device_common.h).  The row pass runs one workgroup per
// group of quartets sharing (i, j): lane (ab, s) owns the bra element ab
// and every S-th quartet, J'_ij stays in registers, rows i and j of K' in
// LDS; both are flushed once per group.  The column pass does J'_kl per
// group of quartets sharing (k, l) the same way.
// ---------------------------------------------------------------------

__global__ void __launch_bounds__(256) jk_rows_kernel(
    const int* __restrict__ shls, const int64_t* __restrict__ off, const int* __restrict__ gq,
    const int* __restrict__ gptr, const double* __restrict__ eri, const int* __restrict__ ao_loc,
    const double* __restrict__ D, int nao, double* __restrict__ Jp, double* __restrict__ Kp) {
  extern __shared__ double lds[];
  const int n0 = gptr[blockIdx.x], n1 = gptr[blockIdx.x + 1];
  const int q0 = gq[n0];
  const int ish = shls[4 * q0], jsh = shls[4 * q0 + 1];
  const int ai = ao_loc[ish], di = ao_loc[ish + 1] - ai;
  const int aj = ao_loc[jsh], dj = ao_loc[jsh + 1] - aj;
  const int64_t N = nao;
  double* Ki = lds;
  double* Kj = lds + di * N;
  double* red = lds + (di + dj) * N;
  const bool wantK = Kp != nullptr, wantJ = Jp != nullptr;
  if (wantK)
    for (int x = threadIdx.x; x < (di + dj) * nao; x += blockDim.x) lds[x] = 0.0;
  __syncthreads();
  const int nab = di * dj;
  const int S = blockDim.x / nab;
  const int ab = threadIdx.x % nab, s = threadIdx.x / nab;
  double jp = 0.0;
  if (s < S) {
    const int a = ab % di, b = ab / di;
    const int64_t P = ai + a, Q = aj + b;
    const double* __restrict__ DP = D + P * N;
    const double* __restrict__ DQ = D + Q * N;
    double* Kia = Ki + a * N;
    double* Kjb = Kj + b * N;
    for (int n = n0 + s; n < n1; n += S) {
      const int q = gq[n];
      const int k = shls[4 * q + 2], l = shls[4 * q + 3];
      const int ak = ao_loc[k], dk = ao_loc[k + 1] - ak;
      const int al = ao_loc[l], dl = ao_loc[l + 1] - al;
      const double f = jk_degeneracy(ish, jsh, k, l);
      const double* __restrict__ blk = eri + off[q] + ab;
      double jq = 0.0;
      for (int c = 0; c < dk; ++c) {
        const int64_t R = ak + c;
        const double* __restrict__ DR = D + R * N;
        const double dqr = DQ[R], dpr = DP[R];
        double tik = 0.0, tjk = 0.0;
        for (int d = 0; d < dl; ++d) {
          const int64_t Sx = al + d;
          const double v = blk[(int64_t)nab * (c + dk * d)];
          jq = fma(v, DR[Sx], jq);
          if (wantK) {
            tik = fma(v, DQ[Sx], tik);
            tjk = fma(v, DP[Sx], tjk);
            atomicAdd(Kia + Sx, f * v * dqr);
            atomicAdd(Kjb + Sx, f * v * dpr);
          }
        }
        if (wantK) {
          atomicAdd(Kia + R, f * tik);
          atomicAdd(Kjb + R, f * tjk);
        }
      }
      jp = fma(2.0 * f, jq, jp);
    }
  }
  red[threadIdx.x] = jp;
  __syncthreads();
  if (wantJ && threadIdx.x < nab) {
    double t = 0.0;
    for (int u = 0; u < S; ++u) t += red[u * nab + threadIdx.x];
    const int a = threadIdx.x % di, b = threadIdx.x / di;
    atomicAdd(Jp + (ai + a) * N + aj + b, t);
  }
  if (wantK) {
    for (int x = threadIdx.x; x < di * nao; x += blockDim.x) {
      const double v = Ki[x];
      if (v != 0.0) atomicAdd(Kp + (ai + x / nao) * N + x % nao, v);
    }
    for (int x = threadIdx.x; x < dj * nao; x += blockDim.x) {
      const double v = Kj[x];
      if (v != 0.0) atomicAdd(Kp + (aj + x / nao) * N + x % nao, v);
    }
  }
}

__global__ void __launch_bounds__(256) jk_cols_kernel(
    const int* __restrict__ shls, const int64_t* __restrict__ off, const int* __restrict__ gq,
    const int* __restrict__ gptr, const double* __restrict__ eri, const int* __restrict__ ao_loc,
    const double* __restrict__ D, int nao, double* __restrict__ Jp) {
  __shared__ double red[256];
  const int n0 = gptr[blockIdx.x], n1 = gptr[blockIdx.x + 1];
  const int q0 = gq[n0];
  const int ksh = shls[4 * q0 + 2], lsh = shls[4 * q0 + 3];
  const int ak = ao_loc[ksh], dk = ao_loc[ksh + 1] - ak;
  const int al = ao_loc[lsh], dl = ao_loc[lsh + 1] - al;
  const int64_t N = nao;
  const int ncd = dk * dl;
  const int S = blockDim.x / ncd;
  const int cd = threadIdx.x % ncd, s = threadIdx.x / ncd;
  double jp = 0.0;
  if (s < S) {
    for (int n = n0 + s; n < n1; n += S) {
      const int q = gq[n];
      const int i = shls[4 * q], j = shls[4 * q + 1];
      const int ai = ao_loc[i], di = ao_loc[i + 1] - ai;
      const int aj = ao_loc[j], dj = ao_loc[j + 1] - aj;
      const int nab = di * dj;
      const double* __restrict__ blk = eri + off[q] + (int64_t)nab * cd;
      double jq = 0.0;
      for (int b = 0; b < dj; ++b) {
        const double* __restrict__ DQ = D + (aj + b);
        for (int a = 0; a < di; ++a) jq = fma(blk[a + di * b], DQ[(ai + a) * N], jq);
      }
      jp = fma(2.0 * jk_degeneracy(i, j, ksh, lsh), jq, jp);
    }
  }
  red[threadIdx.x] = jp;
  __syncthreads();
  if (threadIdx.x < ncd) {
    double t = 0.0;
    for (int u = 0; u < S; ++u) t += red[u * ncd + threadIdx.x];
    const int c = threadIdx.x % dk, d = threadIdx.x / dk;
    atomicAdd(Jp + (ak + c) * N + al + d, t);
  }
}

// out = X + X^T over the full matrix (each unordered pair by one thread)
__global__ void zero_kernel(double* __restrict__ x, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = 0.0;
}

hipError_t launch_zero(double* x, int64_t n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, x, n);
  return hipGetLastError();
}

__global__ void symmetrize_kernel(const double* __restrict__ X, double* __restrict__ out, int n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * n) return;
  const int64_t i = t / n, j = t % n;
  out[t] = X[i * n + j] + X[j * n + i];
}

hipError_t launch_jk_grouped(const int* shls, const int64_t* off, const int* rows_q,
                             const int* rows_ptr, int nrows, size_t rows_lds, const int* cols_q,
                             const int* cols_ptr, int ncols, const double* eri, const int* ao_loc,
                             const double* D, int nao, double* Jp, double* Kp, hipStream_t stream) {
  if (nrows > 0) {
    hipLaunchKernelGGL(jk_rows_kernel, dim3(nrows), dim3(256), rows_lds, stream, shls, off, rows_q,
                       rows_ptr, eri, ao_loc, D, nao, Jp, Kp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (Jp && ncols > 0) {
    hipLaunchKernelGGL(jk_cols_kernel, dim3(ncols), dim3(256), 0, stream, shls, off, cols_q,
                       cols_ptr, eri, ao_loc, D, nao, Jp);
    return hipGetLastError();
  }
  return hipSuccess;
}

hipError_t launch_symmetrize(const double* X, double* out, int n, hipStream_t stream) {
  const int64_t tot = (int64_t)n * n;
  if (tot <= 0) return hipSuccess;
  hipLaunchKernelGGL(symmetrize_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream,
                     X, out, n);
  return hipGetLastError();
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
