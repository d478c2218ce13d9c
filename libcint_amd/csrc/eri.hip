// Device code of libmcint: pair data, the generic Rys kernel (any class,
// plain / i-derivative / 3-centre), and the register-resident fast kernels.
//
// Algorithm (the mathematics of libcint's Rys path, src/g2e.c:4425-4545,
// reorganised for the GPU):
//   per primitive quartet: Rys roots (tables), the Rys coefficients
//   b00/b10/b01/c00/c0p, the 2D vertical recurrence I(e,f) with e built on
//   centre A and f on centre C (src/g2e.c:272-421), and the products
//   [e0|f0] = sum_roots Ix*Iy*Iz accumulated over primitives;
//   per contracted quartet: the horizontal transfers (a,b| <- [e0| and
//   |c,d) <- |f0] done once on contracted data (libcint does them per
//   primitive, src/g2e.c:428-693), the Cartesian -> spherical transform
//   (src/cart2sph.c:5324-5376) and the store in libcint's Fortran order.
#include "device_common.h"
#include "gen_walk.h"

namespace mcint {

// ---------------------------------------------------------------------
// Pair data: one block per ordered virtual-shell pair, one lane per
// primitive pair.  Restates CINTset_pairdata (src/optimizer.c:288-342).
// ---------------------------------------------------------------------
// One workgroup per shell pair.  The records are staged in 'tmp' and then
// stored in ascending order of cce0 (ties by primitive index), so that the
// primitive quartets passing the screening tests of one pair against a
// fixed partner form a prefix of the list.
__global__ void pairdata_kernel(const VShell* __restrict__ vsh, const double* __restrict__ env,
                                int nv, const int* __restrict__ pairs,
                                const int64_t* __restrict__ pair_off, double* __restrict__ pair_ld,
                                PairPrim* __restrict__ tmp, PairPrim* __restrict__ pp,
                                int* __restrict__ pos) {
  const int ia = pairs[2 * blockIdx.x];
  const int ib = pairs[2 * blockIdx.x + 1];
  const VShell a = vsh[ia];
  const VShell b = vsh[ib];
  const double dx = a.r[0] - b.r[0], dy = a.r[1] - b.r[1], dz = a.r[2] - b.r[2];
  const double rr = dx * dx + dy * dy + dz * dz;
  const double* ea = env + a.pexp;
  const double* eb = env + b.pexp;
  const double log_rr0 = 1.7 - 1.5 * log(ea[a.nprim - 1] + eb[b.nprim - 1]);
  if (threadIdx.x == 0) pair_ld[(int64_t)ia * nv + ib] = log(sqrt(rr) + 1.0);
  const int64_t base = pair_off[(int64_t)ia * nv + ib];
  const int n = a.nprim * b.nprim;
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    const int ip = p % a.nprim;
    const int jp = p / a.nprim;
    double mca = 0, mcb = 0;
    for (int c = 0; c < a.nctr; ++c) mca = fmax(mca, fabs(env[a.pcoef0 + c * a.nprim + ip]));
    for (int c = 0; c < b.nctr; ++c) mcb = fmax(mcb, fabs(env[b.pcoef0 + c * b.nprim + jp]));
    const double ai = ea[ip], aj = eb[jp];
    const double aij = ai + aj;
    const double inv = 1.0 / aij;
    const double e = rr * ai * aj * inv;
    PairPrim q;
    const double wj = aj * inv;
    q.px = a.r[0] + wj * (b.r[0] - a.r[0]);
    q.py = a.r[1] + wj * (b.r[1] - a.r[1]);
    q.pz = a.r[2] + wj * (b.r[2] - a.r[2]);
    q.aij = aij;
    q.kab = env[a.pcoef + ip] * env[b.pcoef + jp] * exp(-e);
    q.cce0 = e - log_rr0 - log(mca) - log(mcb);
    q.ai = ai;
    q.aj = aj;
    tmp[base + p] = q;
  }
  __syncthreads();
  // rank sort; a NaN key (never produced by finite inputs) sorts last
  auto key = [](double c) { return c == c ? c : __builtin_huge_val(); };
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    const double c = key(tmp[base + p].cce0);
    int rank = 0;
    for (int m = 0; m < n; ++m) {
      const double cm = key(tmp[base + m].cce0);
      rank += (cm < c) || (cm == c && m < p);
    }
    pp[base + rank] = tmp[base + p];
    pos[base + p] = rank;
  }
}

hipError_t launch_pairdata(const VShell* vsh, const double* env, int nv, const int* pairs,
                           int npairs, const int64_t* pair_off, double* pair_ld, PairPrim* tmp,
                           PairPrim* pp, int* pos, hipStream_t stream) {
  if (npairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(pairdata_kernel, dim3(npairs), dim3(64), 0, stream, vsh, env, nv, pairs,
                     pair_off, pair_ld, tmp, pp, pos);
  return hipGetLastError();
}

// ---------------------------------------------------------------------
// Output addressing shared by all kernels.  Canonical axis A holds the
// original axis (perm >> 2A) & 3; strides are those of the packed
// Fortran block (D0, D0*D1, ...) of the original shell order.
// ---------------------------------------------------------------------
struct OutStrides {
  int64_t s[4];
  int64_t comp;
};

__device__ __forceinline__ OutStrides out_strides(const VShell* vsh, const Task& tk, int perm,
                                                  int naxes, int sph) {
  int dorig[4] = {1, 1, 1, 1};
  for (int A = 0; A < naxes; ++A) {
    const VShell& v = vsh[tk.sh[A]];
    dorig[(perm >> (2 * A)) & 3] = sph ? v.dsph : v.dcart;
  }
  int64_t so[5];
  so[0] = 1;
  for (int o = 0; o < 4; ++o) so[o + 1] = so[o] * dorig[o];
  OutStrides st;
  for (int A = 0; A < 4; ++A) st.s[A] = A < naxes ? so[(perm >> (2 * A)) & 3] : 0;
  st.comp = so[4];
  return st;
}

// ---------------------------------------------------------------------
// Generic kernel: one 64-lane workgroup per contracted quartet.  Roots,
// recurrences and accumulation are spread over lanes through LDS and a
// per-task scratch block; used for every class without a fast kernel.
// ---------------------------------------------------------------------

__device__ __forceinline__ double binom_small(int n, int k) {
  double r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

__device__ __forceinline__ double ipow(double x, int k) {
  double r = 1;
  for (int i = 0; i < k; ++i) r *= x;
  return r;
}

// Horizontal transfer (a,b| = sum_t prod_d C(b_d,t_d) AB_d^(b_d-t_d) [a+t|
// for one a (cart exps ax..az) and b, reading column f of a set whose
// e-list starts at shell set_lo.
__device__ double hrr_bra(const double* __restrict__ X, int nF, int f, int set_lo, int ax, int ay,
                          int az, int bx, int by, int bz, const double* AB) {
  double s = 0;
  for (int tx = 0; tx <= bx; ++tx) {
    const double cx = binom_small(bx, tx) * ipow(AB[0], bx - tx);
    for (int ty = 0; ty <= by; ++ty) {
      const double cy = cx * binom_small(by, ty) * ipow(AB[1], by - ty);
      for (int tz = 0; tz <= bz; ++tz) {
        const double cz = cy * binom_small(bz, tz) * ipow(AB[2], bz - tz);
        const int ex = ax + tx, ey = ay + ty, ez = az + tz;
        const int le = ex + ey + ez;
        const int e = cart_sum(set_lo, le - 1) + cart_idx(le, ex, ez);
        s += cz * X[(int64_t)e * nF + f];
      }
    }
  }
  return s;
}

__global__ void __launch_bounds__(GEN_LANES)
eri_generic_kernel(BasisView B, BucketArgs A) {
  __shared__ double s_u[2 * NRYS_MAX], s_w[2 * NRYS_MAX];
  extern __shared__ double s_V[];  // 3 * roots * (E+1)(F+1), sized at launch
  __shared__ int s_ec[GEN_ECAP];
  __shared__ int s_fc[GEN_FCAP];

  const int64_t t = A.t0 + blockIdx.x;
  if (t >= A.t1) return;
  const int lane = threadIdx.x;
  const Task tk = A.tasks[t];
  const int kind = A.kind;
  const bool ip1 = kind_bra_der(kind);
  const bool ip2 = kind_ket_der(kind);
  const bool three = kind_three(kind);
  const int la = A.la, lb = A.lb, lc = A.lc, ld = three ? 0 : A.ld;
  const GenLayout G = gen_layout(kind, A.sph, la, lb, lc, ld);
  double* X = A.scratch + (t - A.t0) * A.scratch_stride;

  const VShell s0 = B.vsh[tk.sh[0]], s1 = B.vsh[tk.sh[1]], s2 = B.vsh[tk.sh[2]];
  VShell s3 = s2;
  if (!three) s3 = B.vsh[tk.sh[3]];
  const OutStrides OS = out_strides(B.vsh, tk, task_perm(tk), three ? 3 : 4, A.sph);
  double* out = A.out + task_out(tk);

  // component tables of the union e-list and the f-list
  for (int i = lane; i < G.nE; i += GEN_LANES) {
    int l = G.e_lo, k = i;
    while (k >= ncart(l)) { k -= ncart(l); ++l; }
    // k-th component of shell l in libcint order
    int m = 0;
    while ((m + 1) * (m + 2) / 2 <= k) ++m;
    const int lz = k - m * (m + 1) / 2;
    const int lx = l - m;
    const int ly = m - lz;
    s_ec[i] = lx | (ly << 4) | (lz << 8) | (l << 12);
  }
  for (int i = lane; i < G.nF; i += GEN_LANES) {
    int l = lc, k = i;
    while (k >= ncart(l)) { k -= ncart(l); ++l; }
    int m = 0;
    while ((m + 1) * (m + 2) / 2 <= k) ++m;
    const int lz = k - m * (m + 1) / 2;
    s_fc[i] = (l - m) | ((m - lz) << 4) | (lz << 8) | (l << 12);
  }
  for (int64_t i = lane; i < G.oH1; i += GEN_LANES) X[i] = 0;

  const int ncomp_g = G.nE * G.nF;
  const int VS = (G.E + 1) * (G.F + 1);
  auto consume = [&](double ai, double akl, int nrt) {
    const double wai = 2.0 * ai;
    if (ip2) {
      // (e | d/dr_comp k) = k_comp [e|k-1_comp] - 2 a_k [e|k+1_comp], per
      // primitive quartet (the auxiliary exponent a_k is akl here)
      const int F1 = G.F + 1;
      const double mak2 = -2.0 * akl;
      for (int c = lane; c < 3 * ncomp_g; c += GEN_LANES) {
        const int comp = c / ncomp_g;
        const int ie = (c - comp * ncomp_g) / G.nF;
        const int jf = c - comp * ncomp_g - ie * G.nF;
        const int ec = s_ec[ie], fc = s_fc[jf];
        const int e3[3] = {ec & 15, (ec >> 4) & 15, (ec >> 8) & 15};
        const int f3[3] = {fc & 15, (fc >> 4) & 15, (fc >> 8) & 15};
        double g = 0;
        for (int r = 0; r < nrt; ++r) {
          double prod = 1.0;
          for (int d = 0; d < 3; ++d) {
            const double* V = s_V + (d * nrt + r) * VS + e3[d] * F1;
            double v;
            if (d == comp) {
              v = mak2 * V[f3[d] + 1];
              if (f3[d] > 0) v += f3[d] * V[f3[d] - 1];
            } else {
              v = V[f3[d]];
            }
            prod *= v;
          }
          g += prod;
        }
        X[G.oX0 + ((int64_t)comp * G.NE0 + ie) * G.nF + jf] += g;
      }
      return;
    }
    for (int c = lane; c < ncomp_g; c += GEN_LANES) {
      const int ie = c / G.nF;
      const int jf = c - ie * G.nF;
      const int ec = s_ec[ie], fc = s_fc[jf];
      const int ex = ec & 15, ey = (ec >> 4) & 15, ez = (ec >> 8) & 15, le = ec >> 12;
      const int fx = fc & 15, fy = (fc >> 4) & 15, fz = (fc >> 8) & 15;
      const int F1 = G.F + 1;
      double g = 0;
      for (int r = 0; r < nrt; ++r) {
        g += s_V[(0 * nrt + r) * VS + ex * F1 + fx] * s_V[(1 * nrt + r) * VS + ey * F1 + fy] *
             s_V[(2 * nrt + r) * VS + ez * F1 + fz];
      }
      if (le >= G.p_lo && le <= G.p_hi) {
        const int e0 = cart_sum(G.p_lo, le - 1) + cart_idx(le, ex, ez);
        X[G.oX0 + (int64_t)e0 * G.nF + jf] += ip1 ? wai * g : g;
      }
      if (le >= G.m_lo && le <= G.m_hi) {
        const int e1 = cart_sum(G.m_lo, le - 1) + cart_idx(le, ex, ez);
        X[G.oXM + (int64_t)e1 * G.nF + jf] += g;
      }
    }
  };
  const int lij = la + lb + (ip1 ? 1 : 0);
  const int lkl = lc + ld + (ip2 ? 1 : 0);
  const int any = gen_walk(B, tk, s0, s1, s2, s3, kind, la, lb, lc, ld, lij, lkl, G.E, G.F, s_u,
                           s_w, s_V, consume);
  if (A.not0) A.not0[A.task_id ? A.task_id[t] : t] = any;

  if (!any) {
    if (A.dm) return;  // nothing to contract
    const int n = G.ncomp * G.nAo * G.nBo * G.nCo * (three ? 1 : G.nDo);
    for (int i = lane; i < n; i += GEN_LANES) {
      int k = i;
      const int a = k % G.nAo; k /= G.nAo;
      const int b = k % G.nBo; k /= G.nBo;
      const int c = k % G.nCo; k /= G.nCo;
      const int d = three ? 0 : k % G.nDo;
      const int comp = three ? k : k / G.nDo;
      out[comp * OS.comp + a * OS.s[0] + b * OS.s[1] + c * OS.s[2] + (three ? 0 : d * OS.s[3])] = 0;
    }
    return;
  }

  // pass 1: bra transfer (Cartesian), H1[comp][a][b][f]
  const double AB[3] = {s0.r[0] - s1.r[0], s0.r[1] - s1.r[1], s0.r[2] - s1.r[2]};
  {
    const int n = G.ncomp * G.nca * G.ncb * G.nF;
    for (int i = lane; i < n; i += GEN_LANES) {
      int k = i;
      const int f = k % G.nF; k /= G.nF;
      const int b = k % G.ncb; k /= G.ncb;
      const int a = k % G.nca;
      const int comp = k / G.nca;
      // cart exponents of a (shell la) and b (shell lb)
      int m = 0;
      while ((m + 1) * (m + 2) / 2 <= a) ++m;
      const int az = a - m * (m + 1) / 2, ax = la - m, ay = m - az;
      int mb = 0;
      while ((mb + 1) * (mb + 2) / 2 <= b) ++mb;
      const int bz = b - mb * (mb + 1) / 2, bx = lb - mb, by = mb - bz;
      double v;
      if (ip2) {
        v = hrr_bra(X + G.oX0 + (int64_t)comp * G.NE0 * G.nF, G.nF, f, G.p_lo, ax, ay, az, bx, by,
                    bz, AB);
      } else if (!ip1) {
        v = hrr_bra(X + G.oX0, G.nF, f, G.p_lo, ax, ay, az, bx, by, bz, AB);
      } else {
        // (d/dr_comp a b| = a_comp (a-1|b) - [2 a_i weighted] (a+1|b)
        const int dx = comp == 0, dy = comp == 1, dz = comp == 2;
        v = -hrr_bra(X + G.oX0, G.nF, f, G.p_lo, ax + dx, ay + dy, az + dz, bx, by, bz, AB);
        const int ac = comp == 0 ? ax : (comp == 1 ? ay : az);
        if (ac > 0)
          v += ac * hrr_bra(X + G.oXM, G.nF, f, G.m_lo, ax - dx, ay - dy, az - dz, bx, by, bz, AB);
      }
      X[G.oH1 + i] = v;
    }
  }
  __syncthreads();
  // pass 2: bra Cartesian -> spherical, H2[comp][ao][bo][f]
  {
    const int n = G.ncomp * G.nAo * G.nBo * G.nF;
    for (int i = lane; i < n; i += GEN_LANES) {
      int k = i;
      const int f = k % G.nF; k /= G.nF;
      const int bo = k % G.nBo; k /= G.nBo;
      const int ao = k % G.nAo;
      const int comp = k / G.nAo;
      const double* h1 = X + G.oH1 + (int64_t)comp * G.nca * G.ncb * G.nF;
      double v;
      if (!A.sph) {
        v = h1[((int64_t)ao * G.ncb + bo) * G.nF + f];
      } else {
        const double* ca = c2s_coef + c2s_offset[la] + ao * G.nca;
        const double* cb = c2s_coef + c2s_offset[lb] + bo * G.ncb;
        v = 0;
        for (int a = 0; a < G.nca; ++a) {
          if (ca[a] == 0.0) continue;
          double s = 0;
          for (int b = 0; b < G.ncb; ++b)
            if (cb[b] != 0.0) s += cb[b] * h1[((int64_t)a * G.ncb + b) * G.nF + f];
          v += ca[a] * s;
        }
      }
      X[G.oH2 + i] = v;
    }
  }
  __syncthreads();
  const double* H3;
  int n_c;  // Cartesian width of the ket before its spherical transform
  if (!three) {
    // pass 3: ket transfer (Cartesian), H3[comp][ao][bo][c][d]
    const double CD[3] = {s2.r[0] - s3.r[0], s2.r[1] - s3.r[1], s2.r[2] - s3.r[2]};
    const int n = G.ncomp * G.nAo * G.nBo * G.ncc * G.ncd;
    for (int i = lane; i < n; i += GEN_LANES) {
      int k = i;
      const int d = k % G.ncd; k /= G.ncd;
      const int c = k % G.ncc; k /= G.ncc;
      const int row = k;  // comp*nAo*nBo + ao*nBo + bo
      int m = 0;
      while ((m + 1) * (m + 2) / 2 <= c) ++m;
      const int cz = c - m * (m + 1) / 2, cx = lc - m, cy = m - cz;
      int md = 0;
      while ((md + 1) * (md + 2) / 2 <= d) ++md;
      const int dz = d - md * (md + 1) / 2, dx = ld - md, dy = md - dz;
      // ket transfer reads f-columns of one H2 row
      const double* h2 = X + G.oH2 + (int64_t)row * G.nF;
      double s = 0;
      for (int tx = 0; tx <= dx; ++tx) {
        const double wx = binom_small(dx, tx) * ipow(CD[0], dx - tx);
        for (int ty = 0; ty <= dy; ++ty) {
          const double wy = wx * binom_small(dy, ty) * ipow(CD[1], dy - ty);
          for (int tz = 0; tz <= dz; ++tz) {
            const double wz = wy * binom_small(dz, tz) * ipow(CD[2], dz - tz);
            const int fx = cx + tx, fy = cy + ty, fz = cz + tz;
            const int lf = fx + fy + fz;
            const int f = cart_sum(lc, lf - 1) + cart_idx(lf, fx, fz);
            s += wz * h2[f];
          }
        }
      }
      X[G.oH3 + i] = s;
    }
    __syncthreads();
    H3 = X + G.oH3;
    n_c = G.ncc * G.ncd;
  } else {
    H3 = X + G.oH2;
    n_c = G.nF;  // = ncart(lc)
  }
  // pass 4: ket Cartesian -> spherical, then the store (or, for a gradient
  // digest, the contraction with the density)
  {
    const int nD = three ? 1 : G.nDo;
    const int n = G.ncomp * G.nAo * G.nBo * G.nCo * nD;
    // J/K digest of a plain four-centre block (six block updates of
    // jk_degeneracy, device_common.h; one atomic per element and update:
    // the generic path carries few plain classes), or the gradient digest
    const bool jk = A.jk_j != nullptr || A.jk_k != nullptr;
    const bool digest = A.dm != nullptr && !jk;
    const double fjk = jk ? jk_degeneracy(s0.shell, s1.shell, s2.shell, s3.shell) : 0.0;
    double gpart[3] = {0.0, 0.0, 0.0};
    const double fkl = (!three && s2.shell != s3.shell) ? 2.0 : 1.0;
    for (int i = lane; i < n; i += GEN_LANES) {
      int k = i;
      const int ao = k % G.nAo; k /= G.nAo;
      const int bo = k % G.nBo; k /= G.nBo;
      const int co = k % G.nCo; k /= G.nCo;
      const int dd = k % nD;
      const int comp = k / nD;
      const int row = (comp * G.nAo + ao) * G.nBo + bo;
      const double* h3 = H3 + (int64_t)row * n_c;
      double v;
      if (!A.sph) {
        v = three ? h3[co] : h3[co * G.ncd + dd];
      } else if (three) {
        const double* cc = c2s_coef + c2s_offset[lc] + co * G.ncc;
        v = 0;
        for (int c = 0; c < G.ncc; ++c)
          if (cc[c] != 0.0) v += cc[c] * h3[c];
      } else {
        const double* cc = c2s_coef + c2s_offset[lc] + co * G.ncc;
        const double* cd = c2s_coef + c2s_offset[ld] + dd * G.ncd;
        v = 0;
        for (int c = 0; c < G.ncc; ++c) {
          if (cc[c] == 0.0) continue;
          double s = 0;
          for (int d = 0; d < G.ncd; ++d)
            if (cd[d] != 0.0) s += cd[d] * h3[c * G.ncd + d];
          v += cc[c] * s;
        }
      }
      if (jk) {
        const int64_t N = A.nao;
        const int64_t P = s0.ao + ao, Q = s1.ao + bo, R = s2.ao + co, S = s3.ao + dd;
        const double* D = A.dm;
        if (A.jk_j) {
          atomicAdd(A.jk_j + P * N + Q, 2.0 * fjk * v * D[R * N + S]);
          atomicAdd(A.jk_j + R * N + S, 2.0 * fjk * v * D[P * N + Q]);
        }
        if (A.jk_k) {
          atomicAdd(A.jk_k + P * N + R, fjk * v * D[Q * N + S]);
          atomicAdd(A.jk_k + P * N + S, fjk * v * D[Q * N + R]);
          atomicAdd(A.jk_k + Q * N + R, fjk * v * D[P * N + S]);
          atomicAdd(A.jk_k + Q * N + S, fjk * v * D[P * N + R]);
        }
      } else if (digest) {
        // dE/dA = sum (d/dr i j|kl) [-2 D_ij D_kl + (D_ik D_jl + D_il D_jk)/2], times 2
        // when k and l are different shells (the list holds k >= l only)
        const int64_t N = A.nao;
        const int64_t P = s0.ao + ao, Q = s1.ao + bo, R = s2.ao + co, S = s3.ao + dd;
        const double* D = A.dm;
        const double w = fkl * (-2.0 * D[P * N + Q] * D[R * N + S] +
                                0.5 * (D[P * N + R] * D[Q * N + S] + D[P * N + S] * D[Q * N + R]));
        gpart[comp] += v * w;
      } else {
        out[comp * OS.comp + ao * OS.s[0] + bo * OS.s[1] + co * OS.s[2] +
            (three ? 0 : dd * OS.s[3])] = v;
      }
    }
    if (digest) {
      double* gs = A.grad + (int64_t)(blockIdx.x % A.nslot) * A.grad_stride + 3 * s0.atom;
      for (int c = 0; c < 3; ++c) {
        double g = gpart[c];
        for (int off = 1; off < GEN_LANES; off <<= 1) g += __shfl_xor(g, off);
        if (lane == 0) atomicAdd(gs + c, g);
      }
    }
  }
}

// ---------------------------------------------------------------------
// Second-derivative kernel (int2e_ipip1, int2e_ip1ip2): one 64-lane
// workgroup per contracted quartet, runtime class.  The derivative of a
// Cartesian Gaussian of exponent a along axis p is
//   D_p |a> = a_p |a - 1_p> - 2a |a + 1_p>        (src/g2e.c:4550-4697),
// so primitive sums are kept in row sets weighted by powers of 2a and the
// derivatives are formed after contraction, before the transfers:
//   ipip1  (bra sets la-2.., la.., la+2.. weighted 1, 2a_i, 4a_i^2)
//     D_p D_q |a> = a_p a_q |a-1_p-1_q> - a_q 2a|a+1_p-1_q> - a_p 2a|a-1_p+1_q>
//                   + 4a^2 |a+1_p+1_q>                               (p != q)
//     D_p D_p |a> = a_p (a_p-1) |a-2_p> - (2a_p+1) 2a|a> + 4a^2 |a+2_p>
//   ip1ip2 (bra sets la-1.., la+1.. weighted 1, 2a_i; ket sets lc-1..,
//     lc+1.. weighted 1, 2a_k): D_p on the bra, D_q on the ket.
// Component 3p+q of the output as libcint orders it
// (src/autocode/hess.c:1125-1174, 1289-1338).
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(GEN_LANES)
eri_hess_kernel(BasisView B, BucketArgs A) {
  __shared__ double s_u[2 * NRYS_MAX], s_w[2 * NRYS_MAX];
  extern __shared__ double s_V[];  // 3 * roots * (E+1)(F+1), sized at launch
  __shared__ int s_ec[GEN_ECAP];
  __shared__ int s_fc[GEN_FCAP];

  const int64_t t = A.t0 + blockIdx.x;
  if (t >= A.t1) return;
  const int lane = threadIdx.x;
  const Task tk = A.tasks[t];
  const int kind = A.kind;
  const bool ipip = kind == KIND_2E_IPIP1;
  const int la = A.la, lb = A.lb, lc = A.lc, ld = A.ld;
  const GenLayout G = gen_layout(kind, A.sph, la, lb, lc, ld);
  double* X = A.scratch + (t - A.t0) * A.scratch_stride;
  const VShell s0 = B.vsh[tk.sh[0]], s1 = B.vsh[tk.sh[1]], s2 = B.vsh[tk.sh[2]],
               s3 = B.vsh[tk.sh[3]];
  const OutStrides OS = out_strides(B.vsh, tk, task_perm(tk), 4, A.sph);
  double* out = A.out + task_out(tk);

  // Cartesian exponents of the union e-list (shells e_lo..E) and f-list
  for (int i = lane; i < G.nE + G.nF; i += GEN_LANES) {
    const bool isf = i >= G.nE;
    int l = isf ? G.f_lo : G.e_lo, k = isf ? i - G.nE : i;
    while (k >= ncart(l)) { k -= ncart(l); ++l; }
    int m = 0;
    while ((m + 1) * (m + 2) / 2 <= k) ++m;
    const int lz = k - m * (m + 1) / 2;
    const int code = (l - m) | ((m - lz) << 4) | (lz << 8) | (l << 12);
    if (isf) s_fc[i - G.nE] = code;
    else s_ec[i] = code;
  }
  for (int64_t i = lane; i < G.oH1; i += GEN_LANES) X[i] = 0;

  const int VS = (G.E + 1) * (G.F + 1);
  const int F1 = G.F + 1;
  auto consume = [&](double ai, double ak, int nrt) {
    const double wb[3] = {1.0, 2.0 * ai, 4.0 * ai * ai};
    const double wk[2] = {1.0, 2.0 * ak};
    for (int c = lane; c < G.nE * G.nF; c += GEN_LANES) {
      const int ie = c / G.nF;
      const int jf = c - ie * G.nF;
      const int ec = s_ec[ie], fc = s_fc[jf];
      const int ex = ec & 15, ey = (ec >> 4) & 15, ez = (ec >> 8) & 15, le = ec >> 12;
      const int fx = fc & 15, fy = (fc >> 4) & 15, fz = (fc >> 8) & 15, lf = fc >> 12;
      double g = 0;
      for (int r = 0; r < nrt; ++r) {
        g += s_V[(0 * nrt + r) * VS + ex * F1 + fx] * s_V[(1 * nrt + r) * VS + ey * F1 + fy] *
             s_V[(2 * nrt + r) * VS + ez * F1 + fz];
      }
      for (int u = 0; u < 2; ++u) {
        if (G.kset_lo[u] < 0 || lf < G.kset_lo[u] || lf > G.kset_hi[u]) continue;
        const int fu = G.cF[u] + cart_sum(G.kset_lo[u], lf - 1) + cart_idx(lf, fx, fz);
        const double gu = g * wk[u];
        for (int r = 0; r < 3; ++r) {
          if (G.bset_lo[r] < 0 || le < G.bset_lo[r] || le > G.bset_hi[r]) continue;
          const int er = cart_sum(G.bset_lo[r], le - 1) + cart_idx(le, ex, ez);
          X[G.oXs[r] + (int64_t)er * G.KF + fu] += gu * wb[r];
        }
      }
    }
  };
  const int any = gen_walk(B, tk, s0, s1, s2, s3, kind, la, lb, lc, ld,
                           la + lb + kind_bra_raise(kind), lc + ld + kind_ket_raise(kind), G.E,
                           G.F, s_u, s_w, s_V, consume);
  if (A.not0) A.not0[A.task_id ? A.task_id[t] : t] = any;

  if (!any) {
    const int n = 9 * G.nAo * G.nBo * G.nCo * G.nDo;
    for (int i = lane; i < n; i += GEN_LANES) {
      int k = i;
      const int a = k % G.nAo; k /= G.nAo;
      const int b = k % G.nBo; k /= G.nBo;
      const int c = k % G.nCo; k /= G.nCo;
      const int d = k % G.nDo;
      const int comp = k / G.nDo;
      out[comp * OS.comp + a * OS.s[0] + b * OS.s[1] + c * OS.s[2] + d * OS.s[3]] = 0;
    }
    return;
  }

  // pass 1: bra derivatives and transfer (Cartesian), H1[bc][a][b][kf] over
  // every ket column kf; bc = 3p+q (ipip1) or p (ip1ip2)
  const double AB[3] = {s0.r[0] - s1.r[0], s0.r[1] - s1.r[1], s0.r[2] - s1.r[2]};
  {
    const int n = G.nbc * G.nca * G.ncb * G.KF;
    for (int i = lane; i < n; i += GEN_LANES) {
      int k = i;
      const int kf = k % G.KF; k /= G.KF;
      const int b = k % G.ncb; k /= G.ncb;
      const int a = k % G.nca;
      const int bc = k / G.nca;
      int m = 0;
      while ((m + 1) * (m + 2) / 2 <= a) ++m;
      const int az = a - m * (m + 1) / 2;
      int mb = 0;
      while ((mb + 1) * (mb + 2) / 2 <= b) ++mb;
      const int bz = b - mb * (mb + 1) / 2, bx = lb - mb, by = mb - bz;
      const int av[3] = {la - m, m - az, az};
      // transfer of set r at bra exponents av + dv
      auto H = [&](int r, int dx, int dy, int dz) {
        return hrr_bra(X + G.oXs[r], G.KF, kf, G.bset_lo[r], av[0] + dx, av[1] + dy, av[2] + dz,
                       bx, by, bz, AB);
      };
      const int p = ipip ? bc / 3 : bc;
      const int ep[3] = {p == 0, p == 1, p == 2};
      const int ap = av[p];
      double v;
      if (!ipip) {
        v = -H(1, ep[0], ep[1], ep[2]);
        if (ap > 0) v += ap * H(0, -ep[0], -ep[1], -ep[2]);
      } else {
        const int q = bc - 3 * p;
        const int eq[3] = {q == 0, q == 1, q == 2};
        const int aq = av[q];
        if (p == q) {
          v = H(2, 2 * ep[0], 2 * ep[1], 2 * ep[2]) - (2 * ap + 1) * H(1, 0, 0, 0);
          if (ap > 1) v += ap * (ap - 1) * H(0, -2 * ep[0], -2 * ep[1], -2 * ep[2]);
        } else {
          v = H(2, ep[0] + eq[0], ep[1] + eq[1], ep[2] + eq[2]);
          if (aq > 0) v -= aq * H(1, ep[0] - eq[0], ep[1] - eq[1], ep[2] - eq[2]);
          if (ap > 0) v -= ap * H(1, eq[0] - ep[0], eq[1] - ep[1], eq[2] - ep[2]);
          if (ap > 0 && aq > 0)
            v += ap * aq * H(0, -ep[0] - eq[0], -ep[1] - eq[1], -ep[2] - eq[2]);
        }
      }
      X[G.oH1 + i] = v;
    }
  }
  __syncthreads();
  // pass 2: bra Cartesian -> spherical, H2[bc][ao][bo][kf]
  {
    const int n = G.nbc * G.nAo * G.nBo * G.KF;
    for (int i = lane; i < n; i += GEN_LANES) {
      int k = i;
      const int kf = k % G.KF; k /= G.KF;
      const int bo = k % G.nBo; k /= G.nBo;
      const int ao = k % G.nAo;
      const int bc = k / G.nAo;
      const double* h1 = X + G.oH1 + (int64_t)bc * G.nca * G.ncb * G.KF;
      double v;
      if (!A.sph) {
        v = h1[((int64_t)ao * G.ncb + bo) * G.KF + kf];
      } else {
        const double* ca = c2s_coef + c2s_offset[la] + ao * G.nca;
        const double* cb = c2s_coef + c2s_offset[lb] + bo * G.ncb;
        v = 0;
        for (int a = 0; a < G.nca; ++a) {
          if (ca[a] == 0.0) continue;
          double s = 0;
          for (int b = 0; b < G.ncb; ++b)
            if (cb[b] != 0.0) s += cb[b] * h1[((int64_t)a * G.ncb + b) * G.KF + kf];
          v += ca[a] * s;
        }
      }
      X[G.oH2 + i] = v;
    }
  }
  __syncthreads();
  // pass 3: ket derivative (ip1ip2) and transfer, H3[comp][ao][bo][c][d]
  const double CD[3] = {s2.r[0] - s3.r[0], s2.r[1] - s3.r[1], s2.r[2] - s3.r[2]};
  {
    const int n = 9 * G.nAo * G.nBo * G.ncc * G.ncd;
    for (int i = lane; i < n; i += GEN_LANES) {
      int k = i;
      const int d = k % G.ncd; k /= G.ncd;
      const int c = k % G.ncc; k /= G.ncc;
      const int bo = k % G.nBo; k /= G.nBo;
      const int ao = k % G.nAo;
      const int comp = k / G.nAo;
      int m = 0;
      while ((m + 1) * (m + 2) / 2 <= c) ++m;
      const int cz = c - m * (m + 1) / 2;
      const int cv[3] = {lc - m, m - cz, cz};
      int md = 0;
      while ((md + 1) * (md + 2) / 2 <= d) ++md;
      const int dz = d - md * (md + 1) / 2, dx = ld - md, dy = md - dz;
      const int bc = ipip ? comp : comp / 3;
      const double* h2 = X + G.oH2 + ((int64_t)(bc * G.nAo + ao) * G.nBo + bo) * G.KF;
      // transfer of ket set u at ket exponents cv + dv
      auto K = [&](int u, int ox, int oy, int oz) {
        return hrr_bra(h2 + G.cF[u], 1, 0, G.kset_lo[u], cv[0] + ox, cv[1] + oy, cv[2] + oz, dx,
                       dy, dz, CD);
      };
      double v;
      if (ipip) {
        v = K(0, 0, 0, 0);
      } else {
        const int q = comp - 3 * bc;
        const int eq[3] = {q == 0, q == 1, q == 2};
        v = -K(1, eq[0], eq[1], eq[2]);
        if (cv[q] > 0) v += cv[q] * K(0, -eq[0], -eq[1], -eq[2]);
      }
      X[G.oH3 + i] = v;
    }
  }
  __syncthreads();
  // pass 4: ket Cartesian -> spherical and the store
  {
    const int n = 9 * G.nAo * G.nBo * G.nCo * G.nDo;
    const int n_c = G.ncc * G.ncd;
    for (int i = lane; i < n; i += GEN_LANES) {
      int k = i;
      const int ao = k % G.nAo; k /= G.nAo;
      const int bo = k % G.nBo; k /= G.nBo;
      const int co = k % G.nCo; k /= G.nCo;
      const int dd = k % G.nDo;
      const int comp = k / G.nDo;
      const double* h3 = X + G.oH3 + ((int64_t)(comp * G.nAo + ao) * G.nBo + bo) * n_c;
      double v;
      if (!A.sph) {
        v = h3[co * G.ncd + dd];
      } else {
        const double* cc = c2s_coef + c2s_offset[lc] + co * G.ncc;
        const double* cd = c2s_coef + c2s_offset[ld] + dd * G.ncd;
        v = 0;
        for (int c = 0; c < G.ncc; ++c) {
          if (cc[c] == 0.0) continue;
          double s = 0;
          for (int d = 0; d < G.ncd; ++d)
            if (cd[d] != 0.0) s += cd[d] * h3[c * G.ncd + d];
          v += cc[c] * s;
        }
      }
      out[comp * OS.comp + ao * OS.s[0] + bo * OS.s[1] + co * OS.s[2] + dd * OS.s[3]] = v;
    }
  }
}

__global__ void rys_batch_kernel(int n, const double* __restrict__ x, int64_t count,
                                 double* __restrict__ u, double* __restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count * n) return;
  const int64_t k = i / n;
  const int r = (int)(i - k * n);
  double uu, ww;
  rys_root_one(n, r, x[k], uu, ww);
  u[k * n + r] = uu;
  w[k * n + r] = ww;
}

hipError_t launch_rys_batch(int n, const double* x, int64_t count, double* u, double* w,
                            hipStream_t stream) {
  const int64_t tot = count * n;
  if (tot <= 0) return hipSuccess;
  hipLaunchKernelGGL(rys_batch_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream,
                     n, x, count, u, w);
  return hipGetLastError();
}

// Counts primitive quartets that pass libcint's screening (one lane per
// task); used by the benchmark's algorithmic FLOP model.
__global__ void count_prims_kernel(BasisView B, BucketArgs A, unsigned long long* counter) {
  const int64_t t = A.t0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= A.t1) return;
  const Task tk = A.tasks[t];
  const bool three = kind_three(A.kind);
  const bool noscreen = A.kind == KIND_2C2E;
  const VShell s0 = B.vsh[tk.sh[0]], s1 = B.vsh[tk.sh[1]], s2 = B.vsh[tk.sh[2]];
  const int nij = s0.nprim * s1.nprim;
  const int nkl = three ? s2.nprim : s2.nprim * B.vsh[tk.sh[3]].nprim;
  const PairPrim* pij = B.pp + B.pair_off[(int64_t)tk.sh[0] * B.nv + tk.sh[1]];
  const PairPrim* pkl = three ? nullptr : B.pp + B.pair_off[(int64_t)tk.sh[2] * B.nv + tk.sh[3]];
  const double ld_ij = B.pair_ld[(int64_t)tk.sh[0] * B.nv + tk.sh[1]] *
                       (A.la + A.lb + kind_bra_raise(A.kind));
  const double ld_kl = three ? 0.0
                             : B.pair_ld[(int64_t)tk.sh[2] * B.nv + tk.sh[3]] *
                                   (A.lc + A.ld + kind_ket_raise(A.kind));
  const double E = three ? B.expcutoff_3c : B.expcutoff_4c;
  if (noscreen) {
    atomicAdd(counter, (unsigned long long)nij * nkl);
    return;
  }
  int live_ij = 0, live_kl = three;
  for (int p = 0; p < nij; ++p) live_ij |= (pij[p].cce0 - ld_ij) < E;
  if (!three)
    for (int p = 0; p < nkl; ++p) live_kl |= (pkl[p].cce0 - ld_kl) < E;
  if (!(live_ij && live_kl)) return;
  unsigned long long n = 0;
  for (int kl = 0; kl < nkl; ++kl) {
    const double ckl = three ? 0.0 : pkl[kl].cce0 - ld_kl;
    if (ckl >= E) continue;
    for (int ij = 0; ij < nij; ++ij) {
      const double cij = pij[ij].cce0 - ld_ij;
      if (cij < E && cij <= (three ? E : E - ckl)) ++n;
    }
  }
  atomicAdd(counter, n);
}

hipError_t launch_count_prims(const BasisView& B, const BucketArgs& A, unsigned long long* counter,
                              hipStream_t stream) {
  const int64_t n = A.t1 - A.t0;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(count_prims_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, B,
                     A, counter);
  return hipGetLastError();
}

hipError_t launch_generic(const BasisView& B, const BucketArgs& A, hipStream_t stream) {
  const int64_t n = A.t1 - A.t0;
  if (n <= 0) return hipSuccess;
  // grid.x is limited to 2^31-1 workgroups; callers split larger buckets
  const GenLayout G = gen_layout(A.kind, A.sph, A.la, A.lb, A.lc, A.ld);
  const int nr = (A.la + A.lb + A.lc + (kind_three(A.kind) ? 0 : A.ld) + kind_bra_raise(A.kind) +
                  kind_ket_raise(A.kind)) / 2 + 1;
  const int nrt = B.omega < 0 ? 2 * nr : nr;
  const size_t lds = sizeof(double) * 3 * nrt * (G.E + 1) * (G.F + 1);
  if (lds > sizeof(double) * GEN_VCAP || G.nE > GEN_ECAP || G.nF > GEN_FCAP)
    return hipErrorInvalidValue;
  if (A.scratch_stride < G.total) return hipErrorInvalidValue;
  if (kind_hess(A.kind))
    hipLaunchKernelGGL(eri_hess_kernel, dim3((unsigned)n), dim3(GEN_LANES), lds, stream, B, A);
  else
    hipLaunchKernelGGL(eri_generic_kernel, dim3((unsigned)n), dim3(GEN_LANES), lds, stream, B, A);
  return hipGetLastError();
}

}  // namespace mcint

#include "eri_fast.h"

namespace mcint {
int g_wave_kernels = 1;
int8_t g_class_wave[NCLS * 7] = {};
#define MC_WAVE_HAS(a, b, c, d) \
  if (la == a && lb == b && lc == c && ld == d) return WaveCfg<a, b, c, d>::OK ? 1 : 0;
int wave_kernel_exists(int la, int lb, int lc, int ld) {
  MCINT_CHUNK_CLASSES(MC_WAVE_HAS)
  return 0;
}
#undef MC_WAVE_HAS
MCINT_FAST_DISPATCH
}  // namespace mcint
