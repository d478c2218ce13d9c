// Primitive-quartet walk of the one-workgroup-per-quartet kernels (the
// generic kernels of eri.hip and the compile-time wave kernels of
// eri_fast.h): screening, Rys roots and the 2D recurrence into LDS.
#pragma once
#include "device_common.h"

namespace mcint {

constexpr int GEN_LANES = 64;  // lanes of a generic-kernel workgroup (one wave)

// Primitive-quartet walk shared by the generic kernels: for every
// primitive quartet of task tk that passes libcint's screening
// (src/cint2e.c:57-269 with the optimizer's cutoffs; short range as in
// src/cint2e.c:358-380), the Rys roots and weights of the operator and the
// 2D recurrence tables V_d(e, f), e <= E, f <= F, of each root r and
// direction d into LDS at s_V + (d * nrt + r) * (E+1)(F+1); then
// consume(ai, ak, nrt) with the exponents of primitives i and k (the
// auxiliary shell's for three centres), after which the workgroup
// synchronises.  lij and lkl are the angular sums of the bra and ket with
// their derivative raises (the distance terms of the screening estimate).
// Returns libcint's not0: whether any primitive quartet contributed.
template <class Consume>
__device__ int gen_walk(const BasisView& B, const Task& tk, const VShell& s0, const VShell& s1,
                        const VShell& s2, const VShell& s3, int kind, int la, int lb, int lc,
                        int ld, int lij, int lkl, int ER, int FR, double* s_u, double* s_w,
                        double* s_V, Consume&& consume) {
  const int lane = threadIdx.x;
  const bool three = kind_three(kind);
  // (i|k): no screening at all (src/cint2c2e.c:57-147)
  const bool noscreen = kind == KIND_2C2E;
  const int nij = s0.nprim * s1.nprim;
  const int nkl = three ? s2.nprim : s2.nprim * s3.nprim;
  // records visited in libcint's order (ip fastest), so that the sums over
  // primitives round as libcint's do
  const int64_t oij = B.pair_off[(int64_t)tk.sh[0] * B.nv + tk.sh[1]];
  const int64_t okl = three ? 0 : B.pair_off[(int64_t)tk.sh[2] * B.nv + tk.sh[3]];
  const PairPrim* pij = B.pp + oij;
  const PairPrim* pkl = three ? nullptr : B.pp + okl;
  const int* pos_ij = B.pp_pos + oij;
  const int* pos_kl = B.pp_pos + okl;
  const double ld_ij = B.pair_ld[(int64_t)tk.sh[0] * B.nv + tk.sh[1]];
  const double ld_kl = three ? 0.0 : B.pair_ld[(int64_t)tk.sh[2] * B.nv + tk.sh[3]];
  const double E = three ? B.expcutoff_3c : B.expcutoff_4c;
  const int nr = (lij + lkl) / 2 + 1;  // rys_order
  const double omega = B.omega;
  const bool sr = omega < 0;
  // distance terms of the screening estimate cce = cce0 - l * log(d + 1);
  // short range widens them to log(d + theta_pair r_guess + 1) and raises
  // the loop cutoff (src/optimizer.c:301-315, src/cint2e.c:358-380,
  // src/cint3c2e.c:279-296; r_guess = 8)
  double dij_term = lij * ld_ij, dkl_term = lkl * ld_kl, Ecut = E;
  if (sr) {
    const double w2 = omega * omega;
    const double dab = sqrt((s0.r[0] - s1.r[0]) * (s0.r[0] - s1.r[0]) +
                            (s0.r[1] - s1.r[1]) * (s0.r[1] - s1.r[1]) +
                            (s0.r[2] - s1.r[2]) * (s0.r[2] - s1.r[2]));
    const double th_ij = w2 / (w2 + B.env[s0.pexp + s0.nprim - 1] + B.env[s1.pexp + s1.nprim - 1]);
    dij_term = lij > 0 ? lij * log(dab + th_ij * 8.0 + 1.0) : 0.0;
    if (nr > 1 && lij > 0) Ecut += lij * log((dab + th_ij * 8.0 + 1.0) / (dab + 1.0));
    if (three) {
      const double th_k = w2 / (w2 + B.env[s2.pexp + s2.nprim - 1]);
      if (nr > 1 && lkl > 0) Ecut += lkl * log(th_k * 8.0 + 1.0);
    } else {
      const double dcd = sqrt((s2.r[0] - s3.r[0]) * (s2.r[0] - s3.r[0]) +
                              (s2.r[1] - s3.r[1]) * (s2.r[1] - s3.r[1]) +
                              (s2.r[2] - s3.r[2]) * (s2.r[2] - s3.r[2]));
      const double th_kl = w2 / (w2 + B.env[s2.pexp + s2.nprim - 1] + B.env[s3.pexp + s3.nprim - 1]);
      dkl_term = lkl > 0 ? lkl * log(dcd + th_kl * 8.0 + 1.0) : 0.0;
      if (nr > 1 && lkl > 0) Ecut += lkl * log((dcd + th_kl * 8.0 + 1.0) / (dcd + 1.0));
    }
  }

  // pair emptiness (libcint returns early when a pair has no primitive
  // below the cutoff: src/cint2e.c:96-109)
  int live_ij = noscreen ? 1 : 0, live_kl = three ? 1 : 0;
  if (!noscreen)
    for (int p = lane; p < nij; p += GEN_LANES) live_ij |= (pij[p].cce0 - dij_term) < E;
  if (!three)
    for (int p = lane; p < nkl; p += GEN_LANES) live_kl |= (pkl[p].cce0 - dkl_term) < E;
  const int live = __syncthreads_and(__syncthreads_or(live_ij) && __syncthreads_or(live_kl));

  // short range: erfc = 1 - erf, i.e. the full rule on nr roots minus the
  // theta-scaled rule on nr more (libcint does this for rys_order <= 3,
  // src/g2e.c:4455-4476, and a dedicated rule on [sqrt(theta), 1] above;
  // both integrate the polynomial integrand exactly)
  const int nrt = sr ? 2 * nr : nr;
  const int VS = (ER + 1) * (FR + 1);
  const double cfac = PI3_2_SQRTPI * common_fac_sp(la) * common_fac_sp(lb) *
                      common_fac_sp(lc) * (three ? 1.0 : common_fac_sp(ld));
  const double Ax = s0.r[0], Ay = s0.r[1], Az = s0.r[2];
  const double Cx = s2.r[0], Cy = s2.r[1], Cz = s2.r[2];
  const double* ek = B.env + s2.pexp;
  const double* ck = B.env + s2.pcoef;
  int any = 0;

  if (live) {
    for (int p = 0; p < nij * nkl; ++p) {
      const int ij = p % nij;
      const int kl = p / nij;
      const PairPrim P = pij[pos_ij[ij]];
      double cce_kl, akl, qx, qy, qz, kkl;
      if (three) {
        cce_kl = 0;
        akl = ek[kl];
        qx = Cx; qy = Cy; qz = Cz;
        kkl = ck[kl];
      } else {
        const PairPrim Q = pkl[pos_kl[kl]];
        cce_kl = Q.cce0 - dkl_term;
        akl = Q.aij;
        qx = Q.px; qy = Q.py; qz = Q.pz;
        kkl = Q.kab;
        if (cce_kl > Ecut) continue;
      }
      const double cce_ij = noscreen ? -1e300 : P.cce0 - dij_term;
      if (cce_ij > (three ? Ecut : Ecut - cce_kl)) continue;

      const double aij = P.aij;
      const double a1 = aij * akl;
      const double a0 = a1 * rcp_pos(aij + akl);
      const double xpq = P.px - qx, ypq = P.py - qy, zpq = P.pz - qz;
      const double rr = xpq * xpq + ypq * ypq + zpq * zpq;
      double fac1 = sqrt_pos(a0 * rcp_pos(a1 * a1 * a1)) * cfac * P.kab * kkl;
      double x = a0 * rr;
      double theta = 1.0;
      if (sr) {
        // pairs beyond the unwidened cutoff carry no data (centre at 1e18),
        // and tiny erfc tails are dropped: src/g2e.c:4449-4453
        if (cce_ij >= E || cce_kl >= E) continue;
        theta = omega * omega * rcp_pos(omega * omega + a0);
        const double cutoff = (three ? Ecut : Ecut - cce_kl) - cce_ij;
        if (theta * x > cutoff || theta * x > 40.0) continue;  // EXPCUTOFF_SR
        any = 1;
      } else {
        any = 1;
        if (!noscreen && (cce_ij >= E || cce_kl >= E)) continue;  // stored exp factor is 0
        if (omega > 0) {
          theta = omega * omega * rcp_pos(omega * omega + a0);
          x *= theta;
          fac1 *= sqrt_pos(theta);
        }
      }
      if (lane < nrt) {
        double u, w;
        const bool lr_part = sr && lane >= nr;
        rys_root_one(nr, lr_part ? lane - nr : lane, lr_part ? x * theta : x, u, w);
        if (omega > 0 || lr_part) {
          const double ut = u * theta;
          u = ut * rcp_pos(u + 1.0 - ut);
        }
        if (lr_part) w *= -sqrt_pos(theta);
        s_u[lane] = u;
        s_w[lane] = w * fac1;
      }
      __syncthreads();
      if (lane < 3 * nrt) {
        const int d = lane / nrt;
        const int r = lane - d * nrt;
        const double u2 = a0 * s_u[r];
        const double tmp4 = 0.5 * rcp_pos(u2 * (aij + akl) + a1);
        const double tmp5 = u2 * tmp4;
        const double tmp1 = 2.0 * tmp5;
        const double b00 = tmp5;
        const double b10 = tmp5 + tmp4 * akl;
        const double b01 = tmp5 + tmp4 * aij;
        const double pq = d == 0 ? xpq : (d == 1 ? ypq : zpq);
        const double pa = d == 0 ? P.px - Ax : (d == 1 ? P.py - Ay : P.pz - Az);
        const double qc = d == 0 ? qx - Cx : (d == 1 ? qy - Cy : qz - Cz);
        const double c00 = pa - tmp1 * akl * pq;
        const double c0p = qc + tmp1 * aij * pq;
        double* V = s_V + (d * nrt + r) * VS;
        const int F1 = FR + 1;
        V[0] = d == 2 ? s_w[r] : 1.0;
        for (int e = 0; e < ER; ++e) {
          double v = c00 * V[e * F1];
          if (e > 0) v += e * b10 * V[(e - 1) * F1];
          V[(e + 1) * F1] = v;
        }
        for (int f = 0; f < FR; ++f) {
          for (int e = 0; e <= ER; ++e) {
            double v = c0p * V[e * F1 + f];
            if (f > 0) v += f * b01 * V[e * F1 + f - 1];
            if (e > 0) v += e * b00 * V[(e - 1) * F1 + f];
            V[e * F1 + f + 1] = v;
          }
        }
      }
      __syncthreads();
      consume(P.ai, three ? akl : pkl[pos_kl[kl]].ai, nrt);
      __syncthreads();
    }
  }
  return __syncthreads_or(any) && live;
}

}  // namespace mcint
