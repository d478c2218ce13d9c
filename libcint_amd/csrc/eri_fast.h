// Register-resident Rys kernels for the classes whose primitive-contracted
// [e0|f0] block fits in registers (templated on the canonical class).
//
// Lanes -> work: a team of T lanes (T a power of two, chosen per bucket from
// the primitive-quartet count) owns one contracted quartet; lane t of the
// team takes primitive quartets t, t+T, ...  Per primitive quartet a lane
// evaluates the NR Rys roots, runs the 2D recurrence root by root in
// registers and accumulates [e0|f0] (e on centre A, f on centre C) into a
// register array.  The team then sums its lanes with cross-lane shuffles and
// every lane runs the contracted-data epilogue (horizontal transfers,
// Cartesian->spherical, store), each lane writing every T-th element.
#pragma once
#include <type_traits>
#include <utility>

#include "gen_walk.h"

namespace mcint {

namespace ct {
#define C2S_QUAL constexpr
#include "c2s_tables.inc"
#undef C2S_QUAL
}  // namespace ct

#include "chunk_order.inc"

// ---- compile-time component bookkeeping --------------------------------
__host__ __device__ constexpr int csum(int lo, int hi) {
  int s = 0;
  for (int l = lo; l <= hi; ++l) s += (l + 1) * (l + 2) / 2;
  return s;
}
// shell and local index of component k of the list of shells lo, lo+1, ...
__host__ __device__ constexpr int comp_shell(int lo, int k) {
  int l = lo;
  while (k >= (l + 1) * (l + 2) / 2) {
    k -= (l + 1) * (l + 2) / 2;
    ++l;
  }
  return l;
}
__host__ __device__ constexpr int comp_local(int lo, int k) {
  int l = lo;
  while (k >= (l + 1) * (l + 2) / 2) {
    k -= (l + 1) * (l + 2) / 2;
    ++l;
  }
  return k;
}
__host__ __device__ constexpr int local_m(int k) {
  int m = 0;
  while ((m + 1) * (m + 2) / 2 <= k) ++m;
  return m;
}
// Cartesian exponents (libcint order: lx descending, then ly descending)
__host__ __device__ constexpr int cpx(int lo, int k) {
  return comp_shell(lo, k) - local_m(comp_local(lo, k));
}
__host__ __device__ constexpr int cpz(int lo, int k) {
  const int kk = comp_local(lo, k);
  const int m = local_m(kk);
  return kk - m * (m + 1) / 2;
}
__host__ __device__ constexpr int cpy(int lo, int k) {
  return comp_shell(lo, k) - cpx(lo, k) - cpz(lo, k);
}
// index of (x,y,z) in the list of shells lo, lo+1, ...
__host__ __device__ constexpr int cindex(int lo, int x, int y, int z) {
  const int l = x + y + z;
  const int m = l - x;
  return csum(lo, l - 1) + m * (m + 1) / 2 + z;
}
__host__ __device__ constexpr double binomc(int n, int k) {
  double r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}
// Cartesian -> spherical coefficient; identity for l <= 1
__host__ __device__ constexpr double c2sc(int l, int m, int k) {
  return l <= 1 ? (m == k ? 1.0 : 0.0) : ct::c2s_coef[ct::c2s_offset[l] + m * ((l + 1) * (l + 2) / 2) + k];
}

// Class constants.  DER = 0: plain ERIs, rows e = shells LA..LA+LB.
// DER = 1: i-derivative (libcint's int2e_ip1, src/autocode/grad2.c:19-68):
// the recurrence runs to E = LA+LB+1 and the rows are two sets, the raised
// set (shells LA+1..E, each primitive weighted by 2 a_i) followed by the
// lowered set (shells LA-1..E-2, absent for LA = 0), from which the
// epilogue forms (d/dA_x a b| = a_x (a-1_x, b| - 2 a_i (a+1_x, b|.
// DER = 2: three-centre (ij|k) = (ij|k s0) (src/g3c2e.c:21-132): the aux
// shell k sits on C with LC = l_k, LD = 0 is an s function of exponent 0.
// DER = 3: the same with the short-range operator erfc(|omega| r)/r.
template <int LA, int LB, int LC, int LD, int DER = 0>
struct FastClass {
  static constexpr int D1 = DER == 1 ? 1 : 0;
  static constexpr int E = LA + LB + D1;
  static constexpr int F = LC + LD;
  static constexpr int NR = (E + F) / 2 + 1;
  static constexpr int NEP = D1 ? csum(LA + 1, E) : csum(LA, E);
  static constexpr int NEM = (D1 && LA > 0) ? csum(LA - 1, E - 2) : 0;
  static constexpr int NE = NEP + NEM;
  static constexpr int NF = csum(LC, F);
  static constexpr int NA = (LA + 1) * (LA + 2) / 2, NB = (LB + 1) * (LB + 2) / 2;
  static constexpr int NC = (LC + 1) * (LC + 2) / 2, ND = (LD + 1) * (LD + 2) / 2;
  static constexpr int NAo = 2 * LA + 1, NBo = 2 * LB + 1, NCo = 2 * LC + 1, NDo = 2 * LD + 1;
  // first shell of the list that row i belongs to, and its index there
  static constexpr int row_lo(int i) { return D1 ? (i < NEP ? LA + 1 : LA - 1) : LA; }
  static constexpr int row_k(int i) { return i < NEP ? i : i - NEP; }
  static constexpr bool row_scaled(int i) { return D1 && i < NEP; }
};

// Output functions of shell l: spherical (2l+1, real solid harmonics) or
// Cartesian (the identity transform).
constexpr int nout(int l, bool sph) { return sph ? 2 * l + 1 : (l + 1) * (l + 2) / 2; }
constexpr double c2x(bool sph, int l, int o, int k) {
  return sph ? c2sc(l, o, k) : (o == k ? 1.0 : 0.0);
}

// Compile-time loop: f(std::integral_constant<int, 0..N-1>) expanded in place,
// so every index derived from the loop variable is a constant expression.
template <typename F, int... Is>
__device__ __forceinline__ void sfor_impl(F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  sfor_impl(static_cast<F&&>(f), std::make_integer_sequence<int, N>{});
}

// 2D recurrence for one root and one Cartesian direction:
//   V(e+1,0) = c00 V(e,0) + e b10 V(e-1,0)
//   V(e,f+1) = c0p V(e,f) + f b01 V(e,f-1) + e b00 V(e-1,f)
template <int E, int F, typename R>
__device__ __forceinline__ void vrr2d(R (&V)[E + 1][F + 1], R v0, R c00, R c0p, R b10, R b01,
                                      R b00) {
  V[0][0] = v0;
  sfor<E>([&](auto EI) {
    constexpr int e = decltype(EI)::value;
    R v = c00 * V[e][0];
    if constexpr (e > 0) v = fma((R)e * b10, V[e - 1][0], v);
    V[e + 1][0] = v;
  });
  sfor<F>([&](auto FI) {
    constexpr int f = decltype(FI)::value;
    sfor<E + 1>([&](auto EI) {
      constexpr int e = decltype(EI)::value;
      R v = c0p * V[e][f];
      if constexpr (f > 0) v = fma((R)f * b01, V[e][f - 1], v);
      if constexpr (e > 0) v = fma((R)e * b00, V[e - 1][f], v);
      V[e][f + 1] = v;
    });
  });
}

// Length of the prefix of [0, n) on which the monotone predicate ok holds.
template <class F>
__device__ __forceinline__ int prefix_len(int n, F ok) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ok(mid)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// LDS writes of some lanes of a wave made visible to the others.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Row starts of the screening staircase: STAIR_ROWS_PER_LANE * T + 1 ints
// per team of T lanes (teams with more rows test every quartet instead);
// a workgroup of n threads needs at most (STAIR_ROWS_PER_LANE + 1) * n.
constexpr int STAIR_ROWS_PER_LANE = 8;
constexpr int STAIR_LDS_INTS = STAIR_ROWS_PER_LANE * 256 + 256;
constexpr int stair_lds_ints(int threads) { return (STAIR_ROWS_PER_LANE + 1) * threads; }
__device__ __forceinline__ int* stair_rows(int* s, int T) {
  return s + (threadIdx.x / T) * (STAIR_ROWS_PER_LANE * T + 1);
}

#ifndef ROOT_LOOP_DOUBLES
#define ROOT_LOOP_DOUBLES 96
#endif
// Primitive loop of one team: accumulates the [e0|f0] components at
// positions [K0, K1) of the component order over this lane's primitive
// quartets; the order is the flat index (i*NF + j) when ORDW = 0, else
// ChunkOrder<LA,LB,LC,LD,ORDW> (chunks of ORDW components grouped to share
// recurrence entries).  'live' is the team's pair-emptiness verdict, 'any'
// whether this lane saw a primitive quartet that passed libcint's two
// screening tests.
template <int LA, int LB, int LC, int LD, int K0, int K1, typename R = double, int DER = 0,
          int ORDW = 0>
__device__ __forceinline__ void accumulate(const BasisView& Bv, const Task& tk, int T, int tl,
                                           const double* __restrict__ rtab, int* __restrict__ rows,
                                           R (&acc)[K1 - K0], int& any, int& live) {
  using C = FastClass<LA, LB, LC, LD, DER>;
  constexpr int E = C::E, F = C::F, NR = C::NR, NF = C::NF;
  // Roots in a loop rather than unrolled once the accumulators and one
  // root's recurrence tables fill the register file: unrolled, the
  // compiler overlaps several roots' tables and spills.
  constexpr bool ROOT_LOOP = (K1 - K0) + 3 * (E + 1) * (F + 1) > ROOT_LOOP_DOUBLES;
  const VShell& sa = Bv.vsh[tk.sh[0]];
  const VShell& sb = Bv.vsh[tk.sh[1]];
  const VShell& sc = Bv.vsh[tk.sh[2]];
  const VShell& sd = Bv.vsh[tk.sh[3]];
  const double Ax = sa.r[0], Ay = sa.r[1], Az = sa.r[2];
  const double Cx = sc.r[0], Cy = sc.r[1], Cz = sc.r[2];
  constexpr bool THREE = DER == 2 || DER == 3;
  constexpr bool SR = DER == 3 || DER == 4;  // erfc(omega r)/r, omega < 0 (3c, 4c)
  const int nij = sa.nprim * sb.nprim;
  const int nkl = THREE ? sc.nprim : sc.nprim * sd.nprim;
  const int64_t pab = (int64_t)tk.sh[0] * Bv.nv + tk.sh[1];
  const int64_t pcd = (int64_t)tk.sh[2] * Bv.nv + tk.sh[3];
  const PairPrim* __restrict__ pij = Bv.pp + Bv.pair_off[pab];
  const PairPrim* __restrict__ pkl = THREE ? nullptr : Bv.pp + Bv.pair_off[pcd];
  const double ld_ij = Bv.pair_ld[pab] * (LA + LB + C::D1);
  double ld_kl = THREE ? 0.0 : Bv.pair_ld[pcd] * (LC + LD);
  const double Ecut = THREE ? Bv.expcutoff_3c : Bv.expcutoff_4c;
  const double* __restrict__ ek = Bv.env + sc.pexp;   // aux exponents (DER = 2)
  const double* __restrict__ ck = Bv.env + sc.pcoef;  // aux coefficients (DER = 2)
  // 2 pi^3 / sqrt(pi) times the s/p angular factors (src/g1e.c:565-572); the
  // exponent-0 s function of the three-centre case carries none
  constexpr double cfac = 34.986836655249725 *
                          (LA == 0 ? 0.282094791773878143 : LA == 1 ? 0.488602511902919921 : 1.0) *
                          (LB == 0 ? 0.282094791773878143 : LB == 1 ? 0.488602511902919921 : 1.0) *
                          (LC == 0 ? 0.282094791773878143 : LC == 1 ? 0.488602511902919921 : 1.0) *
                          (THREE ? 1.0 : LD == 0 ? 0.282094791773878143
                                         : LD == 1 ? 0.488602511902919921 : 1.0);

  // short range widens the distance term to log(d + theta r_guess + 1) and
  // raises the loop cutoff (src/optimizer.c:301-315, src/cint3c2e.c:279-296;
  // r_guess = 8); the pair-emptiness test keeps the plain cutoff
  double ldij = ld_ij, Eloop = Ecut;
  [[maybe_unused]] double w2 = 0.0;
  if constexpr (SR) {
    constexpr int lij = LA + LB, lk = LC, lkl = LC + LD;
    w2 = Bv.omega * Bv.omega;
    const double dab = sqrt((Ax - sb.r[0]) * (Ax - sb.r[0]) + (Ay - sb.r[1]) * (Ay - sb.r[1]) +
                            (Az - sb.r[2]) * (Az - sb.r[2]));
    const double th_ij =
        w2 / (w2 + Bv.env[sa.pexp + sa.nprim - 1] + Bv.env[sb.pexp + sb.nprim - 1]);
    ldij = lij > 0 ? lij * log(dab + th_ij * 8.0 + 1.0) : 0.0;
    if (NR > 1 && lij > 0) Eloop += lij * log((dab + th_ij * 8.0 + 1.0) / (dab + 1.0));
    if constexpr (THREE) {
      if (NR > 1 && lk > 0) Eloop += lk * log(w2 / (w2 + ek[sc.nprim - 1]) * 8.0 + 1.0);
    } else {
      // four centres: the ket pair is widened the same way
      // (src/cint2e.c:358-380)
      const double dcd = sqrt((Cx - sd.r[0]) * (Cx - sd.r[0]) + (Cy - sd.r[1]) * (Cy - sd.r[1]) +
                              (Cz - sd.r[2]) * (Cz - sd.r[2]));
      const double th_kl =
          w2 / (w2 + Bv.env[sc.pexp + sc.nprim - 1] + Bv.env[sd.pexp + sd.nprim - 1]);
      ld_kl = lkl > 0 ? lkl * log(dcd + th_kl * 8.0 + 1.0) : 0.0;
      if (NR > 1 && lkl > 0) Eloop += lkl * log((dcd + th_kl * 8.0 + 1.0) / (dcd + 1.0));
    }
  }


  // Screening (src/cint2e.c:96-109 and the primitive loops of
  // src/cint2e.c:120-300): a primitive quartet is computed when
  //   cce_kl <= Eloop, cce_ij <= Eloop - cce_kl, cce_ij < Ecut, cce_kl < Ecut
  // with cce = cce0 - distance term.  Each pair's records are stored in
  // ascending cce0, so for a fixed row (one record of the shorter list) the
  // passing records of the other list are a prefix, and the prefix lengths
  // do not grow from row to row: the passing set is a staircase, and the
  // team walks it without visiting the quartets that fail.
  auto ok = [&](double cij, double ckl) {
    return ckl < Ecut && cij < Ecut && !(ckl > Eloop) && !(cij > Eloop - ckl);
  };
  // pair emptiness (the first record is the pair's smallest); the
  // three-centre ket has no pair
  const double cij0 = pij[0].cce0 - ldij;
  const double ckl0 = THREE ? 0.0 : pkl[0].cce0 - ld_kl;
  live = cij0 < Ecut && (THREE || ckl0 < Ecut);

  sfor<K1 - K0>([&](auto I) { acc[decltype(I)::value] = R(0); });
  // a quartet counts as non-zero when some primitive quartet passes the
  // first two tests; for sorted lists that is the first one (short range
  // decides per primitive quartet, below)
  any = SR ? 0 : (live && !(ckl0 > Eloop) && !(cij0 > Eloop - ckl0));
  if (!live) return;

  // rows: the shorter list (the aux primitives for three centres); row r
  // holds the columns [rb, re) of the flat index f
  bool by_kl = true, stair = false;
  int nrow, ncol_u = 0, S;
  if constexpr (THREE) {
    nrow = nkl;
    ncol_u = prefix_len(nij, [&](int c) { return ok(pij[c].cce0 - ldij, 0.0); });
    S = ncol_u * nkl;
  } else {
    by_kl = nkl <= nij;
    nrow = by_kl ? nkl : nij;
    const int ncol = by_kl ? nij : nkl;
    if (nrow > STAIR_ROWS_PER_LANE * T || nrow * ncol <= T) {
      // row starts do not fit the team's LDS, or one pass of the team covers
      // every quartet anyway: every quartet, tested
      ncol_u = ncol;
      S = nrow * ncol;
    } else {
      stair = true;
      // lane tl takes rows [r0, r1); lengths first, then starts in place
      const int rpl = (nrow + T - 1) / T;
      const int r0 = min(tl * rpl, nrow), r1 = min(r0 + rpl, nrow);
      int hi = ncol, sum = 0;
      for (int r = r0; r < r1; ++r) {
        int len;
        if (by_kl) {
          const double ck = pkl[r].cce0 - ld_kl;
          len = prefix_len(hi, [&](int c) { return ok(pij[c].cce0 - ldij, ck); });
        } else {
          const double ci = pij[r].cce0 - ldij;
          len = prefix_len(hi, [&](int c) { return ok(ci, pkl[c].cce0 - ld_kl); });
        }
        hi = len;  // lengths do not grow with r
        rows[r] = len;
        sum += len;
      }
      int incl = sum;
      for (int off = 1; off < T; off <<= 1) {
        const int v = __shfl_up(incl, off, T);
        if (tl >= off) incl += v;
      }
      S = __shfl(incl, T - 1, T);
      int start = incl - sum;
      for (int r = r0; r < r1; ++r) {
        const int len = rows[r];
        rows[r] = start;
        start += len;
      }
      if (tl == 0) rows[nrow] = S;
      wave_lds_sync();
    }
  }

  int r = 0, rb = 0, re = stair ? rows[1] : ncol_u;
  for (int f = tl; f < S; f += T) {
    while (re <= f) {
      ++r;
      rb = re;
      re = stair ? rows[r + 1] : re + ncol_u;
    }
    const int ij = by_kl ? f - rb : r;
    const int kl = by_kl ? r : f - rb;
    PairPrim Q;
    if constexpr (THREE) {
      // ket "pair" of aux primitive kl and the exponent-0 s function
      Q.px = Cx;
      Q.py = Cy;
      Q.pz = Cz;
      Q.aij = ek[kl];
      Q.kab = ck[kl];
      Q.cce0 = 0.0;
    } else {
      Q = pkl[kl];
    }
    const PairPrim P = pij[ij];
    const double cce_kl = Q.cce0 - ld_kl;
    const double cce_ij = P.cce0 - ldij;
    // the staircase holds passing quartets only; the fallback tests each
    if (!stair && !ok(cce_ij, cce_kl)) continue;

    const double aij = P.aij, akl = Q.aij;
    const double a1 = aij * akl;
    const double a0 = a1 * rcp_pos(aij + akl);
    const double xpq = P.px - Q.px, ypq = P.py - Q.py, zpq = P.pz - Q.pz;
    const double rr = xpq * xpq + ypq * ypq + zpq * zpq;
    [[maybe_unused]] double theta = 0.0;
    if constexpr (SR) {
      // pairs beyond the plain cutoff carry no data (centre at 1e18), and
      // negligible erfc tails are dropped (src/g2e.c:4449-4453, EXPCUTOFF_SR)
      theta = w2 * rcp_pos(w2 + a0);
      const double tx = theta * a0 * rr;
      if (tx > Eloop - cce_kl - cce_ij || tx > 40.0) continue;
      any = 1;
    }
    const double fac1 = sqrt_pos(a0 * rcp_pos(a1 * a1 * a1)) * cfac * P.kab * Q.kab;
    const R sai = R(2.0 * P.ai);  // weight of the raised set (DER = 1)
    const double pax = P.px - Ax, pay = P.py - Ay, paz = P.pz - Az;
    const double qcx = Q.px - Cx, qcy = Q.py - Cy, qcz = Q.pz - Cz;
    // one root: coefficients, recurrence, products
    auto root_body = [&](double ur, double wr) {
      const double u2 = a0 * ur;
      const double tmp4 = 0.5 * rcp_pos(u2 * (aij + akl) + a1);
      const double b00 = u2 * tmp4;
      const double b10 = b00 + tmp4 * akl;
      const double b01 = b00 + tmp4 * aij;
      const double tmp2 = 2.0 * b00 * akl;
      const double tmp3 = 2.0 * b00 * aij;
      // coefficients in FP64; the recurrence and the products in R
      R Vx[E + 1][F + 1], Vy[E + 1][F + 1], Vz[E + 1][F + 1];
      vrr2d<E, F, R>(Vx, R(1), R(pax - tmp2 * xpq), R(qcx + tmp3 * xpq), R(b10), R(b01), R(b00));
      vrr2d<E, F, R>(Vy, R(1), R(pay - tmp2 * ypq), R(qcy + tmp3 * ypq), R(b10), R(b01), R(b00));
      vrr2d<E, F, R>(Vz, R(wr * fac1), R(paz - tmp2 * zpq), R(qcz + tmp3 * zpq), R(b10), R(b01),
                     R(b00));
      sfor<K1 - K0>([&](auto KK) {
        constexpr int k = K0 + decltype(KK)::value;
        constexpr int kf = ORDW ? ChunkOrder<LA, LB, LC, LD, ORDW>::at(k) : k;
        constexpr int i = kf / NF, j = kf % NF;
        constexpr int lo = C::row_lo(i), ik = C::row_k(i);
        constexpr int ex = cpx(lo, ik), ey = cpy(lo, ik), ez = cpz(lo, ik);
        constexpr int fx = cpx(LC, j), fy = cpy(LC, j), fz = cpz(LC, j);
        if constexpr (C::row_scaled(i))
          acc[k - K0] = fma(Vx[ex][fx] * Vy[ey][fy], Vz[ez][fz] * sai, acc[k - K0]);
        else
          acc[k - K0] = fma(Vx[ex][fx] * Vy[ey][fy], Vz[ez][fz], acc[k - K0]);
      });
    };
    if constexpr (SR) {
      // erfc = 1 - erf: the full rule on NR roots and the theta-scaled rule
      // on NR more, weights times -sqrt(theta) (src/g2e.c:4455-4476)
      double u[2 * NR], w[2 * NR];
      rys_roots<NR>(a0 * rr, u, w, rtab);
      rys_roots<NR>(theta * a0 * rr, u + NR, w + NR, rtab);
      const double mst = -sqrt_pos(theta);
      sfor<NR>([&](auto RI) {
        constexpr int r = NR + decltype(RI)::value;
        const double ut = u[r] * theta;
        u[r] = ut * rcp_pos(u[r] + 1.0 - ut);
        w[r] *= mst;
      });
      if constexpr (ROOT_LOOP) {
#pragma unroll 1
        for (int r = 0; r < 2 * NR; ++r) root_body(u[r], w[r]);
      } else {
        sfor<2 * NR>([&](auto RI) { root_body(u[decltype(RI)::value], w[decltype(RI)::value]); });
      }
    } else {
      double u[NR], w[NR];
      if (Bv.omega > 0) {
        // long range erf(omega r)/r: the rule at theta x, roots mapped to
        // u theta / (u + 1 - u theta), weights times sqrt(theta)
        // (src/g2e.c:4477-4494)
        const double wl = Bv.omega * Bv.omega;
        const double th = wl * rcp_pos(wl + a0);
        rys_roots<NR>(th * (a0 * rr), u, w, rtab);
        const double sth = sqrt_pos(th);
        sfor<NR>([&](auto RI) {
          constexpr int r = decltype(RI)::value;
          const double ut = u[r] * th;
          u[r] = ut * rcp_pos(u[r] + 1.0 - ut);
          w[r] *= sth;
        });
      } else {
        rys_roots<NR>(a0 * rr, u, w, rtab);
      }
      if constexpr (ROOT_LOOP) {
#pragma unroll 1
        for (int r = 0; r < NR; ++r) root_body(u[r], w[r]);
      } else {
        sfor<NR>([&](auto RI) { root_body(u[decltype(RI)::value], w[decltype(RI)::value]); });
      }
    }
  }
}

// Output strides of the canonical axes: axis a holds the original axis
// (perm >> 2a) & 3 of the packed Fortran block (perm: the task's axis map).
template <bool SPH = true>
__device__ __forceinline__ void canon_strides(const BasisView& Bv, const Task& tk, int64_t& st0,
                                              int64_t& st1, int64_t& st2, int64_t& st3) {
  const int perm = task_perm(tk);
  int dorig[4];
  dorig[perm & 3] = (SPH ? Bv.vsh[tk.sh[0]].dsph : Bv.vsh[tk.sh[0]].dcart);
  dorig[(perm >> 2) & 3] = (SPH ? Bv.vsh[tk.sh[1]].dsph : Bv.vsh[tk.sh[1]].dcart);
  dorig[(perm >> 4) & 3] = (SPH ? Bv.vsh[tk.sh[2]].dsph : Bv.vsh[tk.sh[2]].dcart);
  dorig[(perm >> 6) & 3] = (SPH ? Bv.vsh[tk.sh[3]].dsph : Bv.vsh[tk.sh[3]].dcart);
  int64_t so[4];
  so[0] = 1;
  so[1] = dorig[0];
  so[2] = so[1] * dorig[1];
  so[3] = so[2] * dorig[2];
  st0 = so[perm & 3];
  st1 = so[(perm >> 2) & 3];
  st2 = so[(perm >> 4) & 3];
  st3 = so[(perm >> 6) & 3];
}

// Waves per SIMD the register allocator is asked to keep room for (1 = no
// request), per plain class (chunk = -1: the single-chunk kernel).  Kept
// where measured faster on the c1 workload: (pp|ps) 172 -> 168 VGPRs, three
// waves, 37.2 -> 30.7 ms; (fs|ds) chunks 6.4 -> 5.7 ms.  Requests that cost
// more than ~30 bytes of scratch spill per lane lost ((ps|ps), (ps|ss),
// (ds|ps), (fs|ps), (pp|pp), (fp|ps)).
constexpr int min_waves(int la, int lb, int lc, int ld, int der, int chunk) {
  if (der != 0) return 1;
  const int c = ((la * 4 + lb) * 4 + lc) * 4 + ld;
  if (chunk < 0) {
    switch (c) {
      case ((1 * 4 + 1) * 4 + 1) * 4 + 0: return 3;
      default: return 1;
    }
  }
  switch (c) {
    case ((3 * 4 + 0) * 4 + 2) * 4 + 0: return 3;
    default: return 1;
  }
}

// JK: instead of the store, the J/K digest of the block (jk_degeneracy in
// device_common.h): each lane sums the products of its elements with the
// density into the six block updates, the team adds its lanes' sums, and
// the lanes share the atomic additions to J', K'.
template <int LA, int LB, int LC, int LD, typename R = double, int DER = 0, bool SPH = true,
          bool JK = false>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(min_waves(LA, LB, LC, LD, DER, -1))))
eri_fast_kernel(BasisView Bv, BucketArgs A) {
  using C = FastClass<LA, LB, LC, LD, DER>;
  constexpr int NE = C::NE, NF = C::NF;
  __shared__ double s_rys[RysLds<C::NR>::N];
  __shared__ int s_rows[DER == 2 || DER == 3 ? 1 : STAIR_LDS_INTS];
  const double* rtab = rys_stage_table<C::NR>(s_rys);
  const int T = A.team;
  const int tid = threadIdx.x;
  const int tl = tid & (T - 1);
  const int64_t t = A.t0 + (int64_t)blockIdx.x * (256 / T) + tid / T;
  if (t >= A.t1) return;  // whole teams leave together
  const Task tk = A.tasks[t];
  const VShell& sa = Bv.vsh[tk.sh[0]];
  const VShell& sb = Bv.vsh[tk.sh[1]];
  const VShell& sc = Bv.vsh[tk.sh[2]];
  const VShell& sd = Bv.vsh[tk.sh[3]];
  const double Ax = sa.r[0], Ay = sa.r[1], Az = sa.r[2];
  const double Cx = sc.r[0], Cy = sc.r[1], Cz = sc.r[2];
  R acc[NE * NF];
  int any, live;
  accumulate<LA, LB, LC, LD, 0, NE * NF, R, DER>(Bv, tk, T, tl, rtab, stair_rows(s_rows, T), acc,
                                                 any, live);

  // team sum
  for (int off = 1; off < T; off <<= 1) {
    sfor<NE * NF>([&](auto I) {
      constexpr int i = decltype(I)::value;
      acc[i] += __shfl_xor(acc[i], off);
    });
    any |= __shfl_xor(any, off);
  }
  any = any && live;
  if (tl == 0 && A.not0) A.not0[A.task_id ? A.task_id[t] : t] = any;

  // ---- epilogue --------------------------------------------------------
  int64_t st0, st1, st2, st3;
  canon_strides<SPH>(Bv, tk, st0, st1, st2, st3);
  double* __restrict__ out = A.out + task_out(tk);

  constexpr int NA = C::NA, NB = C::NB, NC = C::NC, ND = C::ND;
  constexpr int NAo = nout(LA, SPH), NBo = nout(LB, SPH), NCo = nout(LC, SPH), NDo = nout(LD, SPH);
  // (the J/K digest keeps screened tasks in the wave: their accumulators
  // are zero and they take part in the wave's reduction)
  if (!JK && !any) {
    sfor<NAo * NBo * NCo * NDo>([&](auto K) {
      constexpr int k = decltype(K)::value;
      constexpr int ao = k / (NBo * NCo * NDo), bo = (k / (NCo * NDo)) % NBo;
      constexpr int co = (k / NDo) % NCo, dd = k % NDo;
      if ((k & (T - 1)) == tl) out[ao * st0 + bo * st1 + co * st2 + dd * st3] = 0.0;
    });
    return;
  }
  // powers of A-B and C-D
  R abp[3][LB + 1], cdp[3][LD + 1];
  {
    const R ab[3] = {R(Ax - sb.r[0]), R(Ay - sb.r[1]), R(Az - sb.r[2])};
    const R cd[3] = {R(Cx - sd.r[0]), R(Cy - sd.r[1]), R(Cz - sd.r[2])};
    sfor<3>([&](auto DI) {
      constexpr int d = decltype(DI)::value;
      abp[d][0] = R(1);
      cdp[d][0] = R(1);
      sfor<LB>([&](auto K) { abp[d][decltype(K)::value + 1] = abp[d][decltype(K)::value] * ab[d]; });
      sfor<LD>([&](auto K) { cdp[d][decltype(K)::value + 1] = cdp[d][decltype(K)::value] * cd[d]; });
    });
  }
  // bra: (a,b| = sum_t C(b,t) (A-B)^(b-t) [a+t|, then spherical over a and b
  R Ys[NAo * NBo * NF];
  sfor<NF>([&](auto JJ) {
    constexpr int j = decltype(JJ)::value;
    R Yc[NA * NB];
    sfor<NA * NB>([&](auto AB) {
      constexpr int a = decltype(AB)::value / NB, b = decltype(AB)::value % NB;
      constexpr int ax = cpx(LA, a), ay = cpy(LA, a), az = cpz(LA, a);
      constexpr int bx = cpx(LB, b), by = cpy(LB, b), bz = cpz(LB, b);
      R s = R(0);
      sfor<(bx + 1) * (by + 1) * (bz + 1)>([&](auto TT) {
        constexpr int tt = decltype(TT)::value;
        constexpr int tx = tt / ((by + 1) * (bz + 1)), ty = (tt / (bz + 1)) % (by + 1), tz = tt % (bz + 1);
        constexpr double cb = binomc(bx, tx) * binomc(by, ty) * binomc(bz, tz);
        constexpr int e = cindex(LA, ax + tx, ay + ty, az + tz);
        const R pw = abp[0][bx - tx] * abp[1][by - ty] * abp[2][bz - tz];
        s = fma(R(cb) * pw, acc[e * NF + j], s);
      });
      Yc[a * NB + b] = s;
    });
    sfor<NAo * NBo>([&](auto OB) {
      constexpr int ao = decltype(OB)::value / NBo, bo = decltype(OB)::value % NBo;
      R s = R(0);
      sfor<NA * NB>([&](auto AB) {
        constexpr int a = decltype(AB)::value / NB, b = decltype(AB)::value % NB;
        constexpr double cc = c2x(SPH, LA, ao, a) * c2x(SPH, LB, bo, b);
        if constexpr (cc != 0.0) s = fma(R(cc), Yc[a * NB + b], s);
      });
      Ys[(ao * NBo + bo) * NF + j] = s;
    });
  });
  if constexpr (JK) {
    static_assert(SPH && (DER == 0), "the J/K digest takes plain spherical blocks");
    constexpr int O1 = NAo * NBo, O2 = O1 + NCo * NDo, O3 = O2 + NAo * NCo;
    constexpr int O4 = O3 + NAo * NDo, O5 = O4 + NBo * NCo, M = O5 + NBo * NDo;
    double g[M];
    sfor<M>([&](auto I) { g[decltype(I)::value] = 0.0; });
    const int64_t N = A.nao;
    const double* __restrict__ D = A.dm;
    const int64_t P0 = sa.ao, Q0 = sb.ao, R0 = sc.ao, S0 = sd.ao;
    sfor<NAo * NBo>([&](auto RW) {
      constexpr int row = decltype(RW)::value;
      constexpr int ao = row / NBo, bo = row % NBo;
      R Zc[NC * ND];
      sfor<NC * ND>([&](auto CD) {
        constexpr int c = decltype(CD)::value / ND, d = decltype(CD)::value % ND;
        constexpr int cx = cpx(LC, c), cy = cpy(LC, c), cz = cpz(LC, c);
        constexpr int dx = cpx(LD, d), dy = cpy(LD, d), dz = cpz(LD, d);
        R s = R(0);
        sfor<(dx + 1) * (dy + 1) * (dz + 1)>([&](auto TT) {
          constexpr int tt = decltype(TT)::value;
          constexpr int tx = tt / ((dy + 1) * (dz + 1)), ty = (tt / (dz + 1)) % (dy + 1), tz = tt % (dz + 1);
          constexpr double cb = binomc(dx, tx) * binomc(dy, ty) * binomc(dz, tz);
          constexpr int f = cindex(LC, cx + tx, cy + ty, cz + tz);
          const R pw = cdp[0][dx - tx] * cdp[1][dy - ty] * cdp[2][dz - tz];
          s = fma(R(cb) * pw, Ys[row * NF + f], s);
        });
        Zc[c * ND + d] = s;
      });
      const double* __restrict__ DP = D + (P0 + ao) * N;
      const double* __restrict__ DQ = D + (Q0 + bo) * N;
      const double dpq = DP[Q0 + bo];
      sfor<NCo * NDo>([&](auto OD) {
        constexpr int co = decltype(OD)::value / NDo, dd = decltype(OD)::value % NDo;
        constexpr int k = (row * NCo + co) * NDo + dd;
        if ((k & (T - 1)) == tl) {
          R s = R(0);
          sfor<NC * ND>([&](auto CD) {
            constexpr int c = decltype(CD)::value / ND, d = decltype(CD)::value % ND;
            constexpr double cc = c2x(SPH, LC, co, c) * c2x(SPH, LD, dd, d);
            if constexpr (cc != 0.0) s = fma(R(cc), Zc[c * ND + d], s);
          });
          const double v = (double)s;
          const int64_t Rr = R0 + co, Ss = S0 + dd;
          g[row] = fma(v, D[Rr * N + Ss], g[row]);
          g[O1 + co * NDo + dd] = fma(v, dpq, g[O1 + co * NDo + dd]);
          g[O2 + ao * NCo + co] = fma(v, DQ[Ss], g[O2 + ao * NCo + co]);
          g[O3 + ao * NDo + dd] = fma(v, DQ[Rr], g[O3 + ao * NDo + dd]);
          g[O4 + bo * NCo + co] = fma(v, DP[Ss], g[O4 + bo * NCo + co]);
          g[O5 + bo * NDo + dd] = fma(v, DP[Rr], g[O5 + bo * NDo + dd]);
        }
      });
    });
    for (int off = 1; off < T; off <<= 1)
      sfor<M>([&](auto I) {
        constexpr int m = decltype(I)::value;
        g[m] += __shfl_xor(g[m], off);
      });
    const double f = jk_degeneracy(sa.shell, sb.shell, sc.shell, sd.shell);
    sfor<M>([&](auto I) {
      constexpr int m = decltype(I)::value;
      g[m] *= m < O2 ? 2.0 * f : f;
    });
    // block r of the six: rows from rs[r], columns from cs[r] of J' or K'
    const int64_t rs[6] = {P0, R0, P0, P0, Q0, Q0};
    const int64_t cs[6] = {Q0, S0, R0, S0, R0, S0};
    // Teams of one wave that update the same block sum their values first
    // (the atomics are the digest's cost): per block, runs of consecutive
    // teams with equal block origin form segments, a segmented suffix sum
    // over the teams leaves each segment's total in its first team, and
    // only first teams add.  Needs every lane of the wave present (the
    // bucket's last wave adds team by team).
    bool head[6] = {true, true, true, true, true, true};
    if (T < 64 && __ballot(1) == ~0ull) {
      const int lane = threadIdx.x & 63;
      sfor<6>([&](auto RR) {
        constexpr int r = decltype(RR)::value;
        constexpr int o0 = r == 0 ? 0 : r == 1 ? O1 : r == 2 ? O2 : r == 3 ? O3 : r == 4 ? O4 : O5;
        constexpr int o1 = r == 0 ? O1 : r == 1 ? O2 : r == 2 ? O3 : r == 3 ? O4 : r == 4 ? O5 : M;
        const int key = (int)(rs[r] * N + cs[r]);
        const int kprev = __shfl_up(key, T);
        head[r] = lane < T || kprev != key;
        // segment number of this lane's team: first teams up to it
        const unsigned long long firsts = __ballot(head[r] && tl == 0);
        const int seg = __popcll(lane == 63 ? firsts : firsts & ((2ull << lane) - 1));
        for (int sh = T; sh < 64; sh <<= 1) {
          const int sd_ = __shfl_down(seg, sh);
          const bool take = lane + sh < 64 && sd_ == seg;
          sfor<o1 - o0>([&](auto I) {
            constexpr int m = o0 + decltype(I)::value;
            const double vd = __shfl_down(g[m], sh);
            if (take) g[m] += vd;
          });
        }
      });
    }
    double* __restrict__ Jp = A.jk_j;
    double* __restrict__ Kp = A.jk_k;
    sfor<M>([&](auto I) {
      constexpr int m = decltype(I)::value;
      constexpr int r = m < O1 ? 0 : m < O2 ? 1 : m < O3 ? 2 : m < O4 ? 3 : m < O5 ? 4 : 5;
      constexpr int o0 = r == 0 ? 0 : r == 1 ? O1 : r == 2 ? O2 : r == 3 ? O3 : r == 4 ? O4 : O5;
      constexpr int ncol = r == 0 ? NBo : (r == 1 || r == 3 || r == 5) ? NDo : NCo;
      constexpr int i = (m - o0) / ncol, j = (m - o0) % ncol;
      double* __restrict__ dst = r < 2 ? Jp : Kp;
      if (head[r] && (m & (T - 1)) == tl && dst && g[m] != 0.0)
        atomicAdd(dst + (rs[r] + i) * N + cs[r] + j, g[m]);
    });
    return;
  }
  // ket, row by row, then the store (lane tl writes every T-th element)
  sfor<NAo * NBo>([&](auto RW) {
    constexpr int row = decltype(RW)::value;
    constexpr int ao = row / NBo, bo = row % NBo;
    R Zc[NC * ND];
    sfor<NC * ND>([&](auto CD) {
      constexpr int c = decltype(CD)::value / ND, d = decltype(CD)::value % ND;
      constexpr int cx = cpx(LC, c), cy = cpy(LC, c), cz = cpz(LC, c);
      constexpr int dx = cpx(LD, d), dy = cpy(LD, d), dz = cpz(LD, d);
      R s = R(0);
      sfor<(dx + 1) * (dy + 1) * (dz + 1)>([&](auto TT) {
        constexpr int tt = decltype(TT)::value;
        constexpr int tx = tt / ((dy + 1) * (dz + 1)), ty = (tt / (dz + 1)) % (dy + 1), tz = tt % (dz + 1);
        constexpr double cb = binomc(dx, tx) * binomc(dy, ty) * binomc(dz, tz);
        constexpr int f = cindex(LC, cx + tx, cy + ty, cz + tz);
        const R pw = cdp[0][dx - tx] * cdp[1][dy - ty] * cdp[2][dz - tz];
        s = fma(R(cb) * pw, Ys[row * NF + f], s);
      });
      Zc[c * ND + d] = s;
    });
    sfor<NCo * NDo>([&](auto OD) {
      constexpr int co = decltype(OD)::value / NDo, dd = decltype(OD)::value % NDo;
      constexpr int k = (row * NCo + co) * NDo + dd;
      if ((k & (T - 1)) == tl) {
        R s = R(0);
        sfor<NC * ND>([&](auto CD) {
          constexpr int c = decltype(CD)::value / ND, d = decltype(CD)::value % ND;
          constexpr double cc = c2x(SPH, LC, co, c) * c2x(SPH, LD, dd, d);
          if constexpr (cc != 0.0) s = fma(R(cc), Zc[c * ND + d], s);
        });
        out[ao * st0 + bo * st1 + co * st2 + dd * st3] = (double)s;
      }
    });
  });
}

// ---------------------------------------------------------------------
// Chunked path for classes whose [e0|f0] block exceeds the register budget:
// chunk kernels (one per compile-time component range) accumulate their
// range in registers and store it to a per-task scratch block; the
// epilogue kernel then runs the transfers and the spherical transform
// cooperatively in LDS.
// ---------------------------------------------------------------------
constexpr int CHUNK_TARGET = 96;
// Plain classes of at most three roots (la+lb+lc+ld <= 5) split finer: their
// roots and recurrences are cheap to recompute per chunk, and a block of
// <= 48 accumulators leaves room for two waves per SIMD (measured on the
// c1 workload: (pp|pp) 48 -> 38 ms, (dp|pp) 43 -> 33 ms, (fp|p s) 31 -> 25 ms).
// Four-root classes with la+lb+lc+ld = 6 gain from <= 64 ((fp|pp) 31 ->
// 26.5 ms, (dd|pp) 7.1 -> 6.1 ms); at 7 and above the recomputation costs
// more than the occupancy gains, and 48 lost for all four-root classes.
constexpr int chunk_target(int la, int lb, int lc, int ld, int der) {
  return (der == 0 && la + lb + lc + ld <= 5)   ? 48
         : (der == 0 && la + lb + lc + ld == 6) ? 64
                                                : CHUNK_TARGET;
}

template <int LA, int LB, int LC, int LD, int DER = 0>
struct ChunkCfg {
  using C = FastClass<LA, LB, LC, LD, DER>;
  static constexpr int THREADS = 256;
  static constexpr int NEF = C::NE * C::NF;
  static constexpr int TGT = chunk_target(LA, LB, LC, LD, DER);
  static constexpr int NCH = (NEF + TGT - 1) / TGT;
  static constexpr int W = (NEF + NCH - 1) / NCH;
  static constexpr int k0(int c) { return c * W; }
  static constexpr int k1(int c) { return (c + 1) * W < NEF ? (c + 1) * W : NEF; }
  // component order of the chunks (flat for the derivative classes, whose
  // rows are the raised and lowered sets, and for classes whose epilogue
  // reads the block from global memory, where gathering through the map
  // costs more than the order saves).  The per-task scratch block holds
  // the components in this order, each chunk's range contiguous; pos() maps
  // a flat index to its position (nullptr: the identity).
  static constexpr bool EPI_GLOBAL =
      NEF + nout(LA, true) * nout(LB, true) * C::NF > 6144;
  static constexpr int ORDW =
      DER != 1 && NCH > 1 && !EPI_GLOBAL && ChunkOrder<LA, LB, LC, LD, W>::HAS ? W : 0;
  __device__ static const short* pos() {
    if constexpr (ORDW != 0) return ChunkOrder<LA, LB, LC, LD, ORDW>::pos();
    return nullptr;
  }
};

template <int LA, int LB, int LC, int LD, int CH, int DER = 0, typename R = double>
__global__ void __launch_bounds__((ChunkCfg<LA, LB, LC, LD, DER>::THREADS))
__attribute__((amdgpu_waves_per_eu(min_waves(LA, LB, LC, LD, DER, CH))))
eri_chunk_kernel(BasisView Bv, BucketArgs A) {
  using G = ChunkCfg<LA, LB, LC, LD, DER>;
  constexpr int K0 = G::k0(CH), K1 = G::k1(CH);
  __shared__ double s_rys[RysLds<G::C::NR>::N];
  __shared__ int s_rows[DER == 2 || DER == 3 ? 1 : stair_lds_ints(G::THREADS)];
  const double* rtab = rys_stage_table<G::C::NR>(s_rys);
  const int T = A.team;
  const int tid = threadIdx.x;
  const int tl = tid & (T - 1);
  const int64_t t = A.t0 + (int64_t)blockIdx.x * (G::THREADS / T) + tid / T;
  if (t >= A.t1) return;
  const Task tk = A.tasks[t];
  R acc[K1 - K0];
  int any, live;
  accumulate<LA, LB, LC, LD, K0, K1, R, DER, G::ORDW>(Bv, tk, T, tl, rtab, stair_rows(s_rows, T), acc,
                                                      any, live);
  for (int off = 1; off < T; off <<= 1) {
    sfor<K1 - K0>([&](auto I) {
      constexpr int i = decltype(I)::value;
      acc[i] += __shfl_xor(acc[i], off);
    });
    any |= __shfl_xor(any, off);
  }
  const int64_t slot = t - A.t0;
  double* __restrict__ dst = A.scratch + slot * G::NEF;
  sfor<K1 - K0>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if ((i & (T - 1)) == tl) dst[K0 + i] = (double)acc[i];
  });
  if constexpr (CH == 0) {
    if (tl == 0) {
      A.flags[slot] = any && live;
      if (A.not0) A.not0[A.task_id ? A.task_id[t] : t] = any && live;
    }
  }
}

// (b, t) pairs of the transfer (a,b| = sum_t C(b,t) (A-B)^(b-t) [a+t| over
// the Cartesian b of shell l, and nonzero Cartesian->spherical coefficients
// of shell l (cost estimates for the epilogue layout).
constexpr int xfer_terms(int l) {
  int n = 0;
  for (int i = 0; i <= l; ++i)
    for (int j = 0; j <= l - i; ++j) n += (i + 1) * (j + 1) * (l - i - j + 1);
  return n;
}
constexpr int nnz_c2s(int l, bool sph = true) {
  int n = 0;
  for (int o = 0; o < nout(l, sph); ++o)
    for (int c = 0; c < (l + 1) * (l + 2) / 2; ++c) n += c2x(sph, l, o, c) != 0.0;
  return n;
}

template <int LA, int LB, int LC, int LD, bool SPH = true>
struct EpiCfg {
  using C = FastClass<LA, LB, LC, LD>;
  static constexpr int NAo = nout(LA, SPH), NBo = nout(LB, SPH), NCo = nout(LC, SPH), NDo = nout(LD, SPH);
  static constexpr int NEF = C::NE * C::NF;
  // LDS budget in doubles: 48 KB, 80 KB for Cartesian output (whose bra
  // result of (ff|ff) alone is 58 KB)
  static constexpr int BUDGET = SPH ? 6144 : 10240;
  // the [e0|f0] block is staged in LDS unless block + bra result exceed the
  // budget; then the bra transfer reads it from the scratch block
  static constexpr bool XG = NEF + NAo * NBo * C::NF > BUDGET;
  static constexpr int PER = (XG ? 0 : NEF) + NAo * NBo * C::NF;  // LDS doubles per task
  static constexpr int TPB_LDS = BUDGET / PER > 0 ? BUDGET / PER : 1;
  // Lanes per task: the bra pass gives each lane whole columns (NF of
  // them), the ket pass whole rows (NAo * NBo).  By default TE covers the
  // larger of the two.  For an (ff| bra (49+ rows, the costliest transfer)
  // whose few columns leave most lanes idle in the bra pass, TE minimises
  // the estimated wave instructions per task, (columns per lane * column
  // cost + rows per lane * row cost) / tasks per wave, when that is 3.5x
  // below the default's: (ff|ps) 8.0 -> 5.7 ms, (ff|ds) 4.9 -> 4.4 ms,
  // (ff|fs) 4.8 -> 4.4 ms.  The estimate ignores occupancy; on other classes
  // the smaller TE lost.
  static constexpr int COL = C::NA * xfer_terms(LB) + nnz_c2s(LA, SPH) * nnz_c2s(LB, SPH);
  static constexpr int ROW = C::NC * xfer_terms(LD) + nnz_c2s(LC, SPH) * nnz_c2s(LD, SPH) + NCo * NDo;
  static constexpr int64_t cost(int te) {
    const int tpb = (256 / te) < TPB_LDS ? 256 / te : TPB_LDS;
    const int lanes = tpb * te < 64 ? tpb * te : 64;
    const int64_t per_lane =
        (int64_t)((C::NF + te - 1) / te) * COL + (int64_t)((NAo * NBo + te - 1) / te) * ROW;
    return per_lane * 64 * te / lanes;  // x 64: tasks per wave = lanes / te
  }
  static constexpr int pick_te() {
    int best = 64;
    for (int te = 32; te >= 1; te /= 2)
      if (cost(te) < cost(best)) best = te;
    return best;
  }
  static constexpr int WORK = C::NF > NAo * NBo ? C::NF : NAo * NBo;
  static constexpr int TE_COVER = WORK > 32 ? 64 : (WORK > 16 ? 32 : (WORK > 8 ? 16 : 8));
  static constexpr int TE = NAo * NBo >= 49 && 2 * cost(TE_COVER) > 7 * cost(pick_te())
                                ? pick_te() : TE_COVER;
  static constexpr int TPB0 = 256 / TE;
  static constexpr int TPB = TPB0 < TPB_LDS ? TPB0 : TPB_LDS;
  static constexpr int THREADS = TPB * TE;
  static_assert(PER <= BUDGET, "epilogue exceeds its LDS budget");
};

// Bra transfer (a,b| = sum_t C(b,t) (A-B)^(b-t) [a+t| and the spherical
// transform over a, b for the columns j = j0, j0+js, ... of the contracted
// block X (NE x NF), written to Ys (NAo*NBo x NF).
template <int LA, int LB, int LC, int LD, bool SPH = true>
__device__ __forceinline__ void epi_bra(const BasisView& Bv, const Task& tk,
                                        const double* __restrict__ X, double* __restrict__ Ys,
                                        int j0, int js, const short* __restrict__ pos = nullptr) {
  using C = FastClass<LA, LB, LC, LD>;
  constexpr int NF = C::NF, NA = C::NA, NB = C::NB, NAo = nout(LA, SPH), NBo = nout(LB, SPH);
  const VShell& sa = Bv.vsh[tk.sh[0]];
  const VShell& sb = Bv.vsh[tk.sh[1]];
  const double ab[3] = {sa.r[0] - sb.r[0], sa.r[1] - sb.r[1], sa.r[2] - sb.r[2]};
  double abp[3][LB + 1];
  sfor<3>([&](auto DI) {
    constexpr int d = decltype(DI)::value;
    abp[d][0] = 1.0;
    sfor<LB>([&](auto K) { abp[d][decltype(K)::value + 1] = abp[d][decltype(K)::value] * ab[d]; });
  });
  for (int j = j0; j < NF; j += js) {
    double Yc[NA * NB];
    sfor<NA * NB>([&](auto AB) {
      constexpr int a = decltype(AB)::value / NB, b = decltype(AB)::value % NB;
      constexpr int ax = cpx(LA, a), ay = cpy(LA, a), az = cpz(LA, a);
      constexpr int bx = cpx(LB, b), by = cpy(LB, b), bz = cpz(LB, b);
      double s = 0.0;
      sfor<(bx + 1) * (by + 1) * (bz + 1)>([&](auto TT) {
        constexpr int tt = decltype(TT)::value;
        constexpr int tx = tt / ((by + 1) * (bz + 1)), ty = (tt / (bz + 1)) % (by + 1), tz = tt % (bz + 1);
        constexpr double cb = binomc(bx, tx) * binomc(by, ty) * binomc(bz, tz);
        constexpr int e = cindex(LA, ax + tx, ay + ty, az + tz);
        const double pw = abp[0][bx - tx] * abp[1][by - ty] * abp[2][bz - tz];
        const int xi = pos ? pos[e * NF + j] : e * NF + j;
        s = fma(cb * pw, X[xi], s);
      });
      Yc[a * NB + b] = s;
    });
    sfor<NAo * NBo>([&](auto OB) {
      constexpr int ao = decltype(OB)::value / NBo, bo = decltype(OB)::value % NBo;
      double s = 0.0;
      sfor<NA * NB>([&](auto AB) {
        constexpr int a = decltype(AB)::value / NB, b = decltype(AB)::value % NB;
        constexpr double cc = c2x(SPH, LA, ao, a) * c2x(SPH, LB, bo, b);
        if constexpr (cc != 0.0) s = fma(cc, Yc[a * NB + b], s);
      });
      Ys[(ao * NBo + bo) * NF + j] = s;
    });
  }
}

// Ket transfer and spherical c, d of the rows r0, r0+rs, ... of Ys, then
// the store with the task's output strides; any == 0 writes zeros.
template <int LA, int LB, int LC, int LD, bool SPH = true>
__device__ __forceinline__ void epi_ket_store(const BasisView& Bv, const Task& tk,
                                              const double* __restrict__ Ys,
                                              double* __restrict__ outbase, int any, int r0,
                                              int rs) {
  using C = FastClass<LA, LB, LC, LD>;
  constexpr int NF = C::NF, NC = C::NC, ND = C::ND;
  constexpr int NAo = nout(LA, SPH), NBo = nout(LB, SPH), NCo = nout(LC, SPH), NDo = nout(LD, SPH);
  int64_t st0, st1, st2, st3;
  canon_strides<SPH>(Bv, tk, st0, st1, st2, st3);
  double* __restrict__ out = outbase + task_out(tk);
  if (!any) {
    for (int k = r0; k < NAo * NBo * NCo * NDo; k += rs) {
      const int ao = k % NAo, bo = (k / NAo) % NBo, co = (k / (NAo * NBo)) % NCo;
      const int dd = k / (NAo * NBo * NCo);
      out[ao * st0 + bo * st1 + co * st2 + dd * st3] = 0.0;
    }
    return;
  }
  const VShell& sc = Bv.vsh[tk.sh[2]];
  const VShell& sd = Bv.vsh[tk.sh[3]];
  const double cd[3] = {sc.r[0] - sd.r[0], sc.r[1] - sd.r[1], sc.r[2] - sd.r[2]};
  double cdp[3][LD + 1];
  sfor<3>([&](auto DI) {
    constexpr int d = decltype(DI)::value;
    cdp[d][0] = 1.0;
    sfor<LD>([&](auto K) { cdp[d][decltype(K)::value + 1] = cdp[d][decltype(K)::value] * cd[d]; });
  });
  for (int row = r0; row < NAo * NBo; row += rs) {
    const int ao = row / NBo, bo = row % NBo;
    const double* yr = Ys + row * NF;
    double Zc[NC * ND];
    sfor<NC * ND>([&](auto CD) {
      constexpr int c = decltype(CD)::value / ND, d = decltype(CD)::value % ND;
      constexpr int cx = cpx(LC, c), cy = cpy(LC, c), cz = cpz(LC, c);
      constexpr int dx = cpx(LD, d), dy = cpy(LD, d), dz = cpz(LD, d);
      double s = 0.0;
      sfor<(dx + 1) * (dy + 1) * (dz + 1)>([&](auto TT) {
        constexpr int tt = decltype(TT)::value;
        constexpr int tx = tt / ((dy + 1) * (dz + 1)), ty = (tt / (dz + 1)) % (dy + 1), tz = tt % (dz + 1);
        constexpr double cb = binomc(dx, tx) * binomc(dy, ty) * binomc(dz, tz);
        constexpr int f = cindex(LC, cx + tx, cy + ty, cz + tz);
        const double pw = cdp[0][dx - tx] * cdp[1][dy - ty] * cdp[2][dz - tz];
        s = fma(cb * pw, yr[f], s);
      });
      Zc[c * ND + d] = s;
    });
    double* orow = out + ao * st0 + bo * st1;
    sfor<NCo * NDo>([&](auto OD) {
      constexpr int co = decltype(OD)::value / NDo, dd = decltype(OD)::value % NDo;
      double s = 0.0;
      sfor<NC * ND>([&](auto CD) {
        constexpr int c = decltype(CD)::value / ND, d = decltype(CD)::value % ND;
        constexpr double cc = c2x(SPH, LC, co, c) * c2x(SPH, LD, dd, d);
        if constexpr (cc != 0.0) s = fma(cc, Zc[c * ND + d], s);
      });
      orow[co * st2 + dd * st3] = s;
    });
  }
}

// The per-task block is read in the component order of the chunks that
// wrote it (ChunkCfg<..., DER>::pos(): the position of each flat index).
template <int LA, int LB, int LC, int LD, bool SPH = true, int DER = 0>
__global__ void __launch_bounds__(256) eri_epi_kernel(BasisView Bv, BucketArgs A) {
  using C = FastClass<LA, LB, LC, LD>;
  using P = EpiCfg<LA, LB, LC, LD, SPH>;
  constexpr int NEF = P::NEF, TE = P::TE;
  const short* __restrict__ pos = ChunkCfg<LA, LB, LC, LD, DER>::pos();
  __shared__ double sm[P::TPB * P::PER];
  const int team = threadIdx.x / TE;
  const int tl = threadIdx.x % TE;
  const int64_t t = A.t0 + (int64_t)blockIdx.x * P::TPB + team;
  const bool valid = t < A.t1;
  double* Ys = sm + team * P::PER + (P::XG ? 0 : NEF);
  const double* X = sm + team * P::PER;
  int any = 0;
  Task tk;
  if (valid) {
    tk = A.tasks[t];
    const int64_t slot = t - A.t0;
    any = A.flags[slot];
    const double* src = A.scratch + slot * NEF;
    if constexpr (P::XG) {
      X = src;
    } else if (any) {
      for (int k = tl; k < NEF; k += TE)
        sm[team * P::PER + k] =
            src[A.col_major ? (k % C::NF) * C::NE + k / C::NF : (pos ? pos[k] : k)];
    }
  }
  __syncthreads();
  if (valid && any) epi_bra<LA, LB, LC, LD, SPH>(Bv, tk, X, Ys, tl, TE, P::XG ? pos : nullptr);
  __syncthreads();
  if (valid) epi_ket_store<LA, LB, LC, LD, SPH>(Bv, tk, Ys, A.out, any, tl, TE);
}

// J/K digest variant of the chunked epilogue (plain spherical blocks): the
// bra pass as in eri_epi_kernel; then the lane that owns a row (ao, bo)
// forms its elements and adds their products with the density into LDS
// accumulators of the task (ds_add): the six block updates of
// jk_degeneracy (device_common.h), J'_ab, J'_cd, K'_ac, K'_ad, K'_bc,
// K'_bd.  Consecutive tasks of the workgroup that update the same block
// are summed in LDS before the one atomic per entry.
template <int LA, int LB, int LC, int LD>
struct EpiJkCfg {
  using P = EpiCfg<LA, LB, LC, LD, true>;
  static constexpr int NAo = P::NAo, NBo = P::NBo, NCo = P::NCo, NDo = P::NDo;
  static constexpr int O1 = NAo * NBo, O2 = O1 + NCo * NDo, O3 = O2 + NAo * NCo;
  static constexpr int O4 = O3 + NAo * NDo, O5 = O4 + NBo * NCo, MJK = O5 + NBo * NDo;
  static constexpr int PER = P::PER + MJK;
  static constexpr int LDS_MAX = 7680;  // doubles (60 KB; the rest: keys and weights)
  static constexpr int TPB = P::TPB * PER <= LDS_MAX ? P::TPB : (LDS_MAX / PER > 0 ? LDS_MAX / PER : 1);
  static constexpr int THREADS = TPB * P::TE;
  static_assert(PER <= LDS_MAX, "J/K epilogue exceeds the LDS of a workgroup");
};

template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void epi_ket_jk(const BasisView& Bv, const Task& tk,
                                           const double* __restrict__ Ys, double* G,
                                           const BucketArgs& A, int r0, int rs) {
  using C = FastClass<LA, LB, LC, LD>;
  using Q = EpiJkCfg<LA, LB, LC, LD>;
  constexpr int NF = C::NF, NC = C::NC, ND = C::ND;
  constexpr int NBo = Q::NBo, NCo = Q::NCo, NDo = Q::NDo;
  const VShell& sa = Bv.vsh[tk.sh[0]];
  const VShell& sb = Bv.vsh[tk.sh[1]];
  const VShell& sc = Bv.vsh[tk.sh[2]];
  const VShell& sd = Bv.vsh[tk.sh[3]];
  const double cd[3] = {sc.r[0] - sd.r[0], sc.r[1] - sd.r[1], sc.r[2] - sd.r[2]};
  double cdp[3][LD + 1];
  sfor<3>([&](auto DI) {
    constexpr int d = decltype(DI)::value;
    cdp[d][0] = 1.0;
    sfor<LD>([&](auto K) { cdp[d][decltype(K)::value + 1] = cdp[d][decltype(K)::value] * cd[d]; });
  });
  const int64_t N = A.nao;
  const double* __restrict__ D = A.dm;
  const int64_t P0 = sa.ao, Q0 = sb.ao, R0 = sc.ao, S0 = sd.ao;
  for (int row = r0; row < Q::NAo * NBo; row += rs) {
    const int ao = row / NBo, bo = row % NBo;
    const double* yr = Ys + row * NF;
    double Zc[NC * ND];
    sfor<NC * ND>([&](auto CD) {
      constexpr int c = decltype(CD)::value / ND, d = decltype(CD)::value % ND;
      constexpr int cx = cpx(LC, c), cy = cpy(LC, c), cz = cpz(LC, c);
      constexpr int dx = cpx(LD, d), dy = cpy(LD, d), dz = cpz(LD, d);
      double s = 0.0;
      sfor<(dx + 1) * (dy + 1) * (dz + 1)>([&](auto TT) {
        constexpr int tt = decltype(TT)::value;
        constexpr int tx = tt / ((dy + 1) * (dz + 1)), ty = (tt / (dz + 1)) % (dy + 1), tz = tt % (dz + 1);
        constexpr double cb = binomc(dx, tx) * binomc(dy, ty) * binomc(dz, tz);
        constexpr int f = cindex(LC, cx + tx, cy + ty, cz + tz);
        const double pw = cdp[0][dx - tx] * cdp[1][dy - ty] * cdp[2][dz - tz];
        s = fma(cb * pw, yr[f], s);
      });
      Zc[c * ND + d] = s;
    });
    const double* __restrict__ DP = D + (P0 + ao) * N;
    const double* __restrict__ DQ = D + (Q0 + bo) * N;
    const double dpq = DP[Q0 + bo];
    double jab = 0.0, kac[NCo], kbc[NCo], kad[NDo], kbd[NDo];
    sfor<NCo>([&](auto I) { kac[decltype(I)::value] = kbc[decltype(I)::value] = 0.0; });
    sfor<NDo>([&](auto I) { kad[decltype(I)::value] = kbd[decltype(I)::value] = 0.0; });
    sfor<NCo * NDo>([&](auto OD) {
      constexpr int co = decltype(OD)::value / NDo, dd = decltype(OD)::value % NDo;
      double v = 0.0;
      sfor<NC * ND>([&](auto CD) {
        constexpr int c = decltype(CD)::value / ND, d = decltype(CD)::value % ND;
        constexpr double cc = c2x(true, LC, co, c) * c2x(true, LD, dd, d);
        if constexpr (cc != 0.0) v = fma(cc, Zc[c * ND + d], v);
      });
      const int64_t Rr = R0 + co, Ss = S0 + dd;
      jab = fma(v, D[Rr * N + Ss], jab);
      atomicAdd(G + Q::O1 + co * NDo + dd, v * dpq);
      kac[co] = fma(v, DQ[Ss], kac[co]);
      kbc[co] = fma(v, DP[Ss], kbc[co]);
      kad[dd] = fma(v, DQ[Rr], kad[dd]);
      kbd[dd] = fma(v, DP[Rr], kbd[dd]);
    });
    G[row] = jab;  // the row's lane alone writes J'_ab of its row
    sfor<NCo>([&](auto I) {
      constexpr int co = decltype(I)::value;
      atomicAdd(G + Q::O2 + ao * NCo + co, kac[co]);
      atomicAdd(G + Q::O4 + bo * NCo + co, kbc[co]);
    });
    sfor<NDo>([&](auto I) {
      constexpr int dd = decltype(I)::value;
      atomicAdd(G + Q::O3 + ao * NDo + dd, kad[dd]);
      atomicAdd(G + Q::O5 + bo * NDo + dd, kbd[dd]);
    });
  }
}

template <int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(256) eri_epi_jk_kernel(BasisView Bv, BucketArgs A) {
  using P = EpiCfg<LA, LB, LC, LD, true>;
  using Q = EpiJkCfg<LA, LB, LC, LD>;
  constexpr int NEF = P::NEF, TE = P::TE, TPB = Q::TPB;
  const short* __restrict__ pos = ChunkCfg<LA, LB, LC, LD, 0>::pos();
  __shared__ double sm[TPB * Q::PER];
  __shared__ int s_key[TPB][6];   // block origins (row AO * nao + column AO)
  __shared__ double s_w[TPB];     // degeneracy factor f of the task
  const int team = threadIdx.x / TE;
  const int tl = threadIdx.x % TE;
  const int64_t t = A.t0 + (int64_t)blockIdx.x * TPB + team;
  const bool valid = t < A.t1;
  double* base = sm + team * Q::PER;
  double* Ys = base + (P::XG ? 0 : NEF);
  double* G = base + P::PER;
  const double* X = base;
  const int64_t N = A.nao;
  int any = 0;
  Task tk;
  int64_t rs[6] = {0, 0, 0, 0, 0, 0}, cs[6] = {0, 0, 0, 0, 0, 0};
  if (valid) {
    tk = A.tasks[t];
    const int64_t slot = t - A.t0;
    any = A.flags[slot];
    const double* src = A.scratch + slot * NEF;
    if constexpr (P::XG) {
      X = src;
    } else if (any) {
      for (int k = tl; k < NEF; k += TE) {
        using C = FastClass<LA, LB, LC, LD>;
        base[k] = src[A.col_major ? (k % C::NF) * C::NE + k / C::NF : (pos ? pos[k] : k)];
      }
    }
    for (int k = tl; k < Q::MJK; k += TE) G[k] = 0.0;
    const VShell& sa = Bv.vsh[tk.sh[0]];
    const VShell& sb = Bv.vsh[tk.sh[1]];
    const VShell& sc = Bv.vsh[tk.sh[2]];
    const VShell& sd = Bv.vsh[tk.sh[3]];
    const int64_t P0 = sa.ao, Q0 = sb.ao, R0 = sc.ao, S0 = sd.ao;
    rs[0] = P0; cs[0] = Q0;
    rs[1] = R0; cs[1] = S0;
    rs[2] = P0; cs[2] = R0;
    rs[3] = P0; cs[3] = S0;
    rs[4] = Q0; cs[4] = R0;
    rs[5] = Q0; cs[5] = S0;
    if (tl == 0) {
      for (int r = 0; r < 6; ++r) s_key[team][r] = (int)(rs[r] * N + cs[r]);
      s_w[team] = jk_degeneracy(sa.shell, sb.shell, sc.shell, sd.shell);
    }
  } else if (tl == 0) {
    for (int r = 0; r < 6; ++r) s_key[team][r] = -1 - team;  // matches no task
    s_w[team] = 0.0;
  }
  __syncthreads();
  if (valid && any) epi_bra<LA, LB, LC, LD, true>(Bv, tk, X, Ys, tl, TE, P::XG ? pos : nullptr);
  __syncthreads();
  if (valid && any) epi_ket_jk<LA, LB, LC, LD>(Bv, tk, Ys, G, A, tl, TE);
  __syncthreads();
  if (!valid) return;
  // the first task of each run of equal block origins adds the run's sum
  for (int m = tl; m < Q::MJK; m += TE) {
    const int r = m < Q::O1 ? 0 : m < Q::O2 ? 1 : m < Q::O3 ? 2 : m < Q::O4 ? 3 : m < Q::O5 ? 4 : 5;
    const int key = s_key[team][r];
    if (team > 0 && s_key[team - 1][r] == key) continue;
    const double jw = r < 2 ? 2.0 : 1.0;
    double v = jw * s_w[team] * G[m];
    for (int u = team + 1; u < TPB && s_key[u][r] == key; ++u)
      v += jw * s_w[u] * sm[u * Q::PER + P::PER + m];
    double* dst = r < 2 ? A.jk_j : A.jk_k;
    if (!dst || v == 0.0) continue;
    const int o0 = r == 0 ? 0 : r == 1 ? Q::O1 : r == 2 ? Q::O2 : r == 3 ? Q::O3 : r == 4 ? Q::O4 : Q::O5;
    const int ncol = r == 0 ? Q::NBo : (r == 1 || r == 3 || r == 5) ? Q::NDo : Q::NCo;
    const int i = (m - o0) / ncol, j = (m - o0) % ncol;
    atomicAdd(dst + (rs[r] + i) * N + cs[r] + j, v);
  }
}

// ---------------------------------------------------------------------
// Epilogue of the i-derivative chunks (int2e_ip1, src/autocode/grad2.c:19-68):
// per task, in LDS,
//   (d/dA_x a b| = a_x (a-1_x, b| - (a+1_x, b|_2a_i        (bra transfers
//     on the lowered and the 2 a_i-weighted raised sets, then spherical a,b)
//   |c d) <- |f0] transfer, spherical c,d;
// then either the store of the three component blocks (comp outermost,
// libcint's layout) or, with A.dm set, the gradient digest
//   g_x += (d/dA_x a b|c d) fkl [-2 D_ab D_cd + (D_ac D_bd + D_ad D_bc)/2]
// summed per block over runs of equal atoms into slot (block % nslot) of
// the gradient accumulators.
// ---------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, bool SPH = true>
struct Ip1EpiCfg {
  using C = FastClass<LA, LB, LC, LD, 1>;
  static constexpr int NAo = nout(LA, SPH), NBo = nout(LB, SPH), NCo = nout(LC, SPH), NDo = nout(LD, SPH);
  // LDS budget in doubles: 48 KB, 64 KB for Cartesian output
  static constexpr int BUDGET = SPH ? 6144 : 8192;
  static constexpr int NEF = C::NE * C::NF;
  static constexpr int ROWS = 3 * NAo * NBo;
  static constexpr int PER = NEF + ROWS * C::NF;  // doubles of LDS per task
  static constexpr int WORK = C::NF > ROWS ? C::NF : ROWS;
  static constexpr int TE_COVER = WORK > 32 ? 64 : (WORK > 16 ? 32 : (WORK > 8 ? 16 : 8));
  static constexpr int TPB_LDS = BUDGET / PER > 0 ? BUDGET / PER : 1;
  // Lanes per task: the bra pass gives each lane whole columns (three
  // directions each), the digest pass whole rows.  TE_COVER covers the
  // larger of the two; a smaller TE is taken when it fills whole waves,
  // keeps at least 8 waves per CU within the LDS, and halves the estimated
  // wave instructions per task (columns per lane * column cost + rows per
  // lane * row cost, over tasks per wave).
  static constexpr int COL = 3 * (C::NA * 2 * xfer_terms(LB) + nnz_c2s(LA, SPH) * nnz_c2s(LB, SPH));
  static constexpr int ROW = C::NC * xfer_terms(LD) + nnz_c2s(LC, SPH) * nnz_c2s(LD, SPH) + 4 * NCo * NDo;
  static constexpr int tpb_for(int te) {
    const int t = (256 / te) < TPB_LDS ? 256 / te : TPB_LDS;
    const int q = te < 64 ? 64 / te : 1;  // tasks per full wave
    return t >= q ? t - t % q : t;
  }
  static constexpr bool usable(int te) {
    const int tpb = tpb_for(te);
    const int threads = tpb * te;
    if (threads % 64) return false;
    const int blocks = 163840 / (tpb * PER * 8 + 1);
    return (blocks < 8 ? blocks : 8) * (threads / 64) >= 8;
  }
  static constexpr int64_t cost(int te) {
    const int64_t per_lane = (int64_t)((C::NF + te - 1) / te) * COL +
                             (int64_t)((ROWS + te - 1) / te) * ROW;
    const int lanes = tpb_for(te) * te < 64 ? tpb_for(te) * te : 64;
    return per_lane * 64 * te / lanes;
  }
  static constexpr int pick_te() {
    int best = TE_COVER;
    for (int te = TE_COVER / 2; te >= 1; te /= 2)
      if (usable(te) && 2 * cost(te) <= cost(TE_COVER) && cost(te) < cost(best)) best = te;
    return best;
  }
  static constexpr int TE = pick_te();
  static constexpr int TPB = TE == TE_COVER ? ((256 / TE) < TPB_LDS ? 256 / TE : TPB_LDS) : tpb_for(TE);
  static constexpr int THREADS = TPB * TE;
  static_assert(PER <= BUDGET, "i-derivative epilogue exceeds its LDS budget");
};

template <int LA, int LB, int LC, int LD, bool SPH = true>
__global__ void __launch_bounds__(256) eri_ip1_epi_kernel(BasisView Bv, BucketArgs A) {
  using C = FastClass<LA, LB, LC, LD, 1>;
  using P = Ip1EpiCfg<LA, LB, LC, LD, SPH>;
  constexpr int NEP = C::NEP, NF = C::NF, NEF = P::NEF, TE = P::TE, ROWS = P::ROWS;
  constexpr int NA = C::NA, NB = C::NB, NC = C::NC, ND = C::ND;
  constexpr int NAo = nout(LA, SPH), NBo = nout(LB, SPH), NCo = nout(LC, SPH), NDo = nout(LD, SPH);
  __shared__ double sm[P::TPB * P::PER];
  __shared__ double s_g[P::TPB][3];
  __shared__ int s_atom[P::TPB];
  const int team = threadIdx.x / TE;
  const int tl = threadIdx.x % TE;
  const int64_t t = A.t0 + (int64_t)blockIdx.x * P::TPB + team;
  const bool valid = t < A.t1;
  const bool digest = A.dm != nullptr;
  double* X = sm + team * P::PER;
  double* Ys = X + NEF;
  int any = 0;
  Task tk;
  if (valid) {
    tk = A.tasks[t];
    const int64_t slot = t - A.t0;
    any = A.flags[slot];
    const double* src = A.scratch + slot * NEF;
    if (any)
      for (int k = tl; k < NEF; k += TE) X[k] = src[k];
  }
  __syncthreads();
  if (valid && any) {
    const VShell& sa = Bv.vsh[tk.sh[0]];
    const VShell& sb = Bv.vsh[tk.sh[1]];
    const double ab[3] = {sa.r[0] - sb.r[0], sa.r[1] - sb.r[1], sa.r[2] - sb.r[2]};
    double abp[3][LB + 1];
    sfor<3>([&](auto DI) {
      constexpr int d = decltype(DI)::value;
      abp[d][0] = 1.0;
      sfor<LB>([&](auto K) { abp[d][decltype(K)::value + 1] = abp[d][decltype(K)::value] * ab[d]; });
    });
    for (int f = tl; f < NF; f += TE) {
      sfor<3>([&](auto DI) {
        constexpr int x = decltype(DI)::value;
        double Yc[NA * NB];
        sfor<NA * NB>([&](auto AB) {
          constexpr int a = decltype(AB)::value / NB, b = decltype(AB)::value % NB;
          constexpr int ax = cpx(LA, a), ay = cpy(LA, a), az = cpz(LA, a);
          constexpr int bx = cpx(LB, b), by = cpy(LB, b), bz = cpz(LB, b);
          constexpr int ux = ax + (x == 0), uy = ay + (x == 1), uz = az + (x == 2);
          constexpr int adir = x == 0 ? ax : (x == 1 ? ay : az);
          double sp = 0.0, sl = 0.0;
          sfor<(bx + 1) * (by + 1) * (bz + 1)>([&](auto TT) {
            constexpr int tt = decltype(TT)::value;
            constexpr int tx = tt / ((by + 1) * (bz + 1)), ty = (tt / (bz + 1)) % (by + 1), tz = tt % (bz + 1);
            constexpr double cb = binomc(bx, tx) * binomc(by, ty) * binomc(bz, tz);
            const double pw = abp[0][bx - tx] * abp[1][by - ty] * abp[2][bz - tz];
            constexpr int ep = cindex(LA + 1, ux + tx, uy + ty, uz + tz);
            sp = fma(cb * pw, X[ep * NF + f], sp);
            if constexpr (adir > 0) {
              constexpr int ex = ax - (x == 0) + tx, ey = ay - (x == 1) + ty, ez = az - (x == 2) + tz;
              constexpr int em = NEP + cindex(LA - 1, ex, ey, ez);
              sl = fma(cb * pw, X[em * NF + f], sl);
            }
          });
          Yc[a * NB + b] = fma(double(adir), sl, -sp);
        });
        sfor<NAo * NBo>([&](auto OB) {
          constexpr int ao = decltype(OB)::value / NBo, bo = decltype(OB)::value % NBo;
          double s = 0.0;
          sfor<NA * NB>([&](auto AB) {
            constexpr int a = decltype(AB)::value / NB, b = decltype(AB)::value % NB;
            constexpr double cc = c2x(SPH, LA, ao, a) * c2x(SPH, LB, bo, b);
            if constexpr (cc != 0.0) s = fma(cc, Yc[a * NB + b], s);
          });
          Ys[((x * NAo + ao) * NBo + bo) * NF + f] = s;
        });
      });
    }
  }
  __syncthreads();
  double g0 = 0.0, g1 = 0.0, g2 = 0.0;
  if (valid) {
    const VShell& sa = Bv.vsh[tk.sh[0]];
    const VShell& sb = Bv.vsh[tk.sh[1]];
    const VShell& sc = Bv.vsh[tk.sh[2]];
    const VShell& sd = Bv.vsh[tk.sh[3]];
    int64_t st0, st1, st2, st3;
    canon_strides<SPH>(Bv, tk, st0, st1, st2, st3);
    const int64_t stc = SPH ? (int64_t)sa.dsph * sb.dsph * sc.dsph * sd.dsph
                            : (int64_t)sa.dcart * sb.dcart * sc.dcart * sd.dcart;
    double* __restrict__ out = A.out + task_out(tk);
    if (!any) {
      if (!digest)
        for (int k = tl; k < 3 * NAo * NBo * NCo * NDo; k += TE) {
          const int ao = k % NAo, bo = (k / NAo) % NBo, co = (k / (NAo * NBo)) % NCo;
          const int dd = (k / (NAo * NBo * NCo)) % NDo, x = k / (NAo * NBo * NCo * NDo);
          out[x * stc + ao * st0 + bo * st1 + co * st2 + dd * st3] = 0.0;
        }
    } else {
      const double cd[3] = {sc.r[0] - sd.r[0], sc.r[1] - sd.r[1], sc.r[2] - sd.r[2]};
      double cdp[3][LD + 1];
      sfor<3>([&](auto DI) {
        constexpr int d = decltype(DI)::value;
        cdp[d][0] = 1.0;
        sfor<LD>([&](auto K) { cdp[d][decltype(K)::value + 1] = cdp[d][decltype(K)::value] * cd[d]; });
      });
      const double fkl = sc.shell != sd.shell ? 2.0 : 1.0;
      const int64_t N = A.nao;
      for (int row = tl; row < ROWS; row += TE) {
        const int x = row / (NAo * NBo);
        const int ao = (row / NBo) % NAo, bo = row % NBo;
        const double* yr = Ys + row * NF;
        double Zc[NC * ND];
        sfor<NC * ND>([&](auto CD) {
          constexpr int c = decltype(CD)::value / ND, d = decltype(CD)::value % ND;
          constexpr int cx = cpx(LC, c), cy = cpy(LC, c), cz = cpz(LC, c);
          constexpr int dx = cpx(LD, d), dy = cpy(LD, d), dz = cpz(LD, d);
          double s = 0.0;
          sfor<(dx + 1) * (dy + 1) * (dz + 1)>([&](auto TT) {
            constexpr int tt = decltype(TT)::value;
            constexpr int tx = tt / ((dy + 1) * (dz + 1)), ty = (tt / (dz + 1)) % (dy + 1), tz = tt % (dz + 1);
            constexpr double cb = binomc(dx, tx) * binomc(dy, ty) * binomc(dz, tz);
            constexpr int f = cindex(LC, cx + tx, cy + ty, cz + tz);
            const double pw = cdp[0][dx - tx] * cdp[1][dy - ty] * cdp[2][dz - tz];
            s = fma(cb * pw, yr[f], s);
          });
          Zc[c * ND + d] = s;
        });
        if (digest) {
          const double* __restrict__ D = A.dm;
          const int64_t Pi = sa.ao + ao, Qi = sb.ao + bo;
          const double* __restrict__ DP = D + Pi * N;
          const double* __restrict__ DQ = D + Qi * N;
          const double dpq = -2.0 * DP[Qi];
          double acc = 0.0;
          sfor<NCo * NDo>([&](auto OD) {
            constexpr int co = decltype(OD)::value / NDo, dd = decltype(OD)::value % NDo;
            double s = 0.0;
            sfor<NC * ND>([&](auto CD) {
              constexpr int c = decltype(CD)::value / ND, d = decltype(CD)::value % ND;
              constexpr double cc = c2x(SPH, LC, co, c) * c2x(SPH, LD, dd, d);
              if constexpr (cc != 0.0) s = fma(cc, Zc[c * ND + d], s);
            });
            const int64_t Ri = sc.ao + co, Si = sd.ao + dd;
            const double w = fma(dpq, D[Ri * N + Si], 0.5 * (DP[Ri] * DQ[Si] + DP[Si] * DQ[Ri]));
            acc = fma(s, w, acc);
          });
          acc *= fkl;
          if (x == 0) g0 += acc;
          else if (x == 1) g1 += acc;
          else g2 += acc;
        } else {
          double* orow = out + x * stc + ao * st0 + bo * st1;
          sfor<NCo * NDo>([&](auto OD) {
            constexpr int co = decltype(OD)::value / NDo, dd = decltype(OD)::value % NDo;
            double s = 0.0;
            sfor<NC * ND>([&](auto CD) {
              constexpr int c = decltype(CD)::value / ND, d = decltype(CD)::value % ND;
              constexpr double cc = c2x(SPH, LC, co, c) * c2x(SPH, LD, dd, d);
              if constexpr (cc != 0.0) s = fma(cc, Zc[c * ND + d], s);
            });
            orow[co * st2 + dd * st3] = s;
          });
        }
      }
    }
  }
  if (digest) {
    for (int off = 1; off < TE; off <<= 1) {
      g0 += __shfl_xor(g0, off);
      g1 += __shfl_xor(g1, off);
      g2 += __shfl_xor(g2, off);
    }
    if (tl == 0) {
      s_g[team][0] = g0;
      s_g[team][1] = g1;
      s_g[team][2] = g2;
      s_atom[team] = (valid && any) ? Bv.vsh[tk.sh[0]].atom : -1;
    }
    __syncthreads();
    grad_block_flush(A, s_g, s_atom, P::TPB);
  }
}

// ---------------------------------------------------------------------
// Wave kernel: LPT = 16, 32 or 64 lanes per quartet (a wave holds 64/LPT
// quartets) for chunked classes whose quartets carry few primitive
// quartets (the team-of-one buckets).  A quartet's lanes screen LPT
// candidate primitive quartets at once (one per lane, in libcint's order),
// compact the survivors and, a batch of PB at a time, evaluate all their
// roots (one (primitive quartet, root) per lane), then all their 2D
// recurrence tables (one (primitive quartet, direction, root) per lane) into
// LDS, so that the memory latency of pair data and root tables is paid once
// per batch; lane l owns the rows e = l (and l + LPT) of [e0|f0] with all NF
// columns in registers, and per root reads the three recurrence rows
// V_d(e_d, 0..F) once for its NF products.  Loop bounds are maxima over the
// wave's quartets, so every barrier is reached by the whole wave.
// Against the chunk kernels: no recurrence recomputed per chunk, one launch
// instead of one per chunk, a few dozen registers instead of the whole
// register file, so that many waves hide the recurrence's latency.  It
// writes the block column-major (coalesced) and the chunked epilogue, told
// so by BucketArgs::col_major, follows.
// ---------------------------------------------------------------------
template <int LA, int LB, int LC, int LD>
struct WaveCfg {
  using C = FastClass<LA, LB, LC, LD>;
  // lanes per quartet: the rows rounded up to 16, 32 or 64 (two rows per
  // lane above 64); a wave holds 64 / LPT quartets
  static constexpr int LPT = C::NE <= 16 ? 16 : (C::NE <= 32 ? 32 : 64);
  static constexpr int TPW = 64 / LPT;
  static constexpr int ROWS = (C::NE + LPT - 1) / LPT;
  static constexpr int VS = (C::E + 1) * (C::F + 1);
  static constexpr int PV = 3 * C::NR * VS;  // recurrence tables of one primitive quartet
  // primitive quartets per quartet whose tables share the LDS at once (20 KB
  // of tables per wave: two waves per SIMD)
  static constexpr int PB0 = 2560 / (TPW * PV);
  static constexpr int PB = PB0 < 1 ? 1 : (PB0 > LPT ? LPT : PB0);
  // compiled: accumulators that leave room for two waves per SIMD
  static constexpr bool OK = C::NE >= 6 && ROWS * C::NF <= 96 &&
                             !EpiCfg<LA, LB, LC, LD, true>::XG;
  // the epilogue in the same kernel (spherical store): the block and the
  // bra result of every quartet of the wave in LDS, at most 32 KB
  static constexpr int PERF = C::NE * C::NF + C::NAo * C::NBo * C::NF;
  static constexpr bool FUSE = TPW * PERF <= 4096;
  // the J/K digest in the same kernel: also the six density-block sums
  static constexpr int PERJ = PERF + EpiJkCfg<LA, LB, LC, LD>::MJK;
  static constexpr bool FUSE_JK = TPW * PERJ <= 4096;
  // used by default (without a tuned table): blocks of 480+ components,
  // measured faster than the chunk kernels on every such c1 bucket ((dd|dd)
  // 1.6 -> 1.0 ms, (fd|fd) 3.5 -> 1.8 ms, (fd|dp) 9.1 -> 6.7 ms)
  static constexpr bool DEFAULT = OK && C::NE * C::NF >= 480;
};

// Per primitive quartet of a wave-kernel batch: the data of its roots and
// recurrence (plain and long-range operators).
struct WavePrim {
  double aij, akl, a0, a1, x, fac, theta;
  double pq[3], pa[3], qc[3];
};

template <int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(64) eri_wave_kernel(BasisView Bv, BucketArgs A) {
  using C = FastClass<LA, LB, LC, LD>;
  using W = WaveCfg<LA, LB, LC, LD>;
  constexpr int NE = C::NE, NF = C::NF, E = C::E, F = C::F, NR = C::NR;
  constexpr int LPT = W::LPT, TPW = W::TPW, ROWS = W::ROWS, VS = W::VS, PV = W::PV, PB = W::PB;
  constexpr int NEF = NE * NF;
  // LDS: the walk's tables (recurrence, roots, primitive data), then, when
  // the epilogue is fused, each quartet's block and bra result
  constexpr int WALK_D = TPW * (PB * PV + 2 * PB * NR + PB * (int)(sizeof(WavePrim) / 8));
  constexpr int EPI_D = W::FUSE_JK ? TPW * W::PERJ : (W::FUSE ? TPW * W::PERF : 0);
  __shared__ double s_mem[WALK_D > EPI_D ? WALK_D : EPI_D];
  const int64_t t_first = A.t0 + (int64_t)blockIdx.x * TPW;
  if (t_first >= A.t1) return;  // the whole workgroup
  const int lane = threadIdx.x;
  const int sub = lane / LPT, sl = lane - sub * LPT;  // quartet of the wave, lane in it
  double* const sV = s_mem + sub * (PB * PV);
  double* const su = s_mem + TPW * PB * PV + sub * (PB * NR);
  double* const sw = s_mem + TPW * PB * PV + TPW * PB * NR + sub * (PB * NR);
  WavePrim* const sp =
      reinterpret_cast<WavePrim*>(s_mem + TPW * (PB * PV + 2 * PB * NR)) + sub * PB;
  const int64_t t = t_first + sub;
  const bool valid = t < A.t1;
  const Task tk = A.tasks[valid ? t : t_first];
  const VShell& sa = Bv.vsh[tk.sh[0]];
  const VShell& sb = Bv.vsh[tk.sh[1]];
  const VShell& sc = Bv.vsh[tk.sh[2]];
  const VShell& sd = Bv.vsh[tk.sh[3]];
  // this lane's rows: Cartesian exponents in the row list (shells LA..E)
  int rx[ROWS], ry[ROWS], rz[ROWS];
  bool has[ROWS];
  sfor<ROWS>([&](auto S) {
    constexpr int r = decltype(S)::value;
    const int e = sl + LPT * r;
    has[r] = e < NE;
    const int ee = has[r] ? e : 0;
    rx[r] = cpx(LA, ee);
    ry[r] = cpy(LA, ee);
    rz[r] = cpz(LA, ee);
  });
  double acc[ROWS][NF];
  sfor<ROWS>([&](auto S) {
    sfor<NF>([&](auto J) { acc[decltype(S)::value][decltype(J)::value] = 0.0; });
  });

  // libcint's screening, as in gen_walk for plain four-centre integrals
  // (src/cint2e.c:57-269 with the optimizer's cutoffs)
  const int nij = sa.nprim * sb.nprim, nkl = sc.nprim * sd.nprim;
  const int64_t pab = (int64_t)tk.sh[0] * Bv.nv + tk.sh[1];
  const int64_t pcd = (int64_t)tk.sh[2] * Bv.nv + tk.sh[3];
  const int64_t oij = Bv.pair_off[pab], okl = Bv.pair_off[pcd];
  const PairPrim* __restrict__ pij = Bv.pp + oij;
  const PairPrim* __restrict__ pkl = Bv.pp + okl;
  const int* __restrict__ pos_ij = Bv.pp_pos + oij;
  const int* __restrict__ pos_kl = Bv.pp_pos + okl;
  const double dij = (LA + LB) * Bv.pair_ld[pab], dkl = (LC + LD) * Bv.pair_ld[pcd];
  const double Ecut = Bv.expcutoff_4c;
  const double omega = Bv.omega;
  constexpr double cfac = 34.986836655249725 *
                          (LA == 0 ? 0.282094791773878143 : LA == 1 ? 0.488602511902919921 : 1.0) *
                          (LB == 0 ? 0.282094791773878143 : LB == 1 ? 0.488602511902919921 : 1.0) *
                          (LC == 0 ? 0.282094791773878143 : LC == 1 ? 0.488602511902919921 : 1.0) *
                          (LD == 0 ? 0.282094791773878143 : LD == 1 ? 0.488602511902919921 : 1.0);
  const unsigned long long own = (LPT == 64 ? ~0ull : ((1ull << LPT) - 1)) << (sub * LPT);
  // pair emptiness (src/cint2e.c:96-109), per quartet of the wave
  int live_ij = 0, live_kl = 0;
  for (int p = sl; p < nij; p += LPT) live_ij |= (pij[p].cce0 - dij) < Ecut;
  for (int p = sl; p < nkl; p += LPT) live_kl |= (pkl[p].cce0 - dkl) < Ecut;
  const bool live = valid && (__ballot(live_ij) & own) != 0 && (__ballot(live_kl) & own) != 0;
  const int npq = live ? nij * nkl : 0;
  // candidate rounds common to the wave's quartets (barriers stay uniform)
  int maxpq = npq;
  for (int off = LPT; off < 64; off <<= 1) maxpq = max(maxpq, __shfl_xor(maxpq, off));
  int any = 0;
  for (int base = 0; base < maxpq; base += LPT) {
    // one candidate per lane, in libcint's order (ip fastest)
    const int p = base + sl;
    bool pass = false;
    WavePrim wp;
    if (p < npq) {
      const PairPrim P = pij[pos_ij[p % nij]];
      const PairPrim Q = pkl[pos_kl[p / nij]];
      const double cce_kl = Q.cce0 - dkl, cce_ij = P.cce0 - dij;
      if (!(cce_kl > Ecut) && !(cce_ij > Ecut - cce_kl)) {
        any = 1;
        pass = cce_ij < Ecut && cce_kl < Ecut;  // else the stored exp factor is 0
      }
      if (pass) {
        wp.aij = P.aij;
        wp.akl = Q.aij;
        wp.a1 = P.aij * Q.aij;
        wp.a0 = wp.a1 * rcp_pos(P.aij + Q.aij);
        wp.pq[0] = P.px - Q.px;
        wp.pq[1] = P.py - Q.py;
        wp.pq[2] = P.pz - Q.pz;
        wp.pa[0] = P.px - sa.r[0];
        wp.pa[1] = P.py - sa.r[1];
        wp.pa[2] = P.pz - sa.r[2];
        wp.qc[0] = Q.px - sc.r[0];
        wp.qc[1] = Q.py - sc.r[1];
        wp.qc[2] = Q.pz - sc.r[2];
        const double rr = wp.pq[0] * wp.pq[0] + wp.pq[1] * wp.pq[1] + wp.pq[2] * wp.pq[2];
        wp.fac = sqrt_pos(wp.a0 * rcp_pos(wp.a1 * wp.a1 * wp.a1)) * cfac * P.kab * Q.kab;
        wp.x = wp.a0 * rr;
        wp.theta = 1.0;
        if (omega > 0) {
          wp.theta = omega * omega * rcp_pos(omega * omega + wp.a0);
          wp.x *= wp.theta;
          wp.fac *= sqrt_pos(wp.theta);
        }
      }
    }
    const unsigned long long mask = __ballot(pass) & own;
    const int nsurv = __popcll(mask);
    const int slot = __popcll(mask & ((1ull << lane) - 1));
    int maxsurv = nsurv;
    for (int off = LPT; off < 64; off <<= 1) maxsurv = max(maxsurv, __shfl_xor(maxsurv, off));
    for (int b0 = 0; b0 < maxsurv; b0 += PB) {
      const int np = nsurv - b0 < 0 ? 0 : (nsurv - b0 < PB ? nsurv - b0 : PB);
      if (pass && slot >= b0 && slot < b0 + PB) sp[slot - b0] = wp;
      __syncthreads();
      // roots and weights of every (primitive quartet, root) of this quartet
      for (int it = sl; it < np * NR; it += LPT) {
        const int q = it / NR, r = it - q * NR;
        double u, w;
        rys_root_one(NR, r, sp[q].x, u, w);
        if (omega > 0) {
          const double ut = u * sp[q].theta;
          u = ut * rcp_pos(u + 1.0 - ut);
        }
        su[it] = u;
        sw[it] = w * sp[q].fac;
      }
      __syncthreads();
      // 2D recurrence of every (primitive quartet, direction, root)
      for (int it = sl; it < np * 3 * NR; it += LPT) {
        const int q = it / (3 * NR), dr = it - q * 3 * NR, d = dr / NR, r = dr - d * NR;
        const WavePrim& P = sp[q];
        const double u2 = P.a0 * su[q * NR + r];
        const double tmp4 = 0.5 * rcp_pos(u2 * (P.aij + P.akl) + P.a1);
        const double tmp5 = u2 * tmp4;
        const double b00 = tmp5, b10 = tmp5 + tmp4 * P.akl, b01 = tmp5 + tmp4 * P.aij;
        const double c00 = P.pa[d] - 2.0 * tmp5 * P.akl * P.pq[d];
        const double c0p = P.qc[d] + 2.0 * tmp5 * P.aij * P.pq[d];
        double* V = sV + q * PV + dr * VS;
        V[0] = d == 2 ? sw[q * NR + r] : 1.0;
        sfor<E>([&](auto EE) {
          constexpr int e = decltype(EE)::value;
          double v = c00 * V[e * (F + 1)];
          if constexpr (e > 0) v += e * b10 * V[(e - 1) * (F + 1)];
          V[(e + 1) * (F + 1)] = v;
        });
        sfor<F>([&](auto FF) {
          constexpr int f = decltype(FF)::value;
          sfor<E + 1>([&](auto EE) {
            constexpr int e = decltype(EE)::value;
            double v = c0p * V[e * (F + 1) + f];
            if constexpr (f > 0) v += f * b01 * V[e * (F + 1) + f - 1];
            if constexpr (e > 0) v += e * b00 * V[(e - 1) * (F + 1) + f];
            V[e * (F + 1) + f + 1] = v;
          });
        });
      }
      __syncthreads();
      // [e0|f0] products of this lane's rows
      for (int q = 0; q < np; ++q) {
        for (int r = 0; r < NR; ++r) {
          const double* Vx = sV + q * PV + (0 * NR + r) * VS;
          const double* Vy = sV + q * PV + (1 * NR + r) * VS;
          const double* Vz = sV + q * PV + (2 * NR + r) * VS;
          sfor<ROWS>([&](auto S) {
            constexpr int rw = decltype(S)::value;
            if (has[rw]) {
              double vx[F + 1], vy[F + 1], vz[F + 1];
              sfor<F + 1>([&](auto K) {
                constexpr int k = decltype(K)::value;
                vx[k] = Vx[rx[rw] * (F + 1) + k];
                vy[k] = Vy[ry[rw] * (F + 1) + k];
                vz[k] = Vz[rz[rw] * (F + 1) + k];
              });
              sfor<NF>([&](auto J) {
                constexpr int j = decltype(J)::value;
                constexpr int fx = cpx(LC, j), fy = cpy(LC, j), fz = cpz(LC, j);
                acc[rw][j] = fma(vx[fx] * vy[fy], vz[fz], acc[rw][j]);
              });
            }
          });
        }
      }
      __syncthreads();
    }
  }
  any = (__ballot(any) & own) != 0 && live;
  if constexpr (W::FUSE_JK) {
    if (A.sph && (A.jk_j || A.jk_k)) {
      // the J/K digest here: block and bra result in LDS as for the store,
      // then the ket rows' products with the density summed into the six
      // block updates in LDS (epi_ket_jk) and added once per entry
      using Q = EpiJkCfg<LA, LB, LC, LD>;
      double* X = s_mem + sub * W::PERJ;
      double* Ys = X + NEF;
      double* G = X + W::PERF;
      if (valid) {
        sfor<ROWS>([&](auto S) {
          constexpr int rw = decltype(S)::value;
          if (has[rw]) {
            const int e = sl + LPT * rw;
            sfor<NF>([&](auto J) {
              constexpr int j = decltype(J)::value;
              X[e * NF + j] = acc[rw][j];
            });
          }
        });
        for (int k = sl; k < Q::MJK; k += LPT) G[k] = 0.0;
      }
      __syncthreads();
      if (valid && any) epi_bra<LA, LB, LC, LD, true>(Bv, tk, X, Ys, sl, LPT, nullptr);
      __syncthreads();
      if (valid && any) epi_ket_jk<LA, LB, LC, LD>(Bv, tk, Ys, G, A, sl, LPT);
      __syncthreads();
      if (valid && any) {
        const double f = jk_degeneracy(sa.shell, sb.shell, sc.shell, sd.shell);
        const int64_t N = A.nao;
        const int64_t P0 = sa.ao, Q0 = sb.ao, R0 = sc.ao, S0 = sd.ao;
        for (int m = sl; m < Q::MJK; m += LPT) {
          const double v = G[m];
          if (v == 0.0) continue;
          const int r = m < Q::O1 ? 0 : m < Q::O2 ? 1 : m < Q::O3 ? 2 : m < Q::O4 ? 3 : m < Q::O5 ? 4 : 5;
          double* dst = r < 2 ? A.jk_j : A.jk_k;
          if (!dst) continue;
          const int o0 = r == 0 ? 0 : r == 1 ? Q::O1 : r == 2 ? Q::O2 : r == 3 ? Q::O3 : r == 4 ? Q::O4 : Q::O5;
          const int ncol = r == 0 ? Q::NBo : (r == 1 || r == 3 || r == 5) ? Q::NDo : Q::NCo;
          const int64_t row0 = r == 0 ? P0 : r == 1 ? R0 : r <= 3 ? P0 : Q0;
          const int64_t col0 = r == 0 ? Q0 : r == 1 ? S0 : (r == 2 || r == 4) ? R0 : S0;
          const int i = (m - o0) / ncol, j = (m - o0) % ncol;
          atomicAdd(dst + (row0 + i) * N + col0 + j, (r < 2 ? 2.0 * f : f) * v);
        }
      }
      return;
    }
  }
  if constexpr (W::FUSE) {
    if (A.sph && !A.jk_j && !A.jk_k) {
      // epilogue here: the block to LDS (flat order), the bra transfer and
      // spherical transform, then the ket and the store, on this quartet's
      // lanes (the walk's LDS is free after its last barrier)
      double* X = s_mem + sub * W::PERF;
      double* Ys = X + NEF;
      if (valid) {
        sfor<ROWS>([&](auto S) {
          constexpr int rw = decltype(S)::value;
          if (has[rw]) {
            const int e = sl + LPT * rw;
            sfor<NF>([&](auto J) {
              constexpr int j = decltype(J)::value;
              X[e * NF + j] = acc[rw][j];
            });
          }
        });
      }
      __syncthreads();
      if (valid && any) epi_bra<LA, LB, LC, LD, true>(Bv, tk, X, Ys, sl, LPT, nullptr);
      __syncthreads();
      if (valid) {
        epi_ket_store<LA, LB, LC, LD, true>(Bv, tk, Ys, A.out, any, sl, LPT);
        if (sl == 0 && A.not0) A.not0[A.task_id ? A.task_id[t] : t] = any;
      }
      return;
    }
  }
  if (!valid) return;
  // column-major ([f][e]): for each column the lanes' rows are contiguous
  const int64_t slot_t = t - A.t0;
  double* __restrict__ dst = A.scratch + slot_t * NEF;
  sfor<ROWS>([&](auto S) {
    constexpr int rw = decltype(S)::value;
    if (has[rw]) {
      const int e = sl + LPT * rw;
      sfor<NF>([&](auto J) {
        constexpr int j = decltype(J)::value;
        dst[j * NE + e] = acc[rw][j];
      });
    }
  });
  if (sl == 0) {
    A.flags[slot_t] = any;
    if (A.not0) A.not0[A.task_id ? A.task_id[t] : t] = any;
  }
}

template <int LA, int LB, int LC, int LD, int CH, int DER>
hipError_t launch_one_chunk(const BasisView& B, const BucketArgs& A, hipStream_t stream, dim3 grid) {
  constexpr int TH = ChunkCfg<LA, LB, LC, LD, DER>::THREADS;
  // the FP32 tier exists for the derivative chunks (gradient digest)
  if constexpr (DER == 1) {
    if (A.fp32) {
      hipLaunchKernelGGL((eri_chunk_kernel<LA, LB, LC, LD, CH, DER, float>), grid, dim3(TH), 0,
                         stream, B, A);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((eri_chunk_kernel<LA, LB, LC, LD, CH, DER>), grid, dim3(TH), 0, stream, B, A);
  return hipGetLastError();
}

template <int LA, int LB, int LC, int LD, int DER, int... CHs>
hipError_t launch_chunks(const BasisView& B, const BucketArgs& A0, hipStream_t stream,
                         std::integer_sequence<int, CHs...>) {
  const int64_t n = A0.t1 - A0.t0;
  const int64_t per = ChunkCfg<LA, LB, LC, LD, DER>::THREADS / A0.team;
  const dim3 grid((unsigned)((n + per - 1) / per));
  hipError_t e = hipSuccess;
  BucketArgs A = A0;
  if constexpr (DER == 0 && WaveCfg<LA, LB, LC, LD>::OK) {
    // team-of-one buckets (few primitive quartets per quartet), plain
    // operator and long range, FP64
    // per class and team size: the tuned choice, else the built-in rule
    // (team-of-one buckets, blocks of 480+ components)
    int tl = 0;
    while ((1 << tl) < A.team && tl < 6) ++tl;
    const int cw = g_class_wave[class_index(LA, LB, LC, LD) * 7 + tl];
    const bool dflt = A.team == 1 && WaveCfg<LA, LB, LC, LD>::DEFAULT;
    if (B.omega >= 0 && !A.fp32 &&
        (g_wave_kernels == 2 ||
         (g_wave_kernels == 1 && (cw ? cw == 2 : dflt)))) {
      A.col_major = 1;
      constexpr int TPW = WaveCfg<LA, LB, LC, LD>::TPW;
      hipLaunchKernelGGL((eri_wave_kernel<LA, LB, LC, LD>), dim3((unsigned)((n + TPW - 1) / TPW)),
                         dim3(64), 0, stream, B, A);
      e = hipGetLastError();
    }
  }
  if (!A.col_major)
    ((e = (e == hipSuccess ? launch_one_chunk<LA, LB, LC, LD, CHs, DER>(B, A, stream, grid) : e)), ...);
  if (e != hipSuccess) return e;
  if constexpr (DER == 0 && WaveCfg<LA, LB, LC, LD>::OK) {
    // the wave kernel stored or digested the blocks itself
    using W = WaveCfg<LA, LB, LC, LD>;
    const bool jk = A.jk_j || A.jk_k;
    if (A.col_major && A.sph && (jk ? W::FUSE_JK : W::FUSE)) return hipSuccess;
  }
  if constexpr (DER == 0) {
    if (A.jk_j || A.jk_k) {
      using Q = EpiJkCfg<LA, LB, LC, LD>;
      if (!A.sph) return hipErrorInvalidValue;
      const dim3 egrid((unsigned)((n + Q::TPB - 1) / Q::TPB));
      hipLaunchKernelGGL((eri_epi_jk_kernel<LA, LB, LC, LD>), egrid, dim3(Q::THREADS), 0, stream, B,
                         A);
      return hipGetLastError();
    }
  }
  if constexpr (DER == 0 || DER == 2) {
    if (A.sph) {
      using P = EpiCfg<LA, LB, LC, LD>;
      const dim3 egrid((unsigned)((n + P::TPB - 1) / P::TPB));
      hipLaunchKernelGGL((eri_epi_kernel<LA, LB, LC, LD, true, DER>), egrid, dim3(P::THREADS), 0,
                         stream, B, A);
    } else {
      using P = EpiCfg<LA, LB, LC, LD, false>;
      const dim3 egrid((unsigned)((n + P::TPB - 1) / P::TPB));
      hipLaunchKernelGGL((eri_epi_kernel<LA, LB, LC, LD, false, DER>), egrid, dim3(P::THREADS), 0,
                         stream, B, A);
    }
  } else if constexpr (DER != 1) {
    using P = EpiCfg<LA, LB, LC, LD>;
    const dim3 egrid((unsigned)((n + P::TPB - 1) / P::TPB));
    hipLaunchKernelGGL((eri_epi_kernel<LA, LB, LC, LD, true, DER>), egrid, dim3(P::THREADS), 0,
                       stream, B, A);
  } else if (A.sph) {
    using P = Ip1EpiCfg<LA, LB, LC, LD>;
    const dim3 egrid((unsigned)((n + P::TPB - 1) / P::TPB));
    hipLaunchKernelGGL((eri_ip1_epi_kernel<LA, LB, LC, LD>), egrid, dim3(P::THREADS), 0, stream, B, A);
  } else {
    using P = Ip1EpiCfg<LA, LB, LC, LD, false>;
    const dim3 egrid((unsigned)((n + P::TPB - 1) / P::TPB));
    hipLaunchKernelGGL((eri_ip1_epi_kernel<LA, LB, LC, LD, false>), egrid, dim3(P::THREADS), 0, stream, B, A);
  }
  return hipGetLastError();
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_chunked_class(const BasisView& B, const BucketArgs& A, hipStream_t stream);

// Classes with a single-chunk fast kernel (canonical: LA>=LB, LC>=LD,
// (LA,LB)>=(LC,LD), [e0|f0] block of at most 100 doubles, beyond which the
// accumulators no longer fit the register file, and (fd|ps) at 138, measured
// faster whole than chunked), spherical output of plain 4c ERIs.  The list is split in groups, one translation unit each
// (fast_<g>.hip).
#define MCINT_FAST_GROUP0(X) X(0, 0, 0, 0) X(1, 0, 0, 0) X(1, 0, 1, 0) X(1, 1, 0, 0) X(2, 0, 0, 0) X(3, 0, 0, 0)
#define MCINT_FAST_GROUP1(X) X(1, 1, 1, 0) X(2, 0, 1, 0) X(2, 1, 0, 0) X(2, 2, 0, 0) X(3, 1, 0, 0)
#define MCINT_FAST_GROUP2(X) X(3, 0, 1, 0) X(3, 2, 0, 0)
#define MCINT_FAST_GROUP3(X) X(2, 1, 1, 0)
#define MCINT_FAST_GROUP4(X) X(2, 0, 2, 0)
#define MCINT_FAST_GROUP5(X) X(3, 0, 3, 0)
#define MCINT_FAST_GROUP6(X) X(2, 1, 2, 0)
#define MCINT_FAST_GROUP7(X)
#define MCINT_FAST_GROUP8(X) X(3, 3, 0, 0) X(3, 2, 1, 0)
#define MCINT_FAST_CLASSES(X)                                                         \
  MCINT_FAST_GROUP0(X) MCINT_FAST_GROUP1(X) MCINT_FAST_GROUP2(X) MCINT_FAST_GROUP3(X) \
  MCINT_FAST_GROUP4(X) MCINT_FAST_GROUP5(X) MCINT_FAST_GROUP6(X) MCINT_FAST_GROUP7(X) \
  MCINT_FAST_GROUP8(X)

// Classes on the chunked path ([e0|f0] blocks above the single-kernel
// budget whose epilogue fits 48 KB of LDS), one translation unit per group
// (chunk_<g>.hip).
#define MCINT_CHUNK_GROUP0(X) X(3, 1, 1, 1) X(3, 1, 3, 0)
#define MCINT_CHUNK_GROUP1(X) X(2, 1, 2, 1) X(2, 2, 2, 0)
#define MCINT_CHUNK_GROUP2(X) X(2, 2, 1, 1) X(3, 3, 1, 0)
#define MCINT_CHUNK_GROUP3(X) X(3, 2, 1, 1) X(3, 2, 2, 0)
#define MCINT_CHUNK_GROUP4(X) X(3, 1, 2, 1) X(3, 0, 2, 2)
#define MCINT_CHUNK_GROUP5(X) X(2, 2, 2, 1) X(3, 3, 2, 0)
#define MCINT_CHUNK_GROUP6(X) X(3, 1, 3, 1) X(3, 2, 3, 0)
#define MCINT_CHUNK_GROUP7(X) X(3, 3, 1, 1)
#define MCINT_CHUNK_GROUP8(X) X(3, 2, 2, 1)
#define MCINT_CHUNK_GROUP9(X) X(3, 3, 3, 0)
#define MCINT_CHUNK_GROUP10(X) X(3, 1, 2, 2)
// larger blocks (961..3404 doubles), f and d shells
#define MCINT_CHUNK_GROUP11(X) X(3, 3, 3, 2)
#define MCINT_CHUNK_GROUP12(X) X(3, 3, 2, 2)
#define MCINT_CHUNK_GROUP13(X) X(3, 2, 3, 2)
#define MCINT_CHUNK_GROUP14(X) X(3, 3, 3, 1)
#define MCINT_CHUNK_GROUP15(X) X(3, 2, 2, 2) X(2, 2, 2, 2)
#define MCINT_CHUNK_GROUP16(X) X(3, 3, 2, 1) X(3, 2, 3, 1)
#define MCINT_CHUNK_GROUP17(X) X(3, 3, 3, 3)
// accumulator blocks of 144..160 doubles: two chunks rather than a spilling
// single kernel
#define MCINT_CHUNK_GROUP18(X) X(2, 1, 1, 1) X(3, 1, 2, 0)
#define MCINT_CHUNK_GROUP19(X) X(3, 0, 2, 1) X(1, 1, 1, 1)
// 54..90 accumulators, at most three roots: two chunks beat one kernel
#define MCINT_CHUNK_GROUP20(X) X(2, 0, 1, 1) X(3, 0, 1, 1)
#define MCINT_CHUNK_GROUP21(X) X(2, 2, 1, 0) X(3, 0, 2, 0) X(3, 1, 1, 0)
#define MCINT_CHUNK_CLASSES(X)                                                              \
  MCINT_CHUNK_GROUP0(X) MCINT_CHUNK_GROUP1(X) MCINT_CHUNK_GROUP2(X) MCINT_CHUNK_GROUP3(X)   \
  MCINT_CHUNK_GROUP4(X) MCINT_CHUNK_GROUP5(X) MCINT_CHUNK_GROUP6(X) MCINT_CHUNK_GROUP7(X)   \
  MCINT_CHUNK_GROUP8(X) MCINT_CHUNK_GROUP9(X) MCINT_CHUNK_GROUP10(X) MCINT_CHUNK_GROUP11(X) \
  MCINT_CHUNK_GROUP12(X) MCINT_CHUNK_GROUP13(X) MCINT_CHUNK_GROUP14(X) MCINT_CHUNK_GROUP15(X) \
  MCINT_CHUNK_GROUP16(X) MCINT_CHUNK_GROUP17(X) MCINT_CHUNK_GROUP18(X) MCINT_CHUNK_GROUP19(X) \
  MCINT_CHUNK_GROUP20(X) MCINT_CHUNK_GROUP21(X)

// i-derivative (int2e_ip1) classes on the chunked path: shells up to d,
// canonical LC >= LD (the derivative shell i stays first), one translation
// unit per group (ip1_<g>.hip).
#define MCINT_IP1_GROUP0(X) X(1, 2, 2, 2)
#define MCINT_IP1_GROUP1(X) X(2, 1, 2, 2) X(0, 1, 1, 0) X(2, 0, 1, 0)
#define MCINT_IP1_GROUP2(X) X(2, 2, 2, 1) X(0, 1, 1, 1) X(2, 0, 2, 0)
#define MCINT_IP1_GROUP3(X) X(0, 2, 2, 2) X(0, 2, 1, 1) X(0, 0, 2, 0) X(1, 1, 0, 0)
#define MCINT_IP1_GROUP4(X) X(1, 1, 2, 2) X(0, 2, 2, 0) X(0, 0, 2, 1) X(1, 1, 1, 0)
#define MCINT_IP1_GROUP5(X) X(1, 2, 2, 1) X(1, 0, 2, 1) X(0, 0, 2, 2) X(1, 2, 0, 0)
#define MCINT_IP1_GROUP6(X) X(2, 2, 1, 1) X(1, 1, 1, 1) X(0, 1, 0, 0) X(2, 0, 0, 0)
#define MCINT_IP1_GROUP7(X) X(2, 1, 2, 1) X(0, 1, 2, 1) X(0, 0, 0, 0) X(0, 1, 2, 0) X(2, 1, 0, 0)
#define MCINT_IP1_GROUP8(X) X(2, 0, 2, 2) X(2, 1, 2, 0) X(0, 0, 1, 0) X(0, 2, 0, 0) X(2, 2, 0, 0)
#define MCINT_IP1_GROUP9(X) X(2, 2, 2, 0) X(2, 2, 1, 0) X(0, 0, 1, 1) X(0, 2, 1, 0)
#define MCINT_IP1_GROUP10(X) X(0, 2, 2, 1) X(0, 1, 2, 2) X(1, 1, 2, 0) X(1, 0, 0, 0)
#define MCINT_IP1_GROUP11(X) X(1, 1, 2, 1) X(1, 0, 2, 2) X(1, 2, 1, 0) X(1, 0, 1, 0)
#define MCINT_IP1_GROUP12(X) X(1, 2, 1, 1) X(1, 2, 2, 0) X(2, 0, 1, 1) X(1, 0, 1, 1)
#define MCINT_IP1_GROUP13(X) X(2, 1, 1, 1) X(2, 0, 2, 1) X(2, 1, 1, 0) X(1, 0, 2, 0)
#define MCINT_IP1_GROUP14(X) X(2, 2, 2, 2)
#define MCINT_IP1_CLASSES(X)                                                        \
  MCINT_IP1_GROUP0(X) MCINT_IP1_GROUP1(X) MCINT_IP1_GROUP2(X) MCINT_IP1_GROUP3(X)   \
  MCINT_IP1_GROUP4(X) MCINT_IP1_GROUP5(X) MCINT_IP1_GROUP6(X) MCINT_IP1_GROUP7(X)   \
  MCINT_IP1_GROUP8(X) MCINT_IP1_GROUP9(X) MCINT_IP1_GROUP10(X) MCINT_IP1_GROUP11(X) \
  MCINT_IP1_GROUP12(X) MCINT_IP1_GROUP13(X) MCINT_IP1_GROUP14(X)

// Three-centre (ij|k) classes (LA >= LB, aux l_k = LC, LD = 0): single
// kernel for blocks <= 100 doubles, chunked above; one translation unit per
// group (fast3_<g>.hip, chunk3_<g>.hip).
#define MCINT_FAST3_GROUP0(X) X(0, 0, 0, 0) X(0, 0, 1, 0) X(0, 0, 2, 0) X(0, 0, 3, 0) X(0, 0, 4, 0) \
  X(1, 0, 0, 0) X(1, 0, 1, 0) X(1, 0, 2, 0)
#define MCINT_FAST3_GROUP1(X) X(1, 0, 3, 0) X(1, 0, 4, 0) X(1, 1, 0, 0) X(1, 1, 1, 0) X(1, 1, 2, 0) \
  X(1, 1, 3, 0) X(2, 0, 0, 0) X(2, 0, 1, 0)
#define MCINT_FAST3_GROUP2(X) X(2, 0, 2, 0) X(2, 0, 3, 0) X(2, 0, 4, 0) X(2, 1, 0, 0) X(2, 1, 1, 0) \
  X(2, 1, 2, 0) X(2, 2, 0, 0) X(2, 2, 1, 0)
#define MCINT_FAST3_GROUP3(X) X(3, 0, 0, 0) X(3, 0, 1, 0) X(3, 0, 2, 0) X(3, 0, 3, 0) X(3, 1, 0, 0) \
  X(3, 1, 1, 0) X(3, 2, 0, 0) X(3, 3, 0, 0)
#define MCINT_FAST3_CLASSES(X) \
  MCINT_FAST3_GROUP0(X) MCINT_FAST3_GROUP1(X) MCINT_FAST3_GROUP2(X) MCINT_FAST3_GROUP3(X)
#define MCINT_CHUNK3_GROUP0(X) X(3, 3, 4, 0) X(1, 1, 4, 0) X(2, 1, 3, 0) X(2, 1, 4, 0)
#define MCINT_CHUNK3_GROUP1(X) X(3, 3, 3, 0) X(3, 2, 4, 0) X(3, 2, 1, 0)
#define MCINT_CHUNK3_GROUP2(X) X(3, 3, 2, 0) X(3, 2, 3, 0) X(2, 2, 4, 0) X(2, 2, 2, 0) X(3, 0, 4, 0)
#define MCINT_CHUNK3_GROUP3(X) X(3, 1, 4, 0) X(2, 2, 3, 0) X(3, 3, 1, 0) X(3, 1, 3, 0) \
  X(3, 2, 2, 0) X(3, 1, 2, 0)
#define MCINT_CHUNK3_CLASSES(X) \
  MCINT_CHUNK3_GROUP0(X) MCINT_CHUNK3_GROUP1(X) MCINT_CHUNK3_GROUP2(X) MCINT_CHUNK3_GROUP3(X)

template <int LA, int LB, int LC, int LD>
hipError_t launch_fast_class(const BasisView& B, const BucketArgs& A, hipStream_t stream);
template <int LA, int LB, int LC, int LD>
hipError_t launch_ip1_class(const BasisView& B, const BucketArgs& A, hipStream_t stream);
template <int LA, int LB, int LC, int LD>
hipError_t launch_fast3_class(const BasisView& B, const BucketArgs& A, hipStream_t stream);
template <int LA, int LB, int LC, int LD>
hipError_t launch_chunk3_class(const BasisView& B, const BucketArgs& A, hipStream_t stream);

// explicit instantiation of one class (used by fast_<g>.hip)
#define MCINT_FAST_DEFINE(a, b, c, d)                                                      \
  template <>                                                                              \
  hipError_t launch_fast_class<a, b, c, d>(const BasisView& B, const BucketArgs& A,        \
                                           hipStream_t stream) {                           \
    const int64_t n = A.t1 - A.t0;                                                         \
    const int64_t per = 256 / A.team;                                                      \
    const dim3 grid((unsigned)((n + per - 1) / per));                                      \
    if (B.omega < 0)                                                                       \
      hipLaunchKernelGGL((eri_fast_kernel<a, b, c, d, double, 4>), grid, dim3(256), 0, stream, B, A); \
    else if (!A.sph)                                                                       \
      hipLaunchKernelGGL((eri_fast_kernel<a, b, c, d, double, 0, false>), grid, dim3(256), 0, stream, B, A); \
    else if ((A.jk_j || A.jk_k) && A.fp32)                                                 \
      hipLaunchKernelGGL((eri_fast_kernel<a, b, c, d, float, 0, true, true>), grid, dim3(256), 0, \
                         stream, B, A);                                                    \
    else if (A.jk_j || A.jk_k)                                                             \
      hipLaunchKernelGGL((eri_fast_kernel<a, b, c, d, double, 0, true, true>), grid, dim3(256), 0, \
                         stream, B, A);                                                    \
    else if (A.fp32)                                                                       \
      hipLaunchKernelGGL((eri_fast_kernel<a, b, c, d, float>), grid, dim3(256), 0, stream, B, A); \
    else                                                                                   \
      hipLaunchKernelGGL((eri_fast_kernel<a, b, c, d, double>), grid, dim3(256), 0, stream, B, A); \
    return hipGetLastError();                                                              \
  }
#define MCINT_CHUNK_DEFINE(a, b, c, d)                                                     \
  template <>                                                                              \
  hipError_t launch_chunked_class<a, b, c, d>(const BasisView& B, const BucketArgs& A,     \
                                              hipStream_t stream) {                        \
    if (B.omega < 0)                                                                       \
      return launch_chunks<a, b, c, d, 4>(                                                 \
          B, A, stream, std::make_integer_sequence<int, ChunkCfg<a, b, c, d, 4>::NCH>{});  \
    return launch_chunks<a, b, c, d, 0>(                                                   \
        B, A, stream, std::make_integer_sequence<int, ChunkCfg<a, b, c, d>::NCH>{});       \
  }
#define MCINT_FAST3_DEFINE(a, b, c, d)                                                     \
  template <>                                                                              \
  hipError_t launch_fast3_class<a, b, c, d>(const BasisView& B, const BucketArgs& A,       \
                                            hipStream_t stream) {                          \
    const int64_t n = A.t1 - A.t0;                                                         \
    const int64_t per = 256 / A.team;                                                      \
    const dim3 grid((unsigned)((n + per - 1) / per));                                      \
    if (B.omega < 0)                                                                       \
      hipLaunchKernelGGL((eri_fast_kernel<a, b, c, d, double, 3>), grid, dim3(256), 0, stream, B, A); \
    else if (!A.sph)                                                                       \
      hipLaunchKernelGGL((eri_fast_kernel<a, b, c, d, double, 2, false>), grid, dim3(256), 0, stream, B, A); \
    else if (A.fp32)                                                                       \
      hipLaunchKernelGGL((eri_fast_kernel<a, b, c, d, float, 2>), grid, dim3(256), 0, stream, B, A); \
    else                                                                                   \
      hipLaunchKernelGGL((eri_fast_kernel<a, b, c, d, double, 2>), grid, dim3(256), 0, stream, B, A); \
    return hipGetLastError();                                                              \
  }
#define MCINT_CHUNK3_DEFINE(a, b, c, d)                                                    \
  template <>                                                                              \
  hipError_t launch_chunk3_class<a, b, c, d>(const BasisView& B, const BucketArgs& A,      \
                                             hipStream_t stream) {                         \
    if (B.omega < 0)                                                                       \
      return launch_chunks<a, b, c, d, 3>(                                                 \
          B, A, stream, std::make_integer_sequence<int, ChunkCfg<a, b, c, d, 3>::NCH>{});  \
    return launch_chunks<a, b, c, d, 2>(                                                   \
        B, A, stream, std::make_integer_sequence<int, ChunkCfg<a, b, c, d, 2>::NCH>{});    \
  }
#define MCINT_IP1_DEFINE(a, b, c, d)                                                       \
  template <>                                                                              \
  hipError_t launch_ip1_class<a, b, c, d>(const BasisView& B, const BucketArgs& A,         \
                                          hipStream_t stream) {                            \
    return launch_chunks<a, b, c, d, 1>(                                                   \
        B, A, stream, std::make_integer_sequence<int, ChunkCfg<a, b, c, d, 1>::NCH>{});    \
  }

#define MCINT_FAST_DECLARE(a, b, c, d)                                                     \
  template <>                                                                              \
  hipError_t launch_fast_class<a, b, c, d>(const BasisView& B, const BucketArgs& A,        \
                                           hipStream_t stream);
#define MCINT_CHUNK_DECLARE(a, b, c, d)                                                    \
  template <>                                                                              \
  hipError_t launch_chunked_class<a, b, c, d>(const BasisView& B, const BucketArgs& A,     \
                                              hipStream_t stream);
#define MCINT_IP1_DECLARE(a, b, c, d)                                                      \
  template <>                                                                              \
  hipError_t launch_ip1_class<a, b, c, d>(const BasisView& B, const BucketArgs& A,         \
                                          hipStream_t stream);
#define MCINT_FAST3_DECLARE(a, b, c, d)                                                    \
  template <>                                                                              \
  hipError_t launch_fast3_class<a, b, c, d>(const BasisView& B, const BucketArgs& A,       \
                                            hipStream_t stream);
#define MCINT_CHUNK3_DECLARE(a, b, c, d)                                                   \
  template <>                                                                              \
  hipError_t launch_chunk3_class<a, b, c, d>(const BasisView& B, const BucketArgs& A,      \
                                             hipStream_t stream);

// dispatch functions (expanded once, in eri.hip)
#define MC_FAST_HAS(a, b, c, d) \
  if (la == a && lb == b && lc == c && ld == d) return 1;
#define MC_CHUNK_HAS(a, b, c, d) \
  if (la == a && lb == b && lc == c && ld == d) return 2;
#define MC_FAST_LAUNCH(a, b, c, d) \
  if (A.la == a && A.lb == b && A.lc == c && A.ld == d) return launch_fast_class<a, b, c, d>(B, A, stream);
#define MC_CHUNK_LAUNCH(a, b, c, d) \
  if (A.la == a && A.lb == b && A.lc == c && A.ld == d) return launch_chunked_class<a, b, c, d>(B, A, stream);
#define MC_CHUNK_NEF(a, b, c, d) \
  if (la == a && lb == b && lc == c && ld == d) return ChunkCfg<a, b, c, d>::NEF;
#define MC_IP1_NEF(a, b, c, d) \
  if (la == a && lb == b && lc == c && ld == d) return ChunkCfg<a, b, c, d, 1>::NEF;
#define MC_IP1_LAUNCH(a, b, c, d) \
  if (A.la == a && A.lb == b && A.lc == c && A.ld == d) return launch_ip1_class<a, b, c, d>(B, A, stream);
#define MC_FAST3_LAUNCH(a, b, c, d) \
  if (A.la == a && A.lb == b && A.lc == c) return launch_fast3_class<a, b, c, d>(B, A, stream);
#define MC_CHUNK3_LAUNCH(a, b, c, d) \
  if (A.la == a && A.lb == b && A.lc == c) return launch_chunk3_class<a, b, c, d>(B, A, stream);
#define MC_CHUNK3_NEF(a, b, c, d) \
  if (la == a && lb == b && lc == c) return ChunkCfg<a, b, c, d, 2>::NEF;
#define MC_FAST3_HAS(a, b, c, d) \
  if (la == a && lb == b && lc == c) return 1;
#define MC_CHUNK3_HAS(a, b, c, d) \
  if (la == a && lb == b && lc == c) return 2;
#define MCINT_FAST_DISPATCH                                                                \
  MCINT_FAST_CLASSES(MCINT_FAST_DECLARE)                                                   \
  MCINT_CHUNK_CLASSES(MCINT_CHUNK_DECLARE)                                                  \
  MCINT_IP1_CLASSES(MCINT_IP1_DECLARE)                                                     \
  MCINT_FAST3_CLASSES(MCINT_FAST3_DECLARE)                                                 \
  MCINT_CHUNK3_CLASSES(MCINT_CHUNK3_DECLARE)                                               \
  int fast_kernel_mode(int kind, int sph, int la, int lb, int lc, int ld, double omega) {  \
    if (!sph && omega < 0) return 0;                                                       \
    if (kind == KIND_2E_IP1) {                                                             \
      MCINT_IP1_CLASSES(MC_CHUNK_HAS)                                                      \
      return 0;                                                                            \
    }                                                                                      \
    if (kind == KIND_3C2E) {                                                               \
      (void)ld;                                                                            \
      MCINT_FAST3_CLASSES(MC_FAST3_HAS)                                                    \
      MCINT_CHUNK3_CLASSES(MC_CHUNK3_HAS)                                                  \
      return 0;                                                                            \
    }                                                                                      \
    if (kind != KIND_2E) return 0;                                                         \
    MCINT_FAST_CLASSES(MC_FAST_HAS)                                                        \
    MCINT_CHUNK_CLASSES(MC_CHUNK_HAS)                                                      \
    return 0;                                                                              \
  }                                                                                        \
  int chunked_block_doubles(int kind, int la, int lb, int lc, int ld) {                    \
    if (kind == KIND_2E_IP1) {                                                             \
      MCINT_IP1_CLASSES(MC_IP1_NEF)                                                        \
      return 0;                                                                            \
    }                                                                                      \
    if (kind == KIND_3C2E) {                                                               \
      MCINT_CHUNK3_CLASSES(MC_CHUNK3_NEF)                                                  \
      return 0;                                                                            \
    }                                                                                      \
    MCINT_CHUNK_CLASSES(MC_CHUNK_NEF)                                                      \
    return 0;                                                                              \
  }                                                                                        \
  int fast_team_lanes_max() { return 64; }                                                 \
  hipError_t launch_fast(const BasisView& B, const BucketArgs& A, hipStream_t stream) {    \
    if (A.t1 <= A.t0) return hipSuccess;                                                   \
    if (A.team < 1 || A.team > 64 || (A.team & (A.team - 1))) return hipErrorInvalidValue; \
    if (A.kind == KIND_3C2E) {                                                             \
      MCINT_FAST3_CLASSES(MC_FAST3_LAUNCH)                                                 \
      return hipErrorNotSupported;                                                         \
    }                                                                                      \
    MCINT_FAST_CLASSES(MC_FAST_LAUNCH)                                                     \
    return hipErrorNotSupported;                                                           \
  }                                                                                        \
  hipError_t launch_chunked(const BasisView& B, const BucketArgs& A, hipStream_t stream) { \
    if (A.t1 <= A.t0) return hipSuccess;                                                   \
    if (A.team < 1 || A.team > 64 || (A.team & (A.team - 1))) return hipErrorInvalidValue; \
    if (A.kind == KIND_2E_IP1) {                                                           \
      MCINT_IP1_CLASSES(MC_IP1_LAUNCH)                                                     \
      return hipErrorNotSupported;                                                         \
    }                                                                                      \
    if (A.kind == KIND_3C2E) {                                                             \
      MCINT_CHUNK3_CLASSES(MC_CHUNK3_LAUNCH)                                               \
      return hipErrorNotSupported;                                                         \
    }                                                                                      \
    MCINT_CHUNK_CLASSES(MC_CHUNK_LAUNCH)                                                   \
    return hipErrorNotSupported;                                                           \
  }

}  // namespace mcint
