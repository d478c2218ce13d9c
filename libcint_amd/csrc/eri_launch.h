// Host-side entry points of the device code (eri.hip).
#pragma once
#include "mcint_internal.h"

namespace mcint {

// Scratch layout of the generic (any class) kernel, per task, in doubles.
// X0: primitive-contracted [e0|f0] (plain), or the 2*a_i weighted set of
//     the i-derivative; XM: the unweighted lowered set of the derivative;
// H1: bra transfer (Cartesian); H2: bra Cartesian->spherical;
// H3: ket transfer (Cartesian, 4-centre only).
// LDS bound (doubles) of the generic kernel's recurrence arrays:
// 3 directions x roots (doubled for short range) x (E+1)(F+1)
constexpr int GEN_VCAP = 3 * 2 * NRYS_MAX * 90;
// component-table bounds of the generic kernels' union e-list and f-list:
// all Cartesian components of l <= 14 (ipip1 of (66|ss) reaches e = 14)
constexpr int GEN_ECAP = 680;
constexpr int GEN_FCAP = 680;

struct GenLayout {
  int e_lo, E, F;         // e shells e_lo..E, f shells f_lo..F
  int f_lo;
  int nE, nF;             // union e-list and f-list lengths
  int p_lo, p_hi;         // shells of the main (or 2a-weighted) set
  int m_lo, m_hi;         // shells of the lowered set (m_hi < m_lo: empty)
  int NE0, NEM;           // lengths of the two e-sets
  int ncomp;
  int nca, ncb, nAo, nBo, ncc, ncd, nCo, nDo;
  int64_t oX0, oXM, oH1, oH2, oH3, total;
  // second derivatives (int2e_ipip1, int2e_ip1ip2): weighted row sets of the
  // bra (role 0: lowered, weight 1; 1: weight 2a_i; 2: weight 4a_i^2) and the
  // ket (role 0: lowered, weight 1; 1: raised, weight 2a_k), -1 when absent;
  // set r covers shells lo..hi, X holds set r's rows (NEs) x all ket columns
  // (KF = sum of NFs, set u's columns from cF[u]); nbc bra components
  int bset_lo[3], bset_hi[3], NEs[3];
  int64_t oXs[3];
  int kset_lo[2], kset_hi[2], NFs[2], cF[2];
  int KF, nbc;
};

// shell range of bra role r / ket role u of a second-derivative class
__host__ __device__ inline void hess_sets(int kind, int la, int lb, int lc, int ld, int* blo,
                                          int* bhi, int* klo, int* khi) {
  for (int r = 0; r < 3; ++r) blo[r] = bhi[r] = -1;
  for (int u = 0; u < 2; ++u) klo[u] = khi[u] = -1;
  if (kind == KIND_2E_IPIP1) {
    if (la >= 2) { blo[0] = la - 2; bhi[0] = la + lb - 2; }
    blo[1] = la; bhi[1] = la + lb;
    blo[2] = la + 2; bhi[2] = la + lb + 2;
    klo[0] = lc; khi[0] = lc + ld;  // one plain ket set, stored in role 0
  } else {
    if (la >= 1) { blo[0] = la - 1; bhi[0] = la + lb - 1; }
    blo[1] = la + 1; bhi[1] = la + lb + 1;
    if (lc >= 1) { klo[0] = lc - 1; khi[0] = lc + ld - 1; }
    klo[1] = lc + 1; khi[1] = lc + ld + 1;
  }
}

__host__ __device__ inline int cart_sum(int lo, int hi) {
  int s = 0;
  for (int l = lo; l <= hi; ++l) s += ncart(l);
  return s;
}

__host__ __device__ inline GenLayout gen_layout(int kind, int sph, int la, int lb, int lc, int ld) {
  GenLayout g;
  if (kind_hess(kind)) {
    int blo[3], bhi[3], klo[2], khi[2];
    hess_sets(kind, la, lb, lc, ld, blo, bhi, klo, khi);
    g.ncomp = 9;
    g.nbc = kind == KIND_2E_IPIP1 ? 9 : 3;
    g.e_lo = 99; g.E = 0; g.f_lo = 99; g.F = 0;
    g.KF = 0;
    for (int u = 0; u < 2; ++u) {
      g.kset_lo[u] = klo[u];
      g.kset_hi[u] = khi[u];
      g.NFs[u] = klo[u] >= 0 ? cart_sum(klo[u], khi[u]) : 0;
      g.cF[u] = g.KF;
      g.KF += g.NFs[u];
      if (klo[u] >= 0) {
        g.f_lo = klo[u] < g.f_lo ? klo[u] : g.f_lo;
        g.F = khi[u] > g.F ? khi[u] : g.F;
      }
    }
    int64_t o = 0;
    for (int r = 0; r < 3; ++r) {
      g.bset_lo[r] = blo[r];
      g.bset_hi[r] = bhi[r];
      g.NEs[r] = blo[r] >= 0 ? cart_sum(blo[r], bhi[r]) : 0;
      g.oXs[r] = o;
      o += (int64_t)g.NEs[r] * g.KF;
      if (blo[r] >= 0) {
        g.e_lo = blo[r] < g.e_lo ? blo[r] : g.e_lo;
        g.E = bhi[r] > g.E ? bhi[r] : g.E;
      }
    }
    g.nE = cart_sum(g.e_lo, g.E);
    g.nF = cart_sum(g.f_lo, g.F);
    g.p_lo = g.p_hi = g.m_lo = g.m_hi = g.NE0 = g.NEM = 0;
    g.nca = ncart(la);
    g.ncb = ncart(lb);
    g.ncc = ncart(lc);
    g.ncd = ncart(ld);
    g.nAo = sph ? nsph(la) : g.nca;
    g.nBo = sph ? nsph(lb) : g.ncb;
    g.nCo = sph ? nsph(lc) : g.ncc;
    g.nDo = sph ? nsph(ld) : g.ncd;
    g.oX0 = g.oXM = 0;
    g.oH1 = o;
    g.oH2 = g.oH1 + (int64_t)g.nbc * g.nca * g.ncb * g.KF;
    g.oH3 = g.oH2 + (int64_t)g.nbc * g.nAo * g.nBo * g.KF;
    g.total = g.oH3 + (int64_t)9 * g.nAo * g.nBo * g.ncc * g.ncd;
    return g;
  }
  g.f_lo = lc;
  g.nbc = g.KF = 0;
  const bool ip1 = kind_bra_der(kind);
  const bool ip2 = kind_ket_der(kind);
  const bool three = kind_three(kind);
  if (three) ld = 0;
  g.ncomp = kind_ncomp(kind);
  if (ip1) {
    g.p_lo = la + 1;
    g.p_hi = la + lb + 1;
    g.m_lo = la - 1;
    g.m_hi = la + lb - 1;
    if (la == 0) { g.m_lo = 1; g.m_hi = 0; }
    g.e_lo = la > 0 ? la - 1 : la + 1;
    g.E = la + lb + 1;
  } else {
    g.p_lo = la;
    g.p_hi = la + lb;
    g.m_lo = 1;
    g.m_hi = 0;
    g.e_lo = la;
    g.E = la + lb;
  }
  // the auxiliary-shell derivative (int3c2e_ip2) runs the recurrence one
  // step further on the ket and keeps the columns of shell lc, formed per
  // primitive quartet as k_x [e|k-1_x] - 2 a_k [e|k+1_x]
  g.F = lc + ld + (ip2 ? 1 : 0);
  g.nE = cart_sum(g.e_lo, g.E);
  g.nF = ip2 ? ncart(lc) : cart_sum(lc, g.F);
  g.NE0 = cart_sum(g.p_lo, g.p_hi);
  g.NEM = g.m_hi >= g.m_lo ? cart_sum(g.m_lo, g.m_hi) : 0;
  g.nca = ncart(la);
  g.ncb = ncart(lb);
  g.ncc = ncart(lc);
  g.ncd = ncart(ld);
  g.nAo = sph ? nsph(la) : g.nca;
  g.nBo = sph ? nsph(lb) : g.ncb;
  g.nCo = sph ? nsph(lc) : g.ncc;
  g.nDo = sph ? nsph(ld) : g.ncd;
  g.oX0 = 0;
  g.oXM = g.oX0 + (int64_t)(ip2 ? 3 : 1) * g.NE0 * g.nF;
  g.oH1 = g.oXM + (int64_t)g.NEM * g.nF;
  g.oH2 = g.oH1 + (int64_t)g.ncomp * g.nca * g.ncb * g.nF;
  g.oH3 = g.oH2 + (int64_t)g.ncomp * g.nAo * g.nBo * g.nF;
  g.total = g.oH3 + (three ? 0 : (int64_t)g.ncomp * g.nAo * g.nBo * g.ncc * g.ncd);
  return g;
}

// Primitive-pair records of the listed shell pairs, each pair's list in
// ascending order of cce0, and pos[k] = position in pp of the k-th record in
// libcint's order; tmp is a staging buffer of the size of pp.
hipError_t launch_pairdata(const VShell* vsh, const double* env, int nv, const int* pairs,
                           int npairs, const int64_t* pair_off, double* pair_ld, PairPrim* tmp,
                           PairPrim* pp, int* pos, hipStream_t stream);

hipError_t launch_rys_batch(int n, const double* x, int64_t count, double* u, double* w,
                            hipStream_t stream);
// Short-range rule of exp(-x t^2) on [lower, 1] for count (x, lower) pairs
// (one wave each); u, w: count * n.
hipError_t launch_sr_rys_batch(int n, const double* x, const double* lower, int64_t count,
                               double* u, double* w, hipStream_t stream);

hipError_t launch_count_prims(const BasisView& B, const BucketArgs& A, unsigned long long* counter,
                              hipStream_t stream);

hipError_t launch_generic(const BasisView& B, const BucketArgs& A, hipStream_t stream);

// grad[i] += sum over the nslot accumulator copies of acc[slot * n + i]
hipError_t launch_grad_reduce(const double* acc, int nslot, int n, double* grad, hipStream_t stream);

// Grouped J/K digest: row pass (groups of quartets sharing (i,j), LDS of
// rows_lds bytes) and column pass (groups sharing (k,l)) into the half
// matrices J', K' (either nullable); symmetrize then forms X + X^T.
hipError_t launch_jk_grouped(const int* shls, const int64_t* off, const int* rows_q,
                             const int* rows_ptr, int nrows, size_t rows_lds, const int* cols_q,
                             const int* cols_ptr, int ncols, const double* eri, const int* ao_loc,
                             const double* D, int nao, double* Jp, double* Kp, hipStream_t stream);
hipError_t launch_symmetrize(const double* X, double* out, int n, hipStream_t stream);
// x[0..n) = 0 (a kernel, so that captured graphs hold no memset node)
hipError_t launch_zero(double* x, int64_t n, hipStream_t stream);

hipError_t launch_jk_digest(const int* shls, const int64_t* off, int64_t nq, const double* eri,
                            const int* ao_loc, const double* D, int nao, double* J, double* K,
                            hipStream_t stream);

// Register-resident kernels for the classes listed in eri_fast.h:
// mode 1 = one kernel with an in-register epilogue, 2 = chunk kernels plus
// an LDS epilogue (scratch of chunked_block_doubles() per task), 0 = none.
// Cartesian output: every kind except short range.
int fast_kernel_mode(int kind, int sph, int la, int lb, int lc, int ld, double omega);
// 1: team-of-one buckets of the chunked classes with a wave kernel run it
// instead of their chunk kernels (mcint_tune MCINT_TUNE_WAVE_KERNELS;
// read at launch, so graphs captured earlier keep their kernels)
extern int g_wave_kernels;
// per-class and team-size choice of the chunked classes' kernels
// (mcint_tune_class_wave): 0 the default rule, 1 chunk kernels, 2 wave
// kernel; indexed by class_index * 7 + log2(team lanes)
constexpr int NCLS = (LMAX + 1) * (LMAX + 1) * (LMAX + 1) * (LMAX + 1);
__host__ __device__ constexpr int class_index(int la, int lb, int lc, int ld) {
  return ((la * (LMAX + 1) + lb) * (LMAX + 1) + lc) * (LMAX + 1) + ld;
}
extern int8_t g_class_wave[NCLS * 7];
// 1 when the canonical plain class (la lb|lc ld) has a wave kernel
int wave_kernel_exists(int la, int lb, int lc, int ld);
int chunked_block_doubles(int kind, int la, int lb, int lc, int ld);
int fast_team_lanes_max();
hipError_t launch_fast(const BasisView& B, const BucketArgs& A, hipStream_t stream);
hipError_t launch_chunked(const BasisView& B, const BucketArgs& A, hipStream_t stream);

}  // namespace mcint
