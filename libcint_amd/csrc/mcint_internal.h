// Internal data structures shared by the HIP kernels and the host driver of
// libmcint (MI355X / gfx950).  Nothing here crosses the public C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcint {

// env / atm / bas slots of the libcint data model (include/cint.h.in:26-67)
constexpr int PTR_EXPCUTOFF = 0;
constexpr int PTR_RANGE_OMEGA = 8;
constexpr int ATOM_OF = 0, ANG_OF = 1, NPRIM_OF = 2, NCTR_OF = 3, PTR_EXP = 5, PTR_COEFF = 6;
constexpr int BAS_SLOTS = 8, ATM_SLOTS = 6, PTR_COORD = 1;

constexpr double EXPCUTOFF = 60.0;      // src/cint_config.h.in:27
constexpr double MIN_EXPCUTOFF = 40.0;  // src/cint_config.h.in:30
constexpr int LMAX = 6;                 // largest shell l handled on the device
constexpr int NRYS_MAX = 9;             // largest Rys root count of the tables

// Integral kinds.  The three-centre kinds and the two-centre one share the
// ket of (ij|k): one shell on C and an exponent-0 s function on D; 2c2e
// also puts an exponent-0 s function (the basis's dummy virtual shell) on B.
enum IntorKind : int {
  KIND_2E = 0,
  KIND_2E_IP1 = 1,    // derivative on shell i (int2e_ip1)
  KIND_3C2E = 2,
  KIND_3C2E_IP1 = 3,  // derivative on shell i (int3c2e_ip1)
  KIND_3C2E_IP2 = 4,  // derivative on the auxiliary shell k (int3c2e_ip2)
  KIND_2C2E = 5,      // (i|k) (int2c2e)
  KIND_2E_IPIP1 = 6,  // second derivative on shell i (int2e_ipip1)
  KIND_2E_IP1IP2 = 7, // derivatives on shells i and k (int2e_ip1ip2)
};
constexpr int NKIND = 8;
__host__ __device__ constexpr bool kind_three(int k) { return k >= KIND_3C2E && k <= KIND_2C2E; }
__host__ __device__ constexpr bool kind_bra_der(int k) {
  return k == KIND_2E_IP1 || k == KIND_3C2E_IP1;
}
__host__ __device__ constexpr bool kind_ket_der(int k) { return k == KIND_3C2E_IP2; }
__host__ __device__ constexpr bool kind_hess(int k) {
  return k == KIND_2E_IPIP1 || k == KIND_2E_IP1IP2;
}
__host__ __device__ constexpr int kind_ncomp(int k) {
  return kind_hess(k) ? 9 : (kind_bra_der(k) || kind_ket_der(k) ? 3 : 1);
}
// angular raises of the bra and the ket pair (the recurrence's extra steps,
// the Rys order and the screening's distance terms)
__host__ __device__ constexpr int kind_bra_raise(int k) {
  return k == KIND_2E_IPIP1 ? 2 : (kind_bra_der(k) || k == KIND_2E_IP1IP2 ? 1 : 0);
}
__host__ __device__ constexpr int kind_ket_raise(int k) {
  return kind_ket_der(k) || k == KIND_2E_IP1IP2 ? 1 : 0;
}
// shells of the caller's tuple: 4, 3 (ij|k) or 2 (i|k)
__host__ __device__ constexpr int kind_axes(int k) {
  return k == KIND_2C2E ? 2 : (kind_three(k) ? 3 : 4);
}

// One "virtual shell": a (shell, contraction column) pair.  Generally
// contracted shells expand into nctr virtual shells, each a plain
// segmented contraction; the output sub-block of column c starts at
// c * nsph inside the real shell's (nsph * nctr) range.
struct VShell {
  double r[3];     // centre (Bohr)
  int l;
  int nprim;
  int pexp;        // offset of the exponents in env
  int pcoef;       // offset of this column's coefficients in env
  int pcoef0;      // offset of the real shell's first coefficient column
  int nctr;        // contraction columns of the real shell
  int dsph;        // output width of the real shell, spherical: (2l+1) * nctr
  int dcart;       // output width of the real shell, Cartesian
  int shell;       // index of the real shell
  int atom;        // atom index (gradient accumulation)
  int ao;          // first spherical AO of this column (CINTshells_spheric_offset + c*(2l+1))
};

// Primitive-pair data of an ordered virtual-shell pair, primitive i fastest.
// cce0 is libcint's screening estimate without the l-dependent distance
// term (src/optimizer.c:288-342):  cce = cce0 - (li+lj) * pair_ld.
struct PairPrim {
  double px, py, pz;  // Gaussian product centre
  double aij;         // ai + aj
  double kab;         // ci * cj * exp(-ai aj / aij |A-B|^2)
  double cce0;
  double ai, aj;
};

// A unit of work: one contracted (virtual) quartet in canonical shell order.
// outp packs the offset of element (0,0,0,0) of the quartet's block (low 56
// bits) and the axis map (high 8 bits): canonical axis A holds the original
// axis (perm >> 2A) & 3, so tasks of every axis order of one class share a
// launch.
struct Task {
  int16_t sh[4];   // virtual shells, canonical order (sh[3] unused for 3c2e)
  int64_t outp;
};
constexpr int64_t TASK_OUT_MASK = (int64_t(1) << 56) - 1;
__host__ __device__ inline int64_t task_out(const Task& t) { return t.outp & TASK_OUT_MASK; }
__host__ __device__ inline int task_perm(const Task& t) { return (int)((uint64_t)t.outp >> 56); }
__host__ __device__ inline int64_t task_pack(int64_t out, int perm) {
  return out | ((int64_t)perm << 56);
}

// Device view of a basis (atm/bas/env) plus its pair data.
struct BasisView {
  const VShell* vsh;
  const double* env;
  const int64_t* pair_off;   // [nv * nv], offset of pair (a,b) in pp[]
  const double* pair_ld;     // [nv * nv], log(|A-B| + 1)
  const PairPrim* pp;        // per pair, in ascending cce0
  const int* pp_pos;         // [record in libcint's order (ip fastest)] -> position in pp
  int nv;
  double expcutoff_4c;  // libcint's cutoff for 4c integrals
  double expcutoff_3c;
  double omega;         // env[PTR_RANGE_OMEGA]
};

// Launch description of one bucket: tasks [t0, t1) share class, kind,
// team size and path (each task carries its own axis map).
struct BucketArgs {
  const Task* tasks;
  int64_t t0, t1;
  double* out;           // output base
  int* not0;             // per task status (nullable), indexed by task id
  const int* task_id;    // original list position of each task (for not0), nullable
  double* scratch;       // generic / chunked path: per-task accumulators
  int* flags;            // chunked path: per-task 'any primitive survived'
  int64_t scratch_stride;
  int la, lb, lc, ld;    // canonical angular momenta
  int kind;
  int sph;
  int team;              // lanes per task (fast kernels)
  int fp32;              // recurrences and contraction in FP32 (fast kernels)
  int col_major;         // chunked classes: per-task blocks stored [f][e] (wave
                         // kernel) instead of in the chunk order
  // gradient digest (i-derivative plans): instead of storing the block,
  // contract it with the density and add to grad[atom of shell i][x]
  const double* dm;      // nao x nao density, row-major (nullable: store mode)
  double* grad;          // nslot blocks of grad_stride (= 3 natm) accumulators
  int nao;
  int nslot;             // gradient accumulator copies (spreads the atomics)
  int grad_stride;
  // J/K digest (plain four-centre plans over unique quartets): instead of
  // storing the block, add its six block updates to the half matrices
  // J' (jk_j) and K' (jk_k), nao x nao row-major, density dm
  double* jk_j;
  double* jk_k;
};

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int nsph(int l) { return 2 * l + 1; }

}  // namespace mcint
