// Host driver and C ABI of libmcint.
//
// Flow of a batch (the GPU counterpart of looping libcint's int2e_* calls,
// examples/time_c2h6.c:719-745):
//   1. mcint_basis_create uploads atm/bas/env and expands generally
//      contracted shells into virtual shells;
//   2. mcint_plan_create canonicalises every shell tuple (permutational
//      symmetry of (ij|kl)), sorts the tuples into buckets of equal
//      (class, team size, path) -- every axis order of a class shares a
//      bucket, each task carrying its own axis map -- and builds the
//      primitive-pair data
//      of every pair the tuples touch (the role of int2e_optimizer,
//      src/optimizer.c:344-425);
//   3. mcint_plan_execute launches one kernel per bucket, captured once into
//      a HIP graph and replayed afterwards.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mcint.h"
#include "eri_launch.h"
#include "host_internal.h"

using namespace mcint;

namespace {

thread_local std::string g_err;

void set_err(const std::string& s) { g_err = s; }

#define HIPCHK(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      set_err(std::string(#call) + " failed: " + hipGetErrorString(e_));               \
      return MCINT_ERR_HIP;                                                            \
    }                                                                                  \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) return hipSuccess;
  return hipMalloc((void**)p, n * sizeof(T));
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

int intor_kind(int intor) {
  switch (intor) {
    case MCINT_INT2E_SPH:
    case MCINT_INT2E_CART: return KIND_2E;
    case MCINT_INT2E_IP1_SPH:
    case MCINT_INT2E_IP1_CART:
    case MCINT_INT2E_IP2_SPH:  // the ip1 kernels with the ket pair first
    case MCINT_INT2E_IP2_CART: return KIND_2E_IP1;
    case MCINT_INT3C2E_SPH:
    case MCINT_INT3C2E_CART: return KIND_3C2E;
    case MCINT_INT3C2E_IP1_SPH:
    case MCINT_INT3C2E_IP1_CART: return KIND_3C2E_IP1;
    case MCINT_INT3C2E_IP2_SPH:
    case MCINT_INT3C2E_IP2_CART: return KIND_3C2E_IP2;
    case MCINT_INT2C2E_SPH:
    case MCINT_INT2C2E_CART: return KIND_2C2E;
    case MCINT_INT2E_IPIP1_SPH:
    case MCINT_INT2E_IPIP1_CART: return KIND_2E_IPIP1;
    case MCINT_INT2E_IP1IP2_SPH:
    case MCINT_INT2E_IP1IP2_CART: return KIND_2E_IP1IP2;
    default: return -1;
  }
}

// spherical output: the even intor codes
int intor_sph(int intor) { return intor >= 0 && intor % 2 == 0; }

// int2e_ip2, (ij|d/dr k l) = (d/dr k l|ij): derivative on the ket's first
// shell, computed by the i-derivative kernels on the axes (k, l, i, j)
int intor_ket_der(int intor) { return intor == MCINT_INT2E_IP2_SPH || intor == MCINT_INT2E_IP2_CART; }

}  // namespace

// ---------------------------------------------------------------------------
// Basis
// ---------------------------------------------------------------------------
struct mcint_basis {
  int natm = 0, nbas = 0, nv = 0;
  std::vector<int> atm, bas;
  std::vector<double> env;
  std::vector<VShell> vsh;
  std::vector<int> vfirst;  // first virtual shell of each real shell
  int dummy = 0;            // virtual shell of the exponent-0 s function (int2c2e's partner)
  size_t envlen = 0;        // doubles of the caller's env copied (env holds 2 more)
  VShell* vsh_d = nullptr;
  double* env_d = nullptr;
  int* ao_loc_d = nullptr;  // spherical AO offsets of the real shells, nbas+1
  double expc4 = EXPCUTOFF, expc3 = EXPCUTOFF, omega = 0;
  int device = 0;
};

// upload = false stops after the host tables (no HIP call): the basis then
// serves only the host stage of plan construction (plan_build_host_only)
static int basis_create_impl(mcint_basis** out, const FINT* atm, FINT natm, const FINT* bas,
                             FINT nbas, const double* env, bool upload) {
  if (!out || !atm || !bas || !env || natm <= 0 || nbas <= 0) {
    set_err("mcint_basis_create: null pointer or empty basis");
    return MCINT_ERR_ARG;
  }
  *out = nullptr;
  auto* b = new mcint_basis();
  b->natm = natm;
  b->nbas = nbas;
  b->atm.assign(atm, atm + (size_t)natm * ATM_SLOTS);
  b->bas.assign(bas, bas + (size_t)nbas * BAS_SLOTS);
  // env has no length argument in libcint's ABI: take the largest slot any
  // atom or shell refers to.
  size_t envlen = 20;
  for (int a = 0; a < natm; ++a) envlen = std::max(envlen, (size_t)atm[a * ATM_SLOTS + PTR_COORD] + 3);
  for (int s = 0; s < nbas; ++s) {
    const int* bs = bas + (size_t)s * BAS_SLOTS;
    if (bs[ATOM_OF] < 0 || bs[ATOM_OF] >= natm || bs[ANG_OF] < 0 || bs[ANG_OF] > LMAX ||
        bs[NPRIM_OF] <= 0 || bs[NCTR_OF] <= 0 || bs[PTR_EXP] < 0 || bs[PTR_COEFF] < 0) {
      delete b;
      char msg[160];
      snprintf(msg, sizeof msg, "mcint_basis_create: shell %d invalid (atom/l/nprim/nctr); l <= %d",
               s, LMAX);
      set_err(msg);
      return MCINT_ERR_ARG;
    }
    envlen = std::max(envlen, (size_t)bs[PTR_EXP] + bs[NPRIM_OF]);
    envlen = std::max(envlen, (size_t)bs[PTR_COEFF] + (size_t)bs[NPRIM_OF] * bs[NCTR_OF]);
  }
  b->env.assign(env, env + envlen);
  if (env[0] != 0) {
    b->expc4 = std::max(MIN_EXPCUTOFF, env[PTR_EXPCUTOFF]) + 1;  // src/g2e.c:57-62
    b->expc3 = std::max(MIN_EXPCUTOFF, env[PTR_EXPCUTOFF]);      // src/g3c2e.c:54-58
  }
  b->omega = env[PTR_RANGE_OMEGA];
  std::vector<int> ao_loc(nbas + 1, 0);
  for (int s = 0; s < nbas; ++s)
    ao_loc[s + 1] = ao_loc[s] + nsph(bas[(size_t)s * BAS_SLOTS + ANG_OF]) * bas[(size_t)s * BAS_SLOTS + NCTR_OF];
  for (int s = 0; s < nbas; ++s) {
    const int* bs = bas + (size_t)s * BAS_SLOTS;
    b->vfirst.push_back((int)b->vsh.size());
    const int l = bs[ANG_OF];
    for (int c = 0; c < bs[NCTR_OF]; ++c) {
      VShell v{};
      const double* r = env + atm[bs[ATOM_OF] * ATM_SLOTS + PTR_COORD];
      v.r[0] = r[0];
      v.r[1] = r[1];
      v.r[2] = r[2];
      v.l = l;
      v.nprim = bs[NPRIM_OF];
      v.pexp = bs[PTR_EXP];
      v.pcoef = bs[PTR_COEFF] + c * bs[NPRIM_OF];
      v.pcoef0 = bs[PTR_COEFF];
      v.nctr = bs[NCTR_OF];
      v.dsph = nsph(l) * bs[NCTR_OF];
      v.dcart = ncart(l) * bs[NCTR_OF];
      v.shell = s;
      v.atom = bs[ATOM_OF];
      v.ao = ao_loc[s] + c * nsph(l);
      b->vsh.push_back(v);
    }
  }
  // the exponent-0 s function that int2c2e pairs with its first shell;
  // coefficient 2 sqrt(pi) cancels the angular factor of an s function
  // (libcint's int2c2e carries none for it, src/g2c2e.c:15-90)
  b->envlen = b->env.size();
  {
    VShell v{};
    v.l = 0;
    v.nprim = 1;
    v.pexp = (int)b->envlen;
    v.pcoef = v.pcoef0 = (int)b->envlen + 1;
    v.nctr = 1;
    v.dsph = v.dcart = 1;
    v.shell = -1;
    v.atom = -1;
    v.ao = 0;
    b->env.push_back(0.0);
    b->env.push_back(3.5449077018110320546);  // 2 sqrt(pi)
    b->dummy = (int)b->vsh.size();
    b->vsh.push_back(v);
  }
  b->nv = (int)b->vsh.size();
  if (b->nv > 32767) {
    delete b;
    set_err("mcint_basis_create: more than 32767 contracted shell columns");
    return MCINT_ERR_UNSUPPORTED;
  }
  if (!upload) {
    *out = b;
    return MCINT_OK;
  }
  hipError_t e = hipGetDevice(&b->device);
  if (e == hipSuccess) e = dalloc(&b->vsh_d, b->vsh.size());
  if (e == hipSuccess) e = dalloc(&b->env_d, b->env.size());
  if (e == hipSuccess) e = dalloc(&b->ao_loc_d, ao_loc.size());
  if (e == hipSuccess)
    e = hipMemcpy(b->ao_loc_d, ao_loc.data(), ao_loc.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->vsh_d, b->vsh.data(), b->vsh.size() * sizeof(VShell), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->env_d, b->env.data(), b->env.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_err(std::string("mcint_basis_create: ") + hipGetErrorString(e));
    dfree(b->vsh_d);
    dfree(b->env_d);
    dfree(b->ao_loc_d);
    delete b;
    return MCINT_ERR_HIP;
  }
  *out = b;
  return MCINT_OK;
}

extern "C" int mcint_basis_create(mcint_basis** out, const FINT* atm, FINT natm, const FINT* bas,
                                  FINT nbas, const double* env) {
  return basis_create_impl(out, atm, natm, bas, nbas, env, true);
}

extern "C" void mcint_basis_destroy(mcint_basis* b) {
  if (!b) return;
  dfree(b->vsh_d);
  dfree(b->env_d);
  dfree(b->ao_loc_d);
  delete b;
}

// ---------------------------------------------------------------------------
// Plan
// ---------------------------------------------------------------------------
namespace {

// Phase timing of plan construction, compiled in only with
// -DMCINT_HOST_TIMING (stderr).
#ifdef MCINT_HOST_TIMING
struct PhaseTimer {
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  void mark(const char* what) {
    const auto n = std::chrono::steady_clock::now();
    fprintf(stderr, "[mcint plan] %-28s %8.3f s\n", what,
            std::chrono::duration<double>(n - t).count());
    t = n;
  }
};
#else
struct PhaseTimer {
  void mark(const char*) {}
};
#endif

// Host threads for plan construction: at most 16 (one GPU's share of a
// node), one per 200k quartets.
int host_threads(size_t n) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  return (int)std::max<size_t>(1, std::min<size_t>({(size_t)16, (size_t)hw, n / 200000 + 1}));
}

// f(thread, begin, end) over nthr contiguous ranges of [0, n)
template <typename F>
void parallel_for(int nthr, size_t n, F&& f) {
  if (nthr <= 1) {
    f(0, (size_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nthr);
  for (int t = 0; t < nthr; ++t) {
    const size_t a = n * (size_t)t / (size_t)nthr, e = n * (size_t)(t + 1) / (size_t)nthr;
    th.emplace_back([&f, t, a, e] { f(t, a, e); });
  }
  for (auto& x : th) x.join();
}

// Shell data of one quartet and its canonical axis order L[A] = original axis.
struct QuartetShape {
  int l[4], w[4], nc[4], np[4], L[4];
  int64_t so[5];
};

// Shell data of one shell tuple s (naxes = 3 for int3c2e) and its
// canonical axis order: 0, or 1 when a shell index is out of range, 2 when
// the class needs more Rys roots than the device tables hold.
int describe_tuple(const int* bas, int nbas, int kind, int sph, const FINT* s, QuartetShape& Q,
                   int ket_der = 0) {
  const int naxes = kind_axes(kind);
  for (int a = 0; a < 4; ++a) {
    Q.l[a] = 0;
    Q.w[a] = Q.nc[a] = Q.np[a] = 1;
  }
  for (int a = 0; a < naxes; ++a) {
    if (s[a] < 0 || s[a] >= nbas) return 1;
    const int* bs = bas + (size_t)s[a] * BAS_SLOTS;
    Q.l[a] = bs[ANG_OF];
    Q.nc[a] = bs[NCTR_OF];
    Q.np[a] = bs[NPRIM_OF];
    Q.w[a] = sph ? nsph(Q.l[a]) : ncart(Q.l[a]);
  }
  const int lsum = Q.l[0] + Q.l[1] + Q.l[2] + Q.l[3] + kind_bra_raise(kind) + kind_ket_raise(kind);
  if (lsum / 2 + 1 > NRYS_MAX) return 2;
  Q.so[0] = 1;
  for (int o = 0; o < 4; ++o) Q.so[o + 1] = Q.so[o] * (int64_t)Q.w[o] * Q.nc[o];
  // canonical axis order L[A] = original axis
  for (int a = 0; a < 4; ++a) Q.L[a] = a;
  int* L = Q.L;
  const int* l = Q.l;
  if (kind == KIND_2E) {
    if (l[L[0]] < l[L[1]]) std::swap(L[0], L[1]);
    if (l[L[2]] < l[L[3]]) std::swap(L[2], L[3]);
    if (l[L[2]] > l[L[0]] || (l[L[2]] == l[L[0]] && l[L[3]] > l[L[1]])) {
      std::swap(L[0], L[2]);
      std::swap(L[1], L[3]);
    }
  } else if (kind == KIND_2E_IP1 && ket_der) {
    // derivative shell k first, then l; the bra pair ordered by l
    L[0] = 2;
    L[1] = 3;
    L[2] = 0;
    L[3] = 1;
    if (l[0] < l[1]) std::swap(L[2], L[3]);
  } else if (kind == KIND_2E_IP1 || kind == KIND_2E_IPIP1) {
    if (l[2] < l[3]) std::swap(L[2], L[3]);
  } else if (kind == KIND_3C2E || kind == KIND_3C2E_IP2) {
    if (l[0] < l[1]) std::swap(L[0], L[1]);
  }
  // the derivative shell of int3c2e_ip1 stays first; int2c2e keeps (i, k);
  // int2e_ip1ip2 keeps (i, j, k, l): both derivative shells stay in place
  return 0;
}

// Canonical class (la, lb, lc, ld) of a described tuple and its axis map
// (canonical axis A holds the original axis (perm >> 2A) & 3).  int2c2e's
// canonical axes are (i, exponent-0 s, k): i is original axis 0, k axis 1,
// and the s function takes the unused axis 2.
void tuple_class(int kind, const QuartetShape& Q, int& la, int& lb, int& lc, int& ld, int& perm) {
  const int* L = Q.L;
  if (kind == KIND_2C2E) {
    la = Q.l[0];
    lb = 0;
    lc = Q.l[1];
    ld = 0;
    perm = 0 | (2 << 2) | (1 << 4) | (3 << 6);
    return;
  }
  la = Q.l[L[0]];
  lb = Q.l[L[1]];
  lc = Q.l[L[2]];
  ld = kind_three(kind) ? 0 : Q.l[L[3]];
  perm = L[0] | (L[1] << 2) | (L[2] << 4) | (L[3] << 6);
}

// The tasks of one described tuple s: one per combination of contraction
// columns, in canonical shell order, output offsets from base, axis map perm.
template <class F>
void for_each_task(const std::vector<int>& vfirst, int dummy, int kind, const QuartetShape& Q,
                   const FINT* s, int64_t base, int perm, F&& emit) {
  const int naxes = kind_axes(kind);
  int cidx[4] = {0, 0, 0, 0};
  for (;;) {
    Task t;
    int64_t o = base;
    for (int a = 0; a < naxes; ++a) o += (int64_t)cidx[a] * Q.w[a] * Q.so[a];
    if (kind == KIND_2C2E) {
      t.sh[0] = (int16_t)(vfirst[s[0]] + cidx[0]);
      t.sh[1] = (int16_t)dummy;
      t.sh[2] = t.sh[3] = (int16_t)(vfirst[s[1]] + cidx[1]);
    } else {
      for (int A = 0; A < 4; ++A) {
        if (A < naxes) {
          const int orig = Q.L[A];
          t.sh[A] = (int16_t)(vfirst[s[orig]] + cidx[orig]);
        } else {
          t.sh[A] = t.sh[naxes - 1];
        }
      }
    }
    t.outp = task_pack(o, perm);
    emit(t);
    int a = 0;
    for (; a < naxes; ++a) {
      if (++cidx[a] < Q.nc[a]) break;
      cidx[a] = 0;
    }
    if (a == naxes) break;
  }
}

// Set of bucket keys (open addressing; keys are < 2^27, ~0u marks empty).
struct KeySet {
  std::vector<uint32_t> slot = std::vector<uint32_t>(1024, ~0u);
  size_t n = 0;
  void insert(uint32_t k) {
    if (2 * (n + 1) > slot.size()) grow();
    size_t m = slot.size() - 1, h = (k * 2654435761u) & m;
    while (slot[h] != ~0u && slot[h] != k) h = (h + 1) & m;
    if (slot[h] == ~0u) {
      slot[h] = k;
      ++n;
    }
  }
  void grow() {
    std::vector<uint32_t> old;
    old.swap(slot);
    slot.assign(old.size() * 2, ~0u);
    n = 0;
    for (uint32_t k : old)
      if (k != ~0u) insert(k);
  }
  std::vector<uint32_t> keys() const {
    std::vector<uint32_t> v;
    for (uint32_t k : slot)
      if (k != ~0u) v.push_back(k);
    return v;
  }
};

// Key -> index in the sorted key list.
struct KeyIndex {
  std::vector<uint32_t> key;
  std::vector<uint32_t> idx;
  explicit KeyIndex(const std::vector<uint32_t>& sorted) {
    size_t cap = 64;
    while (cap < 2 * sorted.size() + 2) cap *= 2;
    key.assign(cap, ~0u);
    idx.assign(cap, 0);
    for (size_t i = 0; i < sorted.size(); ++i) {
      size_t h = (sorted[i] * 2654435761u) & (cap - 1);
      while (key[h] != ~0u) h = (h + 1) & (cap - 1);
      key[h] = sorted[i];
      idx[h] = (uint32_t)i;
    }
  }
  uint32_t find(uint32_t k) const {
    size_t h = (k * 2654435761u) & (key.size() - 1);
    while (key[h] != k) h = (h + 1) & (key.size() - 1);
    return idx[h];
  }
};

struct Bucket {
  int la, lb, lc, ld, team;
  int mode;  // 0 generic, 1 fast (single chunk), 2 chunked
  int fp32;  // recurrences in FP32 (mode 1 only)
  int64_t t0, t1;
  int64_t sstride;  // generic scratch doubles per task
  double cost;      // estimated work (balances the concurrent graph branches)
};

// Cartesian components of shells lo..hi
constexpr int ncart_sum(int lo, int hi) {
  int s = 0;
  for (int l = lo; l <= hi; ++l) s += (l + 1) * (l + 2) / 2;
  return s;
}

// Estimated work of a bucket: per primitive quartet and root the 2D
// recurrence and the [e0|f0] products, per task the transfers; the generic
// kernel counted 15x (its measured slowdown).
double bucket_cost(int kind, int mode, int la, int lb, int lc, int ld, int64_t nprim,
                   int64_t ntask) {
  const int E = la + lb + kind_bra_raise(kind), F = lc + ld + kind_ket_raise(kind);
  const int nr = (E + F) / 2 + 1;
  const double nef = (double)ncart_sum(la, E) * ncart_sum(lc, F);
  const double c = (double)nprim * nr * (2.0 * nef + 9.0 * (E + 1) * (F + 1)) + 4.0 * ntask * nef;
  return mode == 0 ? 15.0 * c : c;
}

// Scratch bound of the generic and chunked paths (doubles, per concurrent
// graph branch); buckets larger than the scratch are launched in slices.
// Slices run one after another, each with fewer waves than the GPU holds,
// so the bound is sized for whole buckets of large workloads: 4 GiB (the
// largest c1 bucket, 695,543 (fd|dp) quartets x 736 doubles, fits), 16 GiB
// over four branches, allocated only up to what the plan needs.
constexpr int64_t GEN_SCRATCH_DOUBLES = 512ll << 20;
// copies of the gradient accumulators (spreads the digest's atomics)
constexpr int GRAD_SLOTS = 64;

// Tuning state (mcint_tune): primitive quartets per lane that the team
// size aims at.
int g_prims_per_lane = 32;
// 1: every quartet on the generic kernel (mcint_tune MCINT_TUNE_GENERIC_ONLY)
int g_generic_only = 0;
// concurrent branches of a captured plan graph (mcint_tune MCINT_TUNE_GRAPH_BRANCHES)
int g_graph_branches = 4;
constexpr int MAX_GRAPH_BRANCHES = 8;
// per-class overrides (mcint_tune_class), 0 = the global value; indexed by
// integral kind and canonical class
int16_t g_class_ppl[NKIND][NCLS] = {};

int pick_team(int64_t nprim_quartets, int ppl) {
  int t = 1;
  while (t < fast_team_lanes_max() && (int64_t)t * ppl < nprim_quartets) t *= 2;
  return t;
}

}  // namespace

struct mcint_plan {
  mcint_basis* b = nullptr;
  int intor = 0, kind = 0, sph = 1, ncomp = 1;
  int64_t nq = 0, ntasks = 0, out_size = 0;
  std::vector<Bucket> buckets;
  Task* tasks_d = nullptr;
  int* task_id_d = nullptr;
  int64_t* pair_off_d = nullptr;
  double* pair_ld_d = nullptr;
  PairPrim* pp_d = nullptr;
  int* pp_pos_d = nullptr;
  // per-task blocks of the chunked and generic paths and the chunked
  // path's flags: sizes fixed at creation, allocated on the first execution
  // that needs them (mcint_plan_execute_noalloc never does)
  double* scratch_d = nullptr;
  int64_t scratch_size = 0;
  int* flags_d = nullptr;
  int64_t flags_size = 0;
  double* gslots_d = nullptr;  // gradient digest accumulators
  double* jk_d = nullptr;      // J', K' half matrices of the fused J/K digest
  int64_t jk_n = 0;            // nao * nao they hold
  int64_t stats[6] = {0, 0, 0, 0, 0, 0};
  int64_t stats_fp32 = 0;
  // per-branch scratch: scratch_d / flags_d hold scratch_nbr regions of
  // scratch_size doubles / flags_size ints
  int scratch_nbr = 0;
  // graph cache
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  double* g_out = nullptr;
  FINT* g_not0 = nullptr;
  hipStream_t g_stream = nullptr;
  int g_nbr = 0;
  // side streams and fork/join events of the branched capture
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> events;
};

static void plan_free(mcint_plan* p) {
  if (!p) return;
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  for (hipStream_t st : p->side) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : p->events) (void)hipEventDestroy(ev);
  dfree(p->tasks_d);
  dfree(p->task_id_d);
  dfree(p->pair_off_d);
  dfree(p->pair_ld_d);
  dfree(p->pp_d);
  dfree(p->pp_pos_d);
  dfree(p->scratch_d);
  dfree(p->flags_d);
  dfree(p->gslots_d);
  dfree(p->jk_d);
  delete p;
}

extern "C" void mcint_plan_destroy(mcint_plan* p) { plan_free(p); }

extern "C" int mcint_tune_class(int intor, int la, int lb, int lc, int ld, int prims_per_lane) {
  const int kind = intor_kind(intor);
  if (kind < 0 || la < 0 || lb < 0 || lc < 0 || ld < 0 || la > LMAX || lb > LMAX || lc > LMAX ||
      ld > LMAX || prims_per_lane < 0 || prims_per_lane > 4096) {
    set_err("mcint_tune_class: bad argument");
    return MCINT_ERR_ARG;
  }
  int16_t& v = g_class_ppl[kind][class_index(la, lb, lc, ld)];
  const int old = v;
  v = (int16_t)prims_per_lane;
  return old;
}

extern "C" int mcint_tune_class_wave(int la, int lb, int lc, int ld, int team, int mode) {
  int tl = 0;
  while ((1 << tl) < team && tl < 7) ++tl;
  if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la > LMAX || lb > LMAX || lc > LMAX || ld > LMAX ||
      team < 1 || team > 64 || (1 << tl) != team || mode < 0 || mode > 2) {
    set_err("mcint_tune_class_wave: bad argument");
    return MCINT_ERR_ARG;
  }
  if (mode == 2 && !wave_kernel_exists(la, lb, lc, ld)) {
    set_err("mcint_tune_class_wave: the class has no wave kernel");
    return MCINT_ERR_UNSUPPORTED;
  }
  int8_t& v = g_class_wave[class_index(la, lb, lc, ld) * 7 + tl];
  const int old = v;
  v = (int8_t)mode;
  return old;
}

extern "C" int mcint_wave_kernel_exists(int la, int lb, int lc, int ld) {
  if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la > LMAX || lb > LMAX || lc > LMAX || ld > LMAX)
    return 0;
  return wave_kernel_exists(la, lb, lc, ld);
}

extern "C" int mcint_tune(int key, int value) {
  if (key == MCINT_TUNE_PRIMS_PER_LANE) {
    if (value < 1 || value > 4096) {
      set_err("mcint_tune: primitive quartets per lane must be in [1, 4096]");
      return MCINT_ERR_ARG;
    }
    const int old = g_prims_per_lane;
    g_prims_per_lane = value;
    return old;
  }
  if (key == MCINT_TUNE_GRAPH_BRANCHES) {
    if (value < 1 || value > MAX_GRAPH_BRANCHES) {
      set_err("mcint_tune: MCINT_TUNE_GRAPH_BRANCHES takes 1 to 8");
      return MCINT_ERR_ARG;
    }
    const int old = g_graph_branches;
    g_graph_branches = value;
    return old;
  }
  if (key == MCINT_TUNE_WAVE_KERNELS) {
    if (value < 0 || value > 2) {
      set_err("mcint_tune: MCINT_TUNE_WAVE_KERNELS takes 0, 1 or 2");
      return MCINT_ERR_ARG;
    }
    const int old = g_wave_kernels;
    g_wave_kernels = value;
    return old;
  }
  if (key == MCINT_TUNE_GENERIC_ONLY) {
    if (value != 0 && value != 1) {
      set_err("mcint_tune: MCINT_TUNE_GENERIC_ONLY takes 0 or 1");
      return MCINT_ERR_ARG;
    }
    const int old = g_generic_only;
    g_generic_only = value;
    return old;
  }
  set_err("mcint_tune: unknown key");
  return MCINT_ERR_ARG;
}

// upload = false returns the plan after its host stage (buckets, task list,
// pair table) without any HIP call; such a plan must never be executed
static int plan_create_impl(mcint_plan** out, mcint_basis* b, int intor, const FINT* shls,
                            size_t nq, const int64_t* out_offsets, const uint8_t* fp32_flags,
                            int64_t* out_size, bool upload = true) {
  if (!out || !b || (!shls && nq)) {
    set_err("mcint_plan_create: null argument");
    return MCINT_ERR_ARG;
  }
  *out = nullptr;
  if (nq > (size_t)INT32_MAX) {
    // task ids (the not0 index) are 32-bit on the device
    set_err("mcint_plan_create: more than 2^31-1 shell tuples in one plan");
    return MCINT_ERR_ARG;
  }
  const int kind = intor_kind(intor);
  if (kind < 0) {
    set_err("mcint_plan_create: unknown integral kind");
    return MCINT_ERR_ARG;
  }
  const int sph = intor_sph(intor);
  const int ncomp = kind_ncomp(kind);
  const int naxes = kind_axes(kind);
  const int* bas = b->bas.data();
  const int nbas = b->nbas;

  auto* p = new mcint_plan();
  p->b = b;
  p->intor = intor;
  p->kind = kind;
  p->sph = sph;
  p->ncomp = ncomp;
  p->nq = (int64_t)nq;

  // ---- per quartet: validation, block size, canonical order, bucket key --
  // Threads take contiguous quartet ranges; tasks are then placed by a
  // stable counting sort on the bucket, so the task order within a bucket
  // is the list order.
  PhaseTimer timer;
  const int nthr = host_threads(nq);
  // uninitialised: first touched by the worker threads
  std::unique_ptr<uint32_t[]> qkey_buf(new uint32_t[nq + 1]);  // bucket key, then index
  std::unique_ptr<uint32_t[]> qnt_buf(new uint32_t[nq + 1]);   // virtual tasks per quartet
  std::unique_ptr<int64_t[]> qoff_buf(new int64_t[nq + 1]);
  uint32_t* qkey = qkey_buf.get();
  uint32_t* qnt = qnt_buf.get();
  int64_t* qoff = qoff_buf.get();
  qoff[0] = 0;
  // path (0 generic, 1 single kernel, 2 chunked) of every class, once
  std::vector<uint8_t> mode_tab((LMAX + 1) * (LMAX + 1) * (LMAX + 1) * (LMAX + 1), 0);
  for (int la = 0; la <= LMAX; ++la)
    for (int lb = 0; lb <= LMAX; ++lb)
      for (int lc = 0; lc <= LMAX; ++lc)
        for (int ld = 0; ld <= LMAX; ++ld)
          mode_tab[((la * (LMAX + 1) + lb) * (LMAX + 1) + lc) * (LMAX + 1) + ld] =
              (uint8_t)fast_kernel_mode(kind, sph, la, lb, lc, ld, b->omega);
  std::vector<int64_t> first_bad(nthr, INT64_MAX);
  std::vector<int64_t> nprim_thr(nthr, 0), ntask_thr(nthr, 0);
  const bool reg_ok = (b->omega >= 0 || kind != KIND_2E_IP1) && !g_generic_only;
  auto describe = [&](size_t q, QuartetShape& Q) -> int {
    return describe_tuple(bas, nbas, kind, sph, shls + 4 * q, Q, intor_ket_der(intor));
  };
  parallel_for(nthr, nq, [&](int th, size_t q0, size_t q1) {
    // per-thread totals kept in registers (the shared arrays are written
    // once: adjacent slots would share a cache line)
    int64_t bad = INT64_MAX, nprim_sum = 0, ntask_sum = 0;
    QuartetShape Q;
    for (size_t q = q0; q < q1; ++q) {
      if (describe(q, Q)) {
        bad = std::min<int64_t>(bad, (int64_t)q);
        qkey[q] = 0;
        qnt[q] = 0;
        continue;
      }
      int la, lb, lc, ld, perm;
      tuple_class(kind, Q, la, lb, lc, ld, perm);
      const int64_t nprimq = (int64_t)Q.np[0] * Q.np[1] * Q.np[2] * (naxes == 4 ? Q.np[3] : 1);
      nprim_sum += nprimq;
      // the short-range derivative integrals run on the generic kernel;
      // everything else has register kernels
      const int mode =
          reg_ok ? mode_tab[((la * (LMAX + 1) + lb) * (LMAX + 1) + lc) * (LMAX + 1) + ld] : 0;
      const int ppl_c = g_class_ppl[kind][class_index(la, lb, lc, ld)];
      const int team = mode ? pick_team(nprimq, ppl_c ? ppl_c : g_prims_per_lane) : 0;
      // FP32 recurrences where a single-chunk kernel or a derivative chunk
      // kernel exists (plain operator, spherical output only); other flagged
      // quartets stay FP64
      const uint32_t lowp = (fp32_flags && fp32_flags[q] && b->omega == 0 && sph &&
                             (mode == 1 || (mode == 2 && kind == KIND_2E_IP1)))
                                ? 1u
                                : 0u;
      int tl = 0;
      while ((1 << tl) < team) ++tl;
      // the axis map (perm) travels with each task, not in the key
      qkey[q] = (uint32_t)la | ((uint32_t)lb << 3) | ((uint32_t)lc << 6) | ((uint32_t)ld << 9) |
                ((uint32_t)tl << 20) | ((uint32_t)mode << 24) | (lowp << 26);
      uint32_t nt = 1;
      for (int a = 0; a < naxes; ++a) nt *= (uint32_t)Q.nc[a];
      qnt[q] = nt;
      ntask_sum += nt;
      qoff[q + 1] = Q.so[4] * ncomp;  // block size for now
    }
    first_bad[th] = bad;
    nprim_thr[th] = nprim_sum;
    ntask_thr[th] = ntask_sum;
  });
  timer.mark(nthr > 1 ? "describe quartets (threads)" : "describe quartets (1 thread)");
  int64_t bad_q = INT64_MAX;
  for (int th = 0; th < nthr; ++th) bad_q = std::min(bad_q, first_bad[th]);
  if (bad_q != INT64_MAX) {
    QuartetShape Q;
    const int why = describe((size_t)bad_q, Q);
    char msg[192];
    if (why == 1) {
      snprintf(msg, sizeof msg, "mcint_plan_create: quartet %lld has a shell index out of range",
               (long long)bad_q);
      set_err(msg);
      plan_free(p);
      return MCINT_ERR_ARG;
    }
    snprintf(msg, sizeof msg,
             "mcint_plan_create: quartet %lld needs more Rys roots than the device tables hold (%d)",
             (long long)bad_q, NRYS_MAX);
    set_err(msg);
    plan_free(p);
    return MCINT_ERR_UNSUPPORTED;
  }
  int64_t nprimq_total = 0, ntasks = 0;
  for (int th = 0; th < nthr; ++th) {
    nprimq_total += nprim_thr[th];
    ntasks += ntask_thr[th];
  }
  if (out_offsets) {
    int64_t mx = 0;
    for (size_t q = 0; q < nq; ++q)
      if (out_offsets[q] < 0) {
        char msg[128];
        snprintf(msg, sizeof msg, "mcint_plan_create: quartet %lld has a negative output offset",
                 (long long)q);
        set_err(msg);
        plan_free(p);
        return MCINT_ERR_ARG;
      }
    for (size_t q = 0; q < nq; ++q) mx = std::max(mx, out_offsets[q] + qoff[q + 1]);
    for (size_t q = 0; q < nq; ++q) qoff[q] = out_offsets[q];
    p->out_size = mx;
  } else {
    for (size_t q = 0; q < nq; ++q) qoff[q + 1] += qoff[q];
    p->out_size = qoff[nq];
  }
  if (out_size) *out_size = p->out_size;
  p->ntasks = ntasks;
  p->stats[4] = nprimq_total;
  p->stats[5] = p->out_size;

  // ---- bucket index of every quartet: the distinct keys (a few hundred),
  // sorted, behind a small open-addressing table
  std::vector<uint32_t> ukeys;
  {
    std::vector<std::vector<uint32_t>> tk(nthr);
    parallel_for(nthr, nq, [&](int th, size_t q0, size_t q1) {
      KeySet ks;
      for (size_t q = q0; q < q1; ++q)
        if (qnt[q]) ks.insert(qkey[q]);
      tk[th] = ks.keys();
    });
    for (auto& v : tk) ukeys.insert(ukeys.end(), v.begin(), v.end());
    std::sort(ukeys.begin(), ukeys.end());
    ukeys.erase(std::unique(ukeys.begin(), ukeys.end()), ukeys.end());
  }
  timer.mark("offsets, distinct keys");
  const KeyIndex kindex(ukeys);
  const size_t nkeys = ukeys.size();
  std::vector<int64_t> tcount((size_t)nthr * nkeys, 0);
  std::vector<int64_t> pcount((size_t)nthr * nkeys, 0);  // primitive quartets
  parallel_for(nthr, nq, [&](int th, size_t q0, size_t q1) {
    int64_t* c = tcount.data() + (size_t)th * nkeys;
    int64_t* pc = pcount.data() + (size_t)th * nkeys;
    for (size_t q = q0; q < q1; ++q) {
      if (!qnt[q]) continue;
      const uint32_t id = kindex.find(qkey[q]);
      qkey[q] = id;
      c[id] += qnt[q];
      int64_t np = qnt[q];
      for (int a = 0; a < naxes; ++a) np *= bas[(size_t)shls[4 * q + a] * BAS_SLOTS + NPRIM_OF];
      pc[id] += np;
    }
  });
  // bucket starts, and each thread's cursor inside every bucket
  std::vector<int64_t> cnt(nkeys + 1, 0);
  std::vector<int64_t> cursor((size_t)nthr * nkeys, 0);
  {
    int64_t run = 0;
    for (size_t i = 0; i < nkeys; ++i) {
      cnt[i] = run;
      for (int th = 0; th < nthr; ++th) {
        cursor[(size_t)th * nkeys + i] = run;
        run += tcount[(size_t)th * nkeys + i];
      }
    }
    cnt[nkeys] = run;
  }
  timer.mark("bucket counts");
  // uninitialised: every slot is written by the placement pass
  std::unique_ptr<Task[]> sorted(new Task[(size_t)std::max<int64_t>(ntasks, 1)]);
  std::unique_ptr<int[]> sorted_q(new int[(size_t)std::max<int64_t>(ntasks, 1)]);
  parallel_for(nthr, nq, [&](int th, size_t q0, size_t q1) {
    int64_t* cur = cursor.data() + (size_t)th * nkeys;
    QuartetShape Q;
    for (size_t q = q0; q < q1; ++q) {
      if (!qnt[q]) continue;
      describe(q, Q);
      int la, lb, lc, ld, perm;
      tuple_class(kind, Q, la, lb, lc, ld, perm);
      const FINT* s = shls + 4 * q;
      int64_t& d = cur[qkey[q]];
      // virtual expansion over contraction columns of the original axes
      for_each_task(b->vfirst, b->dummy, kind, Q, s, qoff[q], perm, [&](const Task& t) {
        sorted[d] = t;
        sorted_q[d] = (int)q;
        ++d;
      });
    }
  });
  timer.mark("task placement");
  // Buckets of one- and two-lane teams hold tasks of 1 to 2*ppl primitive
  // quartets; a wave runs as long as its longest task, so these buckets are
  // ordered by descending primitive-quartet count (stable counting sort:
  // list order within equal counts).  Larger teams split a task's
  // primitives over their lanes and keep the list order.
  {
    const int naxes_s = kind_axes(kind);
    std::vector<size_t> small;
    for (size_t i = 0; i < nkeys; ++i)
      if (((ukeys[i] >> 20) & 15) <= 1 && ((ukeys[i] >> 24) & 3) != 0 && cnt[i + 1] - cnt[i] > 1)
        small.push_back(i);
    std::atomic<size_t> next{0};
    const int nw = std::min<int>(nthr, (int)small.size());
    parallel_for(nw, (size_t)nw, [&](int, size_t, size_t) {
      constexpr int KMAX = 256;
      std::vector<Task> tt;
      std::vector<int> tq;
      std::vector<uint16_t> key;
      for (size_t w; (w = next.fetch_add(1)) < small.size();) {
        const size_t i = small[w];
        const int64_t a = cnt[i], n = cnt[i + 1] - cnt[i];
        key.resize(n);
        int64_t hist[KMAX + 1] = {};
        for (int64_t t = 0; t < n; ++t) {
          const Task& tk = sorted[a + t];
          int64_t np = 1;
          for (int x = 0; x < naxes_s; ++x) np *= b->vsh[tk.sh[x]].nprim;
          key[t] = (uint16_t)(KMAX - std::min<int64_t>(np, KMAX));  // descending
          ++hist[key[t] + 1];
        }
        for (int k = 0; k < KMAX; ++k) hist[k + 1] += hist[k];
        tt.assign(sorted.get() + a, sorted.get() + a + n);
        tq.assign(sorted_q.get() + a, sorted_q.get() + a + n);
        for (int64_t t = 0; t < n; ++t) {
          const int64_t d = hist[key[t]]++;
          sorted[a + d] = tt[t];
          sorted_q[a + d] = tq[t];
        }
      }
    });
  }
  timer.mark("order small-team buckets by primitive count");
  int64_t max_gen = 0, max_flags = 0;
  for (size_t i = 0; i < nkeys; ++i) {
    const uint32_t k = ukeys[i];
    Bucket bk;
    bk.la = k & 7;
    bk.lb = (k >> 3) & 7;
    bk.lc = (k >> 6) & 7;
    bk.ld = (k >> 9) & 7;
    bk.team = 1 << ((k >> 20) & 15);
    bk.mode = (k >> 24) & 3;
    bk.fp32 = (k >> 26) & 1;
    bk.t0 = cnt[i];
    bk.t1 = cnt[i + 1];
    int64_t np = 0;
    for (int th = 0; th < nthr; ++th) np += pcount[(size_t)th * nkeys + i];
    bk.cost = bucket_cost(kind, bk.mode, bk.la, bk.lb, bk.lc, bk.ld, np, bk.t1 - bk.t0);
    if (bk.fp32) p->stats_fp32 += bk.t1 - bk.t0;
    bk.sstride = 0;
    if (bk.mode == 2) {
      const int64_t nef = chunked_block_doubles(kind, bk.la, bk.lb, bk.lc, bk.ld);
      bk.sstride = nef;
      const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(bk.t1 - bk.t0, GEN_SCRATCH_DOUBLES / nef));
      max_gen = std::max(max_gen, chunk * nef);
      max_flags = std::max(max_flags, chunk);
      p->stats[2] += bk.t1 - bk.t0;
    } else if (bk.mode == 0) {
      const GenLayout G = gen_layout(kind, sph, bk.la, bk.lb, bk.lc, bk.ld);
      const int nr = (bk.la + bk.lb + bk.lc + (kind_three(kind) ? 0 : bk.ld) +
                      kind_bra_raise(kind) + kind_ket_raise(kind)) / 2 + 1;
      const int nrt = b->omega < 0 ? 2 * nr : nr;
      if (G.nE > GEN_ECAP || G.nF > GEN_FCAP || 3 * nrt * (G.E + 1) * (G.F + 1) > GEN_VCAP) {
        plan_free(p);
        set_err("mcint_plan_create: angular momentum beyond the generic kernel's tables");
        return MCINT_ERR_UNSUPPORTED;
      }
      bk.sstride = G.total;
      const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(bk.t1 - bk.t0,
                                                   GEN_SCRATCH_DOUBLES / std::max<int64_t>(1, G.total)));
      max_gen = std::max(max_gen, chunk * G.total);
      p->stats[3] += bk.t1 - bk.t0;
    } else {
      p->stats[2] += bk.t1 - bk.t0;
    }
    p->buckets.push_back(bk);
  }
  p->stats[0] = p->ntasks;
  p->stats[1] = (int64_t)p->buckets.size();

  // ---- primitive-pair data of every pair the tasks touch ----------------
  const int nv = b->nv;
  std::vector<uint8_t> used((size_t)nv * nv, 0);
  // read before writing: stores to lines other threads read would bounce
  // the lines between cores
  auto mark = [&](size_t i) {
    if (!__atomic_load_n(&used[i], __ATOMIC_RELAXED)) __atomic_store_n(&used[i], (uint8_t)1, __ATOMIC_RELAXED);
  };
  parallel_for(nthr, (size_t)ntasks, [&](int, size_t t0, size_t t1) {
    for (size_t t = t0; t < t1; ++t) {
      const Task& tk = sorted[t];
      mark((size_t)tk.sh[0] * nv + tk.sh[1]);
      if (!kind_three(kind)) mark((size_t)tk.sh[2] * nv + tk.sh[3]);
    }
  });
  timer.mark("buckets, pair marks");
  std::vector<int64_t> pair_off((size_t)nv * nv, 0);
  std::vector<int> pairs;
  int64_t npp = 0;
  for (int a = 0; a < nv; ++a)
    for (int c = 0; c < nv; ++c) {
      if (!used[(size_t)a * nv + c]) continue;
      pair_off[(size_t)a * nv + c] = npp;
      npp += (int64_t)b->vsh[a].nprim * b->vsh[c].nprim;
      pairs.push_back(a);
      pairs.push_back(c);
    }
  if (!upload) {
    p->scratch_size = max_gen;
    p->flags_size = max_flags;
    *out = p;
    return MCINT_OK;
  }
  int* pairs_d = nullptr;
  PairPrim* pp_tmp = nullptr;
  hipError_t e = dalloc(&p->tasks_d, (size_t)std::max<int64_t>(ntasks, 1));
  if (e == hipSuccess) e = dalloc(&p->task_id_d, (size_t)std::max<int64_t>(ntasks, 1));
  if (e == hipSuccess) e = dalloc(&p->pair_off_d, pair_off.size());
  if (e == hipSuccess) e = dalloc(&p->pair_ld_d, pair_off.size());
  if (e == hipSuccess) e = dalloc(&p->pp_d, (size_t)std::max<int64_t>(npp, 1));
  if (e == hipSuccess) e = dalloc(&p->pp_pos_d, (size_t)std::max<int64_t>(npp, 1));
  if (e == hipSuccess) e = dalloc(&pp_tmp, (size_t)std::max<int64_t>(npp, 1));
  if (e == hipSuccess) e = dalloc(&pairs_d, pairs.size());
  p->scratch_size = max_gen;
  p->flags_size = max_flags;
  if (e == hipSuccess && ntasks > 0)
    e = hipMemcpy(p->tasks_d, sorted.get(), (size_t)ntasks * sizeof(Task), hipMemcpyHostToDevice);
  if (e == hipSuccess && ntasks > 0)
    e = hipMemcpy(p->task_id_d, sorted_q.get(), (size_t)ntasks * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(p->pair_off_d, pair_off.data(), pair_off.size() * sizeof(int64_t),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess && !pairs.empty())
    e = hipMemcpy(pairs_d, pairs.data(), pairs.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_pairdata(b->vsh_d, b->env_d, nv, pairs_d, (int)(pairs.size() / 2), p->pair_off_d,
                        p->pair_ld_d, pp_tmp, p->pp_d, p->pp_pos_d, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  timer.mark("device upload, pair data");
  dfree(pairs_d);
  dfree(pp_tmp);
  if (e != hipSuccess) {
    set_err(std::string("mcint_plan_create: ") + hipGetErrorString(e));
    plan_free(p);
    return MCINT_ERR_HIP;
  }
  *out = p;
  return MCINT_OK;
}

extern "C" int mcint_plan_create(mcint_plan** out, mcint_basis* b, int intor, const FINT* shls,
                                 size_t nq, const int64_t* out_offsets, int64_t* out_size) {
  return plan_create_impl(out, b, intor, shls, nq, out_offsets, nullptr, out_size);
}

extern "C" int mcint_plan_create_mixed(mcint_plan** out, mcint_basis* b, int intor,
                                       const FINT* shls, size_t nq, const int64_t* out_offsets,
                                       const uint8_t* fp32_flags, int64_t* out_size) {
  return plan_create_impl(out, b, intor, shls, nq, out_offsets, fp32_flags, out_size);
}

extern "C" int64_t mcint_plan_fp32_tasks(const mcint_plan* p) { return p ? p->stats_fp32 : 0; }

extern "C" int mcint_plan_stats(const mcint_plan* p, int64_t* n) {
  if (!p || !n) return MCINT_ERR_ARG;
  for (int i = 0; i < 6; ++i) n[i] = p->stats[i];
  return MCINT_OK;
}

static BasisView plan_view(const mcint_plan* p) {
  const mcint_basis* b = p->b;
  BasisView B;
  B.vsh = b->vsh_d;
  B.env = b->env_d;
  B.pair_off = p->pair_off_d;
  B.pair_ld = p->pair_ld_d;
  B.pp = p->pp_d;
  B.pp_pos = p->pp_pos_d;
  B.nv = b->nv;
  B.expcutoff_4c = b->expc4;
  B.expcutoff_3c = b->expc3;
  B.omega = b->omega;
  return B;
}

static BucketArgs bucket_args(const mcint_plan* p, const Bucket& bk, double* out, FINT* not0) {
  BucketArgs A;
  A.tasks = p->tasks_d;
  A.out = out;
  A.not0 = not0;
  A.task_id = p->task_id_d;
  A.scratch = p->scratch_d;
  A.flags = p->flags_d;
  A.scratch_stride = bk.sstride;
  A.la = bk.la;
  A.lb = bk.lb;
  A.lc = bk.lc;
  A.ld = bk.ld;
  A.kind = p->kind;
  A.sph = p->sph;
  A.team = bk.team;
  A.fp32 = bk.fp32;
  A.t0 = bk.t0;
  A.t1 = bk.t1;
  A.dm = nullptr;
  A.grad = nullptr;
  A.nao = 0;
  A.nslot = 1;
  A.grad_stride = 0;
  A.jk_j = nullptr;
  A.jk_k = nullptr;
  A.col_major = 0;
  return A;
}

// Device memory the chunked and generic paths use for per-task blocks
// (doubles) and the chunked path's per-task flags; buckets with more tasks
// than fit run in slices, one after the other on the stream.
struct ScratchView {
  double* blocks = nullptr;
  int64_t nblocks = 0;  // doubles
  int* flags = nullptr;
  int64_t nflags = 0;
};

// region k of the plan's scratch (one per concurrent graph branch)
static ScratchView plan_scratch(const mcint_plan* p, int k = 0) {
  return ScratchView{p->scratch_d ? p->scratch_d + (size_t)k * p->scratch_size : nullptr,
                     p->scratch_size, p->flags_d ? p->flags_d + (size_t)k * p->flags_size : nullptr,
                     p->flags_size};
}

// allocates the plan's scratch on first use, nbr regions; growing it drops
// the captured graph (which holds the old addresses)
static int ensure_scratch(mcint_plan* p, int nbr = 1) {
  if (p->scratch_nbr >= nbr) return MCINT_OK;
  if (p->scratch_size == 0 && p->flags_size == 0) {
    p->scratch_nbr = nbr;
    return MCINT_OK;
  }
  if (p->gexec) {
    HIPCHK(hipDeviceSynchronize());
    (void)hipGraphExecDestroy(p->gexec);
    p->gexec = nullptr;
  }
  if (p->graph) {
    (void)hipGraphDestroy(p->graph);
    p->graph = nullptr;
  }
  if (p->scratch_d) HIPCHK(hipDeviceSynchronize());
  dfree(p->scratch_d);
  dfree(p->flags_d);
  p->scratch_nbr = 0;
  if (p->scratch_size > 0) HIPCHK(dalloc(&p->scratch_d, (size_t)p->scratch_size * nbr));
  if (p->flags_size > 0) HIPCHK(dalloc(&p->flags_d, (size_t)p->flags_size * nbr));
  p->scratch_nbr = nbr;
  return MCINT_OK;
}

// tasks of one bucket that fit the scratch at once (0: not even one)
static int64_t slice_tasks(const Bucket& bk, const ScratchView& S) {
  int64_t n = S.nblocks / std::max<int64_t>(1, bk.sstride);
  if (bk.mode == 2) n = std::min(n, S.nflags);
  return n;
}

static int bucket_launch(mcint_plan* p, const Bucket& bk, double* out, FINT* not0,
                         hipStream_t stream, const ScratchView& S) {
  const BasisView B = plan_view(p);
  BucketArgs A = bucket_args(p, bk, out, not0);
  if (bk.mode == 1) {
    HIPCHK(launch_fast(B, A, stream));
    return MCINT_OK;
  }
  A.scratch = S.blocks;
  A.flags = S.flags;
  const int64_t chunk = slice_tasks(bk, S);
  if (chunk < 1) {
    char msg[200];
    snprintf(msg, sizeof msg,
             "class (%d%d|%d%d): one task needs %lld doubles of scratch, %lld available",
             bk.la, bk.lb, bk.lc, bk.ld, (long long)bk.sstride, (long long)S.nblocks);
    set_err(msg);
    return MCINT_ERR_UNSUPPORTED;
  }
  for (int64_t t = bk.t0; t < bk.t1; t += chunk) {
    A.t0 = t;
    A.t1 = std::min(bk.t1, t + chunk);
    if (bk.mode == 2)
      HIPCHK(launch_chunked(B, A, stream));
    else
      HIPCHK(launch_generic(B, A, stream));
  }
  return MCINT_OK;
}

static int plan_launch(mcint_plan* p, double* out, FINT* not0, hipStream_t stream,
                       const ScratchView& S) {
  for (const Bucket& bk : p->buckets) {
    const int st = bucket_launch(p, bk, out, not0, stream, S);
    if (st) return st;
  }
  return MCINT_OK;
}

// Every bucket of p forked over nbr side streams that join back into
// 'stream' (captured, the graph holds nbr concurrent branches).  Buckets go
// to the least loaded branch in descending estimated cost; each branch runs
// its buckets one after another with its own scratch region:
// launch(bucket, branch stream, branch scratch) launches one bucket.  The
// joins are recorded even when a launch fails, so that a capture can
// always end.
template <class F>
static int launch_branched(mcint_plan* p, hipStream_t stream, int nbr, F&& launch) {
  while ((int)p->side.size() < nbr) {
    hipStream_t st = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    p->side.push_back(st);
  }
  while ((int)p->events.size() < nbr + 1) {
    hipEvent_t ev = nullptr;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    p->events.push_back(ev);
  }
  std::vector<int> order(p->buckets.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return p->buckets[a].cost > p->buckets[b].cost; });
  std::vector<double> load(nbr, 0.0);
  std::vector<std::vector<int>> lists(nbr);
  for (int i : order) {
    const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    lists[k].push_back(i);
    load[k] += p->buckets[i].cost;
  }
  HIPCHK(hipEventRecord(p->events[0], stream));
  int st = MCINT_OK;
  hipError_t e = hipSuccess;
  for (int k = 0; k < nbr; ++k) {
    e = hipStreamWaitEvent(p->side[k], p->events[0], 0);
    if (e != hipSuccess) break;
  }
  if (e == hipSuccess) {
    for (int k = 0; k < nbr && st == MCINT_OK; ++k)
      for (int i : lists[k]) {
        st = launch(p->buckets[i], p->side[k], plan_scratch(p, k));
        if (st) break;
      }
  }
  for (int k = 0; k < nbr; ++k) {
    const hipError_t a = hipEventRecord(p->events[k + 1], p->side[k]);
    const hipError_t b = hipStreamWaitEvent(stream, p->events[k + 1], 0);
    if (e == hipSuccess) e = a != hipSuccess ? a : b;
  }
  if (st != MCINT_OK) return st;
  HIPCHK(e);
  return MCINT_OK;
}

extern "C" int mcint_plan_execute_grad(mcint_plan* p, const double* dm_dev, int nao,
                                       double* grad_dev, void* hip_stream) {
  if (!p || !dm_dev || !grad_dev || nao <= 0) {
    set_err("mcint_plan_execute_grad: null argument");
    return MCINT_ERR_ARG;
  }
  if (p->intor != MCINT_INT2E_IP1_SPH) {
    set_err("mcint_plan_execute_grad: the plan must be int2e_ip1_sph");
    return MCINT_ERR_ARG;
  }
  const int nao_b = CINTtot_cgto_spheric(p->b->bas.data(), p->b->nbas);
  if (nao != nao_b) {
    set_err("mcint_plan_execute_grad: density dimension does not match the basis");
    return MCINT_ERR_ARG;
  }
  hipStream_t stream = (hipStream_t)hip_stream;
  for (const Bucket& bk : p->buckets)
    if (bk.mode == 1) {
      set_err("mcint_plan_execute_grad: internal: derivative bucket on the single-kernel path");
      return MCINT_ERR_UNSUPPORTED;
    }
  // the buckets run over concurrent branches (as a captured plan does);
  // their digests add atomically into the shared accumulators
  const int nbr = std::max(1, std::min<int>(g_graph_branches, (int)p->buckets.size()));
  if (const int st = ensure_scratch(p, nbr)) return st;
  const BasisView B = plan_view(p);
  // per-block partial sums go to GRAD_SLOTS copies of the accumulators
  // (block % GRAD_SLOTS), reduced into grad_dev at the end
  const int n = 3 * p->b->natm;
  if (!p->gslots_d) HIPCHK(dalloc(&p->gslots_d, (size_t)GRAD_SLOTS * n));
  HIPCHK(hipMemsetAsync(p->gslots_d, 0, sizeof(double) * GRAD_SLOTS * n, stream));
  auto launch = [&](const Bucket& bk, hipStream_t s, const ScratchView& S) -> int {
    BucketArgs A = bucket_args(p, bk, nullptr, nullptr);
    A.dm = dm_dev;
    A.grad = p->gslots_d;
    A.nao = nao;
    A.nslot = GRAD_SLOTS;
    A.grad_stride = n;
    A.scratch = S.blocks;
    A.flags = S.flags;
    const int64_t chunk = std::max<int64_t>(1, slice_tasks(bk, S));
    for (int64_t t = bk.t0; t < bk.t1; t += chunk) {
      A.t0 = t;
      A.t1 = std::min(bk.t1, t + chunk);
      if (bk.mode == 2)
        HIPCHK(launch_chunked(B, A, s));
      else
        HIPCHK(launch_generic(B, A, s));
    }
    return MCINT_OK;
  };
  if (nbr > 1) {
    if (const int st = launch_branched(p, stream, nbr, launch)) return st;
  } else {
    for (const Bucket& bk : p->buckets)
      if (const int st = launch(bk, stream, plan_scratch(p))) return st;
  }
  HIPCHK(launch_grad_reduce(p->gslots_d, GRAD_SLOTS, n, grad_dev, stream));
  return MCINT_OK;
}

extern "C" int mcint_plan_execute_jk(mcint_plan* p, const double* dm_dev, int nao,
                                     double* vj_dev, double* vk_dev, void* hip_stream) {
  if (!p || !dm_dev || (!vj_dev && !vk_dev) || nao <= 0) {
    set_err("mcint_plan_execute_jk: null argument");
    return MCINT_ERR_ARG;
  }
  if (p->intor != MCINT_INT2E_SPH) {
    set_err("mcint_plan_execute_jk: the plan must be int2e_sph over unique quartets");
    return MCINT_ERR_ARG;
  }
  if (p->b->omega < 0) {
    set_err("mcint_plan_execute_jk: the short-range operator has no fused digest; "
            "execute the plan and digest its buffer (mcint_jk_plan_execute)");
    return MCINT_ERR_UNSUPPORTED;
  }
  const int nao_b = CINTtot_cgto_spheric(p->b->bas.data(), p->b->nbas);
  if (nao != nao_b) {
    set_err("mcint_plan_execute_jk: density dimension does not match the basis");
    return MCINT_ERR_ARG;
  }
  hipStream_t stream = (hipStream_t)hip_stream;
  const int nbr = std::max(1, std::min<int>(g_graph_branches, (int)p->buckets.size()));
  if (const int st = ensure_scratch(p, nbr)) return st;
  const int64_t nn = (int64_t)nao * nao;
  if (p->jk_n < nn) {
    if (p->jk_d) HIPCHK(hipDeviceSynchronize());
    dfree(p->jk_d);
    HIPCHK(dalloc(&p->jk_d, (size_t)(2 * nn)));
    p->jk_n = nn;
  }
  double* Jp = vj_dev ? p->jk_d : nullptr;
  double* Kp = vk_dev ? p->jk_d + nn : nullptr;
  HIPCHK(launch_zero(p->jk_d, 2 * nn, stream));
  const BasisView B = plan_view(p);
  auto launch = [&](const Bucket& bk, hipStream_t s, const ScratchView& S) -> int {
    BucketArgs A = bucket_args(p, bk, nullptr, nullptr);
    A.dm = dm_dev;
    A.nao = nao;
    A.jk_j = Jp;
    A.jk_k = Kp;
    A.scratch = S.blocks;
    A.flags = S.flags;
    if (bk.mode == 1) {
      HIPCHK(launch_fast(B, A, s));
      return MCINT_OK;
    }
    const int64_t chunk = std::max<int64_t>(1, slice_tasks(bk, S));
    for (int64_t t = bk.t0; t < bk.t1; t += chunk) {
      A.t0 = t;
      A.t1 = std::min(bk.t1, t + chunk);
      if (bk.mode == 2)
        HIPCHK(launch_chunked(B, A, s));
      else
        HIPCHK(launch_generic(B, A, s));
    }
    return MCINT_OK;
  };
  if (nbr > 1) {
    if (const int st = launch_branched(p, stream, nbr, launch)) return st;
  } else {
    for (const Bucket& bk : p->buckets)
      if (const int st = launch(bk, stream, plan_scratch(p))) return st;
  }
  if (vj_dev) HIPCHK(launch_symmetrize(Jp, vj_dev, nao, stream));
  if (vk_dev) HIPCHK(launch_symmetrize(Kp, vk_dev, nao, stream));
  return MCINT_OK;
}

extern "C" int mcint_jk_digest(const FINT* shls_dev, const int64_t* off_dev, size_t nq,
                               const double* eri_dev, const mcint_basis* b, const double* dm_dev,
                               double* vj_dev, double* vk_dev, void* hip_stream) {
  if (!b || (nq && (!shls_dev || !off_dev || !eri_dev)) || !dm_dev) {
    set_err("mcint_jk_digest: null argument");
    return MCINT_ERR_ARG;
  }
  const int nao = CINTtot_cgto_spheric(b->bas.data(), b->nbas);
  HIPCHK(launch_jk_digest(shls_dev, off_dev, (int64_t)nq, eri_dev, b->ao_loc_d, dm_dev, nao, vj_dev,
                          vk_dev, (hipStream_t)hip_stream));
  return MCINT_OK;
}

// ---------------------------------------------------------------------------
// J/K plan: grouped digest of a unique-quartet ERI list
// ---------------------------------------------------------------------------
struct mcint_jk_plan {
  mcint_basis* b = nullptr;
  int nao = 0;
  int64_t nq = 0;
  int* shls_d = nullptr;
  int64_t* off_d = nullptr;
  bool grouped = false;
  int* rows_q_d = nullptr;
  int* rows_ptr_d = nullptr;
  int nrows = 0;
  size_t rows_lds = 0;
  int* cols_q_d = nullptr;
  int* cols_ptr_d = nullptr;
  int ncols = 0;
  double* jp_d = nullptr;
  double* kp_d = nullptr;
};

static void jk_plan_free(mcint_jk_plan* j) {
  if (!j) return;
  dfree(j->shls_d);
  dfree(j->off_d);
  dfree(j->rows_q_d);
  dfree(j->rows_ptr_d);
  dfree(j->cols_q_d);
  dfree(j->cols_ptr_d);
  dfree(j->jp_d);
  dfree(j->kp_d);
  delete j;
}

extern "C" void mcint_jk_plan_destroy(mcint_jk_plan* j) { jk_plan_free(j); }

// quartet indices grouped by a shell-pair key (stable), groups of at most
// JK_GROUP_MAX quartets
static void jk_groups(const std::vector<int64_t>& key, int64_t nkeys, std::vector<int>& order,
                      std::vector<int>& ptr) {
  constexpr int JK_GROUP_MAX = 8192;
  const size_t nq = key.size();
  order.resize(nq);
  std::vector<int64_t> cnt((size_t)nkeys + 1, 0);
  for (size_t q = 0; q < nq; ++q) cnt[key[q] + 1]++;
  for (int64_t k = 0; k < nkeys; ++k) cnt[k + 1] += cnt[k];
  std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
  for (size_t q = 0; q < nq; ++q) order[pos[key[q]]++] = (int)q;
  ptr.clear();
  for (int64_t k = 0; k < nkeys; ++k)
    for (int64_t a = cnt[k]; a < cnt[k + 1]; a += JK_GROUP_MAX) ptr.push_back((int)a);
  ptr.push_back((int)nq);
}

extern "C" int mcint_jk_plan_create(mcint_jk_plan** out, mcint_basis* b, const FINT* shls,
                                    size_t nq, const int64_t* out_offsets) {
  if (!out || !b || (!shls && nq)) {
    set_err("mcint_jk_plan_create: null argument");
    return MCINT_ERR_ARG;
  }
  *out = nullptr;
  if (nq > (size_t)INT32_MAX) {
    set_err("mcint_jk_plan_create: more than 2^31-1 quartets");
    return MCINT_ERR_ARG;
  }
  const int nbas = b->nbas;
  const int* bas = b->bas.data();
  std::vector<int> ao(nbas + 1, 0);
  int dmax = 0;
  for (int i = 0; i < nbas; ++i) {
    const int d = nsph(bas[i * BAS_SLOTS + ANG_OF]) * bas[i * BAS_SLOTS + NCTR_OF];
    ao[i + 1] = ao[i] + d;
    dmax = std::max(dmax, d);
  }
  auto* j = new mcint_jk_plan();
  j->b = b;
  j->nao = ao[nbas];
  j->nq = (int64_t)nq;
  // block offsets: given, or packed in list order (as mcint_plan_create)
  std::vector<int64_t> off(nq, 0);
  int64_t run = 0;
  for (size_t q = 0; q < nq; ++q) {
    const FINT* sq = shls + 4 * q;
    int64_t blk = 1;
    for (int a = 0; a < 4; ++a) {
      if (sq[a] < 0 || sq[a] >= nbas) {
        jk_plan_free(j);
        set_err("mcint_jk_plan_create: shell index out of range");
        return MCINT_ERR_ARG;
      }
      blk *= ao[sq[a] + 1] - ao[sq[a]];
    }
    off[q] = out_offsets ? out_offsets[q] : run;
    if (off[q] < 0) {
      jk_plan_free(j);
      set_err("mcint_jk_plan_create: negative block offset");
      return MCINT_ERR_ARG;
    }
    run += blk;
  }
  const size_t lds = ((size_t)2 * dmax * j->nao + 256) * sizeof(double);
  j->grouped = dmax * dmax <= 256 && lds <= 64 * 1024 && (int64_t)nbas * nbas <= (64ll << 20);
  hipError_t e = dalloc(&j->shls_d, 4 * std::max<size_t>(nq, 1));
  if (e == hipSuccess) e = dalloc(&j->off_d, std::max<size_t>(nq, 1));
  if (e == hipSuccess && nq) e = hipMemcpy(j->shls_d, shls, 4 * nq * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess && nq) e = hipMemcpy(j->off_d, off.data(), nq * sizeof(int64_t), hipMemcpyHostToDevice);
  if (e == hipSuccess && j->grouped) {
    j->rows_lds = lds;
    std::vector<int64_t> key(nq);
    std::vector<int> order, ptr;
    for (size_t q = 0; q < nq; ++q) key[q] = (int64_t)shls[4 * q] * nbas + shls[4 * q + 1];
    jk_groups(key, (int64_t)nbas * nbas, order, ptr);
    j->nrows = (int)ptr.size() - 1;
    e = dalloc(&j->rows_q_d, std::max<size_t>(nq, 1));
    if (e == hipSuccess) e = dalloc(&j->rows_ptr_d, ptr.size());
    if (e == hipSuccess && nq) e = hipMemcpy(j->rows_q_d, order.data(), nq * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(j->rows_ptr_d, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice);
    for (size_t q = 0; q < nq; ++q) key[q] = (int64_t)shls[4 * q + 2] * nbas + shls[4 * q + 3];
    jk_groups(key, (int64_t)nbas * nbas, order, ptr);
    j->ncols = (int)ptr.size() - 1;
    if (e == hipSuccess) e = dalloc(&j->cols_q_d, std::max<size_t>(nq, 1));
    if (e == hipSuccess) e = dalloc(&j->cols_ptr_d, ptr.size());
    if (e == hipSuccess && nq) e = hipMemcpy(j->cols_q_d, order.data(), nq * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(j->cols_ptr_d, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = dalloc(&j->jp_d, (size_t)j->nao * j->nao);
    if (e == hipSuccess) e = dalloc(&j->kp_d, (size_t)j->nao * j->nao);
  }
  if (e != hipSuccess) {
    set_err(std::string("mcint_jk_plan_create: ") + hipGetErrorString(e));
    jk_plan_free(j);
    return MCINT_ERR_HIP;
  }
  *out = j;
  return MCINT_OK;
}

extern "C" int mcint_jk_plan_execute(mcint_jk_plan* j, const double* eri_dev, const double* dm_dev,
                                     double* vj_dev, double* vk_dev, void* hip_stream) {
  if (!j || !dm_dev || (j->nq && !eri_dev)) {
    set_err("mcint_jk_plan_execute: null argument");
    return MCINT_ERR_ARG;
  }
  hipStream_t stream = (hipStream_t)hip_stream;
  const size_t nn = (size_t)j->nao * j->nao * sizeof(double);
  if (!j->grouped) {
    if (vj_dev) HIPCHK(hipMemsetAsync(vj_dev, 0, nn, stream));
    if (vk_dev) HIPCHK(hipMemsetAsync(vk_dev, 0, nn, stream));
    HIPCHK(launch_jk_digest(j->shls_d, j->off_d, j->nq, eri_dev, j->b->ao_loc_d, dm_dev, j->nao,
                            vj_dev, vk_dev, stream));
    return MCINT_OK;
  }
  double* Jp = vj_dev ? j->jp_d : nullptr;
  double* Kp = vk_dev ? j->kp_d : nullptr;
  if (Jp) HIPCHK(hipMemsetAsync(Jp, 0, nn, stream));
  if (Kp) HIPCHK(hipMemsetAsync(Kp, 0, nn, stream));
  if (Jp || Kp)
    HIPCHK(launch_jk_grouped(j->shls_d, j->off_d, j->rows_q_d, j->rows_ptr_d, j->nrows, j->rows_lds,
                             j->cols_q_d, j->cols_ptr_d, j->ncols, eri_dev, j->b->ao_loc_d, dm_dev,
                             j->nao, Jp, Kp, stream));
  if (Jp) HIPCHK(launch_symmetrize(Jp, vj_dev, j->nao, stream));
  if (Kp) HIPCHK(launch_symmetrize(Kp, vk_dev, j->nao, stream));
  return MCINT_OK;
}

extern "C" int mcint_plan_nbuckets(const mcint_plan* p) { return p ? (int)p->buckets.size() : 0; }

extern "C" int mcint_plan_bucket_info(const mcint_plan* p, int i, int64_t* info) {
  if (!p || !info || i < 0 || i >= (int)p->buckets.size()) {
    set_err("mcint_plan_bucket_info: bad bucket index");
    return MCINT_ERR_ARG;
  }
  const Bucket& b = p->buckets[i];
  // perm -1: the bucket holds every axis order of its class
  const int64_t v[8] = {b.la, b.lb, b.lc, b.ld, -1, b.team, b.mode, b.t1 - b.t0};
  for (int k = 0; k < 8; ++k) info[k] = v[k];
  return MCINT_OK;
}

extern "C" int mcint_plan_bucket_tasks(const mcint_plan* p, int i, int64_t first, int64_t n,
                                       int* quartet) {
  if (!p || i < 0 || i >= (int)p->buckets.size() || first < 0 || n < 0 || (n && !quartet)) {
    set_err("mcint_plan_bucket_tasks: bad argument");
    return MCINT_ERR_ARG;
  }
  const Bucket& b = p->buckets[i];
  if (first + n > b.t1 - b.t0) {
    set_err("mcint_plan_bucket_tasks: range beyond the bucket");
    return MCINT_ERR_ARG;
  }
  if (n) HIPCHK(hipMemcpy(quartet, p->task_id_d + b.t0 + first, (size_t)n * sizeof(int),
                          hipMemcpyDeviceToHost));
  return MCINT_OK;
}

extern "C" int mcint_plan_execute_bucket(mcint_plan* p, int i, double* out_dev, FINT* not0_dev,
                                         void* hip_stream) {
  if (!p || i < 0 || i >= (int)p->buckets.size() || !out_dev) {
    set_err("mcint_plan_execute_bucket: bad argument");
    return MCINT_ERR_ARG;
  }
  if (const int st = ensure_scratch(p)) return st;
  return bucket_launch(p, p->buckets[i], out_dev, not0_dev, (hipStream_t)hip_stream,
                       plan_scratch(p));
}

extern "C" int mcint_plan_count_primitives(mcint_plan* p, int64_t* per_bucket) {
  if (!p || !per_bucket) {
    set_err("mcint_plan_count_primitives: null argument");
    return MCINT_ERR_ARG;
  }
  const size_t nb = p->buckets.size();
  unsigned long long* cnt = nullptr;
  HIPCHK(dalloc(&cnt, std::max<size_t>(nb, 1)));
  hipError_t e = hipMemset(cnt, 0, std::max<size_t>(nb, 1) * sizeof(unsigned long long));
  const BasisView B = plan_view(p);
  for (size_t i = 0; i < nb && e == hipSuccess; ++i) {
    BucketArgs A = bucket_args(p, p->buckets[i], nullptr, nullptr);
    e = launch_count_prims(B, A, cnt + i, nullptr);
  }
  std::vector<unsigned long long> h(std::max<size_t>(nb, 1));
  if (e == hipSuccess) e = hipMemcpy(h.data(), cnt, nb * sizeof(unsigned long long), hipMemcpyDeviceToHost);
  dfree(cnt);
  if (e != hipSuccess) {
    set_err(std::string("mcint_plan_count_primitives: ") + hipGetErrorString(e));
    return MCINT_ERR_HIP;
  }
  for (size_t i = 0; i < nb; ++i) per_bucket[i] = (int64_t)h[i];
  return MCINT_OK;
}

extern "C" int mcint_plan_execute_direct(mcint_plan* p, double* out_dev, FINT* not0_dev,
                                         void* hip_stream) {
  if (!p || (!out_dev && p->ntasks)) {
    set_err("mcint_plan_execute: null argument");
    return MCINT_ERR_ARG;
  }
  if (const int st = ensure_scratch(p)) return st;
  return plan_launch(p, out_dev, not0_dev, (hipStream_t)hip_stream, plan_scratch(p));
}

extern "C" int mcint_plan_execute(mcint_plan* p, double* out_dev, FINT* not0_dev, void* hip_stream) {
  if (!p || (!out_dev && p->ntasks)) {
    set_err("mcint_plan_execute: null argument");
    return MCINT_ERR_ARG;
  }
  hipStream_t stream = (hipStream_t)hip_stream;
  if (!stream) {
    if (const int st = ensure_scratch(p)) return st;
    return plan_launch(p, out_dev, not0_dev, stream, plan_scratch(p));
  }
  const int nbr = std::max(1, std::min<int>(g_graph_branches, (int)p->buckets.size()));
  if (const int st = ensure_scratch(p, nbr)) return st;
  if (!p->gexec || p->g_out != out_dev || p->g_not0 != not0_dev || p->g_stream != stream ||
      p->g_nbr != nbr) {
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->gexec = nullptr;
    p->graph = nullptr;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int st = nbr > 1 ? launch_branched(p, stream, nbr,
                                       [&](const Bucket& bk, hipStream_t s, const ScratchView& S) {
                                         return bucket_launch(p, bk, out_dev, not0_dev, s, S);
                                       })
                     : plan_launch(p, out_dev, not0_dev, stream, plan_scratch(p));
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(stream, &g);
    if (st != MCINT_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    HIPCHK(e);
    p->graph = g;
    HIPCHK(hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0));
    p->g_out = out_dev;
    p->g_not0 = not0_dev;
    p->g_stream = stream;
    p->g_nbr = nbr;
  }
  HIPCHK(hipGraphLaunch(p->gexec, stream));
  return MCINT_OK;
}

// Every bucket of p on the stream with caller-supplied scratch (blocks of
// nblocks doubles, nflags ints of flags); rejects, before any launch, a
// class whose single task does not fit.
int mcint::plan_launch_scratch(mcint_plan* p, double* out, FINT* not0, hipStream_t stream,
                               double* blocks, int64_t nblocks, int* flags, int64_t nflags) {
  ScratchView S;
  S.blocks = blocks;
  S.nblocks = nblocks;
  S.flags = flags;
  S.nflags = nflags;
  for (const Bucket& bk : p->buckets) {
    if (bk.mode != 1 && slice_tasks(bk, S) < 1) {
      char msg[200];
      snprintf(msg, sizeof msg,
               "class (%d%d|%d%d) needs %lld doubles of scratch per task, %lld available",
               bk.la, bk.lb, bk.lc, bk.ld, (long long)bk.sstride, (long long)S.nblocks);
      set_err(msg);
      return MCINT_ERR_UNSUPPORTED;
    }
  }
  return plan_launch(p, out, not0, stream, S);
}

int64_t mcint::plan_tasks(const mcint_plan* p) { return p->ntasks; }

int mcint::plan_build_host_only(const FINT* atm, FINT natm, const FINT* bas, FINT nbas,
                                const double* env, int intor, const FINT* shls, size_t nq,
                                const int64_t* out_offsets, const uint8_t* fp32_flags,
                                int64_t* stats, std::vector<int64_t>* bucket_info) {
  mcint_basis* b = nullptr;
  int st = basis_create_impl(&b, atm, natm, bas, nbas, env, false);
  if (st != MCINT_OK) return st;
  mcint_plan* p = nullptr;
  st = plan_create_impl(&p, b, intor, shls, nq, out_offsets, fp32_flags, nullptr, false);
  if (st == MCINT_OK) {
    if (stats) mcint_plan_stats(p, stats);
    if (bucket_info) {
      bucket_info->clear();
      for (int i = 0; i < (int)p->buckets.size(); ++i) {
        int64_t info[8];
        mcint_plan_bucket_info(p, i, info);
        bucket_info->insert(bucket_info->end(), info, info + 8);
      }
    }
    plan_free(p);
  }
  mcint_basis_destroy(b);
  return st;
}

void mcint::set_error(const char* msg) { set_err(msg); }

// ---------------------------------------------------------------------------
// Host-pointer batches
// ---------------------------------------------------------------------------
static int host_batch(int intor, double* out, const size_t* out_offsets, const FINT* shls,
                      size_t nq, const FINT* atm, FINT natm, const FINT* bas, FINT nbas,
                      const double* env, FINT* not0, void* hip_stream) {
  if (!out || (!shls && nq)) {
    set_err("batch: null argument");
    return MCINT_ERR_ARG;
  }
  if (nq == 0) return MCINT_OK;
  hipStream_t stream = (hipStream_t)hip_stream;
  mcint_basis* b = nullptr;
  int st = mcint_basis_create(&b, atm, natm, bas, nbas, env);
  if (st) return st;
  std::vector<int64_t> offs;
  if (out_offsets) offs.assign(out_offsets, out_offsets + nq);
  mcint_plan* p = nullptr;
  int64_t osz = 0;
  st = mcint_plan_create(&p, b, intor, shls, nq, out_offsets ? offs.data() : nullptr, &osz);
  if (st) {
    mcint_basis_destroy(b);
    return st;
  }
  double* od = nullptr;
  FINT* nd = nullptr;
  hipError_t e = dalloc(&od, (size_t)std::max<int64_t>(osz, 1));
  if (e == hipSuccess) e = dalloc(&nd, nq);
  if (e == hipSuccess && out_offsets) {
    // blocks not covered by the list keep the caller's values
    e = hipMemcpyAsync(od, out, (size_t)osz * sizeof(double), hipMemcpyHostToDevice, stream);
  }
  if (e == hipSuccess) {
    st = mcint_plan_execute_direct(p, od, nd, stream);
    if (st == MCINT_OK) {
      e = hipMemcpyAsync(out, od, (size_t)osz * sizeof(double), hipMemcpyDeviceToHost, stream);
      if (e == hipSuccess && not0)
        e = hipMemcpyAsync(not0, nd, nq * sizeof(FINT), hipMemcpyDeviceToHost, stream);
      if (e == hipSuccess) e = hipStreamSynchronize(stream);
    }
  }
  if (e != hipSuccess) {
    set_err(std::string("batch: ") + hipGetErrorString(e));
    st = MCINT_ERR_HIP;
  }
  dfree(od);
  dfree(nd);
  mcint_plan_destroy(p);
  mcint_basis_destroy(b);
  return st;
}

extern "C" int int2e_sph_batch(double* out, const size_t* out_offsets, const FINT* shls, size_t nq,
                               const FINT* atm, FINT natm, const FINT* bas, FINT nbas,
                               const double* env, const CINTOpt* opt, FINT* not0, void* stream) {
  (void)opt;
  return host_batch(MCINT_INT2E_SPH, out, out_offsets, shls, nq, atm, natm, bas, nbas, env, not0,
                    stream);
}
extern "C" int int2e_cart_batch(double* out, const size_t* out_offsets, const FINT* shls,
                                size_t nq, const FINT* atm, FINT natm, const FINT* bas, FINT nbas,
                                const double* env, const CINTOpt* opt, FINT* not0, void* stream) {
  (void)opt;
  return host_batch(MCINT_INT2E_CART, out, out_offsets, shls, nq, atm, natm, bas, nbas, env, not0,
                    stream);
}
extern "C" int int2e_ip1_sph_batch(double* out, const size_t* out_offsets, const FINT* shls,
                                   size_t nq, const FINT* atm, FINT natm, const FINT* bas,
                                   FINT nbas, const double* env, const CINTOpt* opt, FINT* not0,
                                   void* stream) {
  (void)opt;
  return host_batch(MCINT_INT2E_IP1_SPH, out, out_offsets, shls, nq, atm, natm, bas, nbas, env,
                    not0, stream);
}
extern "C" int int3c2e_sph_batch(double* out, const size_t* out_offsets, const FINT* shls,
                                 size_t nq, const FINT* atm, FINT natm, const FINT* bas, FINT nbas,
                                 const double* env, const CINTOpt* opt, FINT* not0, void* stream) {
  (void)opt;
  return host_batch(MCINT_INT3C2E_SPH, out, out_offsets, shls, nq, atm, natm, bas, nbas, env, not0,
                    stream);
}
extern "C" int int2e_ip1_cart_batch(double* out, const size_t* out_offsets, const FINT* shls,
                                    size_t nq, const FINT* atm, FINT natm, const FINT* bas,
                                    FINT nbas, const double* env, const CINTOpt* opt, FINT* not0,
                                    void* stream) {
  (void)opt;
  return host_batch(MCINT_INT2E_IP1_CART, out, out_offsets, shls, nq, atm, natm, bas, nbas, env,
                    not0, stream);
}
extern "C" int int3c2e_cart_batch(double* out, const size_t* out_offsets, const FINT* shls,
                                  size_t nq, const FINT* atm, FINT natm, const FINT* bas,
                                  FINT nbas, const double* env, const CINTOpt* opt, FINT* not0,
                                  void* stream) {
  (void)opt;
  return host_batch(MCINT_INT3C2E_CART, out, out_offsets, shls, nq, atm, natm, bas, nbas, env,
                    not0, stream);
}
#define MC_BATCH(name, code)                                                                    \
  extern "C" int name(double* out, const size_t* out_offsets, const FINT* shls, size_t nq,      \
                      const FINT* atm, FINT natm, const FINT* bas, FINT nbas, const double* env, \
                      const CINTOpt* opt, FINT* not0, void* stream) {                           \
    (void)opt;                                                                                  \
    return host_batch(code, out, out_offsets, shls, nq, atm, natm, bas, nbas, env, not0,        \
                      stream);                                                                  \
  }
MC_BATCH(int3c2e_ip1_sph_batch, MCINT_INT3C2E_IP1_SPH)
MC_BATCH(int3c2e_ip1_cart_batch, MCINT_INT3C2E_IP1_CART)
MC_BATCH(int3c2e_ip2_sph_batch, MCINT_INT3C2E_IP2_SPH)
MC_BATCH(int3c2e_ip2_cart_batch, MCINT_INT3C2E_IP2_CART)
MC_BATCH(int2c2e_sph_batch, MCINT_INT2C2E_SPH)
MC_BATCH(int2c2e_cart_batch, MCINT_INT2C2E_CART)
MC_BATCH(int2e_ip2_sph_batch, MCINT_INT2E_IP2_SPH)
MC_BATCH(int2e_ip2_cart_batch, MCINT_INT2E_IP2_CART)
MC_BATCH(int2e_ipip1_sph_batch, MCINT_INT2E_IPIP1_SPH)
MC_BATCH(int2e_ipip1_cart_batch, MCINT_INT2E_IPIP1_CART)
MC_BATCH(int2e_ip1ip2_sph_batch, MCINT_INT2E_IP1IP2_SPH)
MC_BATCH(int2e_ip1ip2_cart_batch, MCINT_INT2E_IP1IP2_CART)
#undef MC_BATCH

// ---------------------------------------------------------------------------
// libcint-named per-call functions
// ---------------------------------------------------------------------------
namespace {

// Device state of an optimizer (the GPU's counterpart of libcint's CINTOpt,
// src/optimizer.c:183-191,344-425): the basis uploaded once, the primitive
// pair data of every ordered pair of (virtual) shells, and a pool of call
// contexts, one per host thread calling at the same time.  Per-call
// integrals with this optimizer only fill a task, launch its kernel and
// copy the block back.
struct CallCtx {
  hipStream_t stream = nullptr;
  Task* task_h = nullptr;  // pinned staging of the tasks
  Task* task_d = nullptr;
  int* tid_d = nullptr;    // task ids: all zero (one shell tuple per call)
  int ntask_cap = 0;
  double* out_d = nullptr;
  double* out_h = nullptr;  // pinned
  int64_t out_cap = 0;
  FINT* not0_d = nullptr;
  FINT* not0_h = nullptr;  // pinned
  double* scr_d = nullptr;
  int64_t scr_cap = 0;
  int* flags_d = nullptr;
  int flags_cap = 0;
};

void ctx_free(CallCtx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->task_h) (void)hipHostFree(c->task_h);
  if (c->out_h) (void)hipHostFree(c->out_h);
  if (c->not0_h) (void)hipHostFree(c->not0_h);
  dfree(c->task_d);
  dfree(c->tid_d);
  dfree(c->out_d);
  dfree(c->not0_d);
  dfree(c->scr_d);
  dfree(c->flags_d);
  delete c;
}

// grows the context's buffers to nt tasks, blk output doubles, nscr scratch
// doubles and nflag flags
hipError_t ctx_reserve(CallCtx* c, int nt, int64_t blk, int64_t nscr, int nflag) {
  hipError_t e = hipSuccess;
  if (!c->stream) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess && !c->not0_d) e = dalloc(&c->not0_d, 1);
  if (e == hipSuccess && !c->not0_h) e = hipHostMalloc((void**)&c->not0_h, sizeof(FINT));
  if (e == hipSuccess && nt > c->ntask_cap) {
    if (c->task_h) (void)hipHostFree(c->task_h);
    c->task_h = nullptr;
    dfree(c->task_d);
    dfree(c->tid_d);
    c->ntask_cap = 0;
    e = hipHostMalloc((void**)&c->task_h, sizeof(Task) * nt);
    if (e == hipSuccess) e = dalloc(&c->task_d, nt);
    if (e == hipSuccess) e = dalloc(&c->tid_d, nt);
    // on the context's own stream: its kernels run on that non-blocking
    // stream, which does not wait for null-stream work, and read the ids
    if (e == hipSuccess) e = hipMemsetAsync(c->tid_d, 0, sizeof(int) * nt, c->stream);
    if (e == hipSuccess) c->ntask_cap = nt;
  }
  if (e == hipSuccess && blk > c->out_cap) {
    if (c->out_h) (void)hipHostFree(c->out_h);
    c->out_h = nullptr;
    dfree(c->out_d);
    c->out_cap = 0;
    e = hipHostMalloc((void**)&c->out_h, sizeof(double) * blk);
    if (e == hipSuccess) e = dalloc(&c->out_d, blk);
    if (e == hipSuccess) c->out_cap = blk;
  }
  if (e == hipSuccess && nscr > c->scr_cap) {
    dfree(c->scr_d);
    c->scr_cap = 0;
    e = dalloc(&c->scr_d, nscr);
    if (e == hipSuccess) c->scr_cap = nscr;
  }
  if (e == hipSuccess && nflag > c->flags_cap) {
    dfree(c->flags_d);
    c->flags_cap = 0;
    e = dalloc(&c->flags_d, nflag);
    if (e == hipSuccess) c->flags_cap = nflag;
  }
  return e;
}

struct OptDevice {
  mcint_basis* b = nullptr;
  int64_t* pair_off_d = nullptr;
  double* pair_ld_d = nullptr;
  PairPrim* pp_d = nullptr;
  int* pp_pos_d = nullptr;
  std::mutex mu;
  std::vector<CallCtx*> idle, all;
};

void optdev_free(OptDevice* d) {
  if (!d) return;
  for (CallCtx* c : d->all) ctx_free(c);
  dfree(d->pair_off_d);
  dfree(d->pair_ld_d);
  dfree(d->pp_d);
  dfree(d->pp_pos_d);
  mcint_basis_destroy(d->b);
  delete d;
}

// Optimizers whose shells hold more primitives than this (summed over
// contraction columns) keep no device pair table: their pair data would
// exceed ~4 GB, and calls go through the plan path instead (libcint also
// skips its pair data above a primitive count, src/optimizer.h:8).
constexpr int64_t OPT_MAX_TOT_PRIM = 8192;

OptDevice* optdev_create(const FINT* atm, FINT natm, const FINT* bas, FINT nbas,
                         const double* env) {
  auto* d = new OptDevice();
  if (mcint_basis_create(&d->b, atm, natm, bas, nbas, env) != MCINT_OK) {
    delete d;
    return nullptr;
  }
  const mcint_basis* b = d->b;
  const int nv = b->nv;
  int64_t tot = 0;
  for (const VShell& v : b->vsh) tot += v.nprim;
  if (tot > OPT_MAX_TOT_PRIM) {
    optdev_free(d);
    return nullptr;
  }
  std::vector<int64_t> pair_off((size_t)nv * nv);
  std::vector<int> pairs;
  pairs.reserve((size_t)2 * nv * nv);
  int64_t npp = 0;
  for (int a = 0; a < nv; ++a)
    for (int c = 0; c < nv; ++c) {
      pair_off[(size_t)a * nv + c] = npp;
      npp += (int64_t)b->vsh[a].nprim * b->vsh[c].nprim;
      pairs.push_back(a);
      pairs.push_back(c);
    }
  int* pairs_d = nullptr;
  PairPrim* tmp = nullptr;
  hipError_t e = dalloc(&d->pair_off_d, pair_off.size());
  if (e == hipSuccess) e = dalloc(&d->pair_ld_d, pair_off.size());
  if (e == hipSuccess) e = dalloc(&d->pp_d, (size_t)npp);
  if (e == hipSuccess) e = dalloc(&d->pp_pos_d, (size_t)npp);
  if (e == hipSuccess) e = dalloc(&tmp, (size_t)npp);
  if (e == hipSuccess) e = dalloc(&pairs_d, pairs.size());
  if (e == hipSuccess)
    e = hipMemcpy(d->pair_off_d, pair_off.data(), pair_off.size() * sizeof(int64_t),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(pairs_d, pairs.data(), pairs.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_pairdata(b->vsh_d, b->env_d, nv, pairs_d, (int)(pairs.size() / 2), d->pair_off_d,
                        d->pair_ld_d, tmp, d->pp_d, d->pp_pos_d, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  dfree(pairs_d);
  dfree(tmp);
  if (e != hipSuccess) {
    optdev_free(d);
    return nullptr;
  }
  return d;
}

struct McOpt {
  CINTOpt pub;
  uint64_t magic;
  int kind;
  OptDevice* dev;  // null when the device set-up was not possible
};
constexpr uint64_t OPT_MAGIC = 0x6d63696e744f7074ull;

// the optimizer's device state, if opt is one of ours and was built from
// the same atm/bas/env as this call's
OptDevice* opt_device_for(CINTOpt* opt, const FINT* atm, FINT natm, const FINT* bas, FINT nbas,
                          const double* env) {
  if (!opt) return nullptr;
  auto* o = (McOpt*)opt;
  if (o->magic != OPT_MAGIC || !o->dev || !atm || !bas || !env) return nullptr;
  const mcint_basis* b = o->dev->b;
  if (natm != b->natm || nbas != b->nbas) return nullptr;
  if (memcmp(atm, b->atm.data(), sizeof(int) * b->atm.size()) ||
      memcmp(bas, b->bas.data(), sizeof(int) * b->bas.size()) ||
      memcmp(env, b->env.data(), sizeof(double) * b->envlen))
    return nullptr;
  return o->dev;
}

// One shell tuple on the optimizer's device state.  Returns libcint's
// status (any primitive survived) or -1 with mcint_last_error() set; the
// block goes to blk_out (packed, components outermost).
int opt_tuple(OptDevice* d, int intor, const FINT* shls, double* blk_out) {
  const mcint_basis* b = d->b;
  const int kind = intor_kind(intor);
  const int sph = intor_sph(intor);
  const int naxes = kind_axes(kind);
  const int ncomp = kind_ncomp(kind);
  QuartetShape Q;
  const int why = describe_tuple(b->bas.data(), b->nbas, kind, sph, shls, Q, intor_ket_der(intor));
  if (why) {
    set_err(why == 1 ? "shell index out of range" : "more Rys roots than the device tables hold");
    return -1;
  }
  int la, lb, lc, ld, perm;
  tuple_class(kind, Q, la, lb, lc, ld, perm);
  const int64_t nprimq = (int64_t)Q.np[0] * Q.np[1] * Q.np[2] * (naxes == 4 ? Q.np[3] : 1);
  const bool reg_ok = (b->omega >= 0 || kind != KIND_2E_IP1) && !g_generic_only;
  const int mode = reg_ok ? fast_kernel_mode(kind, sph, la, lb, lc, ld, b->omega) : 0;
  const int ppl_c = g_class_ppl[kind][class_index(la, lb, lc, ld)];
  const int team = mode ? pick_team(nprimq, ppl_c ? ppl_c : g_prims_per_lane) : 1;
  int nt = 1;
  for (int a = 0; a < naxes; ++a) nt *= Q.nc[a];
  const int64_t blk = Q.so[4] * ncomp;
  int64_t sstride = 0;
  if (mode == 2) {
    sstride = chunked_block_doubles(kind, la, lb, lc, ld);
  } else if (mode == 0) {
    const GenLayout G = gen_layout(kind, sph, la, lb, lc, ld);
    const int nr = (la + lb + lc + (kind_three(kind) ? 0 : ld) + kind_bra_raise(kind) +
                    kind_ket_raise(kind)) / 2 + 1;
    const int nrt = b->omega < 0 ? 2 * nr : nr;
    if (G.nE > GEN_ECAP || G.nF > GEN_FCAP || 3 * nrt * (G.E + 1) * (G.F + 1) > GEN_VCAP) {
      set_err("angular momentum beyond the generic kernel's tables");
      return -1;
    }
    sstride = G.total;
  }
  CallCtx* c = nullptr;
  {
    std::lock_guard<std::mutex> lk(d->mu);
    if (!d->idle.empty()) {
      c = d->idle.back();
      d->idle.pop_back();
    } else {
      c = new CallCtx();
      d->all.push_back(c);
    }
  }
  auto release = [&] {
    std::lock_guard<std::mutex> lk(d->mu);
    d->idle.push_back(c);
  };
  hipError_t e = ctx_reserve(c, nt, blk, sstride * nt, mode == 2 ? nt : 0);
  if (e != hipSuccess) {
    set_err(std::string("per-call context: ") + hipGetErrorString(e));
    release();
    return -1;
  }
  // virtual tasks over the contraction columns, as the plan's placement
  {
    int t = 0;
    for_each_task(b->vfirst, b->dummy, kind, Q, shls, 0, perm,
                  [&](const Task& tk) { c->task_h[t++] = tk; });
  }
  BasisView B;
  B.vsh = b->vsh_d;
  B.env = b->env_d;
  B.pair_off = d->pair_off_d;
  B.pair_ld = d->pair_ld_d;
  B.pp = d->pp_d;
  B.pp_pos = d->pp_pos_d;
  B.nv = b->nv;
  B.expcutoff_4c = b->expc4;
  B.expcutoff_3c = b->expc3;
  B.omega = b->omega;
  BucketArgs A{};
  A.tasks = c->task_d;
  A.t0 = 0;
  A.t1 = nt;
  A.out = c->out_d;
  A.not0 = c->not0_d;
  A.task_id = c->tid_d;
  A.scratch = c->scr_d;
  A.flags = c->flags_d;
  A.scratch_stride = sstride;
  A.la = la;
  A.lb = lb;
  A.lc = lc;
  A.ld = ld;
  A.kind = kind;
  A.sph = sph;
  A.team = team;
  A.fp32 = 0;
  A.dm = nullptr;
  A.grad = nullptr;
  A.nslot = 1;
  A.jk_j = nullptr;
  A.jk_k = nullptr;
  A.col_major = 0;
  e = hipMemcpyAsync(c->task_d, c->task_h, sizeof(Task) * nt, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = mode == 1 ? launch_fast(B, A, c->stream)
                  : (mode == 2 ? launch_chunked(B, A, c->stream) : launch_generic(B, A, c->stream));
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->out_h, c->out_d, sizeof(double) * blk, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->not0_h, c->not0_d, sizeof(FINT), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  int st = -1;
  if (e == hipSuccess) {
    memcpy(blk_out, c->out_h, sizeof(double) * blk);
    st = *c->not0_h ? 1 : 0;
  } else {
    set_err(std::string("per-call launch: ") + hipGetErrorString(e));
  }
  release();
  return st;
}

CACHE_SIZE_T single_call(int intor, double* out, FINT* dims, FINT* shls, FINT* atm, FINT natm,
                         FINT* bas, FINT nbas, double* env, CINTOpt* opt) {
  const int kind = intor_kind(intor);
  const int sph = intor_sph(intor);
  const int naxes = kind_axes(kind);
  const int ncomp = kind_ncomp(kind);
  if (!shls || !bas) {
    set_err("null shls/bas");
    return 0;
  }
  int cnt[4] = {1, 1, 1, 1};
  for (int a = 0; a < naxes; ++a) {
    if (shls[a] < 0 || shls[a] >= nbas) {
      set_err("shell index out of range");
      return 0;
    }
    const FINT* bs = bas + (size_t)shls[a] * BAS_SLOTS;
    cnt[a] = (sph ? nsph(bs[ANG_OF]) : ncart(bs[ANG_OF])) * bs[NCTR_OF];
  }
  const int64_t blk = (int64_t)cnt[0] * cnt[1] * cnt[2] * cnt[3];
  if (!out) {
    // size hint in doubles, in the spirit of libcint's cache-size query
    return (CACHE_SIZE_T)std::min<int64_t>(blk * ncomp * 2 + 64, 0x7fffffff);
  }
  std::vector<double> tmp((size_t)blk * ncomp);
  FINT q4[4] = {shls[0], shls[1], naxes > 2 ? shls[2] : 0, naxes == 4 ? shls[3] : 0};
  FINT st = 0;
  if (OptDevice* d = opt_device_for(opt, atm, natm, bas, nbas, env)) {
    const int r = opt_tuple(d, intor, q4, tmp.data());
    if (r < 0) return 0;
    st = r;
  } else {
    int rc = host_batch(intor, tmp.data(), nullptr, q4, 1, atm, natm, bas, nbas, env, &st, nullptr);
    if (rc != MCINT_OK) return 0;
  }
  // the caller's dims hold naxes entries
  int d[4] = {1, 1, 1, 1};
  for (int a = 0; a < naxes; ++a) d[a] = dims ? dims[a] : cnt[a];
  const int64_t nout = (int64_t)d[0] * d[1] * d[2] * d[3];
  for (int c = 0; c < ncomp; ++c)
    for (int l = 0; l < cnt[3]; ++l)
      for (int k = 0; k < cnt[2]; ++k)
        for (int j = 0; j < cnt[1]; ++j)
          for (int i = 0; i < cnt[0]; ++i)
            out[c * nout + i + (int64_t)d[0] * (j + (int64_t)d[1] * (k + (int64_t)d[2] * l))] =
                tmp[c * blk + i + (int64_t)cnt[0] * (j + (int64_t)cnt[1] * (k + (int64_t)cnt[2] * l))];
  return st;
}

void make_opt(CINTOpt** opt, int kind, FINT* atm, FINT natm, FINT* bas, FINT nbas, double* env) {
  if (!opt) return;
  auto* o = (McOpt*)calloc(1, sizeof(McOpt));
  o->pub.nbas = nbas;
  o->magic = OPT_MAGIC;
  o->kind = kind;
  o->dev = (atm && bas && env && natm > 0 && nbas > 0) ? optdev_create(atm, natm, bas, nbas, env)
                                                       : nullptr;
  *opt = &o->pub;
}

}  // namespace

#define MC_INTOR(name, code)                                                                   \
  extern "C" CACHE_SIZE_T name(double* out, FINT* dims, FINT* shls, FINT* atm, FINT natm,      \
                               FINT* bas, FINT nbas, double* env, CINTOpt* opt, double* cache) { \
    (void)cache;                                                                               \
    return single_call(code, out, dims, shls, atm, natm, bas, nbas, env, opt);                 \
  }
MC_INTOR(int2e_sph, MCINT_INT2E_SPH)
MC_INTOR(int2e_cart, MCINT_INT2E_CART)
MC_INTOR(int2e_ip1_sph, MCINT_INT2E_IP1_SPH)
MC_INTOR(int2e_ip1_cart, MCINT_INT2E_IP1_CART)
MC_INTOR(int3c2e_sph, MCINT_INT3C2E_SPH)
MC_INTOR(int3c2e_cart, MCINT_INT3C2E_CART)
MC_INTOR(int3c2e_ip1_sph, MCINT_INT3C2E_IP1_SPH)
MC_INTOR(int3c2e_ip1_cart, MCINT_INT3C2E_IP1_CART)
MC_INTOR(int3c2e_ip2_sph, MCINT_INT3C2E_IP2_SPH)
MC_INTOR(int3c2e_ip2_cart, MCINT_INT3C2E_IP2_CART)
MC_INTOR(int2c2e_sph, MCINT_INT2C2E_SPH)
MC_INTOR(int2c2e_cart, MCINT_INT2C2E_CART)
MC_INTOR(int2e_ip2_sph, MCINT_INT2E_IP2_SPH)
MC_INTOR(int2e_ip2_cart, MCINT_INT2E_IP2_CART)
MC_INTOR(int2e_ipip1_sph, MCINT_INT2E_IPIP1_SPH)
MC_INTOR(int2e_ipip1_cart, MCINT_INT2E_IPIP1_CART)
MC_INTOR(int2e_ip1ip2_sph, MCINT_INT2E_IP1IP2_SPH)
MC_INTOR(int2e_ip1ip2_cart, MCINT_INT2E_IP1IP2_CART)
#undef MC_INTOR

extern "C" void int2e_optimizer(CINTOpt** opt, FINT* atm, FINT natm, FINT* bas, FINT nbas,
                                double* env) {
  make_opt(opt, KIND_2E, atm, natm, bas, nbas, env);
}
extern "C" void int2e_ip1_optimizer(CINTOpt** opt, FINT* atm, FINT natm, FINT* bas, FINT nbas,
                                    double* env) {
  make_opt(opt, KIND_2E_IP1, atm, natm, bas, nbas, env);
}
extern "C" void int2e_ip2_optimizer(CINTOpt** opt, FINT* atm, FINT natm, FINT* bas, FINT nbas,
                                    double* env) {
  make_opt(opt, KIND_2E_IP1, atm, natm, bas, nbas, env);
}
extern "C" void int2e_ipip1_optimizer(CINTOpt** opt, FINT* atm, FINT natm, FINT* bas, FINT nbas,
                                      double* env) {
  make_opt(opt, KIND_2E_IPIP1, atm, natm, bas, nbas, env);
}
extern "C" void int2e_ip1ip2_optimizer(CINTOpt** opt, FINT* atm, FINT natm, FINT* bas, FINT nbas,
                                       double* env) {
  make_opt(opt, KIND_2E_IP1IP2, atm, natm, bas, nbas, env);
}
extern "C" void int3c2e_optimizer(CINTOpt** opt, FINT* atm, FINT natm, FINT* bas, FINT nbas,
                                  double* env) {
  make_opt(opt, KIND_3C2E, atm, natm, bas, nbas, env);
}
extern "C" void int3c2e_ip1_optimizer(CINTOpt** opt, FINT* atm, FINT natm, FINT* bas, FINT nbas,
                                      double* env) {
  make_opt(opt, KIND_3C2E_IP1, atm, natm, bas, nbas, env);
}
extern "C" void int3c2e_ip2_optimizer(CINTOpt** opt, FINT* atm, FINT natm, FINT* bas, FINT nbas,
                                      double* env) {
  make_opt(opt, KIND_3C2E_IP2, atm, natm, bas, nbas, env);
}
extern "C" void int2c2e_optimizer(CINTOpt** opt, FINT* atm, FINT natm, FINT* bas, FINT nbas,
                                  double* env) {
  make_opt(opt, KIND_2C2E, atm, natm, bas, nbas, env);
}
extern "C" void CINTdel_optimizer(CINTOpt** opt) {
  if (!opt || !*opt) return;
  auto* o = (McOpt*)(*opt);
  if (o->magic == OPT_MAGIC) optdev_free(o->dev);
  o->magic = 0;
  free(o);
  *opt = nullptr;
}
extern "C" void CINTdel_2e_optimizer(CINTOpt** opt) { CINTdel_optimizer(opt); }

extern "C" int mcint_optimizer_on_device(const CINTOpt* opt) {
  if (!opt) return 0;
  const auto* o = (const McOpt*)opt;
  return o->magic == OPT_MAGIC && o->dev != nullptr;
}

extern "C" FINT cint2e_sph(double* out, FINT* shls, FINT* atm, FINT natm, FINT* bas, FINT nbas,
                           double* env, CINTOpt* opt) {
  return int2e_sph(out, nullptr, shls, atm, natm, bas, nbas, env, opt, nullptr);
}
extern "C" FINT cint2e_cart(double* out, FINT* shls, FINT* atm, FINT natm, FINT* bas, FINT nbas,
                            double* env, CINTOpt* opt) {
  return int2e_cart(out, nullptr, shls, atm, natm, bas, nbas, env, opt, nullptr);
}
extern "C" FINT cint2e_ip1_sph(double* out, FINT* shls, FINT* atm, FINT natm, FINT* bas, FINT nbas,
                               double* env, CINTOpt* opt) {
  return int2e_ip1_sph(out, nullptr, shls, atm, natm, bas, nbas, env, opt, nullptr);
}
extern "C" void cint2e_sph_optimizer(CINTOpt** opt, FINT* atm, FINT natm, FINT* bas, FINT nbas,
                                     double* env) {
  int2e_optimizer(opt, atm, natm, bas, nbas, env);
}

// ---------------------------------------------------------------------------
// Basis helpers (src/misc.c:72-93, src/cint_bas.c:36-160)
// ---------------------------------------------------------------------------
extern "C" double CINTgto_norm(FINT n, double a) {
  const double n1 = n + 1.5;
  return 1.0 / std::sqrt(std::exp(std::lgamma(n1)) / (2.0 * std::pow(2.0 * a, n1)));
}
extern "C" FINT CINTcgto_spheric(const FINT bas_id, const FINT* bas) {
  return (bas[bas_id * BAS_SLOTS + ANG_OF] * 2 + 1) * bas[bas_id * BAS_SLOTS + NCTR_OF];
}
extern "C" FINT CINTcgto_cart(const FINT bas_id, const FINT* bas) {
  const int l = bas[bas_id * BAS_SLOTS + ANG_OF];
  return (l + 1) * (l + 2) / 2 * bas[bas_id * BAS_SLOTS + NCTR_OF];
}
extern "C" FINT CINTtot_cgto_spheric(const FINT* bas, const FINT nbas) {
  FINT s = 0;
  for (int i = 0; i < nbas; ++i) s += CINTcgto_spheric(i, bas);
  return s;
}
extern "C" FINT CINTtot_cgto_cart(const FINT* bas, const FINT nbas) {
  FINT s = 0;
  for (int i = 0; i < nbas; ++i) s += CINTcgto_cart(i, bas);
  return s;
}
extern "C" void CINTshells_spheric_offset(FINT ao_loc[], const FINT* bas, const FINT nbas) {
  ao_loc[0] = 0;
  for (int i = 1; i < nbas; ++i) ao_loc[i] = ao_loc[i - 1] + CINTcgto_spheric(i - 1, bas);
}
extern "C" void CINTshells_cart_offset(FINT ao_loc[], const FINT* bas, const FINT nbas) {
  ao_loc[0] = 0;
  for (int i = 1; i < nbas; ++i) ao_loc[i] = ao_loc[i - 1] + CINTcgto_cart(i - 1, bas);
}

// ---------------------------------------------------------------------------
// Rys roots on the device
// ---------------------------------------------------------------------------
extern "C" int mcint_rys_roots_batch(int nroots, const double* x, size_t n, double* u, double* w) {
  if (nroots < 1 || nroots > NRYS_MAX || (!x && n)) {
    set_err("mcint_rys_roots_batch: nroots must be 1..9");
    return MCINT_ERR_ARG;
  }
  if (n == 0) return MCINT_OK;
  for (size_t i = 0; i < n; ++i)
    if (!(x[i] >= 0) || std::isinf(x[i])) {
      set_err("mcint_rys_roots_batch: x must be finite and >= 0");
      return MCINT_ERR_ARG;
    }
  double *xd = nullptr, *ud = nullptr, *wd = nullptr;
  hipError_t e = dalloc(&xd, n);
  if (e == hipSuccess) e = dalloc(&ud, n * nroots);
  if (e == hipSuccess) e = dalloc(&wd, n * nroots);
  if (e == hipSuccess) e = hipMemcpy(xd, x, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_rys_batch(nroots, xd, (int64_t)n, ud, wd, nullptr);
  if (e == hipSuccess) e = hipMemcpy(u, ud, n * nroots * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(w, wd, n * nroots * sizeof(double), hipMemcpyDeviceToHost);
  dfree(xd);
  dfree(ud);
  dfree(wd);
  if (e != hipSuccess) {
    set_err(std::string("mcint_rys_roots_batch: ") + hipGetErrorString(e));
    return MCINT_ERR_HIP;
  }
  return MCINT_OK;
}

// libcint's signature (src/rys_roots.h:3) returns nothing; on an invalid
// nroots or x, or a device error, u and w are left untouched and
// mcint_last_error() says why.
extern "C" void CINTrys_roots(int nroots, double x, double* u, double* w) {
  (void)mcint_rys_roots_batch(nroots, &x, 1, u, w);
}

extern "C" int mcint_sr_rys_roots_batch(int nroots, const double* x, const double* lower,
                                        size_t n, double* u, double* w) {
  if (nroots < 1 || nroots > NRYS_MAX || ((!x || !lower || !u || !w) && n)) {
    set_err("mcint_sr_rys_roots_batch: nroots must be 1..9 and the arrays non-null");
    return MCINT_ERR_ARG;
  }
  if (n == 0) return MCINT_OK;
  if (n > (size_t)INT32_MAX) {
    set_err("mcint_sr_rys_roots_batch: more than 2^31-1 rules in one call");
    return MCINT_ERR_ARG;
  }
  for (size_t i = 0; i < n; ++i)
    if (!(x[i] >= 0) || std::isinf(x[i]) || !(lower[i] >= 0) || !(lower[i] < 1)) {
      set_err("mcint_sr_rys_roots_batch: x must be finite and >= 0, lower in [0, 1)");
      return MCINT_ERR_ARG;
    }
  double *xd = nullptr, *ld = nullptr, *ud = nullptr, *wd = nullptr;
  hipError_t e = dalloc(&xd, n);
  if (e == hipSuccess) e = dalloc(&ld, n);
  if (e == hipSuccess) e = dalloc(&ud, n * nroots);
  if (e == hipSuccess) e = dalloc(&wd, n * nroots);
  if (e == hipSuccess) e = hipMemcpy(xd, x, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(ld, lower, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_sr_rys_batch(nroots, xd, ld, (int64_t)n, ud, wd, nullptr);
  if (e == hipSuccess) e = hipMemcpy(u, ud, n * nroots * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(w, wd, n * nroots * sizeof(double), hipMemcpyDeviceToHost);
  dfree(xd);
  dfree(ld);
  dfree(ud);
  dfree(wd);
  if (e != hipSuccess) {
    set_err(std::string("mcint_sr_rys_roots_batch: ") + hipGetErrorString(e));
    return MCINT_ERR_HIP;
  }
  return MCINT_OK;
}

// libcint's signature (src/rys_roots.h:4) returns nothing; on an invalid
// argument or a device error u and w are left untouched and
// mcint_last_error() says why (libcint prints and calls exit()).
extern "C" void CINTsr_rys_roots(int nroots, double x, double lower, double* u, double* w) {
  (void)mcint_sr_rys_roots_batch(nroots, &x, &lower, 1, u, w);
}

// ---------------------------------------------------------------------------
// Schwarz bounds
// ---------------------------------------------------------------------------
extern "C" int mcint_schwarz(mcint_basis* b, int nshell, double* q) {
  if (!b || !q) {
    set_err("mcint_schwarz: null argument");
    return MCINT_ERR_ARG;
  }
  const int nb = b->nbas;
  if (nshell <= 0 || nshell > nb) nshell = nb;
  std::vector<FINT> shls;
  for (int i = 0; i < nshell; ++i)
    for (int j = 0; j <= i; ++j) {
      shls.push_back(i);
      shls.push_back(j);
      shls.push_back(i);
      shls.push_back(j);
    }
  const size_t nq = shls.size() / 4;
  mcint_plan* p = nullptr;
  int64_t osz = 0;
  int st = mcint_plan_create(&p, b, MCINT_INT2E_SPH, shls.data(), nq, nullptr, &osz);
  if (st) return st;
  double* od = nullptr;
  hipError_t e = dalloc(&od, (size_t)osz);
  std::vector<double> h((size_t)osz);
  if (e == hipSuccess) {
    st = mcint_plan_execute_direct(p, od, nullptr, nullptr);
    if (st == MCINT_OK) e = hipMemcpy(h.data(), od, (size_t)osz * sizeof(double), hipMemcpyDeviceToHost);
  }
  dfree(od);
  mcint_plan_destroy(p);
  if (st) return st;
  if (e != hipSuccess) {
    set_err(std::string("mcint_schwarz: ") + hipGetErrorString(e));
    return MCINT_ERR_HIP;
  }
  for (size_t i = 0; i < (size_t)nb * nb; ++i) q[i] = 0;
  size_t off = 0, k = 0;
  for (int i = 0; i < nshell; ++i)
    for (int j = 0; j <= i; ++j, ++k) {
      const int di = CINTcgto_spheric(i, b->bas.data()), dj = CINTcgto_spheric(j, b->bas.data());
      const size_t n = (size_t)di * dj * di * dj;
      double m = 0;
      for (size_t t = 0; t < n; ++t) m = std::max(m, std::fabs(h[off + t]));
      off += n;
      q[(size_t)i * nb + j] = q[(size_t)j * nb + i] = std::sqrt(m);
    }
  return MCINT_OK;
}

extern "C" const char* mcint_last_error(void) { return g_err.c_str(); }
