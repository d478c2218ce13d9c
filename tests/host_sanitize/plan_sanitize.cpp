// Host-sanitizer driver for libmcint's plan builder and host-side entry
// points.  Test infrastructure: the Makefile beside this file compiles
// libcint_amd/csrc/api.hip with AddressSanitizer and UndefinedBehavior-
// Sanitizer on its host code only (device code is not instrumented) and
// links it with this driver against libmcint.so for the kernels.
//
// The driver builds a synthetic basis (s..g shells, contracted and
// general-contraction shells, a long-range and a short-range copy) and
// drives every host path that allocates, sorts, hashes or indexes: plan
// construction over a multi-threaded quartet list (counting sort, class
// buckets, pair table), custom and packed output offsets, mixed-precision
// flags, every integral kind, empty lists, invalid shells, plan execution
// (graph, direct, no-alloc, per bucket), the fused J/K and gradient digests,
// the J/K plan over a stored buffer, the optimizer-backed per-call entry
// points and the host-pointer batches.  Any sanitizer finding aborts with a
// nonzero status; a failed status check exits with 1.  It runs the leak
// check itself and prints "plan_sanitize ok" at the end.
#include <hip/hip_runtime.h>
#include <sanitizer/lsan_interface.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <random>
#include <vector>

#include "mcint.h"
#include "synth_basis.h"

namespace {

using synth::SynthBasis;
using synth::make_basis;
using synth::nao_sph;
using synth::unique_quartets;


void check(int st, const char* what) {
  if (st != MCINT_OK) {
    std::fprintf(stderr, "FAIL %s: status %d (%s)\n", what, st, mcint_last_error());
    std::exit(1);
  }
}

void expect(int st, int want, const char* what) {
  if (st != want) {
    std::fprintf(stderr, "FAIL %s: status %d, expected %d (%s)\n", what, st, want, mcint_last_error());
    std::exit(1);
  }
}

void hipcheck(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "FAIL %s: %s\n", what, hipGetErrorString(e));
    std::exit(1);
  }
}

// create, execute (graph, direct, per bucket, no-alloc) and destroy one plan
// may_refuse: a class beyond the device tables (Rys roots, generic-kernel
// components) may make the plan builder refuse the list with
// MCINT_ERR_UNSUPPORTED; the refusal path is exercised too
void run_plan(mcint_basis* b, int intor, const std::vector<int>& q, const int64_t* offs,
              const uint8_t* flags, const char* name, bool may_refuse = false) {
  const size_t nq = q.size() / 4;
  mcint_plan* p = nullptr;
  int64_t out_size = 0;
  const int* sh = nq ? q.data() : nullptr;
  const int cst = flags ? mcint_plan_create_mixed(&p, b, intor, sh, nq, offs, flags, &out_size)
                        : mcint_plan_create(&p, b, intor, sh, nq, offs, &out_size);
  if (may_refuse && cst == MCINT_ERR_UNSUPPORTED) {
    std::printf("  %-34s nq %8zu refused: %s\n", name, nq, mcint_last_error());
    return;
  }
  check(cst, name);
  int64_t st[6];
  check(mcint_plan_stats(p, st), "mcint_plan_stats");
  double* out = nullptr;
  int* not0 = nullptr;
  hipcheck(hipMalloc(&out, sizeof(double) * (size_t)std::max<int64_t>(out_size, 1)), "hipMalloc out");
  hipcheck(hipMalloc(&not0, sizeof(int) * std::max<size_t>(nq, 1)), "hipMalloc not0");
  check(mcint_plan_execute(p, out, not0, nullptr), "mcint_plan_execute");
  check(mcint_plan_execute(p, out, not0, nullptr), "mcint_plan_execute replay");
  check(mcint_plan_execute_direct(p, out, not0, nullptr), "mcint_plan_execute_direct");
  const int nb = mcint_plan_nbuckets(p);
  std::vector<int64_t> prims(std::max(nb, 1));
  check(mcint_plan_count_primitives(p, prims.data()), "mcint_plan_count_primitives");
  for (int k = 0; k < nb; ++k) {
    int64_t info[8];
    check(mcint_plan_bucket_info(p, k, info), "mcint_plan_bucket_info");
    const int64_t n = std::min<int64_t>(info[7], 5);
    std::vector<int> ids(std::max<int64_t>(n, 1));
    check(mcint_plan_bucket_tasks(p, k, info[7] - n, n, ids.data()), "mcint_plan_bucket_tasks");
    for (int64_t t = 0; t < n; ++t)
      if (ids[t] < 0 || (size_t)ids[t] >= nq) {
        std::fprintf(stderr, "FAIL %s: bucket %d task id %d\n", name, k, ids[t]);
        std::exit(1);
      }
    if (k % 7 == 0) check(mcint_plan_execute_bucket(p, k, out, not0, nullptr), "execute_bucket");
  }
  int st_na = mcint_plan_execute_noalloc(p, out, not0, nullptr);
  if (st_na != MCINT_OK && st_na != MCINT_ERR_UNSUPPORTED) check(st_na, "execute_noalloc");
  hipcheck(hipDeviceSynchronize(), name);
  // a finite sample of the output
  std::vector<double> h(std::min<int64_t>(out_size, 4096));
  if (!h.empty()) {
    hipcheck(hipMemcpy(h.data(), out, h.size() * sizeof(double), hipMemcpyDeviceToHost), "copy");
    for (double v : h)
      if (!std::isfinite(v)) {
        std::fprintf(stderr, "FAIL %s: non-finite output\n", name);
        std::exit(1);
      }
  }
  hipcheck(hipFree(out), "hipFree");
  hipcheck(hipFree(not0), "hipFree");
  mcint_plan_destroy(p);
  std::printf("  %-34s nq %8zu buckets %4d tasks %9lld out %lld\n", name, nq, nb,
              (long long)st[0], (long long)out_size);
  std::fflush(stdout);
}

}  // namespace

int main() {
  std::setvbuf(stdout, nullptr, _IOLBF, 0);  // progress lines survive an abort
  std::mt19937 rng(7);
  for (double omega : {0.0, 0.4, -0.4}) {
    SynthBasis B = make_basis(omega == 0.0 ? 10 : 4, omega, 11);
    std::printf("basis: %d atoms, %d shells, omega %.1f\n", B.natm, B.nbas, omega);
    mcint_basis* b = nullptr;
    check(mcint_basis_create(&b, B.atm.data(), B.natm, B.bas.data(), B.nbas, B.env.data()),
          "mcint_basis_create");
    // short range: a class whose doubled root count exceeds the device
    // tables may make the builder refuse a whole list
    const bool sr = omega < 0;
    std::vector<int> uq = unique_quartets(B.nbas);
    const size_t nq = uq.size() / 4;

    // int2e over every unique quartet: a list long enough for the
    // multi-threaded counting sort when omega == 0
    run_plan(b, MCINT_INT2E_SPH, uq, nullptr, nullptr, "int2e_sph unique", sr);

    // a shuffled ragged sub-list with scattered output offsets, cartesian
    std::vector<int> sub;
    std::vector<int64_t> offs;
    {
      std::vector<size_t> pick;
      for (size_t q = 0; q < nq; q += 1 + rng() % 5) pick.push_back(q);
      std::shuffle(pick.begin(), pick.end(), rng);
      int64_t off = 0;
      for (size_t q : pick) {
        sub.insert(sub.end(), uq.begin() + 4 * q, uq.begin() + 4 * q + 4);
        offs.push_back(off);
        int n = 1;
        for (int c = 0; c < 4; ++c) {
          const int s = uq[4 * q + c], l = B.bas[8 * s + 1];
          n *= (l + 1) * (l + 2) / 2 * B.bas[8 * s + 3];
        }
        off += n + (int64_t)(rng() % 3);  // gaps between blocks
      }
      // reverse so offsets are not monotone in list order
      std::reverse(offs.begin(), offs.end());
      std::vector<int> rev(sub.size());
      for (size_t q = 0; q < pick.size(); ++q)
        std::copy(sub.begin() + 4 * q, sub.begin() + 4 * q + 4, rev.begin() + 4 * (pick.size() - 1 - q));
      sub.swap(rev);
    }
    run_plan(b, MCINT_INT2E_CART, sub, offs.data(), nullptr, "int2e_cart scattered offsets", sr);

    // mixed precision: random FP32 flags
    std::vector<uint8_t> flags(nq);
    for (auto& f : flags) f = (uint8_t)(rng() % 2);
    run_plan(b, MCINT_INT2E_SPH, uq, nullptr, flags.data(), "int2e_sph mixed", sr);

    // every other kind on a prefix of all ordered tuples
    const int ns = std::min(B.nbas, 9);
    std::vector<int> ordered;
    for (int i = 0; i < ns; ++i)
      for (int j = 0; j < ns; ++j)
        for (int k = 0; k < ns; ++k)
          for (int l = 0; l < ns; l += 2) ordered.insert(ordered.end(), {i, j, k, l});
    const struct { int intor; const char* name; } kinds[] = {
        {MCINT_INT2E_IP1_SPH, "int2e_ip1_sph"},     {MCINT_INT2E_IP1_CART, "int2e_ip1_cart"},
        {MCINT_INT2E_IP2_SPH, "int2e_ip2_sph"},     {MCINT_INT2E_IPIP1_SPH, "int2e_ipip1_sph"},
        {MCINT_INT2E_IP1IP2_CART, "int2e_ip1ip2_cart"}, {MCINT_INT3C2E_SPH, "int3c2e_sph"},
        {MCINT_INT3C2E_IP1_SPH, "int3c2e_ip1_sph"}, {MCINT_INT3C2E_IP2_CART, "int3c2e_ip2_cart"},
        {MCINT_INT2C2E_SPH, "int2c2e_sph"},         {MCINT_INT2C2E_CART, "int2c2e_cart"}};
    // shells of l <= 2 among the first 14: every class within the tables
    std::vector<int> low_sh, ordered_low;
    for (int i = 0; i < std::min(B.nbas, 14); ++i)
      if (B.bas[8 * i + 1] <= 2) low_sh.push_back(i);
    for (int i : low_sh)
      for (int j : low_sh)
        for (int k : low_sh)
          for (size_t l = 0; l < low_sh.size(); l += 3) ordered_low.insert(ordered_low.end(), {i, j, k, low_sh[l]});
    for (auto& k : kinds) {
      run_plan(b, k.intor, ordered, nullptr, nullptr, k.name, true);
      run_plan(b, k.intor, ordered_low, nullptr, nullptr, k.name, sr);
    }

    // empty list
    run_plan(b, MCINT_INT2E_SPH, {}, nullptr, nullptr, "int2e_sph empty");

    // invalid input is refused with a status
    {
      mcint_plan* p = nullptr;
      int64_t n = 0;
      std::vector<int> bad = {0, 0, B.nbas, 0};
      if (mcint_plan_create(&p, b, MCINT_INT2E_SPH, bad.data(), 1, nullptr, &n) == MCINT_OK) {
        std::fprintf(stderr, "FAIL: shell index nbas accepted\n");
        return 1;
      }
      bad = {0, -1, 0, 0};
      if (mcint_plan_create(&p, b, MCINT_INT2E_SPH, bad.data(), 1, nullptr, &n) == MCINT_OK) {
        std::fprintf(stderr, "FAIL: shell index -1 accepted\n");
        return 1;
      }
      if (mcint_plan_create(&p, b, 99, uq.data(), 1, nullptr, &n) == MCINT_OK) {
        std::fprintf(stderr, "FAIL: unknown intor accepted\n");
        return 1;
      }
      const int64_t neg[2] = {0, -8};
      expect(mcint_plan_create(&p, b, MCINT_INT2E_SPH, uq.data(), 2, neg, &n), MCINT_ERR_ARG,
             "negative output offset");
      mcint_jk_plan* j = nullptr;
      expect(mcint_jk_plan_create(&j, b, uq.data(), 2, neg), MCINT_ERR_ARG,
             "negative J/K block offset");
    }

    // J/K: fused, and over a stored buffer with a J/K plan
    const int nao = nao_sph(B);
    std::vector<double> dm((size_t)nao * nao);
    for (int i = 0; i < nao; ++i)
      for (int j = 0; j <= i; ++j) dm[(size_t)i * nao + j] = dm[(size_t)j * nao + i] = 0.1 * std::cos(i + 2.0 * j);
    double *dm_d, *vj, *vk, *eri, *grad;
    hipcheck(hipMalloc(&dm_d, dm.size() * sizeof(double)), "hipMalloc dm");
    hipcheck(hipMemcpy(dm_d, dm.data(), dm.size() * sizeof(double), hipMemcpyHostToDevice), "dm");
    hipcheck(hipMalloc(&vj, dm.size() * sizeof(double)), "hipMalloc vj");
    hipcheck(hipMalloc(&vk, dm.size() * sizeof(double)), "hipMalloc vk");
    hipcheck(hipMalloc(&grad, 3 * B.natm * sizeof(double)), "hipMalloc grad");
    mcint_plan* jp = nullptr;
    int64_t n_eri = 0;
    const int jst = mcint_plan_create(&jp, b, MCINT_INT2E_SPH, uq.data(), nq, nullptr, &n_eri);
    if (!(sr && jst == MCINT_ERR_UNSUPPORTED)) check(jst, "jk plan");
    if (jp) {
      mcint_plan* p = jp;
      const int64_t n = n_eri;
      // the fused digest takes the plain and long-range operators only
      const int want = omega < 0 ? MCINT_ERR_UNSUPPORTED : MCINT_OK;
      expect(mcint_plan_execute_jk(p, dm_d, nao, vj, vk, nullptr), want, "execute_jk");
      expect(mcint_plan_execute_jk(p, dm_d, nao, vj, nullptr, nullptr), want, "execute_jk J only");
      // (the operator is checked before the density's dimension)
      expect(mcint_plan_execute_jk(p, dm_d, nao + 1, vj, vk, nullptr), sr ? want : MCINT_ERR_ARG,
             "execute_jk wrong nao");
      hipcheck(hipMalloc(&eri, sizeof(double) * (size_t)std::max<int64_t>(n, 1)), "hipMalloc eri");
      check(mcint_plan_execute(p, eri, nullptr, nullptr), "eri for jk plan");
      mcint_jk_plan* j = nullptr;
      check(mcint_jk_plan_create(&j, b, uq.data(), nq, nullptr), "mcint_jk_plan_create");
      check(mcint_jk_plan_execute(j, eri, dm_d, vj, vk, nullptr), "mcint_jk_plan_execute");
      hipcheck(hipDeviceSynchronize(), "jk");
      mcint_jk_plan_destroy(j);
      mcint_plan_destroy(p);
      hipcheck(hipFree(eri), "hipFree eri");
    }
    if (omega >= 0) {
      // gradient digest over every (i, j, k>=l) of the first 40 shells
      std::vector<int> gq;
      const int ng = std::min(B.nbas, 40);
      for (int i = 0; i < ng; ++i)
        for (int j = 0; j < ng; ++j)
          for (int k = 0; k < ng; ++k)
            for (int l = 0; l <= k; ++l) gq.insert(gq.end(), {i, j, k, l});
      mcint_plan* p = nullptr;
      int64_t n = 0;
      check(mcint_plan_create(&p, b, MCINT_INT2E_IP1_SPH, gq.data(), gq.size() / 4, nullptr, &n),
            "gradient plan");
      hipcheck(hipMemset(grad, 0, 3 * B.natm * sizeof(double)), "memset grad");
      check(mcint_plan_execute_grad(p, dm_d, nao, grad, nullptr), "mcint_plan_execute_grad");
      hipcheck(hipDeviceSynchronize(), "grad");
      mcint_plan_destroy(p);
      std::printf("  %-34s nq %8zu\n", "int2e_ip1_sph gradient digest", gq.size() / 4);
    }
    hipcheck(hipFree(dm_d), "hipFree");
    hipcheck(hipFree(vj), "hipFree");
    hipcheck(hipFree(vk), "hipFree");
    hipcheck(hipFree(grad), "hipFree");
    mcint_basis_destroy(b);

    // libcint-named per-call functions with and without an optimizer, and
    // the host-pointer batch
    {
      int* atm = B.atm.data();
      int* bas = B.bas.data();
      double* env = B.env.data();
      CINTOpt* opt = nullptr;
      int2e_optimizer(&opt, atm, B.natm, bas, B.nbas, env);
      std::vector<double> buf(1 << 16);
      for (int t = 0; t < 40; ++t) {
        int sh[4] = {(int)(rng() % B.nbas), (int)(rng() % B.nbas), (int)(rng() % B.nbas),
                     (int)(rng() % B.nbas)};
        int2e_sph(buf.data(), nullptr, sh, atm, B.natm, bas, B.nbas, env, t % 2 ? opt : nullptr, nullptr);
        int2e_ip1_sph(buf.data(), nullptr, sh, atm, B.natm, bas, B.nbas, env, nullptr, nullptr);
        int dims[4] = {0, 0, 0, 0};
        for (int c = 0; c < 4; ++c) dims[c] = (2 * bas[8 * sh[c] + 1] + 1) * bas[8 * sh[c] + 3];
        int2e_sph(buf.data(), dims, sh, atm, B.natm, bas, B.nbas, env, opt, nullptr);
        int3c2e_sph(buf.data(), nullptr, sh, atm, B.natm, bas, B.nbas, env, nullptr, nullptr);
      }
      CINTdel_optimizer(&opt);
      // Rys rules: full range and short range, single and batched
      double ur[9], wr[9];
      CINTrys_roots(5, 3.0, ur, wr);
      CINTsr_rys_roots(5, 3.0, 0.4, ur, wr);
      const double xs[3] = {0.0, 2.5, 80.0}, lows[3] = {0.0, 0.3, 0.9};
      std::vector<double> ub(27), wb(27);
      check(mcint_sr_rys_roots_batch(9, xs, lows, 3, ub.data(), wb.data()), "sr roots batch");
      expect(mcint_sr_rys_roots_batch(9, xs, xs, 3, ub.data(), wb.data()), MCINT_ERR_ARG,
             "sr roots with lower >= 1");
      const size_t nb = std::min<size_t>(nq, 2000);
      std::vector<size_t> boff(nb);
      size_t tot = 0;
      for (size_t q = 0; q < nb; ++q) {
        boff[q] = tot;
        int n = 1;
        for (int c = 0; c < 4; ++c) {
          const int s = uq[4 * q + c];
          n *= (2 * bas[8 * s + 1] + 1) * bas[8 * s + 3];
        }
        tot += n;
      }
      std::vector<double> host(tot);
      std::vector<int> not0(nb);
      check(int2e_sph_batch(host.data(), boff.data(), uq.data(), nb, atm, B.natm, bas, B.nbas, env,
                            nullptr, not0.data(), nullptr),
            "int2e_sph_batch");
      check(int2e_sph_batch(host.data(), nullptr, uq.data(), nb, atm, B.natm, bas, B.nbas, env,
                            nullptr, nullptr, nullptr),
            "int2e_sph_batch packed");
      std::printf("  %-34s nq %8zu\n", "per-call functions and batch", nb);
    }
  }
  // Leak check now, then leave without the ROCm runtime's exit-time
  // teardown: under ASan that teardown recycles quarantined device
  // allocations after the runtime has unloaded and trips an internal
  // ASan CHECK (sanitizer_allocator_device.h, "dev_runtime_unloaded_").
  std::fflush(stdout);
  __lsan_do_leak_check();
  std::printf("plan_sanitize ok\n");
  std::fflush(stdout);
  std::fflush(stderr);
  _exit(0);
}
