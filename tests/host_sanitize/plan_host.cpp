// CPU-only host-sanitizer driver for libmcint's plan builder (test
// infrastructure; needs no GPU).  The Makefile beside this file compiles
// libcint_amd/csrc/api.hip with AddressSanitizer and UndefinedBehavior-
// Sanitizer on its host code and links it with this driver.  The driver
// runs the host stage of plan construction (mcint::plan_build_host_only:
// basis tables, validation, canonical classes, threaded describe pass,
// counting sort into buckets, pair table) over synthetic lists and checks
// what that stage must produce:
//   - tasks = sum over tuples of the product of the shells' contraction
//     counts (one task per contraction column);
//   - output doubles = sum of ncomp x block sizes (packed), or the largest
//     offset + block (given offsets);
//   - the bucket task counts add up to the tasks, and every bucket carries a
//     valid class, team and path;
//   - the same list gives the same plan twice;
//   - invalid shells, intors, offsets and classes beyond the tables are
//     refused with the documented status.
// Any sanitizer finding aborts; a failed check exits with 1.  Prints
// "plan_host ok" at the end.
#include <sanitizer/lsan_interface.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../libcint_amd/csrc/host_internal.h"
#include "mcint.h"
#include "synth_basis.h"

namespace {

using synth::SynthBasis;

int fails = 0;

void fail(const char* what, long long got, long long want) {
  std::printf("FAIL %s: got %lld, expected %lld (%s)\n", what, got, want, mcint_last_error());
  ++fails;
}

// shells per tuple and components of an integral code
int tuple_shells(int intor) { return intor >= 4 && intor <= 9 ? 3 : (intor == 10 || intor == 11) ? 2 : 4; }
int ncomp(int intor) {
  if (intor >= 14) return 9;
  if (intor == 2 || intor == 3 || (intor >= 6 && intor <= 9) || intor == 12 || intor == 13) return 3;
  return 1;
}

int shell_dim(const SynthBasis& B, int s, bool sph) {
  const int l = B.bas[8 * s + 1];
  return (sph ? 2 * l + 1 : (l + 1) * (l + 2) / 2) * B.bas[8 * s + 3];
}

struct Expect {
  long long tasks = 0, out = 0;
};

Expect expected(const SynthBasis& B, int intor, const std::vector<int>& q, const int64_t* offs) {
  Expect e;
  const int ns = tuple_shells(intor);
  const bool sph = intor % 2 == 0;
  for (size_t t = 0; t < q.size() / 4; ++t) {
    long long nt = 1, blk = ncomp(intor);
    for (int c = 0; c < ns; ++c) {
      nt *= B.bas[8 * q[4 * t + c] + 3];
      blk *= shell_dim(B, q[4 * t + c], sph);
    }
    e.tasks += nt;
    e.out = offs ? std::max<long long>(e.out, offs[t] + blk) : e.out + blk;
  }
  return e;
}

// builds the host stage of one plan and checks it; returns the status
int build(const SynthBasis& B, int intor, const std::vector<int>& q, const int64_t* offs,
          const uint8_t* flags, const char* name, int want_status = MCINT_OK) {
  int64_t st[6] = {0, 0, 0, 0, 0, 0};
  std::vector<int64_t> info;
  const size_t nq = q.size() / 4;
  const int rc = mcint::plan_build_host_only(B.atm.data(), B.natm, B.bas.data(), B.nbas,
                                             B.env.data(), intor, nq ? q.data() : nullptr, nq,
                                             offs, flags, st, &info);
  if (rc != want_status) {
    fail(name, rc, want_status);
    return rc;
  }
  if (rc != MCINT_OK) {
    std::printf("  %-34s nq %8zu refused as expected: %s\n", name, nq, mcint_last_error());
    return rc;
  }
  const Expect e = expected(B, intor, q, offs);
  if (st[0] != e.tasks) fail(name, st[0], e.tasks);
  if (st[5] != e.out) fail(name, st[5], e.out);
  const int nb = (int)(info.size() / 8);
  if (st[1] != nb) fail(name, st[1], nb);
  long long sum = 0;
  for (int k = 0; k < nb; ++k) {
    const int64_t* b = info.data() + 8 * k;
    bool ok = b[5] >= 1 && b[5] <= 64 && (b[5] & (b[5] - 1)) == 0 && b[6] >= 0 && b[6] <= 2 &&
              b[7] > 0;
    for (int a = 0; a < 4; ++a) ok = ok && b[a] >= 0 && b[a] <= 4;
    if (!ok) fail(name, k, -1);
    sum += b[7];
  }
  if (sum != st[0]) fail(name, sum, st[0]);
  // the same list again: the same plan
  int64_t st2[6];
  std::vector<int64_t> info2;
  mcint::plan_build_host_only(B.atm.data(), B.natm, B.bas.data(), B.nbas, B.env.data(), intor,
                              nq ? q.data() : nullptr, nq, offs, flags, st2, &info2);
  if (!std::equal(st, st + 6, st2) || info != info2) fail(name, 0, 1);
  std::printf("  %-34s nq %8zu buckets %4d tasks %9lld out %lld\n", name, nq, nb, (long long)st[0],
              (long long)st[5]);
  return rc;
}

}  // namespace

int main() {
  std::setvbuf(stdout, nullptr, _IOLBF, 0);
  std::mt19937 rng(5);
  for (double omega : {0.0, 0.4, -0.4}) {
    const SynthBasis B = synth::make_basis(omega == 0.0 ? 10 : 4, omega, 11);
    std::printf("basis: %d atoms, %d shells, omega %.1f\n", B.natm, B.nbas, omega);
    const std::vector<int> uq = synth::unique_quartets(B.nbas);
    const size_t nq = uq.size() / 4;
    // long enough (3.1M tuples at omega 0) for the threaded describe pass
    build(B, MCINT_INT2E_SPH, uq, nullptr, nullptr, "int2e_sph unique");
    std::vector<uint8_t> flags(nq);
    for (auto& f : flags) f = (uint8_t)(rng() % 2);
    build(B, MCINT_INT2E_SPH, uq, nullptr, flags.data(), "int2e_sph mixed");

    // shuffled sub-list, cartesian, with gaps and decreasing offsets
    std::vector<int> sub;
    std::vector<int64_t> offs;
    {
      std::vector<size_t> pick;
      for (size_t t = 0; t < nq; t += 1 + rng() % 5) pick.push_back(t);
      std::shuffle(pick.begin(), pick.end(), rng);
      if (pick.size() > 20000) pick.resize(20000);
      int64_t off = 0;
      for (size_t t : pick) {
        offs.push_back(off);
        long long blk = 1;
        for (int c = 0; c < 4; ++c) blk *= shell_dim(B, uq[4 * t + c], false);
        off += blk + (int64_t)(rng() % 3);
      }
      // list order reversed against offset order
      std::reverse(pick.begin(), pick.end());
      std::reverse(offs.begin(), offs.end());
      for (size_t t : pick) sub.insert(sub.end(), uq.begin() + 4 * t, uq.begin() + 4 * t + 4);
    }
    build(B, MCINT_INT2E_CART, sub, offs.data(), nullptr, "int2e_cart scattered offsets");

    // every integral code on low-l tuples and on tuples up to (gg|gf)
    std::vector<int> low_sh, ordered_low, ordered;
    for (int i = 0; i < std::min(B.nbas, 14); ++i)
      if (B.bas[8 * i + 1] <= 2) low_sh.push_back(i);
    for (int i : low_sh)
      for (int j : low_sh)
        for (int k : low_sh)
          for (size_t l = 0; l < low_sh.size(); l += 3) ordered_low.insert(ordered_low.end(), {i, j, k, low_sh[l]});
    const int ns = std::min(B.nbas, 9);
    for (int i = 0; i < ns; ++i)
      for (int j = 0; j < ns; ++j)
        for (int k = 0; k < ns; ++k)
          for (int l = 0; l < ns; l += 2) ordered.insert(ordered.end(), {i, j, k, l});
    for (int intor = 0; intor <= 17; ++intor) {
      char name[64];
      std::snprintf(name, sizeof name, "code %d, l <= 2", intor);
      build(B, intor, ordered_low, nullptr, nullptr, name);
      std::snprintf(name, sizeof name, "code %d, up to g", intor);
      build(B, intor, ordered, nullptr, nullptr, name);
    }
    // second derivatives of (gg|gg) need 10 Rys roots, one more than the
    // device tables hold
    build(B, MCINT_INT2E_IPIP1_SPH, {7, 7, 7, 7}, nullptr, nullptr, "ipip1 (gg|gg)", MCINT_ERR_UNSUPPORTED);
    build(B, MCINT_INT2E_IP1IP2_CART, {7, 7, 7, 7}, nullptr, nullptr, "ip1ip2 (gg|gg)", MCINT_ERR_UNSUPPORTED);
    build(B, MCINT_INT2E_SPH, {}, nullptr, nullptr, "empty list");

    // refusals
    build(B, MCINT_INT2E_SPH, {0, 0, B.nbas, 0}, nullptr, nullptr, "shell index nbas", MCINT_ERR_ARG);
    build(B, MCINT_INT2E_SPH, {0, -1, 0, 0}, nullptr, nullptr, "shell index -1", MCINT_ERR_ARG);
    build(B, 99, {0, 0, 0, 0}, nullptr, nullptr, "unknown code", MCINT_ERR_ARG);
    const int64_t neg[2] = {0, -8};
    build(B, MCINT_INT2E_SPH, {0, 0, 0, 0, 1, 0, 0, 0}, neg, nullptr, "negative offset", MCINT_ERR_ARG);
  }
  {
    // a shell beyond the largest angular momentum is refused by the basis
    SynthBasis B = synth::make_basis(2, 0.0, 3);
    B.bas[8 * 1 + 1] = 9;
    build(B, MCINT_INT2E_SPH, {0, 0, 0, 0}, nullptr, nullptr, "basis with l = 9", MCINT_ERR_ARG);
  }
  std::fflush(stdout);
  __lsan_do_leak_check();
  if (fails) {
    std::printf("%d checks failed\n", fails);
    std::fflush(stdout);
    _exit(1);
  }
  std::printf("plan_host ok\n");
  std::fflush(stdout);
  _exit(0);
}
