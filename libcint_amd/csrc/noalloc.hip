// Plan execution that allocates no device memory, for callers that cannot
// call hipMalloc (code under HIP graph capture, tight serving loops).
//
// The chunked and generic paths need per-task blocks and per-task flags
// beyond the caller's output.  Here they live in one file-scope device
// array, 'scratch', of fixed size: MCINT_NOALLOC_SCRATCH_DOUBLES doubles,
// the first NOALLOC_BLOCK_DOUBLES of which hold the blocks and the rest the
// flags (as ints).  Each GPU holds its own copy until the process ends.
// Host code never passes the array's name to a kernel (on the host it is a
// stand-in, not a device address): the device address comes from
// hipGetSymbolAddress, once per device, on the first call with that device
// current.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (other streams, other host threads) race.  Nothing in the array has
// to start at zero: the chunk kernels write every flag of their slice and
// the generic kernel clears its own block before accumulating.
#include <hip/hip_runtime.h>

#include <mutex>

#include "host_internal.h"

namespace {

constexpr int64_t NOALLOC_BLOCK_DOUBLES = 7ll << 20;  // 56 MiB of per-task blocks
constexpr int64_t NOALLOC_FLAG_INTS =
    2 * (MCINT_NOALLOC_SCRATCH_DOUBLES - NOALLOC_BLOCK_DOUBLES);  // 8 MiB of flags
static_assert(NOALLOC_FLAG_INTS > 0, "scratch too small for the flags");

__device__ __align__(256) double scratch[MCINT_NOALLOC_SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of 'scratch' on the current device (looked up once).
int scratch_address(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    mcint::set_error(hipGetErrorString(e));
    return MCINT_ERR_HIP;
  }
  if (dev < 0 || dev >= MAX_DEVICES) {
    mcint::set_error("mcint_plan_execute_noalloc: device index beyond the supported range");
    return MCINT_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      mcint::set_error(e != hipSuccess ? hipGetErrorString(e)
                                       : "mcint_plan_execute_noalloc: no device address for scratch");
      return MCINT_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return MCINT_OK;
}

}  // namespace

extern "C" int mcint_plan_execute_noalloc(mcint_plan* p, double* out_dev, FINT* not0_dev,
                                          void* hip_stream) {
  if (!p || (!out_dev && mcint::plan_tasks(p))) {
    mcint::set_error("mcint_plan_execute_noalloc: null argument");
    return MCINT_ERR_ARG;
  }
  double* base = nullptr;
  if (const int st = scratch_address(&base)) return st;
  int* flags = reinterpret_cast<int*>(base + NOALLOC_BLOCK_DOUBLES);
  return mcint::plan_launch_scratch(p, out_dev, not0_dev, (hipStream_t)hip_stream, base,
                                    NOALLOC_BLOCK_DOUBLES, flags, NOALLOC_FLAG_INTS);
}

extern "C" int64_t mcint_noalloc_scratch_doubles(void) { return MCINT_NOALLOC_SCRATCH_DOUBLES; }
