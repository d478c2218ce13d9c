// Register-resident Rys kernels (placeholder: no class has a fast kernel yet).
#pragma once

namespace mcint {

bool fast_kernel_available(int, int, int, int, int, int) { return false; }
int fast_team_lanes_max() { return 64; }
hipError_t launch_fast(const BasisView&, const BucketArgs&, hipStream_t) {
  return hipErrorNotSupported;
}

}  // namespace mcint
