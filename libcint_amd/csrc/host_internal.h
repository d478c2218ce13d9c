// Host-driver functions shared between the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/mcint.h"

namespace mcint {

// Launches every bucket of plan p on the stream, the chunked and generic
// paths using the caller's scratch: blocks of nblocks doubles and nflags
// ints of per-task flags.  Returns MCINT_ERR_UNSUPPORTED, before any
// launch, when one task of some class needs more than nblocks doubles.
int plan_launch_scratch(mcint_plan* p, double* out, FINT* not0, hipStream_t stream,
                        double* blocks, int64_t nblocks, int* flags, int64_t nflags);

// Tasks (contracted quartets after the virtual-shell expansion) of plan p.
int64_t plan_tasks(const mcint_plan* p);

// The host stage of plan construction alone, with no HIP call: the basis
// tables and the plan's validation, canonical classes, buckets, counting
// sort and pair table for the list, as mcint_basis_create followed by
// mcint_plan_create[_mixed] would build them before the device upload.
// Writes the statistics of mcint_plan_stats (6 entries, nullable) and the
// 8 entries of mcint_plan_bucket_info per bucket (nullable), then frees
// everything.  For host-side test builds (sanitizers) on machines without
// a GPU.  Returns the status plan creation would return.
int plan_build_host_only(const FINT* atm, FINT natm, const FINT* bas, FINT nbas, const double* env,
                         int intor, const FINT* shls, size_t nq, const int64_t* out_offsets,
                         const uint8_t* fp32_flags, int64_t* stats,
                         std::vector<int64_t>* bucket_info);

// Sets the calling thread's mcint_last_error() text.
void set_error(const char* msg);

}  // namespace mcint
