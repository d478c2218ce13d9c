// Host-driver functions shared between the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

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

// Sets the calling thread's mcint_last_error() text.
void set_error(const char* msg);

}  // namespace mcint
