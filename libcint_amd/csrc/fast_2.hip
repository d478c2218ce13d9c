// Fast-kernel instantiations, group 2 (see eri_fast.h).
#include "device_common.h"
#include "eri_fast.h"

namespace mcint {
MCINT_FAST_GROUP2(MCINT_FAST_DEFINE)
}  // namespace mcint
