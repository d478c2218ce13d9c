// Fast-kernel instantiations, group 1 (see eri_fast.h).
#include "device_common.h"
#include "eri_fast.h"

namespace mcint {
MCINT_FAST_GROUP1(MCINT_FAST_DEFINE)
}  // namespace mcint
