// Three-centre single-kernel instantiations, group 1 (see eri_fast.h).
#include "device_common.h"
#include "eri_fast.h"

namespace mcint {
MCINT_FAST3_GROUP1(MCINT_FAST3_DEFINE)
}  // namespace mcint
