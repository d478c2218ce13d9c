// Three-centre single-kernel instantiations, group 3 (see eri_fast.h).
#include "device_common.h"
#include "eri_fast.h"

namespace mcint {
MCINT_FAST3_GROUP3(MCINT_FAST3_DEFINE)
}  // namespace mcint
