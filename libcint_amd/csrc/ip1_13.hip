// i-derivative chunked-path instantiations, group 13 (see eri_fast.h).
#include "device_common.h"
#include "eri_fast.h"

namespace mcint {
MCINT_IP1_GROUP13(MCINT_IP1_DEFINE)
}  // namespace mcint
