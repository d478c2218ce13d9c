// i-derivative chunked-path instantiations, group 14 (see eri_fast.h).
#include "device_common.h"
#include "eri_fast.h"

namespace mcint {
MCINT_IP1_GROUP14(MCINT_IP1_DEFINE)
}  // namespace mcint
