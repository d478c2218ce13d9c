// Three-centre chunked-path instantiations, group 3 (see eri_fast.h).
#include "device_common.h"
#include "eri_fast.h"

namespace mcint {
MCINT_CHUNK3_GROUP3(MCINT_CHUNK3_DEFINE)
}  // namespace mcint
