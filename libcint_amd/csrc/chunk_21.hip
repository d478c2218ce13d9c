// Chunked-path instantiations, group 21 (see eri_fast.h).
#include "device_common.h"
#include "eri_fast.h"

namespace mcint {
MCINT_CHUNK_GROUP21(MCINT_CHUNK_DEFINE)
}  // namespace mcint
