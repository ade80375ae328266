// Step-pass launchers for lines of 192, 384, 768, 1536 points (one of the per-length
// instantiation units, compiled in parallel).
#include "xpfc_lenops.hpp"

namespace xpfc {

XPFC_INSTANTIATE_LEN(192)
XPFC_INSTANTIATE_LEN(384)
XPFC_INSTANTIATE_LEN(768)
XPFC_INSTANTIATE_LEN(1536)

}  // namespace xpfc
