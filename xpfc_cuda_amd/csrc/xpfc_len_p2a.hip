// Step-pass launchers for lines of 4, 8, 16, 32, 64 points (one of the per-length
// instantiation units, compiled in parallel).
#include "xpfc_lenops.hpp"

namespace xpfc {

XPFC_INSTANTIATE_LEN(4)
XPFC_INSTANTIATE_LEN(8)
XPFC_INSTANTIATE_LEN(16)
XPFC_INSTANTIATE_LEN(32)
XPFC_INSTANTIATE_LEN(64)

}  // namespace xpfc
